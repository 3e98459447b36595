/*
 * oracle.c -- FP64 CPU restatement of the IS-SPA MD hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle the HIP kernels are
 * compared against.  It is never linked into the product library.  See
 * oracle.h for the list of reference files each function follows.
 *
 * Build: gcc -O2 -fPIC -shared -fopenmp -ffp-contract=off (oracle/Makefile).
 * -ffp-contract=off matters: the FP32 decision values below must be formed
 * with exactly the roundings the device uses (explicit fmaf where the device
 * uses v_fma_f32, plain * and + elsewhere).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10.  Round function and Weyl key schedule as published by     */
/* Salmon, Moraes, Dror and Shaw, "Parallel Random Numbers: As Easy as 1,   */
/* 2, 3" (SC'11).                                                            */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Counter layout shared with the device (isspa_amd.h):
 *   ctr = { index, (stream << 24) | draw, step, system }, key = { seed_lo, seed_hi } */
#define ORC_STREAM_POINTS   0u
#define ORC_STREAM_ANDERSEN 1u
#define ORC_MC_MAX_ATTEMPTS 64
/* accept u in [-1,1)^3 iff |u|^2 < 0.99; on the 2^-24 integer lattice this is
 * sum (2v-2^24)^2 < 0.99 * 2^48, i.e. <= floor(0.99 * 2^48). */
#define ORC_MC_ACCEPT_MAX 278660226943549ULL

static void philox_draw(uint64_t seed, uint32_t index, uint32_t stream, uint32_t draw,
                        uint32_t step, uint32_t sys, uint32_t out[4])
{
    uint32_t ctr[4] = {index, (stream << 24) | draw, step, sys};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------ */
/* constants (src/constants.h:6-13), evaluated in FP32 like the reference   */
/* ------------------------------------------------------------------------ */
void orc_default_consts(orc_consts *c)
{
    volatile float pi = 3.1415926535f, rho = 0.0075f, ep = 2.3473f;
    volatile float p0 = 4.173705894203977f;
    float e0 = 3.0f / (4.0f * pi * rho * p0) * (1.0f - 1.0f / ep);
    c->e0 = e0;
    c->p0 = p0;
    c->q0 = 1.6899960102882601f;
    c->o0 = 9.3658298730779350f;
    c->ep = ep;
}

/* single-image minimum image, src/cuda_vector_routines.h:28-47 */
static inline float mi1f(float r, float L, float h)
{
    if (r > h) return r - L;
    if (r < -h) return r + L;
    return r;
}
static inline double mi1d(double r, double L, double h)
{
    if (r > h) return r - L;
    if (r < -h) return r + L;
    return r;
}
/* a - b in FP64 with the image the device takes: the shift is decided on the
 * FP32 difference of the FP32 coordinates against the FP32 half box, so a
 * pair at exactly half a box (where the single-image rule is discontinuous)
 * gets the same image on both sides; the separation itself is FP64. */
static inline double mi_dev(double a, double b, double L)
{
    const float rf = (float)a - (float)b, hf = 0.5f * (float)L;
    const double k = rf > hf ? 1.0 : (rf < -hf ? -1.0 : 0.0);
    return a - b - k * L;
}

/* FP32 pair separation and squared distance exactly as the device forms them. */
static inline float sep_f32(const float *a, const float *b, float L, float h, float r[3])
{
    r[0] = mi1f(a[0] - b[0], L, h);
    r[1] = mi1f(a[1] - b[1], L, h);
    r[2] = mi1f(a[2] - b[2], L, h);
    return fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
}

/* ------------------------------------------------------------------------ */
/* A.1  MC points, src/isspa_force_cuda.cu:73-107 (curand -> Philox)        */
/* ------------------------------------------------------------------------ */
void orc_mc_points(int n, int nmc, const float *xyzq, const int *itype, const float *rmax,
                   uint64_t seed, uint32_t step, uint32_t sys, float *pts)
{
    const float two_m24 = 5.9604644775390625e-08f; /* 2^-24 */
    for (int k = 0; k < n; ++k) {
        const float R = rmax[itype[k]];
        for (int s = 0; s < nmc; ++s) {
            uint32_t m = (uint32_t)k * (uint32_t)nmc + (uint32_t)s;
            int32_t iv[3] = {0, 0, 0};
            for (int a = 0; a < ORC_MC_MAX_ATTEMPTS; ++a) {
                uint32_t o[4];
                philox_draw(seed, m, ORC_STREAM_POINTS, (uint32_t)a, step, sys, o);
                int64_t v0 = 2 * (int64_t)(o[0] >> 8) - (1 << 24);
                int64_t v1 = 2 * (int64_t)(o[1] >> 8) - (1 << 24);
                int64_t v2 = 2 * (int64_t)(o[2] >> 8) - (1 << 24);
                uint64_t r2 = (uint64_t)(v0 * v0 + v1 * v1 + v2 * v2);
                if (r2 <= ORC_MC_ACCEPT_MAX) {
                    iv[0] = (int32_t)v0; iv[1] = (int32_t)v1; iv[2] = (int32_t)v2;
                    break;
                }
            }
            float *p = pts + 4 * (size_t)m;
            for (int d = 0; d < 3; ++d) {
                float u = (float)iv[d] * two_m24; /* exact */
                float du = R * u;                 /* one rounding */
                p[d] = xyzq[4 * k + d] + du;      /* one rounding */
            }
            p[3] = 0.0f;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* A.2  field pass, src/isspa_force_cuda.cu:110-206                          */
/* ------------------------------------------------------------------------ */
void orc_isspa_field(int n, int nmc, const float *xyzq, const int *itype, const orc_tables *t,
                     const float *pts, float lbox, const orc_consts *c, double *g, int *cnt,
                     double *E, double *E0)
{
    orc_isspa_field_subset(n, (long)n * nmc, nmc, NULL, xyzq, itype, t, pts, lbox, c, g, cnt, E,
                           E0);
}

void orc_isspa_field_subset(int n, long P, int nmc, const int *owner, const float *xyzq,
                            const int *itype, const orc_tables *t, const float *pts, float lbox,
                            const orc_consts *c, double *g, int *cnt, double *E, double *E0)
{
    const float hbox = 0.5f * lbox;
    const double e0 = c->e0;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < P; ++m) {
        const int k = owner ? owner[m] : (int)(m / nmc);
        const int tk = itype[k];
        const float R2 = t->rmax[tk] * t->rmax[tk];
        double gm = 1.0, Em[3] = {0, 0, 0}, E0m[3] = {0, 0, 0};
        int cm = 0;
        for (int j = 0; j < n; ++j) {
            float rf[3];
            float d2f = sep_f32(pts + 4 * m, xyzq + 4 * j, lbox, hbox, rf);
            double r[3] = {rf[0], rf[1], rf[2]};
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            double d = sqrt(d2);
            double s = -e0 * (double)xyzq[4 * j + 3] / (d2 * d);
            for (int a = 0; a < 3; ++a) Em[a] += s * r[a];
            if (d2f <= R2) {
                const int tj = itype[j];
                cm += 1;
                int b = (int)((d - t->g_r0) / t->g_dr);
                if (b < 0) {
                    gm *= 0.0;
                } else if (b < t->n_g - 1) {
                    double ph = (d - ((double)t->g_r0 + b * (double)t->g_dr)) / t->g_dr;
                    const float *G = t->g + (size_t)tj * t->n_g;
                    gm *= G[b] * (1.0 - ph) + G[b + 1] * ph;
                    double phe = (d - ((double)t->e_r0 + b * (double)t->e_dr)) / t->e_dr;
                    const float *Et = t->e + (size_t)tj * t->n_e;
                    double et = Et[b] * (1.0 - phe) + Et[b + 1] * phe;
                    for (int a = 0; a < 3; ++a) Em[a] += et * r[a] / d;
                }
            } else {
                for (int a = 0; a < 3; ++a) E0m[a] += s * r[a];
            }
        }
        g[m] = gm;
        cnt[m] = cm;
        for (int a = 0; a < 3; ++a) {
            E[3 * m + a] = Em[a];
            E0[3 * m + a] = E0m[a];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Langevin moments <P_l(cos)> of a dipole in a field of strength x:         */
/*   <P1> = coth x - 1/x,  <P2> = 1 - 3<P1>/x,  <P3> = <P1> - 5<P2>/x        */
/* which are the reference's c1, 1.5c2 - 0.5 and 2.5c3 - 1.5c1              */
/* (src/isspa_force_cuda.cu:270-287).  The upward recurrence cancels for     */
/* small x, so below x = 5 the moments are taken as ratios of modified       */
/* spherical Bessel functions, <P_l> = i_l(x)/i_0(x), whose power series     */
/*   i_l(x) = x^l sum_k (x^2/2)^k / (k! (2l+2k+1)!!)                         */
/* have only positive terms (30 terms reach 1e-22 at x = 5).  Above 5 the    */
/* recurrence loses at most a factor of ~3.                                 */
/* ------------------------------------------------------------------------ */
void orc_langevin_moments(double x, double m[3])
{
    if (x < 5.0) {
        const double h = 0.5 * x * x;
        double s[4];
        double first = 1.0;
        for (int l = 0; l < 4; ++l) {
            if (l > 0) first /= (double)(2 * l + 1);
            double t = first, sum = first;
            for (int k = 0; k < 30; ++k) {
                t *= h / ((double)(k + 1) * (double)(2 * l + 2 * k + 3));
                sum += t;
            }
            s[l] = sum;
        }
        m[0] = x * s[1] / s[0];
        m[1] = x * x * s[2] / s[0];
        m[2] = x * x * x * s[3] / s[0];
        return;
    }
    double coth = 1.0 / tanh(x);
    m[0] = coth - 1.0 / x;
    m[1] = 1.0 - 3.0 * m[0] / x;
    m[2] = m[0] - 5.0 * m[1] / x;
}

/* ------------------------------------------------------------------------ */
/* A.3  force pass, src/isspa_force_cuda.cu:208-327                          */
/* ------------------------------------------------------------------------ */
void orc_isspa_force(int n, int nmc, const float *xyzq, const int *itype, const orc_tables *t,
                     const float *pts, const double *g, const int *cnt, const double *E,
                     const double *E0, float lbox, const orc_consts *c, int i0, int i1,
                     double *f_lj, double *f_pol, double *f_cdd)
{
    const long P = (long)n * nmc;
    const float hbox = 0.5f * lbox;
    const double p0 = c->p0, q0 = c->q0, o0 = c->o0;
    /* per-point hoisted quantities */
    double *hat = (double *)malloc(sizeof(double) * 3 * P);
    double *mom = (double *)malloc(sizeof(double) * 3 * P);
#pragma omp parallel for schedule(static)
    for (long m = 0; m < P; ++m) {
        const double *e = E + 3 * m;
        double x = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        for (int a = 0; a < 3; ++a) hat[3 * m + a] = x > 0 ? e[a] / x : 0.0;
        orc_langevin_moments(x, mom + 3 * m);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = i0; i < i1; ++i) {
        const int ti = itype[i];
        const float R2 = t->rmax[ti] * t->rmax[ti];
        const double vtot = t->vtot[ti];
        const double q = xyzq[4 * i + 3];
        const float *F = t->f + (size_t)ti * t->n_f;
        double lj[3] = {0, 0, 0}, pol[3] = {0, 0, 0}, cdd[3] = {0, 0, 0};
        for (long m = 0; m < P; ++m) {
            float rf[3];
            float d2f = sep_f32(pts + 4 * m, xyzq + 4 * i, lbox, hbox, rf);
            double r[3] = {rf[0], rf[1], rf[2]};
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            double d = sqrt(d2);
            const double *eh = hat + 3 * m;
            const double *M = mom + 3 * m;
            double nu = g[m] * vtot / cnt[m];
            double Rz = (eh[0] * r[0] + eh[1] * r[1] + eh[2] * r[2]) / d;
            double dp1 = 3.0 * Rz;
            double dp2 = 7.5 * Rz * Rz - 1.5;
            double dp3 = (17.5 * Rz * Rz - 7.5) * Rz;
            double s1 = -q * p0 * M[0] * nu / (d2 * d);
            double s2 = -q * q0 * M[1] * nu / (d2 * d2);
            double s3 = -q * o0 * M[2] * nu / (d2 * d2 * d);
            for (int a = 0; a < 3; ++a)
                pol[a] += s1 * (r[a] * dp1 / d - eh[a]) + s2 * (r[a] * dp2 / d - dp1 * eh[a]) +
                          s3 * (r[a] * dp3 / d - dp2 * eh[a]);
            if (d2f <= R2) {
                int b = (int)((d - t->f_r0) / t->f_dr + 0.5);
                double fs = 0.0;
                if (b >= 0 && b < t->n_f) {
                    double ph = (d - ((double)t->f_r0 + b * (double)t->f_dr)) / t->f_dr;
                    double f1 = F[b];
                    double f2 = (b + 1 < t->n_f) ? F[b + 1] : 0.0; /* Q4: zero past the table */
                    fs = (f1 * (1.0 - ph) + f2 * ph) * nu;
                }
                for (int a = 0; a < 3; ++a) lj[a] += -fs * r[a] / d;
            } else {
                const double *e0v = E0 + 3 * m;
                double ev[3] = {e0v[0] / 3.0, e0v[1] / 3.0, e0v[2] / 3.0};
                double fs = -q * p0 / (d2 * d);
                double pdotr = 3.0 * (ev[0] * r[0] + ev[1] * r[1] + ev[2] * r[2]) / d2;
                double w = vtot / cnt[m];
                for (int a = 0; a < 3; ++a) cdd[a] += fs * (pdotr * r[a] - ev[a]) * w;
            }
        }
        for (int a = 0; a < 3; ++a) {
            if (f_lj) f_lj[3 * i + a] = lj[a];
            if (f_pol) f_pol[3 * i + a] = pol[a];
            if (f_cdd) f_cdd[3 * i + a] = cdd[a];
        }
    }
    free(hat);
    free(mom);
}

/* ------------------------------------------------------------------------ */
/* A.4  nonbond, src/nonbond_force_cuda.cu:31-174                            */
/* ------------------------------------------------------------------------ */
/* exclusion test exactly as the reference scans the one-directional list:
 * the lower-index atom's slice of the prefix-summed list, with the early exit
 * on the first entry larger than the partner. */
static int excluded_pair(int a, int b, const int *pre, const int *list)
{
    int lo = a < b ? a : b, hi = a < b ? b : a;
    if (lo == hi) return 0;
    int start = lo == 0 ? 0 : pre[lo - 1];
    int stop = pre[lo];
    for (int e = start; e < stop; ++e) {
        if (list[e] == hi) return 1;
        if (list[e] > hi) break;
    }
    return 0;
}

void orc_nonbond(int n, const double *xyzq, int n_lj_types, const int *ityp, const int *nbparm,
                 const double *lj_a, const double *lj_b, const int *nexcl_prefix,
                 const int *excl_list, const int *isspa_type, const float *rmax, float lbox,
                 const orc_consts *c, double *f_lj, double *f_coul, double *f_pair,
                 double *energy)
{
    const double L = lbox, h = 0.5 * (double)lbox;
    const double fep = 1.0 - 1.0 / (double)c->ep;
    double elj = 0.0, ecoul = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : elj, ecoul)
    for (int i = 0; i < n; ++i) {
        double lj[3] = {0, 0, 0}, co[3] = {0, 0, 0}, pr[3] = {0, 0, 0};
        const double qi = xyzq[4 * i + 3];
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double r[3];
            for (int a = 0; a < 3; ++a)
                r[a] = mi_dev((double)xyzq[4 * i + a], (double)xyzq[4 * j + a], L);
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            double d = sqrt(d2);
            double qq = qi * (double)xyzq[4 * j + 3];
            double R = rmax[isspa_type[j]];
            double fdir;
            if (d > 2.0 * R)
                fdir = -2.0 * qq * fep / (3.0 * d2 * d);
            else
                fdir = -qq * fep * (8.0 * R - 3.0 * d) / (24.0 * R * R * R * R);
            for (int a = 0; a < 3; ++a) pr[a] += fdir * r[a];
            if (!excluded_pair(i, j, nexcl_prefix, excl_list)) {
                int p = nbparm[n_lj_types * ityp[i] + ityp[j]];
                double A = lj_a[p], B = lj_b[p];
                double r6 = 1.0 / (d2 * d2 * d2);
                double flj = r6 * (12.0 * A * r6 - 6.0 * B) / d2;
                double fc = qq / (d2 * d);
                for (int a = 0; a < 3; ++a) {
                    lj[a] += flj * r[a];
                    co[a] += fc * r[a];
                }
                if (j > i) {
                    elj += A * r6 * r6 - B * r6;
                    ecoul += qq / d;
                }
            }
        }
        for (int a = 0; a < 3; ++a) {
            if (f_lj) f_lj[3 * i + a] = lj[a];
            if (f_coul) f_coul[3 * i + a] = co[a];
            if (f_pair) f_pair[3 * i + a] = pr[a];
        }
    }
    if (energy) {
        energy[0] = elj;
        energy[1] = ecoul;
    }
}

/* ------------------------------------------------------------------------ */
/* A.5  bonded terms                                                         */
/* ------------------------------------------------------------------------ */
static void sep_d(const double *xyzq, int a, int b, double L, double h, double r[3])
{
    (void)h;
    for (int k = 0; k < 3; ++k) r[k] = mi_dev((double)xyzq[4 * a + k], (double)xyzq[4 * b + k], L);
}

/* src/bond_force_cuda.cu:15-53: F = 2k (r0/|r| - 1) r, E = k (|r| - r0)^2 */
void orc_bonds(int nb, const int *bond_atoms, const float *bond_params, const double *xyzq,
               float lbox, double *f, double *energy)
{
    double L = lbox, h = 0.5 * (double)lbox, e = 0.0;
    for (int q = 0; q < nb; ++q) {
        int i = bond_atoms[4 * q], j = bond_atoms[4 * q + 1], ty = bond_atoms[4 * q + 2];
        double k2 = bond_params[2 * ty], r0 = bond_params[2 * ty + 1];
        double r[3];
        sep_d(xyzq, i, j, L, h, r);
        double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double fb = k2 * (r0 / d - 1.0);
        for (int a = 0; a < 3; ++a) {
            f[3 * i + a] += fb * r[a];
            f[3 * j + a] -= fb * r[a];
        }
        e += 0.5 * k2 * (d - r0) * (d - r0);
    }
    if (energy) *energy = e;
}

/* src/angle_force_cuda.cu:12-64: E = k (theta - theta0)^2 with params.x = 2k */
void orc_angles(int na, const int *angle_atoms, const float *angle_params, const double *xyzq,
                float lbox, double *f, double *energy)
{
    double L = lbox, h = 0.5 * (double)lbox, e = 0.0;
    for (int q = 0; q < na; ++q) {
        int i = angle_atoms[4 * q], j = angle_atoms[4 * q + 1], k = angle_atoms[4 * q + 2];
        int ty = angle_atoms[4 * q + 3];
        double k2 = angle_params[2 * ty], th0 = angle_params[2 * ty + 1];
        double r1[3], r2[3];
        sep_d(xyzq, i, j, L, h, r1);
        sep_d(xyzq, j, k, L, h, r2);
        double c11 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
        double c22 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
        double c12 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
        double b = -c12 / sqrt(c11 * c22);
        double th;
        if (b >= 1.0) th = 1.0e-16;
        else if (b <= -1.0) th = 3.1415926535f;
        else th = acos(b);
        double fa = k2 * (th - th0) / sqrt(c11 * c22 - c12 * c12);
        for (int a = 0; a < 3; ++a) {
            f[3 * i + a] += fa * (c12 / c11 * r1[a] - r2[a]);
            f[3 * j + a] += fa * ((1.0 + c12 / c22) * r2[a] - (1.0 + c12 / c11) * r1[a]);
            f[3 * k + a] += fa * (r1[a] - c12 / c22 * r2[a]);
        }
        e += 0.5 * k2 * (th - th0) * (th - th0);
    }
    if (energy) *energy = e;
}

/* src/dih_force_cuda.cu:13-136: E = V (1 + cos(n phi - phase)) + scaled 1-4 */
void orc_dihedrals(int nd, const int *dih_atoms, const int *dih_types, const float *dih_params,
                   const float *sc14, const double *xyzq, int n_lj_types, const int *ityp,
                   const int *nbparm, const double *lj_a, const double *lj_b, float lbox,
                   double *f, double *energy)
{
    double L = lbox, h = 0.5 * (double)lbox;
    double et = 0.0, e14c = 0.0, e14v = 0.0;
    for (int q = 0; q < nd; ++q) {
        int ai = dih_atoms[4 * q], aj = dih_atoms[4 * q + 1];
        int ak = dih_atoms[4 * q + 2], al = dih_atoms[4 * q + 3];
        int ty = dih_types[q];
        if (ak > 0 && al > 0) { /* Q20 mirrored: index 0 never takes the 1-4 path */
            int p = nbparm[n_lj_types * ityp[ai] + ityp[al]];
            double A = lj_a[p], B = lj_b[p];
            double scee = sc14[2 * ty], scnb = sc14[2 * ty + 1];
            double r[3];
            sep_d(xyzq, ai, al, L, h, r);
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            double r6 = 1.0 / (d2 * d2 * d2);
            double qq = (double)xyzq[4 * ai + 3] * (double)xyzq[4 * al + 3];
            double fe = qq / (d2 * sqrt(d2) * scee);
            double fv = r6 * (12.0 * A * r6 - 6.0 * B) / (scnb * d2);
            for (int a = 0; a < 3; ++a) {
                f[3 * ai + a] += (fe + fv) * r[a];
                f[3 * al + a] -= (fe + fv) * r[a];
            }
            e14c += qq / (sqrt(d2) * scee);
            e14v += (A * r6 * r6 - B * r6) / scnb;
        }
        ak = abs(ak);
        al = abs(al);
        double r12[3], r23[3], r34[3];
        sep_d(xyzq, ai, aj, L, h, r12);
        sep_d(xyzq, aj, ak, L, h, r23);
        sep_d(xyzq, ak, al, L, h, r34);
#define DOT(u, v) (u[0] * v[0] + u[1] * v[1] + u[2] * v[2])
        double c11 = DOT(r12, r12), c22 = DOT(r23, r23), c12 = DOT(r12, r23);
        double c33 = DOT(r34, r34), c23 = DOT(r23, r34), c13 = DOT(r12, r34);
#undef DOT
        double t1 = c13 * c22 - c12 * c23;
        double t2 = c11 * c23 - c12 * c13;
        double t3 = c12 * c12 - c11 * c22;
        double t4 = c22 * c33 - c23 * c23;
        double t5 = c13 * c23 - c12 * c33;
        double t6 = -t1;
        double b = sqrt(-t3 * t4);
        double a = t6 / b;
        double pn = dih_params[4 * ty], V = dih_params[4 * ty + 1], ph = dih_params[4 * ty + 2];
        double fd = 0.0;
        if (a > -1.0 && a < 1.0) {
            double phi = acos(a);
            fd = pn * V * sin(pn * phi - ph) * c22 / (sin(phi) * b);
            et += V * (1.0 + cos(pn * phi - ph));
        } else {
            double phi = a >= 1.0 ? 0.0 : M_PI;
            et += V * (1.0 + cos(pn * phi - ph));
        }
        double f1[3], f4[3];
        for (int x = 0; x < 3; ++x) {
            f1[x] = fd / t3 * (t1 * r12[x] + t2 * r23[x] + t3 * r34[x]);
            f4[x] = -fd / t4 * (t4 * r12[x] + t5 * r23[x] + t6 * r34[x]);
        }
        for (int x = 0; x < 3; ++x) {
            f[3 * ai + x] += f1[x];
            f[3 * aj + x] += -(1.0 + c12 / c22) * f1[x] + c23 / c22 * f4[x];
            f[3 * ak + x] += c12 / c22 * f1[x] - (1.0 + c23 / c22) * f4[x];
            f[3 * al + x] += f4[x];
        }
    }
    if (energy) {
        energy[0] = et;
        energy[1] = e14c;
        energy[2] = e14v;
    }
}

/* ------------------------------------------------------------------------ */
/* A.15 leap-frog + Andersen, src/leapfrog_cuda.cu:33-67 (curand -> Philox) */
/* ------------------------------------------------------------------------ */
static uint32_t andersen_threshold(float pnu)
{
    double t = (double)pnu * 4294967296.0;
    if (t <= 0.0) return 0u;
    if (t >= 4294967295.0) return 4294967295u;
    return (uint32_t)t;
}

void orc_leapfrog(int n, const float *pos, const float *vel, const double *force, float T,
                  float dt, float pnu, float lbox, uint64_t seed, uint32_t step, uint32_t sys,
                  double *pos_out, double *vel_out, int *collided)
{
    const uint32_t thr = andersen_threshold(pnu);
    const double two_m24 = 5.9604644775390625e-08;
    for (int i = 0; i < n; ++i) {
        double m = vel[4 * i + 3];
        double v[3] = {vel[4 * i], vel[4 * i + 1], vel[4 * i + 2]};
        double x[3] = {pos[4 * i], pos[4 * i + 1], pos[4 * i + 2]};
        uint32_t o[4];
        philox_draw(seed, (uint32_t)i, ORC_STREAM_ANDERSEN, 0u, step, sys, o);
        int hit = o[0] < thr;
        if (hit) {
            uint32_t z[4];
            philox_draw(seed, (uint32_t)i, ORC_STREAM_ANDERSEN, 1u, step, sys, z);
            double u1a = ((z[0] >> 8) + 1) * two_m24, u2a = (z[1] >> 8) * two_m24;
            double u1b = ((z[2] >> 8) + 1) * two_m24, u2b = (z[3] >> 8) * two_m24;
            double ra = sqrt(-2.0 * log(u1a)), rb = sqrt(-2.0 * log(u1b));
            double nrm[3] = {ra * cos(2.0 * M_PI * u2a), ra * sin(2.0 * M_PI * u2a),
                             rb * cos(2.0 * M_PI * u2b)};
            double s = sqrt((double)T / m);
            for (int a = 0; a < 3; ++a) v[a] = nrm[a] * s + force[3 * i + a] * dt / (2.0 * m);
        } else {
            for (int a = 0; a < 3; ++a) v[a] += force[3 * i + a] * dt / m;
        }
        for (int a = 0; a < 3; ++a) {
            x[a] += v[a] * dt;
            if (x[a] > lbox) x[a] -= lbox;
            else if (x[a] < 0.0) x[a] += lbox;
            pos_out[3 * i + a] = x[a];
            vel_out[3 * i + a] = v[a];
        }
        if (collided) collided[i] = hit;
    }
}

/* ------------------------------------------------------------------------ */
/* A.16 umbrella restraint, src/us_force_cuda.cu:14-47, src/us_class.cpp:136-157 */
/* ------------------------------------------------------------------------ */
void orc_us_force(int nbias, const int *atom_list, const float *mass, const float *kumb,
                  float x0, const double *xyzq, float lbox, double *f, double *com,
                  double *energy)
{
    double L = lbox, h = 0.5 * (double)lbox;
    double c[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int b = 0; b < nbias; ++b) {
        int a = atom_list[2 * b], gr = atom_list[2 * b + 1];
        for (int k = 0; k < 3; ++k) c[gr][k] += (double)mass[b] * (double)xyzq[4 * a + k];
    }
    double dv[3];
    for (int k = 0; k < 3; ++k) dv[k] = mi1d(c[0][k] - c[1][k], L, h);
    double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    for (int b = 0; b < nbias; ++b) {
        int a = atom_list[2 * b], gr = atom_list[2 * b + 1];
        double pre = (double)kumb[gr] * (d - (double)x0) * (double)mass[b] / d;
        for (int k = 0; k < 3; ++k) f[3 * a + k] += pre * dv[k];
    }
    if (com)
        for (int k = 0; k < 3; ++k) {
            com[k] = c[0][k];
            com[3 + k] = c[1][k];
        }
    if (energy) *energy = 0.5 * (double)kumb[1] * (d - x0) * (d - x0);
}

/* ------------------------------------------------------------------------ */
/* Restatement of cpu_ispa_force, src/isspa_force.omp.c:11-105.  The float  */
/* and double roundings below follow the original's expression types so that */
/* at one thread the result is identical to the original file's.             */
/* ------------------------------------------------------------------------ */
void orc_srand(unsigned int s) { srand(s); }

static inline float unit_draw(void) /* 1 - 2 U, U = rand()/RAND_MAX in float */
{
    return 1.0f - 2.0f * (float)rand() / (float)RAND_MAX;
}

void orc_cpu_ispa_force(float *xyz, float *f, float *w, float *x0, float *g0, float *gr2,
                        float *alpha, float *lj_A, float *lj_B, int *ityp, int nAtoms, int nMC,
                        int nThreads)
{
    long total = (long)nAtoms * nMC;
#pragma omp parallel for num_threads(nThreads)
    for (long idx = 0; idx < total; ++idx) {
        int atom = (int)(idx % nAtoms);
        int it = ityp[atom];
        /* depth: p(u) ~ 1 - u^2 on [-1,1] by rejection */
        float u = unit_draw();
        float prob = u * u;
        float att = (float)rand() / (float)RAND_MAX;
        while (att < prob) {
            u = unit_draw();
            prob = u * u;
            att = (float)rand() / (float)RAND_MAX;
        }
        float rdep = w[it] * u + x0[it];
        /* direction: Marsaglia's method */
        float a1 = unit_draw(), a2 = unit_draw();
        float s2 = a1 * a1 + a2 * a2;
        while (s2 > 1.0f) {
            a1 = unit_draw();
            a2 = unit_draw();
            s2 = a1 * a1 + a2 * a2;
        }
        float dir[3];
        dir[0] = rdep * (1.0f - 2.0f * s2);
        s2 = 2.0f * sqrt(1.0f - s2);
        dir[1] = rdep * a1 * s2;
        dir[2] = rdep * a2 * s2;
        float pt[3] = {dir[0] + xyz[3 * atom], dir[1] + xyz[3 * atom + 1],
                       dir[2] + xyz[3 * atom + 2]};
        float gnow = 1.0f;
        int inside = 0;
        for (int j = 0; j < nAtoms; ++j) {
            if (j == atom) continue;
            int jt = ityp[j];
            float dd = 0.0f;
            for (int k = 0; k < 3; ++k) {
                float tmp = pt[k] - xyz[3 * j + k];
                dd += tmp * tmp;
            }
            if (dd < gr2[2 * jt]) {
                inside = 1;
                break;
            } else if (dd < gr2[2 * jt + 1]) {
                float tmp = sqrt(dd) - x0[jt];
                gnow *= (-alpha[jt] * tmp * tmp + g0[jt]);
            }
        }
        if (!inside) {
            float ri = 1.0f / rdep;
            float ri2 = ri * ri;
            float ri6 = ri2 * ri2 * ri2;
            float fs = gnow * ri6 * (lj_B[it] - lj_A[it] * ri6);
#pragma omp atomic
            f[3 * atom] += fs * dir[0];
#pragma omp atomic
            f[3 * atom + 1] += fs * dir[1];
#pragma omp atomic
            f[3 * atom + 2] += fs * dir[2];
        }
    }
}
