/*
 * oracle_diff.c -- FP64 CPU restatement of the differentiable IS-SPA +
 * nonbond force model and its adjoint (TEST INFRASTRUCTURE ONLY: the parity
 * oracle of the device adjoint; the product path never calls it).
 *
 * The model is the plain path (src/isspa_force_cuda.cu:110-327 field and
 * force passes, src/nonbond_force_cuda.cu:31-174 all-pairs term) with the
 * per-type tables replaced by the analytic functions the tables are sampled
 * from, so that forces are differentiable in the per-type parameters
 * theta_t = (sigma, eps, x0, g0, alpha) and the per-atom charges q_i:
 *
 *   sigma_ts = (sigma + sigma_s)/2, eps_ts = sqrt(eps eps_s)
 *   r_lo = max(x0 - sqrt(g0/alpha), core sigma_ts),  r_hi = x0 + sqrt((g0-1)/alpha)
 *   S(d) = smoothstep((d - r_lo)/delta + 1)            (0 below r_lo - delta, 1 above r_lo)
 *   g(d) = S(d) max(0, g0 - alpha (d - x0)^2)   for d < r_hi,   1 beyond
 *   E(d) = c exp(-(d - x0)^2 / 2) S(d)
 *   F(d) = 24 eps_ts / d (2 (sigma_ts/d)^12 - (sigma_ts/d)^6) S(d)
 *   nonbond: A_ij = 4 eps_ij sigma_ij^12, B_ij = 4 eps_ij sigma_ij^6 (Lorentz-Berthelot)
 *
 * The switch S makes g, E and F continuous in d and theta where the table
 * generator's core floor cuts them off, so central differences are well
 * defined.  MC points and in-sphere decisions depend on positions and rmax
 * only, which are fixed; the same FP32 decisions as the device are used.
 *
 * The force pass is written in the reference's three-term form (s1, s2, s3
 * of src/isspa_force_cuda.cu:270-318), not in the device's factorised form,
 * and the adjoint is reverse-mode differentiation of exactly this code.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline float mi1f(float r, float L, float h)
{
    if (r > h) return r - L;
    if (r < -h) return r + L;
    return r;
}

static inline float sep_f32(const float *a, const float *b, float L, float h, float r[3])
{
    r[0] = mi1f(a[0] - b[0], L, h);
    r[1] = mi1f(a[1] - b[1], L, h);
    r[2] = mi1f(a[2] - b[2], L, h);
    return fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0]));
}

/* Per-type analytic functions at distance d: val = {g, E, F} and their
 * derivatives dval[5*k + p] for function k and parameter p in
 * (sigma, eps, x0, g0, alpha). */
void orc_diff_type_fn(const orc_dparams *P, int t, double d, double val[3], double dval[15])
{
    const double sig = P->sigma[t], eps = P->eps[t], x0 = P->x0[t], g0 = P->g0[t];
    const double al = P->alpha[t], c = P->camp[t];
    const double s_ts = 0.5 * (sig + P->sig_s), e_ts = sqrt(eps * P->eps_s);
    const double w = sqrt(g0 / al);
    const double r_hi = x0 + sqrt((g0 - 1.0) / al);
    double rlo, drlo[5] = {0, 0, 0, 0, 0};
    if (x0 - w >= P->core * s_ts) {
        rlo = x0 - w;
        drlo[2] = 1.0;
        drlo[3] = -1.0 / (2.0 * w * al);
        drlo[4] = w / (2.0 * al);
    } else {
        rlo = P->core * s_ts;
        drlo[0] = 0.5 * P->core;
    }
    const double z = (d - rlo) / P->delta + 1.0;
    double S, dSdz;
    if (z <= 0.0) {
        S = 0.0;
        dSdz = 0.0;
    } else if (z >= 1.0) {
        S = 1.0;
        dSdz = 0.0;
    } else {
        S = z * z * (3.0 - 2.0 * z);
        dSdz = 6.0 * z * (1.0 - z);
    }
    double dS[5];
    for (int p = 0; p < 5; ++p) dS[p] = -dSdz / P->delta * drlo[p];
    memset(dval, 0, 15 * sizeof(double));
    /* g */
    if (d >= r_hi) {
        val[0] = 1.0;
    } else {
        const double u = d - x0;
        const double Pv = g0 - al * u * u;
        if (Pv <= 0.0) {
            val[0] = 0.0;
        } else {
            const double dP[5] = {0.0, 0.0, 2.0 * al * u, 1.0, -u * u};
            val[0] = S * Pv;
            for (int p = 0; p < 5; ++p) dval[p] = dS[p] * Pv + S * dP[p];
        }
    }
    /* E */
    {
        const double u = d - x0;
        const double Er = c * exp(-0.5 * u * u);
        val[1] = Er * S;
        for (int p = 0; p < 5; ++p) dval[5 + p] = Er * dS[p];
        dval[5 + 2] += Er * u * S;
    }
    /* F */
    {
        const double sr6 = pow(s_ts / d, 6.0);
        const double FLJ = 24.0 * e_ts / d * (2.0 * sr6 * sr6 - sr6);
        const double dFds = 24.0 * e_ts / d * (24.0 * sr6 * sr6 - 6.0 * sr6) / s_ts;
        val[2] = FLJ * S;
        for (int p = 0; p < 5; ++p) dval[10 + p] = FLJ * dS[p];
        dval[10 + 0] += 0.5 * dFds * S;
        dval[10 + 1] += FLJ / (2.0 * eps) * S;
    }
}

/* Langevin moments and their x-derivatives (FP64): the all-positive Bessel
 * series ratio below x = 5, the coth recurrence above. */
static void moments_and_derivs(double x, double m[3], double dm[3])
{
    if (x < 5.0) {
        const double h = 0.5 * x * x;
        double s[4], ds[4];
        double first = 1.0;
        for (int l = 0; l < 4; ++l) {
            if (l > 0) first /= (double)(2 * l + 1);
            double t = first, sum = first, dsum = 0.0;
            for (int k = 0; k < 30; ++k) {
                t *= h / ((double)(k + 1) * (double)(2 * l + 2 * k + 3));
                sum += t;
                dsum += (double)(k + 1) * t / h; /* d/dh of c h^(k+1) */
            }
            s[l] = sum;
            ds[l] = h > 0 ? dsum : (l == 0 ? 0.0 : 0.0);
        }
        if (h == 0.0) {
            /* d/dh of the series at 0 is its k=1 coefficient */
            for (int l = 0; l < 4; ++l) {
                double f = 1.0;
                for (int q = 1; q <= l; ++q) f /= (double)(2 * q + 1);
                ds[l] = f / (double)(2 * l + 3);
            }
        }
        const double s0 = s[0];
        for (int l = 1; l <= 3; ++l) {
            const double xl = pow(x, (double)l);
            m[l - 1] = xl * s[l] / s0;
            /* d/dx (x^l s_l/s_0) with dh/dx = x */
            dm[l - 1] = (l * pow(x, (double)(l - 1)) * s[l] / s0) +
                        xl * x * (ds[l] * s0 - s[l] * ds[0]) / (s0 * s0);
        }
        return;
    }
    const double coth = 1.0 / tanh(x);
    const double sh = sinh(x);
    m[0] = coth - 1.0 / x;
    m[1] = 1.0 - 3.0 * m[0] / x;
    m[2] = m[0] - 5.0 * m[1] / x;
    dm[0] = 1.0 / (x * x) - 1.0 / (sh * sh);
    dm[1] = -3.0 * dm[0] / x + 3.0 * m[0] / (x * x);
    dm[2] = dm[0] - 5.0 * dm[1] / x + 5.0 * m[1] / (x * x);
}

void orc_langevin_derivs(double x, double m[3], double dm[3]) { moments_and_derivs(x, m, dm); }

static int excluded(int a, int b, const int *pre, const int *list)
{
    /* the lower-index atom's slice, early exit on a larger entry
     * (src/nonbond_force_cuda.cu:71-108) */
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const int start = lo == 0 ? 0 : pre[lo - 1], stop = pre[lo];
    for (int e = start; e < stop; ++e) {
        if (list[e] == hi) return 1;
        if (list[e] > hi) return 0;
    }
    return 0;
}

/* per-point field of the analytic model; psi[8m..] = E(3), E0(3), g, cnt */
static void diff_field(int n, int nmc, const float *xyz4, const double *q, const int *itype,
                       const float *pts, float lbox, const orc_dparams *P, double *psi)
{
    const long Pn = (long)n * nmc;
    const float hb = 0.5f * lbox;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < Pn; ++m) {
        const int k = (int)(m / nmc);
        const float R = P->rmax[itype[k]];
        const float R2 = R * R;
        double E[3] = {0, 0, 0}, E0[3] = {0, 0, 0}, g = 1.0;
        int cnt = 0;
        for (int j = 0; j < n; ++j) {
            float rf[3];
            const float d2f = sep_f32(pts + 4 * m, xyz4 + 4 * j, lbox, hb, rf);
            const double r[3] = {rf[0], rf[1], rf[2]};
            const double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            const double s = -P->e0 * q[j] / (d * d * d);
            for (int a = 0; a < 3; ++a) E[a] += s * r[a];
            if (d2f <= R2) {
                double v[3], dv[15];
                orc_diff_type_fn(P, itype[j], d, v, dv);
                cnt++;
                g *= v[0];
                for (int a = 0; a < 3; ++a) E[a] += v[1] * r[a] / d;
            } else {
                for (int a = 0; a < 3; ++a) E0[a] += s * r[a];
            }
        }
        double *o = psi + 8 * m;
        for (int a = 0; a < 3; ++a) {
            o[a] = E[a];
            o[3 + a] = E0[a];
        }
        o[6] = g;
        o[7] = cnt;
    }
}

void orc_diff_forward(int n, int nmc, const float *xyz4, const double *q, const int *itype,
                      const float *pts, float lbox, const orc_dparams *P, const int *pre,
                      const int *list, double *f_isspa, double *f_nb, double *psi_out)
{
    const long Pn = (long)n * nmc;
    double *psi = psi_out ? psi_out : (double *)malloc(sizeof(double) * 8 * Pn);
    diff_field(n, nmc, xyz4, q, itype, pts, lbox, P, psi);
    double *hat = (double *)malloc(sizeof(double) * 3 * Pn);
    double *mom = (double *)malloc(sizeof(double) * 3 * Pn);
    for (long m = 0; m < Pn; ++m) {
        const double *e = psi + 8 * m;
        const double x = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        for (int a = 0; a < 3; ++a) hat[3 * m + a] = x > 0 ? e[a] / x : 0.0;
        orc_langevin_moments(x, mom + 3 * m);
    }
    const float hb = 0.5f * lbox;
    const double p0 = P->p0, q0 = P->q0, o0 = P->o0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        const int ti = itype[i];
        const float R2 = P->rmax[ti] * P->rmax[ti];
        const double vt = P->vtot[ti], qi = q[i];
        double f[3] = {0, 0, 0};
        for (long m = 0; m < Pn; ++m) {
            float rf[3];
            const float d2f = sep_f32(pts + 4 * m, xyz4 + 4 * i, lbox, hb, rf);
            const double r[3] = {rf[0], rf[1], rf[2]};
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], d = sqrt(d2);
            const double *eh = hat + 3 * m, *M = mom + 3 * m, *ps = psi + 8 * m;
            const double nu = ps[6] * vt / ps[7];
            const double Rz = (eh[0] * r[0] + eh[1] * r[1] + eh[2] * r[2]) / d;
            const double dp1 = 3.0 * Rz, dp2 = 7.5 * Rz * Rz - 1.5, dp3 = (17.5 * Rz * Rz - 7.5) * Rz;
            const double s1 = -qi * p0 * M[0] * nu / (d2 * d);
            const double s2 = -qi * q0 * M[1] * nu / (d2 * d2);
            const double s3 = -qi * o0 * M[2] * nu / (d2 * d2 * d);
            for (int a = 0; a < 3; ++a)
                f[a] += s1 * (r[a] * dp1 / d - eh[a]) + s2 * (r[a] * dp2 / d - dp1 * eh[a]) +
                        s3 * (r[a] * dp3 / d - dp2 * eh[a]);
            if (d2f <= R2) {
                double v[3], dv[15];
                orc_diff_type_fn(P, ti, d, v, dv);
                for (int a = 0; a < 3; ++a) f[a] += -v[2] * nu * r[a] / d;
            } else {
                const double ev[3] = {ps[3] / 3.0, ps[4] / 3.0, ps[5] / 3.0};
                const double pr = 3.0 * (ev[0] * r[0] + ev[1] * r[1] + ev[2] * r[2]) / d2;
                const double fs = -qi * p0 / (d2 * d) * vt / ps[7];
                for (int a = 0; a < 3; ++a) f[a] += fs * (pr * r[a] - ev[a]);
            }
        }
        for (int a = 0; a < 3; ++a) f_isspa[3 * i + a] = f[a];
    }
    /* nonbond */
    const double L = lbox, h = 0.5 * L, fep = 1.0 - 1.0 / P->ep;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        double f[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double r[3];
            for (int a = 0; a < 3; ++a) {
                double v = (double)xyz4[4 * i + a] - (double)xyz4[4 * j + a];
                if (v > h) v -= L;
                else if (v < -h) v += L;
                r[a] = v;
            }
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], d = sqrt(d2);
            double s = 0.0;
            if (!excluded(i, j, pre, list)) {
                const int a = itype[i], b = itype[j];
                const double sg = 0.5 * (P->sigma[a] + P->sigma[b]), ep = sqrt(P->eps[a] * P->eps[b]);
                const double A = 4.0 * ep * pow(sg, 12.0), B = 4.0 * ep * pow(sg, 6.0);
                const double r6 = 1.0 / (d2 * d2 * d2);
                s += (12.0 * A * r6 * r6 - 6.0 * B * r6) / d2 + q[i] * q[j] / (d2 * d);
            }
            const double R = P->rmax[itype[j]];
            if (d > 2.0 * R)
                s += -2.0 * q[i] * q[j] * fep / (3.0 * d2 * d);
            else
                s += -q[i] * q[j] * fep * (8.0 * R - 3.0 * d) / (24.0 * R * R * R * R);
            for (int a = 0; a < 3; ++a) f[a] += s * r[a];
        }
        for (int a = 0; a < 3; ++a) f_nb[3 * i + a] = f[a];
    }
    free(hat);
    free(mom);
    if (!psi_out) free(psi);
}

/* Adjoint: given lam = dL/dF (F = f_isspa + f_nb, double[N*3]), the
 * gradients gq[N] = dL/dq and gt[T*5] = dL/dtheta_t. */
void orc_diff_vjp(int n, int nmc, const float *xyz4, const double *q, const int *itype,
                  const float *pts, float lbox, const orc_dparams *P, const int *pre,
                  const int *list, const double *lam, double *gq, double *gt)
{
    const long Pn = (long)n * nmc;
    const int T = P->n_types;
    double *psi = (double *)malloc(sizeof(double) * 8 * Pn);
    diff_field(n, nmc, xyz4, q, itype, pts, lbox, P, psi);
    double *hat = (double *)malloc(sizeof(double) * 3 * Pn);
    double *mom = (double *)malloc(sizeof(double) * 3 * Pn);
    double *dmom = (double *)malloc(sizeof(double) * 3 * Pn);
    double *eps_m = (double *)malloc(sizeof(double) * Pn);
    for (long m = 0; m < Pn; ++m) {
        const double *e = psi + 8 * m;
        const double x = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        eps_m[m] = x;
        for (int a = 0; a < 3; ++a) hat[3 * m + a] = x > 0 ? e[a] / x : 0.0;
        moments_and_derivs(x, mom + 3 * m, dmom + 3 * m);
    }
    double *ga = (double *)calloc((size_t)n * 5, sizeof(double)); /* per-atom type-param bars */
    for (int i = 0; i < n; ++i) gq[i] = 0.0;
    const float hb = 0.5f * lbox;
    const double p0 = P->p0, q0 = P->q0, o0 = P->o0;

    /* (1) receiver side of the force pass: dq_i direct, F-function params */
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        const int ti = itype[i];
        const float R2 = P->rmax[ti] * P->rmax[ti];
        const double vt = P->vtot[ti];
        const double *lm = lam + 3 * i;
        double gqi = 0.0, gpi[5] = {0, 0, 0, 0, 0};
        for (long m = 0; m < Pn; ++m) {
            float rf[3];
            const float d2f = sep_f32(pts + 4 * m, xyz4 + 4 * i, lbox, hb, rf);
            const double r[3] = {rf[0], rf[1], rf[2]};
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], d = sqrt(d2);
            const double *eh = hat + 3 * m, *M = mom + 3 * m, *ps = psi + 8 * m;
            const double nu = ps[6] * vt / ps[7];
            const double Rz = (eh[0] * r[0] + eh[1] * r[1] + eh[2] * r[2]) / d;
            const double dp1 = 3.0 * Rz, dp2 = 7.5 * Rz * Rz - 1.5, dp3 = (17.5 * Rz * Rz - 7.5) * Rz;
            const double lr = lm[0] * r[0] + lm[1] * r[1] + lm[2] * r[2];
            const double le = lm[0] * eh[0] + lm[1] * eh[1] + lm[2] * eh[2];
            const double L1 = lr * dp1 / d - le, L2 = lr * dp2 / d - dp1 * le,
                         L3 = lr * dp3 / d - dp2 * le;
            gqi += -p0 * M[0] * nu / (d2 * d) * L1 - q0 * M[1] * nu / (d2 * d2) * L2 -
                   o0 * M[2] * nu / (d2 * d2 * d) * L3;
            if (d2f <= R2) {
                double v[3], dv[15];
                orc_diff_type_fn(P, ti, d, v, dv);
                for (int p = 0; p < 5; ++p) gpi[p] += -dv[10 + p] * nu * lr / d;
            } else {
                const double ev[3] = {ps[3] / 3.0, ps[4] / 3.0, ps[5] / 3.0};
                const double pr = 3.0 * (ev[0] * r[0] + ev[1] * r[1] + ev[2] * r[2]) / d2;
                const double lev = lm[0] * ev[0] + lm[1] * ev[1] + lm[2] * ev[2];
                gqi += -p0 / (d2 * d) * vt / ps[7] * (pr * lr - lev);
            }
        }
        gq[i] += gqi;
        for (int p = 0; p < 5; ++p) ga[5 * i + p] += gpi[p];
    }

    /* (2) point side of the force pass: bars of E-hat, moments, g, E0 per point,
     * then back through the epilogue to dL/dE and dL/dE0 */
    double *Ebar = (double *)malloc(sizeof(double) * 3 * Pn);
    double *E0bar = (double *)malloc(sizeof(double) * 3 * Pn);
    double *gbar = (double *)malloc(sizeof(double) * Pn);
#pragma omp parallel for schedule(static)
    for (long m = 0; m < Pn; ++m) {
        const double *eh = hat + 3 * m, *M = mom + 3 * m, *ps = psi + 8 * m;
        double hb_[3] = {0, 0, 0}, Mb[3] = {0, 0, 0}, e0b[3] = {0, 0, 0}, gb = 0.0;
        for (int i = 0; i < n; ++i) {
            const int ti = itype[i];
            const float R2 = P->rmax[ti] * P->rmax[ti];
            const double vt = P->vtot[ti], qi = q[i];
            const double *lm = lam + 3 * i;
            float rf[3];
            const float d2f = sep_f32(pts + 4 * m, xyz4 + 4 * i, lbox, hb, rf);
            const double r[3] = {rf[0], rf[1], rf[2]};
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], d = sqrt(d2);
            const double nu_g = vt / ps[7]; /* nu / g */
            const double nu = ps[6] * nu_g;
            const double Rz = (eh[0] * r[0] + eh[1] * r[1] + eh[2] * r[2]) / d;
            const double dp1 = 3.0 * Rz, dp2 = 7.5 * Rz * Rz - 1.5, dp3 = (17.5 * Rz * Rz - 7.5) * Rz;
            const double lr = lm[0] * r[0] + lm[1] * r[1] + lm[2] * r[2];
            const double le = lm[0] * eh[0] + lm[1] * eh[1] + lm[2] * eh[2];
            const double L1 = lr * dp1 / d - le, L2 = lr * dp2 / d - dp1 * le,
                         L3 = lr * dp3 / d - dp2 * le;
            const double c1 = -qi * p0 / (d2 * d), c2 = -qi * q0 / (d2 * d2),
                         c3 = -qi * o0 / (d2 * d2 * d);
            const double s1 = c1 * M[0] * nu, s2 = c2 * M[1] * nu, s3 = c3 * M[2] * nu;
            Mb[0] += c1 * nu * L1;
            Mb[1] += c2 * nu * L2;
            Mb[2] += c3 * nu * L3;
            gb += (c1 * M[0] * L1 + c2 * M[1] * L2 + c3 * M[2] * L3) * nu_g;
            const double dRz = s1 * 3.0 * lr / d + s2 * (15.0 * Rz * lr / d - 3.0 * le) +
                               s3 * ((52.5 * Rz * Rz - 7.5) * lr / d - 15.0 * Rz * le);
            const double dle = -(s1 + s2 * dp1 + s3 * dp2);
            for (int a = 0; a < 3; ++a) hb_[a] += dRz * r[a] / d + dle * lm[a];
            if (d2f <= R2) {
                double v[3], dv[15];
                orc_diff_type_fn(P, ti, d, v, dv);
                gb += -v[2] * nu_g * lr / d;
            } else {
                /* cdd = -q p0/d^3 (3 (ev.r) r/d^2 - ev) vt/cnt, ev = E0/3 */
                const double k = -qi * p0 / (d2 * d) * vt / ps[7] / 3.0;
                for (int a = 0; a < 3; ++a) e0b[a] += k * (3.0 * lr * r[a] / d2 - lm[a]);
            }
        }
        /* epilogue backward: E-hat = E/|E|, M_l(|E|) */
        const double x = eps_m[m];
        const double *dM = dmom + 3 * m;
        const double xb = Mb[0] * dM[0] + Mb[1] * dM[1] + Mb[2] * dM[2];
        const double hdot = hb_[0] * eh[0] + hb_[1] * eh[1] + hb_[2] * eh[2];
        for (int a = 0; a < 3; ++a) {
            Ebar[3 * m + a] = xb * eh[a] + (x > 0 ? (hb_[a] - hdot * eh[a]) / x : 0.0);
            E0bar[3 * m + a] = e0b[a];
        }
        gbar[m] = gb;
    }

    /* (3) field pass backward, atom side: dq_j and the g / E params of t_j */
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < n; ++j) {
        const int tj = itype[j];
        double gqj = 0.0, gpj[5] = {0, 0, 0, 0, 0};
        for (long m = 0; m < Pn; ++m) {
            const int k = (int)(m / nmc);
            const float R = P->rmax[itype[k]];
            const float R2 = R * R;
            float rf[3];
            const float d2f = sep_f32(pts + 4 * m, xyz4 + 4 * j, lbox, hb, rf);
            const double r[3] = {rf[0], rf[1], rf[2]};
            const double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            const double *Eb = Ebar + 3 * m, *E0b = E0bar + 3 * m;
            const double Er = Eb[0] * r[0] + Eb[1] * r[1] + Eb[2] * r[2];
            gqj += -P->e0 / (d * d * d) * Er;
            if (d2f <= R2) {
                double v[3], dv[15];
                orc_diff_type_fn(P, tj, d, v, dv);
                for (int p = 0; p < 5; ++p) gpj[p] += Er / d * dv[5 + p];
                const double gm = psi[8 * m + 6];
                if (gm != 0.0 && v[0] != 0.0)
                    for (int p = 0; p < 5; ++p) gpj[p] += gbar[m] * gm / v[0] * dv[p];
            } else {
                gqj += -P->e0 / (d * d * d) * (E0b[0] * r[0] + E0b[1] * r[1] + E0b[2] * r[2]);
            }
        }
        gq[j] += gqj;
        for (int p = 0; p < 5; ++p) ga[5 * j + p] += gpj[p];
    }

    /* (4) nonbond: per receiver i over partners j, with (lam_i - lam_j).r_ij for
     * the symmetric LJ/Coulomb part and the general form for the pair term */
    const double L = lbox, h = 0.5 * L, fep = 1.0 - 1.0 / P->ep;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        const int a = itype[i];
        double gqi = 0.0, gsi = 0.0, gei = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double r[3];
            for (int c = 0; c < 3; ++c) {
                double v = (double)xyz4[4 * i + c] - (double)xyz4[4 * j + c];
                if (v > h) v -= L;
                else if (v < -h) v += L;
                r[c] = v;
            }
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], d = sqrt(d2);
            const double li = lam[3 * i] * r[0] + lam[3 * i + 1] * r[1] + lam[3 * i + 2] * r[2];
            const double lj = lam[3 * j] * r[0] + lam[3 * j + 1] * r[1] + lam[3 * j + 2] * r[2];
            const double cij = li - lj;
            if (!excluded(i, j, pre, list)) {
                const int b = itype[j];
                const double sg = 0.5 * (P->sigma[a] + P->sigma[b]), ep = sqrt(P->eps[a] * P->eps[b]);
                const double r6 = 1.0 / (d2 * d2 * d2);
                const double A = 4.0 * ep * pow(sg, 12.0), B = 4.0 * ep * pow(sg, 6.0);
                const double dA = 48.0 * ep * pow(sg, 11.0), dB = 24.0 * ep * pow(sg, 5.0);
                gsi += cij * 0.5 * (12.0 * dA * r6 * r6 - 6.0 * dB * r6) / d2;
                gei += cij * (12.0 * A * r6 * r6 - 6.0 * B * r6) / d2 / (2.0 * P->eps[a]);
                gqi += cij * q[j] / (d2 * d);
            }
            /* pair term s_ij = q_i q_j phi(d, R_j): d/dq_i of sum over ordered
             * pairs = q_j [lam_i.r phi(d,R_j) - lam_j.r phi(d,R_i)] */
            const double Rj = P->rmax[itype[j]], Ri = P->rmax[a];
            const double phj = d > 2.0 * Rj ? -2.0 * fep / (3.0 * d2 * d)
                                            : -fep * (8.0 * Rj - 3.0 * d) / (24.0 * Rj * Rj * Rj * Rj);
            const double phi = d > 2.0 * Ri ? -2.0 * fep / (3.0 * d2 * d)
                                            : -fep * (8.0 * Ri - 3.0 * d) / (24.0 * Ri * Ri * Ri * Ri);
            gqi += q[j] * (li * phj - lj * phi);
        }
        gq[i] += gqi;
        ga[5 * i + 0] += gsi;
        ga[5 * i + 1] += gei;
    }
    for (int k = 0; k < 5 * T; ++k) gt[k] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < 5; ++p) gt[5 * itype[i] + p] += ga[5 * i + p];
    free(ga);
    free(Ebar);
    free(E0bar);
    free(gbar);
    free(hat);
    free(mom);
    free(dmom);
    free(eps_m);
    free(psi);
}
