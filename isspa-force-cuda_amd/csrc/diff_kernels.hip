// Differentiable IS-SPA + nonbond model: parameter preparation, forward
// reduction and the hand-written adjoint (vector-Jacobian product) for gfx950.
//
// Forward (reusing the plain kernels with analytic per-type functions):
//   points -> field -> epilogue -> force partials -> k_diff_reduce
//   F_i = -q_i vtot_i S_e,i - vtot_i S_lj,i  (+ nonbond)
// Adjoint for a cotangent lambda = dL/dF, in reverse order:
//   k_vjp_points   lanes = MC points, receivers streamed: bars of the point
//                  record (E-hat, A1..A3, E0c, w) from every receiver
//   k_vjp_epilogue per point, FP64: back through |E|, E-hat, the Langevin
//                  moments and w = g/cnt to dL/dE, dL/dE0, g dL/dg
//   k_vjp_atoms    lanes = atoms, points streamed: dL/dq_j through the field,
//                  and the per-atom bars of the type parameters through g, E
//                  (field side) and F (receiver side)
//   k_vjp_nonbond  lanes = receivers, partners streamed
//   k_vjp_finish   per atom: sum the chunk partials in order, add the direct
//                  charge term lambda_i . (-vtot_i S_e,i)
//   k_type_reduce  per type, FP64, fixed order
// Every accumulation is an FP32 tile sum flushed into FP64, summed over
// partials in a fixed order: results are deterministic.
#include "device_common.hpp"
#include "diff_device.hpp"
#include "kernels.hpp"

namespace isspa {

constexpr int kB = 256;
constexpr int kTile = 64;

static inline int nblk(long long n) { return (int)((n + kB - 1) / kB); }

// ---------------------------------------------------------------------------
// parameters -> per-type constants and LJ pair tables
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_diff_prep_types(DiffPrepArgs a)
{
    const int T = a.T;
    for (int t = threadIdx.x; t < T; t += kB) {
        const double sig = a.theta[5 * t], eps = a.theta[5 * t + 1], x0 = a.theta[5 * t + 2];
        const double g0 = a.theta[5 * t + 3], al = a.theta[5 * t + 4];
        const double sts = 0.5 * (sig + a.sig_s), ets = sqrt(eps * a.eps_s);
        const double w = sqrt(g0 / al);
        DiffType d;
        d.sts = (float)sts;
        d.ets = (float)ets;
        d.x0 = (float)x0;
        d.g0 = (float)g0;
        d.al = (float)al;
        d.c = (float)a.camp[t];
        d.rhi = (float)(x0 + sqrt((g0 - 1.0) / al));
        d.idel = (float)(1.0 / a.delta);
        d.eps = (float)eps;
        for (int p = 0; p < 5; ++p) d.drlo[p] = 0.0f;
        if (x0 - w >= a.core * sts) {
            d.rlo = (float)(x0 - w);
            d.drlo[2] = 1.0f;
            d.drlo[3] = (float)(-1.0 / (2.0 * w * al));
            d.drlo[4] = (float)(w / (2.0 * al));
        } else {
            d.rlo = (float)(a.core * sts);
            d.drlo[0] = (float)(0.5 * a.core);
        }
        d.pad = 0.0f;
        a.types[t] = d;
    }
    for (int k = threadIdx.x; k < T * T; k += kB) {
        const int i = k / T, j = k % T;
        const double s = 0.5 * (a.theta[5 * i] + a.theta[5 * j]);
        const double e = sqrt(a.theta[5 * i + 1] * a.theta[5 * j + 1]);
        const double s2 = s * s, s6 = s2 * s2 * s2;
        const double A = 4.0 * e * s6 * s6, B = 4.0 * e * s6;
        a.nb_lj[k] = make_float2((float)(12.0 * A), (float)(6.0 * B));
        // {A, B, dA/dsigma_ij, dB/dsigma_ij}
        a.nb_ld[k] = make_float4((float)A, (float)B, (float)(12.0 * A / s), (float)(6.0 * B / s));
    }
}

// positions with the charges as FP32 .w; receiver records for the adjoint
__global__ __launch_bounds__(kB) void k_diff_prep_atoms(DiffAtomArgs a)
{
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.n) return;
    const float4 x = a.xyz[i];
    const double q = a.q[i];
    a.xq[i] = make_float4(x.x, x.y, x.z, (float)q);
    if (!a.lam) return;
    const int t = a.itype[i];
    const double vt = a.vtot[t];
    const double l0 = a.lam[3 * i], l1 = a.lam[3 * i + 1], l2 = a.lam[3 * i + 2];
    const double cq = -q * vt;
    float4 *r = a.arec + 4 * (size_t)i;
    r[0] = make_float4(x.x, x.y, x.z, a.rmax2[t]);
    r[1] = make_float4((float)(cq * l0), (float)(cq * l1), (float)(cq * l2), __int_as_float(t));
    r[2] = make_float4((float)(-vt * l0), (float)(-vt * l1), (float)(-vt * l2), (float)q);
    r[3] = make_float4((float)l0, (float)l1, (float)l2, 0.0f);
}

// F_i = -q_i vtot_i S_e - vtot_i S_lj (overwrites f), S_e kept for the adjoint
__global__ __launch_bounds__(kB) void k_diff_reduce(DiffReduceArgs a)
{
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.n) return;
    const WinDesc W = a.wins[a.atom_win[i]];
    const int il = i - W.a0;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < W.chunks; ++c) {
        const long long o = W.ppo + (long long)c * W.n + il;
        const double2 *e = reinterpret_cast<const double2 *>(a.part + o * force_part_stride(false));
        for (int h = 0; h < 3; ++h) {
            const double2 d = e[h];
            s[2 * h] += d.x;
            s[2 * h + 1] += d.y;
        }
    }
    const double vt = a.vtot[a.itype[i]];
    const double qv = -a.q[i] * vt;
    for (int c = 0; c < 3; ++c) {
        a.se[3 * i + c] = s[c];
        a.f[3 * i + c] = qv * s[c] - vt * s[3 + c];
    }
}

// ---------------------------------------------------------------------------
// adjoint, point side: lanes own MC points, receivers stream (wave-uniform)
// ---------------------------------------------------------------------------
// per-type records of the analytic model in LDS for the in-sphere branches
// when they fit (at most kDtLdsFloats floats); the whole block calls this
constexpr int kDtLdsFloats = 4096;
__host__ __device__ inline int dtype_floats(int ntypes)
{
    return ntypes * (int)(sizeof(DiffType) / sizeof(float));
}

__device__ __forceinline__ const DiffType *stage_dtypes(const VjpArgs &a)
{
    extern __shared__ float lds_dt[];
    const int nf = dtype_floats(a.ntypes);
    if (nf > kDtLdsFloats) return a.dtypes;
    const float *src = reinterpret_cast<const float *>(a.dtypes);
    for (int t = threadIdx.x; t < nf; t += kB) lds_dt[t] = src[t];
    __syncthreads();
    return reinterpret_cast<const DiffType *>(lds_dt);
}

struct PointBar {
    float hx, hy, hz;   // E-hat bar
    float a1, a2, a3;   // A1..A3 bar
    float cx, cy, cz;   // E0c bar
    float w;            // w bar
};

template <bool PBC>
__device__ __forceinline__ void vjp_points_tile(const VjpArgs &a, const DiffType *dtypes, int j0,
                                                int j1, float L, float h, float px, float py,
                                                float pz, float Ex, float Ey, float Ez, float A1,
                                                float A2, float A3, PointBar &b)
{
    const float4 *__restrict__ ar = a.arec;
    for (int j = j0; j < j1; ++j) {
        const float4 r0 = ar[4 * j + 0];  // wave-uniform
        const float4 r1 = ar[4 * j + 1];
        const float4 r2 = ar[4 * j + 2];
        float rx = px - r0.x, ry = py - r0.y, rz = pz - r0.z;
        if (PBC) {
            rx = min_image1(rx, L, h);
            ry = min_image1(ry, L, h);
            rz = min_image1(rz, L, h);
        }
        const float d2 = dist2_f32(rx, ry, rz);
        const float rinv = rsqrt_f32(d2);
        const float u = __builtin_fmaf(Ez, rz, __builtin_fmaf(Ey, ry, Ex * rx)) * rinv;
        const float u2 = u * u;
        const float P2 = __builtin_fmaf(7.5f, u2, -1.5f);
        const float P3 = __builtin_fmaf(17.5f, u2, -7.5f) * u;
        const float dP2 = 15.0f * u;
        const float dP3 = __builtin_fmaf(52.5f, u2, -7.5f);
        const float rinv2 = rinv * rinv;
        const float rinv3 = rinv2 * rinv;
        const float rinv4 = rinv2 * rinv2;
        const float rinv5 = rinv4 * rinv;
        const float rinv6 = rinv4 * rinv2;
        const float la = __builtin_fmaf(r1.z, rz, __builtin_fmaf(r1.y, ry, r1.x * rx));  // Lambda.r
        const float lb = __builtin_fmaf(r1.z, Ez, __builtin_fmaf(r1.y, Ey, r1.x * Ex));  // Lambda.E-hat
        const float u3 = 3.0f * u;
        b.a1 += la * u3 * rinv4 - lb * rinv3;
        b.a2 += la * rinv5 * P2 - lb * u3 * rinv4;
        b.a3 += la * rinv6 * P3 - lb * rinv5 * P2;
        const float dal = 3.0f * A1 * rinv4 + A2 * dP2 * rinv5 + A3 * dP3 * rinv6;
        const float dk = 3.0f * A2 * rinv4 + dP2 * A3 * rinv5;
        const float k = u3 * A2 * rinv4 + P2 * A3 * rinv5;
        const float cu = (la * dal - lb * dk) * rinv;
        const float ce = -(rinv3 * A1 + k);
        b.hx += cu * rx + ce * r1.x;
        b.hy += cu * ry + ce * r1.y;
        b.hz += cu * rz + ce * r1.z;
        if (d2 > r0.w) {
            const float c5 = rinv5 * la, c3 = rinv3 * (1.0f / 3.0f);
            b.cx += c5 * rx - c3 * r1.x;
            b.cy += c5 * ry - c3 * r1.y;
            b.cz += c5 * rz - c3 * r1.z;
        } else {
            const float lp = __builtin_fmaf(r2.z, rz, __builtin_fmaf(r2.y, ry, r2.x * rx));
            const int ti = __float_as_int(r1.w);
            b.w += diff_F(dtypes[ti], d2 * rinv, rinv) * lp * rinv;
        }
    }
}

__global__ __launch_bounds__(kB) void k_vjp_points(VjpArgs a)
{
    const int4 bt = a.fblk[blockIdx.x];
    const WinDesc W = a.wins[bt.x];
    const int P = W.n * a.nmc;
    const int m = bt.y * kB + threadIdx.x;
    const int split = bt.z;
    const int j0 = W.a0 + split * W.aps;
    const int j1 = min(W.a0 + W.n, j0 + W.aps);
    const bool valid = m < P;
    const long long gp = (long long)W.a0 * a.nmc + (valid ? m : 0);
    const float4 q0 = a.rec4[4 * gp + 0];  // p, Ehat.x
    const float4 q1 = a.rec4[4 * gp + 1];  // Ehat.yz
    const float4 q3 = a.rec4[4 * gp + 3];  // A2, 3 A2, A3, w
    const float A1 = a.aux4[gp].w;
    const bool pbc = a.pbc_flag[bt.x] != 0;
    const DiffType *dtypes = stage_dtypes(a);
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int jt = j0; jt < j1; jt += kTile) {
        const int je = min(j1, jt + kTile);
        PointBar b{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (pbc)
            vjp_points_tile<true>(a, dtypes, jt, je, W.L, W.h, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, A1,
                                  q3.x, q3.z, b);
        else
            vjp_points_tile<false>(a, dtypes, jt, je, W.L, W.h, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, A1,
                                   q3.x, q3.z, b);
        acc[0] += b.hx; acc[1] += b.hy; acc[2] += b.hz;
        acc[3] += b.a1; acc[4] += b.a2; acc[5] += b.a3;
        acc[6] += b.cx; acc[7] += b.cy; acc[8] += b.cz;
        acc[9] += b.w;
    }
    if (!valid) return;
    const long long o = W.fpo + (long long)split * P + m;
    for (int k = 0; k < 10; ++k) a.ppart[k * a.F + o] = acc[k];
}

// Langevin moments and their derivatives in FP64 (series ratio below 5, coth
// recurrence above), as the forward epilogue
__device__ void moments_derivs(double x, double m[3], double dm[3])
{
    if (x < 5.0) {
        const double h = 0.5 * x * x;
        double s[4], ds[4];
        for (int l = 0; l < 4; ++l) {
            // terms T_0 = 1/(2l+1)!!, T_(k+1) = T_k h / ((k+1)(2l+2k+3)), so
            // d T_(k+1) / dh = T_k / (2l+2k+3); stop once a term is below
            // 1e-17 of the sum
            double t = kLangevin.first[l], sum = t, dsum = 0.0;
            for (int k = 0; k < kLgTerms && t > 1e-17 * sum; ++k) {
                dsum += t * kLangevin.inv[l][k];
                t *= h * kLangevin.step[l][k];
                sum += t;
            }
            s[l] = sum;
            ds[l] = dsum;
        }
        const double s0 = s[0];
        double xl = 1.0;
        for (int l = 1; l <= 3; ++l) {
            const double xl1 = xl;  // x^(l-1)
            xl *= x;                // x^l
            m[l - 1] = xl * s[l] / s0;
            dm[l - 1] = (double)l * xl1 * s[l] / s0 + xl * x * (ds[l] * s0 - s[l] * ds[0]) / (s0 * s0);
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

__global__ __launch_bounds__(kB) void k_vjp_epilogue(VjpArgs a)
{
    const long long gp = (long long)blockIdx.x * kB + threadIdx.x;
    if (gp >= a.Ptot) return;
    const WinDesc W = a.wins[a.atom_win[gp / a.nmc]];
    const int P = W.n * a.nmc;
    const long long m = gp - (long long)W.a0 * a.nmc;
    double b[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < W.splits; ++s) {
        const long long o = W.fpo + (long long)s * P + m;
        for (int k = 0; k < 10; ++k) b[k] += a.ppart[k * a.F + o];
    }
    const double E[3] = {a.E_out[3 * gp], a.E_out[3 * gp + 1], a.E_out[3 * gp + 2]};
    const double g = a.g_out[gp];
    const double cnt = (double)a.cnt_out[gp];
    const double x = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    double eh[3];
    for (int c = 0; c < 3; ++c) eh[c] = x > 0.0 ? E[c] / x : 0.0;
    double mm[3], dm[3];
    moments_derivs(x, mm, dm);
    const double w = g / cnt, cp = (double)a.p0 / cnt;
    const double p0 = a.p0, q0 = a.q0, o0 = a.o0;
    const double wbar = b[9] + p0 * mm[0] * b[3] + q0 * mm[1] * b[4] + o0 * mm[2] * b[5];
    const double gbar = wbar / cnt;
    const double xbar = w * (p0 * dm[0] * b[3] + q0 * dm[1] * b[4] + o0 * dm[2] * b[5]);
    const double hdot = b[0] * eh[0] + b[1] * eh[1] + b[2] * eh[2];
    double Eb[3], E0b[3];
    for (int c = 0; c < 3; ++c) {
        Eb[c] = xbar * eh[c] + (x > 0.0 ? (b[c] - hdot * eh[c]) / x : 0.0);
        E0b[c] = cp * b[6 + c];
    }
    const float4 p = a.pts[gp];
    float4 *r = a.prec + 4 * gp;
    r[0] = p;
    r[1] = make_float4((float)Eb[0], (float)Eb[1], (float)Eb[2], (float)(g * gbar));
    r[2] = make_float4((float)E0b[0], (float)E0b[1], (float)E0b[2], (float)w);
    r[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// adjoint, atom side: lanes own atoms, point records stream (wave-uniform)
// ---------------------------------------------------------------------------
template <bool PBC>
__device__ __forceinline__ void vjp_atoms_tile(const VjpArgs &a, const float4 *__restrict__ pr,
                                               const DiffType &Tj, int m0, int m1, float L,
                                               float h, float4 xj, float R2j, float lpx,
                                               float lpy, float lpz, float &qb, double tb[5])
{
    const float e0 = a.e0;
    for (int m = m0; m < m1; ++m) {
        const float4 p = pr[4 * m + 0];  // wave-uniform
        const float4 eb = pr[4 * m + 1];
        const float4 e0b = pr[4 * m + 2];
        float rx = p.x - xj.x, ry = p.y - xj.y, rz = p.z - xj.z;
        if (PBC) {
            rx = min_image1(rx, L, h);
            ry = min_image1(ry, L, h);
            rz = min_image1(rz, L, h);
        }
        const float d2 = dist2_f32(rx, ry, rz);
        const float rinv = rsqrt_f32(d2);
        const float er = __builtin_fmaf(eb.z, rz, __builtin_fmaf(eb.y, ry, eb.x * rx));
        const float e0r = __builtin_fmaf(e0b.z, rz, __builtin_fmaf(e0b.y, ry, e0b.x * rx));
        const bool in_own = d2 <= p.w;
        const bool in_rec = d2 <= R2j;
        const float s = -e0 * rinv * rinv * rinv;
        qb = __builtin_fmaf(s, in_own ? er : er + e0r, qb);
        if (in_own || in_rec) {
            const float d = d2 * rinv;
            float v[3], dv[15];
            diff_all(Tj, d, rinv, v, dv);
            if (in_own) {
                const float ce = er * rinv;
                const float cg = (eb.w != 0.0f && v[0] != 0.0f) ? eb.w / v[0] : 0.0f;
                for (int k = 0; k < 5; ++k) tb[k] += (double)(ce * dv[5 + k] + cg * dv[k]);
            }
            if (in_rec) {
                const float lp = __builtin_fmaf(lpz, rz, __builtin_fmaf(lpy, ry, lpx * rx));
                const float cf = e0b.w * lp * rinv;
                for (int k = 0; k < 5; ++k) tb[k] += (double)(cf * dv[10 + k]);
            }
        }
    }
}

__global__ __launch_bounds__(kB) void k_vjp_atoms(VjpArgs a)
{
    const int4 bt = a.pblk[blockIdx.x];
    const WinDesc W = a.wins[bt.x];
    const int il = bt.y * kB + threadIdx.x;
    const int chunk = bt.z;
    const long long pbase = (long long)W.a0 * a.nmc;
    const float4 *__restrict__ pr = a.prec + 4 * pbase;  // this window's point bars
    const int m0 = chunk * W.ppc;                        // window-local point range
    const int m1 = min(W.n * a.nmc, m0 + W.ppc);
    const bool valid = il < W.n;
    const int j = W.a0 + (valid ? il : 0);
    const float4 xj = a.xq[j];
    const int tj = a.itype[j];
    const float R2j = valid ? a.rmax2[tj] : -1.0f;
    const float4 lp = a.arec[4 * (size_t)j + 2];  // Lambda' = -vtot lambda
    const bool pbc = a.pbc_flag[bt.x] != 0;
    const DiffType &Tj = stage_dtypes(a)[tj];
    // a wave with no atom of the window stops after the block's barrier
    if (bt.y * kB + (int)(threadIdx.x & ~(kWave - 1)) >= W.n) return;
    double qacc = 0.0, tb[5] = {0, 0, 0, 0, 0};
    for (int mt = m0; mt < m1; mt += kTile) {
        const int me = min(m1, mt + kTile);
        float qb = 0.0f;
        if (pbc)
            vjp_atoms_tile<true>(a, pr, Tj, mt, me, W.L, W.h, xj, R2j, lp.x, lp.y, lp.z, qb, tb);
        else
            vjp_atoms_tile<false>(a, pr, Tj, mt, me, W.L, W.h, xj, R2j, lp.x, lp.y, lp.z, qb, tb);
        qacc += (double)qb;
    }
    if (!valid) return;
    const long long o = W.ppo + (long long)chunk * W.n + il;
    a.apart[o] = qacc;
    for (int k = 0; k < 5; ++k) a.apart[(1 + k) * a.C + o] = tb[k];
}

// ---------------------------------------------------------------------------
// adjoint of the all-pairs term: lanes own receivers i, partners stream
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_vjp_nonbond(VjpNbArgs a)
{
    extern __shared__ float4 lds_ld[];
    for (int t = threadIdx.x; t < a.T * a.T; t += kB) lds_ld[t] = a.ld[t];
    __syncthreads();
    const int4 bt = a.blocks[blockIdx.x];
    const WinDesc W = a.wins[bt.x];
    const int il = bt.y * kB + threadIdx.x;
    const int j0 = W.a0 + bt.z * W.apc;
    const int j1 = min(W.a0 + W.n, j0 + W.apc);
    const bool valid = il < W.n;
    const int i = W.a0 + (valid ? il : W.n - 1);
    const float4 xi = a.xq[i];
    const float4 auxi = a.aux[i];
    const int ti = __float_as_int(auxi.w);
    const float4 li4 = a.arec[4 * (size_t)i + 3];
    uint32_t ex[4];
    for (int k = 0; k < 4; ++k) ex[k] = a.excl[4 * i + k];
    const bool pbc = a.pbc_flag[bt.x] != 0;
    const float L = W.L, h = W.h;
    const int wave_lo = W.a0 + bt.y * kB + (threadIdx.x & ~(kWave - 1));
    const float4 *ldrow = lds_ld + ti * a.T;
    double qacc = 0.0, sacc = 0.0, eacc = 0.0;
    for (int jt = j0; jt < j1; jt += kTile) {
        const int je = min(j1, jt + kTile);
        const bool near = (jt < wave_lo + kWave + 64) && (je > wave_lo - 64);
        float qb = 0.0f, sb = 0.0f, eb = 0.0f;
        for (int j = jt; j < je; ++j) {
            const float4 xj = a.xq[j];  // wave-uniform
            const float4 auxj = a.aux[j];
            const float4 lj4 = a.arec[4 * (size_t)j + 3];
            float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            if (pbc) {
                rx = min_image1(rx, L, h);
                ry = min_image1(ry, L, h);
                rz = min_image1(rz, L, h);
            }
            const float d2 = dist2_f32(rx, ry, rz);
            const float inv = rsqrt_f32(d2);
            const float inv2 = inv * inv;
            const float d = d2 * inv;
            const float lri = __builtin_fmaf(li4.z, rz, __builtin_fmaf(li4.y, ry, li4.x * rx));
            const float lrj = __builtin_fmaf(lj4.z, rz, __builtin_fmaf(lj4.y, ry, lj4.x * rx));
            const float c = lri - lrj;
            bool excl = false;
            if (near) {
                const int w = j - i + 64;
                if (w >= 0 && w < 128) excl = (ex[w >> 5] >> (w & 31)) & 1u;
            }
            if (a.excl_ptr) {
                const int e0 = a.excl_ptr[i], e1 = a.excl_ptr[i + 1];
                for (int e = e0; e < e1; ++e) excl |= (a.excl_idx[e] == j);
            }
            const bool self = (j == i);
            // pair term phi(d, R): far -2 fep/(3 d^3), near cA + cB d
            const float phj = d2 > auxj.x ? a.cfar * inv2 * inv : __builtin_fmaf(auxj.z, d, auxj.y);
            const float phi = d2 > auxi.x ? a.cfar * inv2 * inv : __builtin_fmaf(auxi.z, d, auxi.y);
            float dq = xj.w * (lri * phj - lrj * phi);
            float ds = 0.0f, de = 0.0f;
            if (!excl) {
                const float4 pt = ldrow[__float_as_int(auxj.w)];
                const float r6 = inv2 * inv2 * inv2;
                ds = 0.5f * c * (12.0f * pt.z * r6 * r6 - 6.0f * pt.w * r6) * inv2;
                de = c * (12.0f * pt.x * r6 * r6 - 6.0f * pt.y * r6) * inv2;
                dq = __builtin_fmaf(c * xj.w, inv2 * inv, dq);
            }
            qb += self ? 0.0f : dq;
            sb += (self || excl) ? 0.0f : ds;
            eb += (self || excl) ? 0.0f : de;
        }
        qacc += (double)qb;
        sacc += (double)sb;
        eacc += (double)eb;
    }
    if (!valid) return;
    const long long o = W.npo + (long long)bt.z * W.n + il;
    a.part[o] = qacc;
    a.part[a.B + o] = sacc;
    a.part[2 * a.B + o] = eacc;
}

// per atom: chunk partials in order + the direct charge term
__global__ __launch_bounds__(kB) void k_vjp_finish(VjpFinishArgs a)
{
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.n) return;
    const WinDesc W = a.wins[a.atom_win[i]];
    const WinDesc Wn = a.nb_wins[a.atom_win[i]];
    const int il = i - W.a0;
    double q = 0.0, t[5] = {0, 0, 0, 0, 0};
    for (int c = 0; c < W.chunks; ++c) {
        const long long o = W.ppo + (long long)c * W.n + il;
        q += a.apart[o];
        for (int k = 0; k < 5; ++k) t[k] += a.apart[(1 + k) * a.C + o];
    }
    double nq = 0.0, ns = 0.0, ne = 0.0;
    for (int c = 0; c < Wn.nbc; ++c) {
        const long long o = Wn.npo + (long long)c * Wn.n + il;
        nq += a.nbpart[o];
        ns += a.nbpart[a.B + o];
        ne += a.nbpart[2 * a.B + o];
    }
    const int ti = a.itype[i];
    const double vt = a.vtot[ti];
    const double direct = -vt * (a.lam[3 * i] * a.se[3 * i] + a.lam[3 * i + 1] * a.se[3 * i + 1] +
                                 a.lam[3 * i + 2] * a.se[3 * i + 2]);
    a.gq[i] = q + nq + direct;
    a.gatom[5 * i + 0] = t[0] + ns;
    a.gatom[5 * i + 1] = t[1] + ne / (2.0 * a.theta[5 * ti + 1]);
    for (int k = 2; k < 5; ++k) a.gatom[5 * i + k] = t[k];
}

// per-type sums of the per-atom bars, FP64, fixed order (one block per type)
__global__ __launch_bounds__(kB) void k_type_reduce(const double *__restrict__ gatom,
                                                    const int *__restrict__ itype, int n,
                                                    double *gtheta)
{
    __shared__ double red[5][kB];
    const int t = blockIdx.x;
    double s[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += kB)
        if (itype[i] == t)
            for (int k = 0; k < 5; ++k) s[k] += gatom[5 * i + k];
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) gtheta[5 * t + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_diff_prep_types(const DiffPrepArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_diff_prep_types, dim3(1), dim3(kB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_diff_prep_atoms(const DiffAtomArgs &a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_diff_prep_atoms, dim3(nblk(a.n)), dim3(kB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_diff_reduce(const DiffReduceArgs &a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_diff_reduce, dim3(nblk(a.n)), dim3(kB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_points(const VjpArgs &a, int nblocks, hipStream_t s)
{
    if (nblocks == 0) return hipSuccess;
    const int nf = dtype_floats(a.ntypes);
    const size_t lds = nf <= kDtLdsFloats ? (size_t)nf * sizeof(float) : 0;
    hipLaunchKernelGGL(k_vjp_points, dim3(nblocks), dim3(kB), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_epilogue(const VjpArgs &a, hipStream_t s)
{
    if (a.Ptot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vjp_epilogue, dim3(nblk(a.Ptot)), dim3(kB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_atoms(const VjpArgs &a, int nblocks, hipStream_t s)
{
    if (nblocks == 0) return hipSuccess;
    const int nf = dtype_floats(a.ntypes);
    const size_t lds = nf <= kDtLdsFloats ? (size_t)nf * sizeof(float) : 0;
    hipLaunchKernelGGL(k_vjp_atoms, dim3(nblocks), dim3(kB), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_nonbond(const VjpNbArgs &a, int nblocks, hipStream_t s)
{
    if (nblocks == 0) return hipSuccess;
    const size_t lds = sizeof(float4) * a.T * a.T;
    hipLaunchKernelGGL(k_vjp_nonbond, dim3(nblocks), dim3(kB), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_vjp_finish(const VjpFinishArgs &a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vjp_finish, dim3(nblk(a.n)), dim3(kB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_type_reduce(const double *gatom, const int *itype, int n, int T,
                              double *gtheta, hipStream_t s)
{
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_type_reduce, dim3(T), dim3(kB), 0, s, gatom, itype, n, gtheta);
    return hipGetLastError();
}

}  // namespace isspa
