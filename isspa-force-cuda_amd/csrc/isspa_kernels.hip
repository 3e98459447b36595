// IS-SPA solvent kernels for gfx950 (CDNA4): MC points, field pass, per-point
// epilogue, force pass, chunk reduction.
//
// Semantics follow src/isspa_force_cuda.cu:73-327 of the reference; the
// decomposition does not.  The reference launches one thread per (point,
// atom) pair and folds results with wave32 shuffles and float atomics.  Here:
//
//  * field pass: one lane owns one MC point; the atom stream is wave-uniform
//    (scalar loads), so each pair costs ~14 VALU ops + one v_rsq.  In-sphere
//    pairs (<2% of wave iterations) take a skipped-when-empty branch.
//  * epilogue: per point, in FP64: |E|, E-hat, Langevin moments by series or
//    recurrence (no tanh cancellation, quirk Q22), and the hoisted constants
//    of the force pass packed in a 64-byte record.
//  * force pass: one lane owns one receiving atom; the point-record stream is
//    wave-uniform.  Per-atom factors (-q_i vtot_i, -vtot_i) are applied once
//    in the reduction.  The CDD term is taken for every pair in the main path
//    and undone in the rare in-sphere branch, so the main path has no select.
//  * no float atomics: lanes keep FP32 partials over short tiles, flush them
//    into FP64, and write per-chunk partials that a reduction kernel sums in a
//    fixed order (quirk Q10).
//  * every launch covers any number of independent windows (kernels.hpp).
//  * the field and force kernels are templated on AN: table lookups (the
//    plain path) or the analytic per-type functions of the differentiable
//    model (diff_device.hpp); only the rare in-sphere branches differ.
#include <algorithm>

#include "device_common.hpp"
#include "diff_device.hpp"
#include "kernels.hpp"
#include "pbc_device.hpp"
#include "points_device.hpp"
#include "nonbond_device.hpp"
#include "reduce_device.hpp"

namespace isspa {

constexpr int kBlock = 256;
constexpr int kFieldTile = 128;
constexpr int kForceTile = 128;
// per-type tables up to this many floats (16 KB) are staged in LDS for the
// in-sphere branches; 8 resident blocks per CU then hold 128 of its 160 KB
constexpr int kTabLdsFloats = 4096;

__host__ __device__ inline int field_tab_floats(const TableView &t)
{
    return t.ntypes * (t.ld_g + t.ld_e);
}
__host__ __device__ inline int force_tab_floats(const TableView &t) { return t.ntypes * t.ld_f; }
// per-type records of the analytic model (one DiffType per type)
__host__ __device__ inline int diff_tab_floats(const TableView &t)
{
    return t.ntypes * (int)(sizeof(DiffType) / sizeof(float));
}

// copy n floats into LDS with the whole block (n is block-uniform); eight
// loads per thread are issued before their stores, so a small system's
// kernel waits for the table once rather than once per 256 floats
__device__ __forceinline__ void stage_lds(float *dst, const float *__restrict__ src, int n)
{
    constexpr int kU = 8;
    for (int t0 = threadIdx.x; t0 < n; t0 += kU * kBlock) {
        float v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int t = t0 + u * kBlock;
            v[u] = t < n ? src[t] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int t = t0 + u * kBlock;
            if (t < n) dst[t] = v[u];
        }
    }
}

// ---------------------------------------------------------------------------
// A.1 MC points: one lane per point (mc_point).  The first nflag blocks (one
// per window, none when nflag = 0) compute instead the per-window
// minimum-image flags that the field and force passes read, which saves a
// launch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mc_points(PointsArgs a, int nflag, float rmax_all,
                                                     int *flag)
{
    if ((int)blockIdx.x < nflag) {
        const WinDesc W = a.wins[blockIdx.x];
        pbc_flag_block(W, a.xyzq, rmax_all, flag + blockIdx.x);
        return;
    }
    const long long gp = (long long)(blockIdx.x - nflag) * kBlock + threadIdx.x;
    if (gp >= (long long)a.ntot * a.nmc) return;
    mc_point(a, gp);
}

// ---------------------------------------------------------------------------
// A.2 field pass (src/isspa_force_cuda.cu:110-206)
// ---------------------------------------------------------------------------
// in-sphere state of one point: g product and count over the split, and the
// E_in sum of the current tile (flushed into FP64 per tile)
struct FieldRare {
    float g;
    int cnt;
    float ex, ey, ez;
};

__device__ __forceinline__ void field_in_sphere(const TableView &tb, const float *Gtab,
                                                const float *Etab, int tj, float d2, float inv,
                                                float s, float rx, float ry, float rz, float e0,
                                                FieldRare &st)
{
    const float d = d2 * inv;
    const float x = (d - tb.g_r0) * tb.g_inv;
    const int b = (int)x;  // truncation, as the reference's int()
    float fac = 1.0f, et = 0.0f;
    if (b < 0) {
        fac = 0.0f;
    } else if (b < tb.n_g - 1) {
        const float ph = (d - (tb.g_r0 + (float)b * tb.g_dr)) * tb.g_inv;
        const float *G = Gtab + tj * tb.ld_g;
        fac = __builtin_fmaf(G[b + 1], ph, G[b] * (1.0f - ph));
        const float phe = (d - (tb.e_r0 + (float)b * tb.e_dr)) * tb.e_inv;
        const float *Et = Etab + tj * tb.ld_e;
        et = __builtin_fmaf(Et[b + 1], phe, Et[b] * (1.0f - phe));
    }
    st.g *= fac;
    st.cnt += 1;
    const float c = __builtin_fmaf(-e0, s, et * inv);
    st.ex = __builtin_fmaf(c, rx, st.ex);
    st.ey = __builtin_fmaf(c, ry, st.ey);
    st.ez = __builtin_fmaf(c, rz, st.ez);
}

// the same with the analytic g and E of the differentiable model
__device__ __forceinline__ void field_in_sphere_an(const DiffType &T, float d2, float inv, float s,
                                                   float rx, float ry, float rz, float e0,
                                                   FieldRare &st)
{
    const float d = d2 * inv;
    st.g *= diff_g(T, d);
    st.cnt += 1;
    const float c = __builtin_fmaf(-e0, s, diff_E(T, d) * inv);
    st.ex = __builtin_fmaf(c, rx, st.ex);
    st.ey = __builtin_fmaf(c, ry, st.ey);
    st.ez = __builtin_fmaf(c, rz, st.ez);
}

template <bool PBC, bool AN>
__device__ __forceinline__ void field_tile(const FieldArgs &a, const float *Gtab,
                                           const float *Etab, const DiffType *dtypes, int j0,
                                           int j1, float L, float h,
                                           float px, float py, float pz, float R2, float &ax,
                                           float &ay, float &az, FieldRare &st)
{
    // the tile's atoms by a pointer that advances 16 bytes an atom (an index
    // recomputed from j costs the scalar unit a sign extension and a shift
    // before every load)
    const auto *xq = read_only(reinterpret_cast<const float *>(a.xyzq + j0));
    const auto *ity = read_only(a.itype);
#pragma unroll 2
    for (int j = j0; j < j1; ++j, xq += 4) {
        // wave-uniform
        const float4 q = make_float4(xq[0], xq[1], xq[2], xq[3]);
        float rx = px - q.x, ry = py - q.y, rz = pz - q.z;
        if (PBC) {
            rx = min_image1(rx, L, h);
            ry = min_image1(ry, L, h);
            rz = min_image1(rz, L, h);
        }
        const float d2 = dist2_f32(rx, ry, rz);
        const float inv = rsqrt_f32(d2);
        const float s = q.w * (inv * inv * inv);
        const bool in = d2 <= R2;
        const float so = in ? 0.0f : s;
        ax = __builtin_fmaf(so, rx, ax);
        ay = __builtin_fmaf(so, ry, ay);
        az = __builtin_fmaf(so, rz, az);
        if (in) {
            if (AN)
                field_in_sphere_an(dtypes[ity[j]], d2, inv, s, rx, ry, rz, a.e0, st);
            else
                field_in_sphere(a.tab, Gtab, Etab, ity[j], d2, inv, s, rx, ry, rz, a.e0,
                                st);
        }
    }
}

// a point's sums over atoms [j0, j1) of its split
struct FieldSums {
    double sx = 0.0, sy = 0.0, sz = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
    FieldRare st{1.0f, 0, 0.0f, 0.0f, 0.0f};
};

template <bool AN>
__device__ __forceinline__ void field_span(const FieldArgs &a, const float *Gtab,
                                           const float *Etab, const DiffType *dtypes, int j0,
                                           int j1, const WinDesc &W, bool pbc, float4 p,
                                           FieldSums &fs)
{
    FieldRare &st = fs.st;
    for (int jt = j0; jt < j1; jt += kFieldTile) {
        const int je = min(j1, jt + kFieldTile);
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        st.ex = st.ey = st.ez = 0.0f;
        if (pbc)
            field_tile<true, AN>(a, Gtab, Etab, dtypes, jt, je, W.L, W.h, p.x, p.y, p.z, p.w, ax,
                                 ay, az, st);
        else
            field_tile<false, AN>(a, Gtab, Etab, dtypes, jt, je, W.L, W.h, p.x, p.y, p.z, p.w, ax,
                                  ay, az, st);
        fs.sx += (double)ax;
        fs.sy += (double)ay;
        fs.sz += (double)az;
        fs.ex += (double)st.ex;
        fs.ey += (double)st.ey;
        fs.ez += (double)st.ez;
    }
}

// A block takes fg splits (W.fg: 1, 2 or 4) of the same 256 / fg points, one
// split for each 256 / fg work-items, and adds its splits' sums in split order
// through LDS: it leaves one partial per group of fg splits, so the epilogue
// reads fg times fewer.
struct FieldXchg {
    double s[6];
    float g;
    int cnt;
};
constexpr int kFieldXchgBytes = 3 * 64 * (int)sizeof(FieldXchg);  // fg = 4: 3 splits of 64 points

// NB: the launch also carries the all-pairs blocks after the field blocks (a
// separate instantiation, so the plain field loop keeps its own registers)
template <bool AN, bool NB>
__global__ __launch_bounds__(kBlock) void k_field(FieldArgs a)
{
    extern __shared__ float lds_tab[];
    if (NB && (int)blockIdx.x >= a.nfield) {
        // an all-pairs block riding on this launch (its LJ table in the LDS)
        nonbond_block<false, false>(a.nb, blockIdx.x - a.nfield, reinterpret_cast<float2 *>(lds_tab));
        return;
    }
    const int4 bt = a.blocks[blockIdx.x];
    const WinDesc W = a.wins[bt.x];
    const int P = W.n * a.nmc;
    const int fg = W.fg, np = kBlock / fg;
    // the block's split for this work-item (a whole number of waves each),
    // read from the first lane so the compiler knows it is wave-uniform and
    // the atom range stays on scalar loads
    const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x / np);
    const int lt = (int)threadIdx.x - sub * np;
    const int m = bt.y * np + lt;  // point within the window
    const int split = bt.z * fg + sub;
    const int j0 = W.a0 + split * W.aps;
    const int j1 = min(W.a0 + W.n, j0 + W.aps);
    const bool valid = m < P;
    const long long gp = (long long)W.a0 * a.nmc + m;
    const float4 p = valid ? a.pts[gp] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    const bool pbc = a.pbc_flag[bt.x] != 0;
    // Warm-up: the 128-byte lines of the split's atoms (8 atoms a line, up to
    // two lines a work-item) are requested at block start, so that the loop's
    // scalar atom loads find them in L2 instead of waiting for each line in
    // turn (the positions were last written by another XCD).  The values are
    // only kept alive by the empty asm after the loop.
    float w0 = 0.0f, w1 = 0.0f;
    if (j1 > j0) {
        const int l0 = (j0 >> 3) + lt, last = (j1 - 1) >> 3;
        if (l0 <= last) w0 = a.xyzq[8 * l0].x;
        if (l0 + np <= last) w1 = a.xyzq[8 * (l0 + np)].x;
    }
    // g and E tables (or the analytic records) for the in-sphere branch: LDS
    // when they fit.  Separate call sites give each table pointer one address
    // space, so the LDS ones compile to LDS reads.
    FieldSums fs;
    if (!AN && field_tab_floats(a.tab) <= kTabLdsFloats) {
        const int ng = a.tab.ntypes * a.tab.ld_g;
        stage_lds(lds_tab, a.tab.g, ng);
        stage_lds(lds_tab + ng, a.tab.e, a.tab.ntypes * a.tab.ld_e);
        __syncthreads();
        field_span<AN>(a, lds_tab, lds_tab + ng, a.dtypes, j0, j1, W, pbc, p, fs);
    } else if (AN && diff_tab_floats(a.tab) <= kTabLdsFloats) {
        stage_lds(lds_tab, reinterpret_cast<const float *>(a.dtypes), diff_tab_floats(a.tab));
        __syncthreads();
        field_span<AN>(a, a.tab.g, a.tab.e, reinterpret_cast<const DiffType *>(lds_tab), j0, j1,
                       W, pbc, p, fs);
    } else {
        field_span<AN>(a, a.tab.g, a.tab.e, a.dtypes, j0, j1, W, pbc, p, fs);
    }
    asm volatile("" ::"v"(w0), "v"(w1));
    FieldRare &st = fs.st;
    double sx = fs.sx, sy = fs.sy, sz = fs.sz, ex = fs.ex, ey = fs.ey, ez = fs.ez;
    // the g factors of a split multiply in FP32; splits multiply in FP64 (a
    // product over several splits can leave the FP32 range)
    double g = (double)st.g;
    if (fg > 1) {
        // the group's splits in split order, through LDS (the tables are done)
        FieldXchg *xc = reinterpret_cast<FieldXchg *>(lds_tab);
        __syncthreads();
        if (sub > 0) {
            FieldXchg &e = xc[(sub - 1) * np + lt];
            e.s[0] = sx; e.s[1] = sy; e.s[2] = sz;
            e.s[3] = ex; e.s[4] = ey; e.s[5] = ez;
            e.g = st.g;
            e.cnt = st.cnt;
        }
        __syncthreads();
        if (sub > 0) return;
        for (int k = 1; k < fg; ++k) {
            const FieldXchg &e = xc[(k - 1) * np + lt];
            sx += e.s[0]; sy += e.s[1]; sz += e.s[2];
            ex += e.s[3]; ey += e.s[4]; ez += e.s[5];
            g *= (double)e.g;
            st.cnt += e.cnt;
        }
    }
    if (!valid) return;
    const long long o = W.fpo + (long long)bt.z * P + m;
    a.part_g[o] = g;
    a.part_cnt[o] = st.cnt;
    a.part_s0[o] = sx;
    a.part_s0[a.F + o] = sy;
    a.part_s0[2 * a.F + o] = sz;
    a.part_ein[o] = ex;
    a.part_ein[a.F + o] = ey;
    a.part_ein[2 * a.F + o] = ez;
}

// ---------------------------------------------------------------------------
// per-point epilogue in FP64: Langevin moments <P1>,<P2>,<P3>
// (= c1, 1.5 c2 - 0.5, 2.5 c3 - 1.5 c1 of src/isspa_force_cuda.cu:270-287)
// ---------------------------------------------------------------------------
// Below x = 5: <P_l> = i_l(x)/i_0(x) from the all-positive power series
// i_l(x) = x^l sum_k (x^2/2)^k / (k! (2l+2k+1)!!), 30 terms; above 5 the
// upward recurrence from coth, which then cancels by at most ~3x.
__device__ __forceinline__ void langevin_moments(double x, double inv_x, double &m1, double &m2,
                                                 double &m3)
{
    if (x < 5.0) {
        // all 30 terms of the four series, advanced together as four
        // independent FP64 chains with compile-time coefficients: past the
        // first few terms they add nothing (below 1e-17 of the sum for small
        // x), but a fixed count keeps the loop free of branches and loads
        const double h = 0.5 * x * x;
        double t[4], s[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            t[l] = kLangevin.first[l];
            s[l] = t[l];
        }
#pragma unroll
        for (int k = 0; k < kLgTerms; ++k)
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                t[l] *= h * kLangevin.step[l][k];
                s[l] += t[l];
            }
        const double r = 1.0 / s[0];
        m1 = x * s[1] * r;
        m2 = x * x * s[2] * r;
        m3 = x * x * x * s[3] * r;
        return;
    }
    const double coth = 1.0 / tanh(x);
    m1 = coth - inv_x;
    m2 = 1.0 - 3.0 * m1 * inv_x;
    m3 = m1 - 5.0 * m2 * inv_x;
}

// One point's epilogue: its field partials (one per group of W.fg splits)
// summed in group order (B groups' loads in flight at a time), then the FP64
// Langevin moments, the 16-float force record {p, Ehat, V, D, A2, 3 A2, A3,
// w} and the aux entry {cp E0, A1}.  The epilogue kernel writes them to
// memory; the fused force kernel of small systems keeps them in LDS.  Both
// run this same code, so the records are the same bit for bit.
struct PointRecord {
    float4 r[4];
    float4 aux;
    double g;
    int cnt;
    double E[3], E0[3];
};

template <int B>
__device__ __forceinline__ void epilogue_point(const EpilogueArgs &a, const WinDesc &W,
                                               long long gp, PointRecord &out)
{
    const int P = W.n * a.nmc;
    const long long m = gp - (long long)W.a0 * a.nmc;
    const float4 p = a.pts[gp];  // requested with the partials, used at the end
    double g = 1.0, s0[3] = {0, 0, 0}, ei[3] = {0, 0, 0};
    int cnt = 0;
    const int ngr = W.splits / W.fg;  // one partial per group of splits
    for (int sp = 0; sp < ngr; sp += B) {
        double pg[B];
        int pc[B];
        double ps[B][3], pe[B][3];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            pg[u] = 1.0;
            pc[u] = 0;
            for (int c = 0; c < 3; ++c) ps[u][c] = pe[u][c] = 0.0;
            if (sp + u < ngr) {
                const long long o = W.fpo + (long long)(sp + u) * P + m;
                pg[u] = a.part_g[o];
                pc[u] = a.part_cnt[o];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ps[u][c] = a.part_s0[c * a.F + o];
                    pe[u][c] = a.part_ein[c * a.F + o];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (sp + u >= ngr) break;
            g *= pg[u];
            cnt += pc[u];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s0[c] += ps[u][c];
                ei[c] += pe[u][c];
            }
        }
    }
    double E0[3], E[3];
    for (int c = 0; c < 3; ++c) {
        E0[c] = -(double)a.e0 * s0[c];
        E[c] = E0[c] + ei[c];
    }
    const double x = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    const double inv_x = x > 0.0 ? 1.0 / x : 0.0;
    double eh[3];
    for (int c = 0; c < 3; ++c) eh[c] = E[c] * inv_x;
    double m1, m2, m3;
    langevin_moments(x, inv_x, m1, m2, m3);
    // cnt >= 1 always: the owner atom lies inside its own sphere
    const double inv_cnt = 1.0 / (double)(cnt > 0 ? cnt : 1);
    const double w = g * inv_cnt;
    const double cp = (double)a.p0 * inv_cnt;
    const double A1 = (double)a.p0 * m1 * w;
    const double A2 = (double)a.q0 * m2 * w;
    const double A3 = (double)a.o0 * m3 * w;
    double c3[3];  // cp E0 / 3
    for (int c = 0; c < 3; ++c) c3[c] = cp * E0[c] * (1.0 / 3.0);
    out.r[0] = make_float4(p.x, p.y, p.z, (float)eh[0]);
    out.r[1] = make_float4((float)eh[1], (float)eh[2], (float)(A1 * eh[0] + c3[0]),
                           (float)(A1 * eh[1] + c3[1]));
    out.r[2] = make_float4((float)(A1 * eh[2] + c3[2]),
                           (float)(3.0 * A1 * eh[0] + cp * E0[0]),
                           (float)(3.0 * A1 * eh[1] + cp * E0[1]),
                           (float)(3.0 * A1 * eh[2] + cp * E0[2]));
    out.r[3] = make_float4((float)A2, (float)(3.0 * A2), (float)A3, (float)w);
    out.aux = make_float4((float)(cp * E0[0]), (float)(cp * E0[1]), (float)(cp * E0[2]), (float)A1);
    out.g = g;
    out.cnt = cnt;
    for (int c = 0; c < 3; ++c) {
        out.E[c] = E[c];
        out.E0[c] = E0[c];
    }
}

// B groups' loads in flight at a time: 8 for launches with fewer blocks than
// CUs, where each batch costs a full memory latency, and 2 for larger ones,
// where the 8-wide batch's registers would cost occupancy in an HBM-bound
// pass
constexpr int kEpiWideGrid = 256;

template <int B>
__global__ __launch_bounds__(kBlock) void k_field_epilogue(EpilogueArgs a)
{
    const long long gp = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (gp >= a.Ptot) return;
    // one window: no per-point window lookup
    const WinDesc W = a.nwin == 1 ? a.wins[0] : a.wins[a.atom_win[gp / a.nmc]];
    PointRecord pr;
    epilogue_point<B>(a, W, gp, pr);
    float4 *r4 = reinterpret_cast<float4 *>(a.rec + gp * kRecFloats);
    for (int k = 0; k < 4; ++k) r4[k] = pr.r[k];
    a.aux[gp] = pr.aux;
    if (a.g_out) a.g_out[gp] = pr.g;
    if (a.cnt_out) a.cnt_out[gp] = pr.cnt;
    if (a.E_out)
        for (int c = 0; c < 3; ++c) a.E_out[3 * gp + c] = pr.E[c];
    if (a.E0_out)
        for (int c = 0; c < 3; ++c) a.E0_out[3 * gp + c] = pr.E0[c];
}

// ---------------------------------------------------------------------------
// A.3 force pass (src/isspa_force_cuda.cu:208-327)
//
// For receiver i and point m (r = p_m - x_i, d = |r|, u = Ehat.r/d,
// P2 = 7.5u^2-1.5, P3 = (17.5u^2-7.5)u), the reference's three multipole
// terms plus the CDD term, with -q_i vtot_i factored out, are
//     S += alpha r - d^-3 V - k Ehat
//     alpha = d^-4 (3 A1 u + d^-1 (A2 P2 + d^-1 A3 P3)) + d^-5 (E0c.r)
//     k     = d^-4 (3 u A2 + d^-1 P2 A3)
// with V = A1 Ehat + cp E0/3 and E0c = cp E0.  With e = Ehat.r (u = e/d),
// t = e/d^2 and P3' = 17.5u^2-7.5, both coefficients collapse onto d^-5:
//     alpha = d^-5 (D.r + A2 P2 + A3 P3' t),   D = 3 A1 Ehat + E0c
//     k     = d^-5 (3 A2 e + A3 P2)
// which is 37 VALU operations per pair (one of them v_rsq).  Inside the
// receiver's sphere the CDD part (d^-5 (E0c.r) r - d^-3 E0c/3) is removed
// again and the LJ table term is added instead (with -vtot_i factored out).
// ---------------------------------------------------------------------------
struct ForceAcc {
    double ex, ey, ez;  // electrostatic (pol + CDD) sum, or pol in split mode
    double lx, ly, lz;  // LJ sum
    double cx, cy, cz;  // CDD sum (split mode only)
};

// d^-5 (E0c.r) r - d^-3 E0c/3 for one pair, E0c = c.xyz
__device__ __forceinline__ float3 cdd_term(float4 c, float rx, float ry, float rz, float rinv3,
                                           float rinv5)
{
    const float ec = __builtin_fmaf(c.z, rz, __builtin_fmaf(c.y, ry, c.x * rx));
    const float ac = ec * rinv5;
    const float bc = rinv3 * (1.0f / 3.0f);
    return make_float3(__builtin_fmaf(ac, rx, -bc * c.x), __builtin_fmaf(ac, ry, -bc * c.y),
                       __builtin_fmaf(ac, rz, -bc * c.z));
}

// FP32 sums of one tile: the electrostatic (or, split, polarisation) sum s and
// the in-sphere LJ sum l; flushed into the FP64 ForceAcc per tile
struct TileSums {
    float sx, sy, sz;
    float lx, ly, lz;
};

template <bool PBC, bool SPLIT, bool AN>
__device__ __forceinline__ void force_tile(const ForceArgs &a, const float4 *__restrict__ rec,
                                           const float4 *__restrict__ aux, int m0, int m1,
                                           float L, float h, float xi, float yi, float zi,
                                           float R2, const float *Ftab,
                                           const DiffFParams &fp, TileSums &ts, ForceAcc &acc)
{
    float &sx = ts.sx, &sy = ts.sy, &sz = ts.sz;
    // the next point's record (wave-uniform) is fetched while this one is
    // evaluated, which hides the scalar load latency when few waves share a
    // SIMD; 32-bit point indices within the window keep the loop test on the
    // SALU
    // the next record's pointers advance one record an iteration and stop at
    // the last one (an index recomputed from m costs the scalar unit a clamp,
    // a sign extension and shifts before every load)
    const float4 *nr = rec + 4 * m0, *nx = aux + m0;
    float4 n0 = nr[0], n1 = nr[1], n2 = nr[2], n3 = nr[3], na = nx[0];
#pragma unroll 2
    for (int m = m0; m < m1; ++m) {
        const float4 q0 = n0, q1 = n1, q2 = n2, q3 = n3, c4 = na;
        if (m + 1 < m1) {
            nr += 4;
            nx += 1;
        }
        n0 = nr[0];
        n1 = nr[1];
        n2 = nr[2];
        n3 = nr[3];
        na = nx[0];
        const float Ex = q0.w, Ey = q1.x, Ez = q1.y;
        const float Vx = q1.z, Vy = q1.w, Vz = q2.x;
        const float Dx = q2.y, Dy = q2.z, Dz = q2.w;
        const float A2 = q3.x, A2x3 = q3.y, A3 = q3.z, w = q3.w;
        float rx = q0.x - xi, ry = q0.y - yi, rz = q0.z - zi;
        if (PBC) {
            rx = min_image1(rx, L, h);
            ry = min_image1(ry, L, h);
            rz = min_image1(rz, L, h);
        }
        const float d2 = dist2_f32(rx, ry, rz);
        const float rinv = rsqrt_f32(d2);
        const float e = __builtin_fmaf(Ez, rz, __builtin_fmaf(Ey, ry, Ex * rx));
        const float rinv2 = rinv * rinv;
        const float t = e * rinv2;
        const float u2 = e * t;
        const float P2 = __builtin_fmaf(7.5f, u2, -1.5f);
        const float P3p = __builtin_fmaf(17.5f, u2, -7.5f);
        const float rinv3 = rinv2 * rinv;
        const float rinv5 = rinv3 * rinv2;
        const float k = __builtin_fmaf(A2x3, e, A3 * P2) * rinv5;
        const bool in = d2 <= R2;
        if (!SPLIT) {
            const float dr = __builtin_fmaf(Dz, rz, __builtin_fmaf(Dy, ry, Dx * rx));
            const float alpha = __builtin_fmaf(A2, P2, __builtin_fmaf(A3, P3p * t, dr)) * rinv5;
            sx = __builtin_fmaf(alpha, rx, sx);
            sy = __builtin_fmaf(alpha, ry, sy);
            sz = __builtin_fmaf(alpha, rz, sz);
            sx = __builtin_fmaf(-rinv3, Vx, sx);
            sy = __builtin_fmaf(-rinv3, Vy, sy);
            sz = __builtin_fmaf(-rinv3, Vz, sz);
            sx = __builtin_fmaf(-k, Ex, sx);
            sy = __builtin_fmaf(-k, Ey, sy);
            sz = __builtin_fmaf(-k, Ez, sz);
        } else {
            // polarisation alone: alpha_p r - (d^-3 A1 + k) Ehat, alpha_p
            // being alpha without the E0c.r part
            const float A1 = c4.w;
            const float ap =
                __builtin_fmaf(A2, P2, __builtin_fmaf(A3, P3p * t, (3.0f * A1) * e)) * rinv5;
            const float bp = __builtin_fmaf(rinv3, A1, k);
            sx = __builtin_fmaf(ap, rx, __builtin_fmaf(-bp, Ex, sx));
            sy = __builtin_fmaf(ap, ry, __builtin_fmaf(-bp, Ey, sy));
            sz = __builtin_fmaf(ap, rz, __builtin_fmaf(-bp, Ez, sz));
            if (!in) {
                const float3 c = cdd_term(c4, rx, ry, rz, rinv3, rinv5);
                acc.cx += (double)c.x;
                acc.cy += (double)c.y;
                acc.cz += (double)c.z;
            }
        }
        if (in) {
            const float d = d2 * rinv;
            float fs = 0.0f;
            if (AN) {
                fs = diff_F(fp, d, rinv) * w;
            } else {
                const int b = (int)((d - a.tab.f_r0) * a.tab.f_inv + 0.5f);
                if (b >= 0 && b < a.tab.n_f) {
                    const float ph = (d - (a.tab.f_r0 + (float)b * a.tab.f_dr)) * a.tab.f_inv;
                    fs = __builtin_fmaf(Ftab[b + 1], ph, Ftab[b] * (1.0f - ph)) * w;
                }
            }
            const float fl = fs * rinv;
            ts.lx = __builtin_fmaf(fl, rx, ts.lx);
            ts.ly = __builtin_fmaf(fl, ry, ts.ly);
            ts.lz = __builtin_fmaf(fl, rz, ts.lz);
            if (!SPLIT) {
                // remove the CDD part taken in the main path
                const float3 c = cdd_term(c4, rx, ry, rz, rinv3, rinv5);
                sx -= c.x;
                sy -= c.y;
                sz -= c.z;
            }
        }
    }
}

// points [m0, m1) of one chunk for one receiver, in tiles of kForceTile
template <bool SPLIT, bool AN>
__device__ __forceinline__ void force_chunk(const ForceArgs &a, const float4 *__restrict__ rec,
                                            const float4 *__restrict__ aux, int m0, int m1,
                                            const WinDesc &W, bool pbc, float4 xi, float R2,
                                            const float *Ftab, const DiffFParams &fp,
                                            ForceAcc &acc)
{
    for (int mt = m0; mt < m1; mt += kForceTile) {
        const int me = min(m1, mt + kForceTile);
        TileSums ts{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        // tile-local point indices from the tile's first record
        if (pbc)
            force_tile<true, SPLIT, AN>(a, rec + 4 * mt, aux + mt, 0, me - mt, W.L, W.h, xi.x,
                                        xi.y, xi.z, R2, Ftab, fp, ts, acc);
        else
            force_tile<false, SPLIT, AN>(a, rec + 4 * mt, aux + mt, 0, me - mt, W.L, W.h, xi.x,
                                         xi.y, xi.z, R2, Ftab, fp, ts, acc);
        acc.ex += (double)ts.sx;
        acc.ey += (double)ts.sy;
        acc.ez += (double)ts.sz;
        acc.lx += (double)ts.lx;
        acc.ly += (double)ts.ly;
        acc.lz += (double)ts.lz;
    }
}

// Replicas 1..r-1 of the block's receivers leave their sums in LDS; replica
// 0 adds them in replica order.  Every wave of the block reaches the barrier
// (with r > 1 every wave holds receivers).  Returns false on the waves whose
// sums were handed over.
template <bool SPLIT>
__device__ __forceinline__ bool replica_sum(double *red, int r, int q, int slot, int nslots,
                                            ForceAcc &acc)
{
    constexpr int nc = SPLIT ? 9 : 6;
    if (q > 0) {
        double *d = red + ((long long)(q - 1) * nslots + slot) * nc;
        d[0] = acc.ex; d[1] = acc.ey; d[2] = acc.ez;
        d[3] = acc.lx; d[4] = acc.ly; d[5] = acc.lz;
        if (SPLIT) { d[6] = acc.cx; d[7] = acc.cy; d[8] = acc.cz; }
    }
    __syncthreads();
    if (q > 0) return false;
    for (int k = 1; k < r; ++k) {
        const double *d = red + ((long long)(k - 1) * nslots + slot) * nc;
        acc.ex += d[0]; acc.ey += d[1]; acc.ez += d[2];
        acc.lx += d[3]; acc.ly += d[4]; acc.lz += d[5];
        if (SPLIT) { acc.cx += d[6]; acc.cy += d[7]; acc.cz += d[8]; }
    }
    return true;
}

// REP: some window of the launch has receiver replicas (the replica code is
// compiled only into that instantiation, so the large-system kernel keeps its
// registers)
template <bool SPLIT, bool AN, bool REP>
__global__ __launch_bounds__(kBlock) void k_force(ForceArgs a)
{
    const int4 bt = a.blocks[a.xcd_runs ? xcd_run_entry(blockIdx.x, gridDim.x) : blockIdx.x];
    const WinDesc W = a.wins[bt.x];
    // receiver replicas: wave wv serves receivers (wv % nrw) * 64 + lane of
    // the block on the q-th share of the chunk's points
    const int r = REP ? force_replicas(W.n) : 1, nrw = 4 / r;
    // the wave index read from the first lane, so the compiler knows it is
    // wave-uniform and the replica's point range stays on scalar loads
    const int wv = REP ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave) : 0;
    const int q = REP ? wv / nrw : 0;
    const int slot = REP ? (wv % nrw) * kWave + (int)threadIdx.x % kWave : (int)threadIdx.x;
    const int il = bt.y * kBlock + slot;  // receiver within the window
    const int chunk = bt.z;
    const long long pbase = (long long)W.a0 * a.nmc;
    const float4 *__restrict__ rec = a.rec4 + 4 * pbase;  // this window's records
    const float4 *__restrict__ aux = a.aux4 + pbase;
    const int m0 = chunk * W.ppc;                        // window-local point range
    const int m1 = min(W.n * a.nmc, m0 + W.ppc);
    const int sub = (m1 - m0 + r - 1) / r;               // this replica's share
    const int ms = min(m1, m0 + q * sub), me = min(m1, ms + sub);
    const bool valid = il < W.n;
    const int i = W.a0 + (valid ? il : 0);
    const float4 xi = a.xyzq[i];
    const int ti = a.itype[i];
    const float R2 = valid ? a.rmax2[ti] : -1.0f;
    // Warm-up: the chunk's records and aux entries are requested at block
    // start, one 128-byte line per work-item, the lines dealt over the
    // chunk's receiver blocks, so that all of them are in flight at once.
    // Right after the epilogue has written the records they sit in another
    // XCD's L2 or in memory, and the loop's one-record-ahead scalar loads
    // would otherwise wait for each line in turn.  The loaded values are only
    // kept alive (an empty asm takes them after the loop), never used.
    float w0 = 0.0f, w1 = 0.0f;
    {
        const int nrb = (W.n + kBlock - 1) / kBlock;  // receiver blocks of the chunk
        const int np = m1 - m0;
        const int nrl = (16 * np + 31) / 32;          // record lines, then aux lines
        const int nl = nrl + (4 * np + 31) / 32;
        const float *rb = reinterpret_cast<const float *>(rec + 4 * m0);
        const float *ab = reinterpret_cast<const float *>(aux + m0);
        const int l0 = bt.y + nrb * (int)threadIdx.x, l1 = l0 + nrb * kBlock;
        if (l0 < nl) {
            const float *p = l0 < nrl ? rb + 32 * l0 : ab + 32 * (l0 - nrl);
            w0 = *p;
        }
        if (l1 < nl) {
            const float *p = l1 < nrl ? rb + 32 * l1 : ab + 32 * (l1 - nrl);
            w1 = *p;
        }
    }
    // the receiver type's F table for the in-sphere branch: LDS when it fits;
    // the analytic model keeps its four F constants in registers
    extern __shared__ float lds_tab[];
    DiffFParams fp{0.0f, 0.0f, 0.0f, 0.0f};
    if (AN) fp = diff_f_params(a.dtypes[ti]);
    const bool pbc = a.pbc_flag[bt.x] != 0;
    ForceAcc acc{0, 0, 0, 0, 0, 0, 0, 0, 0};
    // two call sites, so the table pointer of each has one address space and
    // the LDS one compiles to LDS reads
    // a wave with no receiver of the window (small windows) stops after the
    // block's one barrier instead of running the point loop for nothing
    // (with replicas every wave holds receivers, so none leaves before the
    // replica barrier)
    const bool wave_idle = bt.y * kBlock + (slot & ~(kWave - 1)) >= W.n;
    if (!AN && force_tab_floats(a.tab) <= kTabLdsFloats) {
        stage_lds(lds_tab, a.tab.f, force_tab_floats(a.tab));
        __syncthreads();
        if (wave_idle) return;
        force_chunk<SPLIT, AN>(a, rec, aux, ms, me, W, pbc, xi, R2, lds_tab + ti * a.tab.ld_f, fp,
                               acc);
    } else {
        if (wave_idle) return;
        force_chunk<SPLIT, AN>(a, rec, aux, ms, me, W, pbc, xi, R2,
                               AN ? nullptr : a.tab.f + ti * a.tab.ld_f, fp, acc);
    }
    asm volatile("" ::"v"(w0), "v"(w1));
    if (REP && r > 1 &&
        !replica_sum<SPLIT>(reinterpret_cast<double *>(lds_tab + a.red_off), r, q, slot,
                            nrw * kWave, acc))
        return;
    if (!valid) return;
    const long long o = W.ppo + (long long)chunk * W.n + il;
    double2 *e = reinterpret_cast<double2 *>(a.part + o * force_part_stride(SPLIT));
    e[0] = make_double2(acc.ex, acc.ey);
    e[1] = make_double2(acc.ez, acc.lx);
    e[2] = make_double2(acc.ly, acc.lz);
    if (SPLIT) {
        e[3] = make_double2(acc.cx, acc.cy);
        e[4] = make_double2(acc.cz, 0.0);
    }
}

// Small systems (every window within one receiver block, at most
// kFuseMaxPts points a chunk, at most kFuseMaxBlocks blocks): the epilogue
// and the force pass in one launch.  The block runs the epilogue of its
// chunk's points (no other block needs them: the chunk has one receiver
// block) into LDS, then the force loop reads the records from there.
// Records and force sums are those of the two launches bit for bit; the
// records are not written to memory.  The epilogue keeps all (up to eight)
// split groups' partial loads in flight, and its FP64 registers then allow one
// block per CU, so the launch is fused only when its blocks fit one round of
// them: beyond that the separate launches are faster.
constexpr int kFuseMaxPts = 256;
constexpr int kFuseMaxBlocks = 256;

__host__ __device__ inline int fused_lds_floats(const TableView &t)
{
    return ((force_tab_floats(t) + 3) & ~3) + kFuseMaxPts * (kRecFloats + 4);
}

__global__ __launch_bounds__(kBlock) void k_force_epi(ForceArgs a, EpilogueArgs ea)
{
    const int4 bt = a.blocks[a.xcd_runs ? xcd_run_entry(blockIdx.x, gridDim.x) : blockIdx.x];
    const WinDesc W = a.wins[bt.x];
    const int r = force_replicas(W.n), nrw = 4 / r;
    // the wave index read from the first lane, so the compiler knows it is
    // wave-uniform and the replica's point range stays on scalar loads
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave), q = wv / nrw;
    const int slot = (wv % nrw) * kWave + (int)threadIdx.x % kWave;
    const int il = bt.y * kBlock + slot;
    const int chunk = bt.z;
    const long long pbase = (long long)W.a0 * a.nmc;
    const int m0 = chunk * W.ppc;
    const int np = min(W.n * a.nmc, m0 + W.ppc) - m0;
    const int sub = (np + r - 1) / r;
    const int ms = min(np, q * sub), me = min(np, ms + sub);  // chunk-local
    const bool valid = il < W.n;
    const int i = W.a0 + (valid ? il : 0);
    const float4 xi = a.xyzq[i];
    const int ti = a.itype[i];
    const float R2 = valid ? a.rmax2[ti] : -1.0f;
    extern __shared__ float lds_tab[];
    const int nt = force_tab_floats(a.tab);
    float4 *lrec = reinterpret_cast<float4 *>(lds_tab + ((nt + 3) & ~3));
    float4 *laux = lrec + 4 * kFuseMaxPts;
    for (int t = threadIdx.x; t < np; t += kBlock) {
        PointRecord pr;
        epilogue_point<8>(ea, W, pbase + m0 + t, pr);
        for (int k = 0; k < 4; ++k) lrec[4 * t + k] = pr.r[k];
        laux[t] = pr.aux;
    }
    stage_lds(lds_tab, a.tab.f, nt);
    __syncthreads();
    const bool wave_idle = bt.y * kBlock + (wv % nrw) * kWave >= W.n;
    if (wave_idle) return;
    const bool pbc = a.pbc_flag[bt.x] != 0;
    const DiffFParams fp{0.0f, 0.0f, 0.0f, 0.0f};
    ForceAcc acc{0, 0, 0, 0, 0, 0, 0, 0, 0};
    force_chunk<false, false>(a, lrec, laux, ms, me, W, pbc, xi, R2, lds_tab + ti * a.tab.ld_f,
                              fp, acc);
    if (r > 1 && !replica_sum<false>(reinterpret_cast<double *>(lds_tab + a.red_off), r, q, slot,
                                     nrw * kWave, acc))
        return;
    if (!valid) return;
    const long long o = W.ppo + (long long)chunk * W.n + il;
    double2 *e = reinterpret_cast<double2 *>(a.part + o * force_part_stride(false));
    e[0] = make_double2(acc.ex, acc.ey);
    e[1] = make_double2(acc.ez, acc.lx);
    e[2] = make_double2(acc.ly, acc.lz);
}

// F_i = -q_i vtot_i S_e - vtot_i S_lj over the chunk partials, summed in the
// fixed slice order of isspa_slice_sum (kSumAtoms atoms x kSumSlices chunk
// slices a block).
template <bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_force_reduce(ReduceArgs a)
{
    constexpr int comps = SPLIT ? 9 : 6;
    __shared__ double red[kSumSlices][comps][kSumAtoms];
    const int al = threadIdx.x % kSumAtoms, sl = threadIdx.x / kSumAtoms;
    const int i = blockIdx.x * kSumAtoms + al;
    const bool valid = i < a.ntot;
    double s[comps];
    for (int k = 0; k < comps; ++k) s[k] = 0.0;
    if (valid) isspa_slice_sum<comps>(a, i, sl, s);
    combine_slices<comps>(red, al, sl, s);
    if (sl != 0 || !valid) return;
    if (!SPLIT) {
        const float3 fi = isspa_atom_force(a, i, s);
        float4 f = a.f[i];
        f.x += fi.x;
        f.y += fi.y;
        f.z += fi.z;
        a.f[i] = f;
        if (a.f_alone) a.f_alone[i] = make_float4(fi.x, fi.y, fi.z, 0.0f);
    } else {
        const float4 x = a.xyzq[i];
        const double vt = (double)a.vtot[a.itype[i]];
        const double qv = -(double)x.w * vt;
        for (int c = 0; c < 3; ++c) {
            a.f_pol[3 * i + c] = qv * s[c];
            a.f_lj[3 * i + c] = -vt * s[3 + c];
            a.f_cdd[3 * i + c] = qv * s[6 + c];
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

hipError_t launch_mc_points(const WinDesc *wins, const int *atom_win, const float4 *xyzq,
                            const int *itype, const float *rmax, const float *rmax2, int ntot,
                            int nmc, uint64_t seed, uint32_t step, const uint32_t *step_d,
                            float4 *pts, hipStream_t s, int nwin, float rmax_all, int *flag)
{
    const long long P = (long long)ntot * nmc;
    const int nflag = flag ? nwin : 0;
    if (P == 0 && nflag == 0) return hipSuccess;
    const PointsArgs a{wins, atom_win, xyzq, itype, rmax, rmax2, ntot, nmc, seed, step, step_d, pts};
    hipLaunchKernelGGL(k_mc_points, dim3(blocks_for(P) + nflag), dim3(kBlock), 0, s, a, nflag,
                       rmax_all, flag);
    return hipGetLastError();
}

hipError_t launch_field(const FieldArgs &a0, int nblocks, hipStream_t s)
{
    FieldArgs a = a0;
    a.nfield = nblocks;
    if (a.dtypes) {
        if (nblocks == 0) return hipSuccess;
        const int nt = diff_tab_floats(a.tab);
        const size_t lds = std::max<size_t>(nt <= kTabLdsFloats ? (size_t)nt * sizeof(float) : 0,
                                            a.max_fg > 1 ? kFieldXchgBytes : 0);
        hipLaunchKernelGGL((k_field<true, false>), dim3(nblocks), dim3(kBlock), lds, s, a);
    } else {
        if (nblocks + a.nb_blocks == 0) return hipSuccess;
        const int nt = field_tab_floats(a.tab);
        size_t lds = std::max<size_t>(nt <= kTabLdsFloats ? (size_t)nt * sizeof(float) : 0,
                                      a.max_fg > 1 ? kFieldXchgBytes : 0);
        if (a.nb_blocks) {
            lds = std::max(lds, sizeof(float2) * a.nb.T * a.nb.T);
            hipLaunchKernelGGL((k_field<false, true>), dim3(nblocks + a.nb_blocks), dim3(kBlock),
                               lds, s, a);
        } else {
            hipLaunchKernelGGL((k_field<false, false>), dim3(nblocks), dim3(kBlock), lds, s, a);
        }
    }
    return hipGetLastError();
}

hipError_t launch_epilogue(const EpilogueArgs &a, hipStream_t s)
{
    if (a.Ptot == 0) return hipSuccess;
    const int eb = (int)blocks_for(a.Ptot);
    if (eb < kEpiWideGrid)
        hipLaunchKernelGGL(k_field_epilogue<8>, dim3(eb), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_field_epilogue<2>, dim3(eb), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

bool force_epi_fits(const TableView &t, int max_window_atoms, int max_chunk_points, int nblocks)
{
    return max_window_atoms <= kBlock && max_chunk_points <= kFuseMaxPts &&
           nblocks <= kFuseMaxBlocks && force_tab_floats(t) <= kTabLdsFloats;
}

hipError_t launch_force_epi(const ForceArgs &a, const EpilogueArgs &ea, int nblocks, hipStream_t s)
{
    if (nblocks == 0) return hipSuccess;
    ForceArgs b = a;
    b.red_off = fused_lds_floats(a.tab);  // a multiple of 4 floats
    const int red = a.replicas > 1 ? 2 * force_replica_doubles(false) : 0;
    const size_t lds = (size_t)(b.red_off + red) * sizeof(float);
    hipLaunchKernelGGL(k_force_epi, dim3(nblocks), dim3(kBlock), lds, s, b, ea);
    return hipGetLastError();
}

hipError_t launch_force(const ForceArgs &a, int nblocks, bool split, hipStream_t s)
{
    if (nblocks == 0) return hipSuccess;
    // dynamic LDS: the F table when it fits (table model), then the
    // replica-sum area when some window has receiver replicas
    ForceArgs b = a;
    const int nt = force_tab_floats(a.tab);
    b.red_off = (!a.dtypes && nt <= kTabLdsFloats) ? (nt + 3) & ~3 : 0;
    const int red = a.replicas > 1 ? 2 * force_replica_doubles(split) : 0;
    const size_t lds = (size_t)(b.red_off + red) * sizeof(float);
    const dim3 g(nblocks), t(kBlock);
    const int v = (a.dtypes ? 4 : 0) + (split ? 2 : 0) + (a.replicas > 1 ? 1 : 0);
    switch (v) {
    case 0: hipLaunchKernelGGL((k_force<false, false, false>), g, t, lds, s, b); break;
    case 1: hipLaunchKernelGGL((k_force<false, false, true>), g, t, lds, s, b); break;
    case 2: hipLaunchKernelGGL((k_force<true, false, false>), g, t, lds, s, b); break;
    case 3: hipLaunchKernelGGL((k_force<true, false, true>), g, t, lds, s, b); break;
    case 4: hipLaunchKernelGGL((k_force<false, true, false>), g, t, lds, s, b); break;
    case 5: hipLaunchKernelGGL((k_force<false, true, true>), g, t, lds, s, b); break;
    case 6: hipLaunchKernelGGL((k_force<true, true, false>), g, t, lds, s, b); break;
    default: hipLaunchKernelGGL((k_force<true, true, true>), g, t, lds, s, b); break;
    }
    return hipGetLastError();
}

hipError_t launch_force_reduce(const ReduceArgs &a, bool split, hipStream_t s)
{
    if (a.ntot == 0) return hipSuccess;
    const int nb = (a.ntot + kSumAtoms - 1) / kSumAtoms;
    if (split)
        hipLaunchKernelGGL(k_force_reduce<true>, dim3(nb), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_force_reduce<false>, dim3(nb), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}


}  // namespace isspa
