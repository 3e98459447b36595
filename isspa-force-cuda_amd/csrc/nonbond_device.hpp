// A.4 all-pairs nonbond as a per-block device function, shared by the
// all-pairs kernel and the field launch of the MD engine (which runs the
// all-pairs blocks beside the field blocks when it has no second stream).
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace isspa {

// A.4 all-pairs nonbond (src/nonbond_force_cuda.cu:31-174):
//   non-excluded pairs: (12A/d^12 - 6B/d^6)/d^2 + q_i q_j/d^3
//   all pairs:          IS-SPA pair electrostatics fdir(d, R_j)
// F_i += s r with r = min_image(x_i - x_j), j over the receiver's window.
// Tiles come in three kinds:
//   kNbNear  within 64 indices of the wave's receivers: the receiver itself
//            and its excluded partners are tested per pair;
//   kNbFar   farther tiles; only the exclusions beyond the 128-wide windows
//            (a per-atom list) can fall there;
//   kNbFast  farther tiles of a total-force evaluation with no such list:
//            no pair is excluded, and the pair term takes its far form
//            everywhere, s = flj + qq d^-3 (1 + cfar), its near form being
//            swapped in by a branch that only waves holding a partner within
//            R_j of one of their receivers take.
constexpr int kNbNear = 0, kNbFar = 1, kNbFast = 2;

template <bool PBC, bool SPLIT, int KIND>
__device__ __forceinline__ void nonbond_tile(const NonbondArgs &a, int i, int j0, int j1,
                                             float L, float h, float4 xi, const uint32_t *ex,
                                             const float2 *__restrict__ ljrow, double *acc)
{
    static_assert(KIND != kNbFast || !SPLIT, "the fast tile sums the total force only");
    constexpr bool NEAR = KIND == kNbNear;
    const float cfar1 = 1.0f + a.cfar;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;           // total (or LJ in split mode)
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;           // Coulomb (split)
    float p0 = 0.f, p1 = 0.f, p2 = 0.f;           // pair term (split)
    // the next partner (wave-uniform position and aux) and the receiver's LJ
    // entry for its type are fetched while this pair is evaluated
    float4 nx = a.xyzq[j0], na = a.aux[j0];
    float2 nlj = ljrow[__float_as_int(na.w)];
    for (int j = j0; j < j1; ++j) {
        const float4 xj = nx, aj = na;
        const float2 lj = nlj;
        const int jn = min(j + 1, j1 - 1);
        nx = a.xyzq[jn];
        na = a.aux[jn];
        nlj = ljrow[__float_as_int(na.w)];
        float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        if (PBC) {
            rx = min_image1(rx, L, h);
            ry = min_image1(ry, L, h);
            rz = min_image1(rz, L, h);
        }
        const float d2 = dist2_f32(rx, ry, rz);
        const float inv = rsqrt_f32(d2);
        const float inv2 = inv * inv;
        const float qq = xi.w * xj.w;
        const float r6 = inv2 * inv2 * inv2;
        float flj = __builtin_fmaf(lj.x, r6, -lj.y) * r6 * inv2;
        float fc = qq * (inv2 * inv);
        if (KIND == kNbFast) {
            float s = __builtin_fmaf(fc, cfar1, flj);
            const bool in = d2 <= aj.x;
            // a wave-uniform test, marked unlikely, so the rare fix is a
            // skipped branch rather than straight-line selects on every pair
            if (__builtin_expect(wave_any(in), 0)) {
                // partner within R_j: the pair term's near form for its far one
                const float nearv = qq * __builtin_fmaf(aj.z, d2 * inv, aj.y);
                s = in ? s + (nearv - fc * a.cfar) : s;
            }
            s0 = __builtin_fmaf(s, rx, s0);
            s1 = __builtin_fmaf(s, ry, s1);
            s2 = __builtin_fmaf(s, rz, s2);
            continue;
        }
        const float far = fc * a.cfar;  // qq cfar / d^3
        const float nearv = qq * __builtin_fmaf(aj.z, d2 * inv, aj.y);
        float fdir = d2 > aj.x ? far : nearv;
        if (NEAR) {
            // tiles within 64 indices of the wave's receivers: the receiver
            // itself and its excluded partners
            bool excl = false;
            const int w = j - i + 64;
            if (w >= 0 && w < 128) excl = (ex[w >> 5] >> (w & 31)) & 1u;
            if (a.excl_ptr) {  // fallback: exclusions beyond the 128-wide window
                const int e0 = a.excl_ptr[i], e1 = a.excl_ptr[i + 1];
                for (int e = e0; e < e1; ++e) excl |= (a.excl_idx[e] == j);
            }
            const bool self = (j == i);
            flj = (excl || self) ? 0.0f : flj;
            fc = (excl || self) ? 0.0f : fc;
            fdir = self ? 0.0f : fdir;
        } else if (KIND == kNbFar && a.excl_ptr) {
            // far tiles can hold only the exclusions beyond the windows
            bool excl = false;
            const int e0 = a.excl_ptr[i], e1 = a.excl_ptr[i + 1];
            for (int e = e0; e < e1; ++e) excl |= (a.excl_idx[e] == j);
            flj = excl ? 0.0f : flj;
            fc = excl ? 0.0f : fc;
        }
        if (!SPLIT) {
            const float s = flj + fc + fdir;
            s0 = __builtin_fmaf(s, rx, s0);
            s1 = __builtin_fmaf(s, ry, s1);
            s2 = __builtin_fmaf(s, rz, s2);
        } else {
            s0 = __builtin_fmaf(flj, rx, s0);
            s1 = __builtin_fmaf(flj, ry, s1);
            s2 = __builtin_fmaf(flj, rz, s2);
            c0 = __builtin_fmaf(fc, rx, c0);
            c1 = __builtin_fmaf(fc, ry, c1);
            c2 = __builtin_fmaf(fc, rz, c2);
            p0 = __builtin_fmaf(fdir, rx, p0);
            p1 = __builtin_fmaf(fdir, ry, p1);
            p2 = __builtin_fmaf(fdir, rz, p2);
        }
    }
    acc[0] += s0; acc[1] += s1; acc[2] += s2;
    if (SPLIT) {
        acc[3] += c0; acc[4] += c1; acc[5] += c2;
        acc[6] += p0; acc[7] += p1; acc[8] += p2;
    }
}

// ---------------------------------------------------------------------------
// Each unordered pair once (total force of windows of kSymMinAtoms or more).
//
// The window's atoms are cut into blocks of kNbBlock = 256.  A super-tile
// pairs receiver block rb with partner block pb >= rb + 2: its 65,536 pairs
// are all more than 256 indices apart, so none is excluded (the windows hold
// every exclusion) and none is a self pair.  Each pair is evaluated once and
// its force added to both atoms.  The pairs of a block with itself and with
// its two neighbours are evaluated in both orders by ordered blocks (one per
// receiver block and partner block), which also test the exclusions.  Atom
// a of block A thus has one partial entry per block q: from super-tile
// (q, A) for q <= A - 2, from the ordered block (A, q) for |q - A| <= 1, and
// from super-tile (A, q) for q >= A + 2.
//
// Inside a super-tile each 16-lane row of a wave holds 16 receivers (one a
// lane, in VGPRs) and, in round k, the 16 partners of group (row + k) mod 16
// (one a lane).  Step s of a round pairs lane c with the partner of lane
// (c + s) mod 16, read by a DPP row rotation, so no pair operand is
// wave-uniform and the pair loop reads no SGPR.  The partner-side force
// accumulators travel with their partners: each step moves them one lane
// (a DPP move) and adds the step's term to them.  After
// 16 rounds every row has met every partner group once; the rows' partner
// sums go to LDS slots in round order, one barrier a round.
//
// Pair terms are those of the ordered kernel's far tiles: the pair
// electrostatics in the far form, the near form swapped in by a wave-uniform
// branch, separately for the two directions (the receiver-side force uses
// the partner's sphere R_j, the partner-side force the receiver's R_i).
// With the symmetric LJ table each direction's term equals bit for bit what
// the ordered kernel computes for that ordered pair; only the grouping of
// the sums differs.
// ---------------------------------------------------------------------------
// position of a missing partner (past the window's end): every pair term
// with it underflows to zero
constexpr float kNbAbsent = 1.0e18f;

// lane c of each 16-lane row reads lane (c + S) mod 16 of the same row
// (row_ror:N hands lane c the value of lane c - N)
template <int S>
__device__ __forceinline__ int row_rot_i(int v)
{
    if constexpr ((S & 15) == 0)
        return v;
    else
        return __builtin_amdgcn_mov_dpp(v, 0x120 + (16 - (S & 15)), 0xF, 0xF, true);
}
template <int S>
__device__ __forceinline__ float row_rot(float v)
{
    return __int_as_float(row_rot_i<S>(__float_as_int(v)));
}

struct SymLane {
    float xi, yi, zi, qi, ri2, cai, cbi;  // the lane's receiver
    float cfar, cfar1, L, h;              // in VGPRs
    const char *ljrow;                    // receiver's LJ row in LDS
    float px, py, pz, pq, pr2, pca, pcb;  // the lane's partner of this round
    int pt8;                              // its LJ column, in bytes
    float2 lj;                            // the LJ pair entry of the next step
    float sx, sy, sz;                     // receiver sums of the round
    float jx, jy, jz;                     // partner sums travelling with the partners
};

template <int S>
__device__ __forceinline__ float2 sym_lj(const SymLane &l)
{
    return *reinterpret_cast<const float2 *>(l.ljrow + row_rot_i<S>(l.pt8));
}

template <bool PBC, int S>
__device__ __forceinline__ void sym_step(SymLane &l)
{
    float rx = l.xi - row_rot<S>(l.px);
    float ry = l.yi - row_rot<S>(l.py);
    float rz = l.zi - row_rot<S>(l.pz);
    if (PBC) {
        rx = min_image1(rx, l.L, l.h);
        ry = min_image1(ry, l.L, l.h);
        rz = min_image1(rz, l.L, l.h);
    }
    const float d2 = dist2_f32(rx, ry, rz);
    const float inv = rsqrt_f32(d2);
    const float inv2 = inv * inv;
    const float qq = l.qi * row_rot<S>(l.pq);
    const float r6 = inv2 * inv2 * inv2;
    // this step's LJ entry was read during the previous step; read the next one
    const float2 lj = l.lj;
    if constexpr (S + 1 < 16) l.lj = sym_lj<S + 1>(l);
    const float flj = __builtin_fmaf(lj.x, r6, -lj.y) * r6 * inv2;
    const float fc = qq * (inv2 * inv);
    float si = __builtin_fmaf(fc, l.cfar1, flj);
    float sj = si;
    const float rj2 = row_rot<S>(l.pr2);
    const bool in_i = d2 <= rj2;   // inside the partner's sphere: near form on the receiver
    const bool in_j = d2 <= l.ri2;  // inside the receiver's sphere: near form on the partner
    // wave-uniform and rare
    if (__builtin_expect(wave_any(in_i || in_j), 0)) {
        const float d = d2 * inv;
        const float far = fc * l.cfar;
        const float ni = qq * __builtin_fmaf(row_rot<S>(l.pcb), d, row_rot<S>(l.pca));
        const float nj = qq * __builtin_fmaf(l.cbi, d, l.cai);
        si = in_i ? si + (ni - far) : si;
        sj = in_j ? sj + (nj - far) : sj;
    }
    l.sx = __builtin_fmaf(si, rx, l.sx);
    l.sy = __builtin_fmaf(si, ry, l.sy);
    l.sz = __builtin_fmaf(si, rz, l.sz);
    // the receiver sums are due now: left to itself the compiler sinks the
    // 16 steps' multiply-adds below the last rare branch of the round and
    // keeps every step's r and s alive until then (over 100 VGPRs)
    asm volatile("" : "+v"(l.sx), "+v"(l.sy), "+v"(l.sz));
    // the partner sums move one lane down, then take this step's term
    // (-sj r: the partner sees the pair from the other side)
    if constexpr (S == 0) {
        l.jx = -(sj * rx);
        l.jy = -(sj * ry);
        l.jz = -(sj * rz);
    } else {
        l.jx = __builtin_fmaf(-sj, rx, row_rot<1>(l.jx));
        l.jy = __builtin_fmaf(-sj, ry, row_rot<1>(l.jy));
        l.jz = __builtin_fmaf(-sj, rz, row_rot<1>(l.jz));
    }
}

template <bool PBC, int S>
__device__ __forceinline__ void sym_steps(SymLane &l)
{
    sym_step<PBC, S>(l);
    if constexpr (S + 1 < 16) sym_steps<PBC, S + 1>(l);
}

// super-tile (rb, pb) of window W; jl: kSymLdsDoubles zeroed LDS doubles
template <bool PBC>
__device__ __forceinline__ void nonbond_supertile(const NonbondArgs &a, const WinDesc &W, int rb,
                                                  int pb, const float2 *lds_lj, double *jl)
{
    const int t = threadIdx.x, c = t & 15, row = t >> 4;
    const int i = W.a0 + rb * kNbBlock + t;  // receiver blocks of super-tiles are full
    const float4 xi = a.xyzq[i];
    const float4 ai = a.aux[i];
    SymLane l;
    l.xi = xi.x; l.yi = xi.y; l.zi = xi.z; l.qi = xi.w;
    l.ri2 = ai.x; l.cai = ai.y; l.cbi = ai.z;
    // the far-form coefficients as lane values, so that the pair loop's
    // multiplies by them read VGPRs (a wave-uniform SGPR operand halves the
    // issue rate of an FP32 operation)
    asm volatile("v_mov_b32 %0, %1" : "=v"(l.cfar) : "s"(a.cfar));
    asm volatile("v_mov_b32 %0, %1" : "=v"(l.L) : "s"(W.L));
    asm volatile("v_mov_b32 %0, %1" : "=v"(l.h) : "s"(W.h));
    l.cfar1 = 1.0f + l.cfar;
    l.ljrow = reinterpret_cast<const char *>(lds_lj + __float_as_int(ai.w) * a.T);
    const int pbase = W.a0 + pb * kNbBlock, pend = W.a0 + W.n;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < 16; ++k) {
        const int grp = (row + k) & 15;
        const int p = pbase + 16 * grp + c;
        float4 xp = make_float4(kNbAbsent, kNbAbsent, kNbAbsent, 0.0f);
        float4 ap = make_float4(-1.0f, 0.0f, 0.0f, 0.0f);  // LJ type 0
        if (p < pend) {
            xp = a.xyzq[p];
            ap = a.aux[p];
        }
        l.px = xp.x; l.py = xp.y; l.pz = xp.z; l.pq = xp.w;
        l.pr2 = ap.x; l.pca = ap.y; l.pcb = ap.z;
        l.pt8 = __float_as_int(ap.w) * (int)sizeof(float2);
        l.sx = l.sy = l.sz = 0.0f;
        l.lj = sym_lj<0>(l);
        sym_steps<PBC, 0>(l);
        sx += (double)l.sx;
        sy += (double)l.sy;
        sz += (double)l.sz;
        // after step 15 lane c holds partner c - 1: one more rotation
        // aligns lane c with partner c of the group
        const float jx = row_rot<1>(l.jx), jy = row_rot<1>(l.jy), jz = row_rot<1>(l.jz);
        // every row has a different group this round: no two lanes share a slot
        double *slot = jl + 3 * (16 * grp + c);
        slot[0] += (double)jx;
        slot[1] += (double)jy;
        slot[2] += (double)jz;
        __syncthreads();
    }
    // receiver sums: entry pb of receiver i; partner sums: entry rb of partner pbase + t
    const long long oi = W.spo + (long long)pb * W.n + (i - W.a0);
    a.part[oi] = sx;
    a.part[a.B + oi] = sy;
    a.part[2 * a.B + oi] = sz;
    if (pbase + t < pend) {
        const long long oj = W.spo + (long long)rb * W.n + (pbase + t - W.a0);
        a.part[oj] = jl[3 * t];
        a.part[a.B + oj] = jl[3 * t + 1];
        a.part[2 * a.B + oj] = jl[3 * t + 2];
    }
}

// block blk of the all-pairs grid; lds_lj holds T*T float2 (dynamic LDS),
// followed, with SYM, by kSymLdsDoubles doubles (8-byte aligned).  SPLIT
// writes the ordered layout (npo, one entry per partner chunk); the total
// force writes the total-mode layout (spo, nte entries per atom).  SYM
// compiles the super-tile kinds in (the field launch that carries the
// all-pairs blocks of small systems has none).
template <bool SPLIT, bool SYM>
__device__ __forceinline__ void nonbond_block(const NonbondArgs &a, int blk, float2 *lds_lj)
{
    static_assert(!(SPLIT && SYM), "the split components take every pair in both orders");
    for (int t = threadIdx.x; t < a.T * a.T; t += kNbBlock) lds_lj[t] = a.lj[t];
    const int4 bt = a.blocks[blk];
    const WinDesc W = a.wins[bt.x];
    if (SYM && bt.w == kNbKindSym) {
        double *jl = reinterpret_cast<double *>(lds_lj + ((a.T * a.T + 1) & ~1));
        for (int t = threadIdx.x; t < kSymLdsDoubles; t += kNbBlock) jl[t] = 0.0;
        __syncthreads();
        if (a.pbc_flag[bt.x] != 0)
            nonbond_supertile<true>(a, W, bt.y, bt.z, lds_lj, jl);
        else
            nonbond_supertile<false>(a, W, bt.y, bt.z, lds_lj, jl);
        return;
    }
    __syncthreads();
    const bool near_kind = SYM && bt.w == kNbKindNear;
    // partners: a chunk, or (near kind) 256-atom block bt.z, the receivers'
    // own block or a neighbour; both write entry bt.z
    const int j0 = W.a0 + bt.z * (near_kind ? kNbBlock : W.apc);
    const int j1 = min(W.a0 + W.n, j0 + (near_kind ? kNbBlock : W.apc));
    const int entry = bt.z;
    const int il = bt.y * kNbBlock + threadIdx.x;
    const bool valid = il < W.n;
    const int i = W.a0 + (valid ? il : W.n - 1);  // idle lanes mirror the last atom, results dropped
    const float4 xi = a.xyzq[i];
    const int ti = __float_as_int(a.aux[i].w);
    uint32_t ex[4];
    for (int k = 0; k < 4; ++k) ex[k] = a.excl[4 * i + k];
    const bool pbc = a.pbc_flag[bt.x] != 0;
    // first receiver of this wave, for the near-diagonal test
    const int wave_lo = W.a0 + bt.y * kNbBlock + (threadIdx.x & ~(kWave - 1));
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const float2 *ljrow = lds_lj + ti * a.T;
    constexpr int kTile = 64;
    for (int jt = j0; jt < j1; jt += kTile) {
        const int je = min(j1, jt + kTile);
        const bool near = (jt < wave_lo + kWave + 64) && (je > wave_lo - 64);
        if (near) {
            if (pbc)
                nonbond_tile<true, SPLIT, kNbNear>(a, i, jt, je, W.L, W.h, xi, ex, ljrow, acc);
            else
                nonbond_tile<false, SPLIT, kNbNear>(a, i, jt, je, W.L, W.h, xi, ex, ljrow, acc);
        } else if (SPLIT || a.excl_ptr) {
            if (pbc)
                nonbond_tile<true, SPLIT, kNbFar>(a, i, jt, je, W.L, W.h, xi, ex, ljrow, acc);
            else
                nonbond_tile<false, SPLIT, kNbFar>(a, i, jt, je, W.L, W.h, xi, ex, ljrow, acc);
        } else {
            constexpr int kFast = SPLIT ? kNbFar : kNbFast;
            if (pbc)
                nonbond_tile<true, SPLIT, kFast>(a, i, jt, je, W.L, W.h, xi, ex, ljrow, acc);
            else
                nonbond_tile<false, SPLIT, kFast>(a, i, jt, je, W.L, W.h, xi, ex, ljrow, acc);
        }
    }
    if (!valid) return;
    const long long o = (SPLIT ? W.npo : W.spo) + (long long)entry * W.n + il;
    const int comps = SPLIT ? 9 : 3;
    for (int k = 0; k < comps; ++k) a.part[k * a.B + o] = acc[k];
}

}  // namespace isspa
