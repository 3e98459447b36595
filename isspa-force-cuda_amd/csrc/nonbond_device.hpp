// A.4 all-pairs nonbond as a per-block device function, shared by the
// all-pairs kernel and the field launch of the MD engine (which runs the
// all-pairs blocks beside the field blocks when it has no second stream).
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace isspa {

constexpr int kNbBlock = 256;

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

// block blk of the all-pairs grid; lds_lj holds T*T float2 (dynamic LDS)
template <bool SPLIT>
__device__ __forceinline__ void nonbond_block(const NonbondArgs &a, int blk, float2 *lds_lj)
{
    for (int t = threadIdx.x; t < a.T * a.T; t += kNbBlock) lds_lj[t] = a.lj[t];
    __syncthreads();
    const int4 bt = a.blocks[blk];
    const WinDesc W = a.wins[bt.x];
    const int il = bt.y * kNbBlock + threadIdx.x;
    const int j0 = W.a0 + bt.z * W.apc;
    const int j1 = min(W.a0 + W.n, j0 + W.apc);
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
    const long long o = W.npo + (long long)bt.z * W.n + il;
    const int comps = SPLIT ? 9 : 3;
    for (int k = 0; k < comps; ++k) a.part[k * a.B + o] = acc[k];
}

}  // namespace isspa
