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
template <bool PBC, bool SPLIT>
__device__ __forceinline__ void nonbond_tile(const NonbondArgs &a, int i, int j0, int j1,
                                             float L, float h, float4 xi, const uint32_t *ex,
                                             bool near, const float2 *__restrict__ ljrow,
                                             double *acc)
{
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;           // total (or LJ in split mode)
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;           // Coulomb (split)
    float p0 = 0.f, p1 = 0.f, p2 = 0.f;           // pair term (split)
    for (int j = j0; j < j1; ++j) {
        const float4 xj = a.xyzq[j];  // wave-uniform
        const float4 aj = a.aux[j];
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
        const int tj = __float_as_int(aj.w);
        const float2 lj = ljrow[tj];
        const float r6 = inv2 * inv2 * inv2;
        const float flj = __builtin_fmaf(lj.x, r6, -lj.y) * r6 * inv2;
        const float fc = qq * inv2 * inv;
        const float far = qq * a.cfar * inv2 * inv;
        const float nearv = qq * __builtin_fmaf(aj.z, d2 * inv, aj.y);
        float fdir = d2 > aj.x ? far : nearv;
        bool excl = false;
        if (near) {
            const int w = j - i + 64;
            if (w >= 0 && w < 128) excl = (ex[w >> 5] >> (w & 31)) & 1u;
        }
        if (a.excl_ptr) {  // fallback: exclusions beyond the 128-wide window
            const int e0 = a.excl_ptr[i], e1 = a.excl_ptr[i + 1];
            for (int e = e0; e < e1; ++e) excl |= (a.excl_idx[e] == j);
        }
        const bool self = (j == i);
        const float fnb = (excl || self) ? 0.0f : flj + fc;
        fdir = self ? 0.0f : fdir;
        if (!SPLIT) {
            const float s = fnb + fdir;
            s0 = __builtin_fmaf(s, rx, s0);
            s1 = __builtin_fmaf(s, ry, s1);
            s2 = __builtin_fmaf(s, rz, s2);
        } else {
            const float fl = (excl || self) ? 0.0f : flj;
            const float fcc = (excl || self) ? 0.0f : fc;
            s0 = __builtin_fmaf(fl, rx, s0);
            s1 = __builtin_fmaf(fl, ry, s1);
            s2 = __builtin_fmaf(fl, rz, s2);
            c0 = __builtin_fmaf(fcc, rx, c0);
            c1 = __builtin_fmaf(fcc, ry, c1);
            c2 = __builtin_fmaf(fcc, rz, c2);
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
        if (pbc)
            nonbond_tile<true, SPLIT>(a, i, jt, je, W.L, W.h, xi, ex, near, ljrow, acc);
        else
            nonbond_tile<false, SPLIT>(a, i, jt, je, W.L, W.h, xi, ex, near, ljrow, acc);
    }
    if (!valid) return;
    const long long o = W.npo + (long long)bt.z * W.n + il;
    const int comps = SPLIT ? 9 : 3;
    for (int k = 0; k < comps; ++k) a.part[k * a.B + o] = acc[k];
}

}  // namespace isspa
