// A.1 MC points (src/isspa_force_cuda.cu:73-107), Philox instead of curand,
// as a per-point device function shared by the points kernel and the MD
// engine's step prologue.
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace isspa {


// x + R * (v 2^-24) with two roundings, never fused into one FMA, so the
// oracle forms the same bits (v 2^-24 itself is exact)
__device__ __forceinline__ float point_coord(float x, float R, int v)
{
#pragma clang fp contract(off)
    const float u = (float)v * 5.9604644775390625e-08f;
    const float du = R * u;
    return x + du;
}

// Point gp (global index; window w owns points [a0_w nMC, (a0_w + n_w) nMC),
// point m of the window belongs to atom m / nMC) is drawn with counter
// {m, draw, step, sys_w}: the first draw whose 24-bit signed lattice vector
// lies in the unit ball.  .w of each point carries the owner's rmax^2.
__device__ __forceinline__ void mc_point(const PointsArgs &a, long long gp)
{
    const uint32_t step = a.step_d ? *a.step_d : a.step;
    const int k = (int)(gp / a.nmc);
    const WinDesc W = a.wins[a.atom_win[k]];
    const uint32_t m = (uint32_t)(gp - (long long)W.a0 * a.nmc);
    const int t = a.itype[k];
    int v0 = 0, v1 = 0, v2 = 0;
    for (int d = 0; d < kMcMaxAttempts; ++d) {
        const uint4 o = philox_draw(a.seed, m, kStreamPoints, (uint32_t)d, step, W.sys);
        const int c0 = 2 * (int)(o.x >> 8) - (1 << 24);
        const int c1 = 2 * (int)(o.y >> 8) - (1 << 24);
        const int c2 = 2 * (int)(o.z >> 8) - (1 << 24);
        const uint64_t r2 = (uint64_t)((int64_t)c0 * c0 + (int64_t)c1 * c1 + (int64_t)c2 * c2);
        if (r2 <= kMcAcceptMax) {
            v0 = c0; v1 = c1; v2 = c2;
            break;
        }
    }
    const float R = a.rmax[t];
    const float4 x = a.xyzq[k];
    float4 p;
    p.x = point_coord(x.x, R, v0);
    p.y = point_coord(x.y, R, v1);
    p.z = point_coord(x.z, R, v2);
    p.w = a.rmax2[t];
    a.pts[gp] = p;
}

}  // namespace isspa
