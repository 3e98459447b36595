// Per-window minimum-image flag, computed by one 256-thread block.
#pragma once

#include "kernels.hpp"

namespace isspa {

// flag = 1 when some pair separation of window W may exceed half a box
// (minimum image needed); 0 when (extent + rmax + margin) < L/2 on every axis,
// in which case the single-image minimum image is the identity for every
// atom-atom and point-atom pair and the kernels take their shift-free loops.
// Every thread of the block must call it.
__device__ inline void pbc_flag_block(const WinDesc &W, const float4 *__restrict__ xyzq,
                                      float rmax_all, int *flag)
{
    __shared__ float lo[3][256], hi[3][256];
    float l[3] = {3.0e38f, 3.0e38f, 3.0e38f}, u[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int i = W.a0 + (int)threadIdx.x; i < W.a0 + W.n; i += 256) {
        const float4 x = xyzq[i];
        l[0] = fminf(l[0], x.x); u[0] = fmaxf(u[0], x.x);
        l[1] = fminf(l[1], x.y); u[1] = fmaxf(u[1], x.y);
        l[2] = fminf(l[2], x.z); u[2] = fmaxf(u[2], x.z);
    }
    for (int k = 0; k < 3; ++k) {
        lo[k][threadIdx.x] = l[k];
        hi[k][threadIdx.x] = u[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) {
                lo[k][threadIdx.x] = fminf(lo[k][threadIdx.x], lo[k][threadIdx.x + w]);
                hi[k][threadIdx.x] = fmaxf(hi[k][threadIdx.x], hi[k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int need = 0;
        for (int k = 0; k < 3; ++k) {
            const float ext = hi[k][0] - lo[k][0];
            if (!(ext + rmax_all + 1.0e-2f < W.h)) need = 1;  // NaN -> need
        }
        *flag = need;
    }
}

}  // namespace isspa
