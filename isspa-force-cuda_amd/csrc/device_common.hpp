// Device helpers shared by every gfx950 kernel of the IS-SPA engine.
//
// Arithmetic contract with the CPU oracle: decision values (squared pair
// distances for the in-sphere / LJ-versus-CDD tests, minimum-image shifts,
// MC point coordinates) are formed with the exact FP32 operation sequence
// written here -- explicit fmaf where a fused op is used, __fmul_rn /
// __fadd_rn where contraction must not happen -- so the oracle can form the
// same values bit for bit and both take the same branch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace isspa {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Coefficients of the Langevin-moment series below x = 5:
//   i_l(x) / x^l = sum_k T_k,  T_0 = 1/(2l+1)!!,  T_(k+1) = T_k h / ((k+1)(2l+2k+3)),
// h = x^2/2.  The per-term divisors are compile-time constants, so the FP64
// loops multiply by their reciprocals instead of dividing.
// ---------------------------------------------------------------------------
constexpr int kLgTerms = 30;
struct LangevinCoef {
    double first[4];           // 1 / (2l+1)!!
    double step[4][kLgTerms];  // 1 / ((k+1)(2l+2k+3))
    double inv[4][kLgTerms];   // 1 / (2l+2k+3)  (derivative of the series)
};
constexpr LangevinCoef make_langevin_coef()
{
    LangevinCoef c{};
    double f = 1.0;
    for (int l = 0; l < 4; ++l) {
        if (l > 0) f /= (double)(2 * l + 1);
        c.first[l] = f;
        for (int k = 0; k < kLgTerms; ++k) {
            const double d = (double)(2 * l + 2 * k + 3);
            c.step[l][k] = 1.0 / ((double)(k + 1) * d);
            c.inv[l][k] = 1.0 / d;
        }
    }
    return c;
}
__constant__ constexpr LangevinCoef kLangevin = make_langevin_coef();

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter layout:
//   ctr = { index, (stream << 24) | draw, step, system }, key = { seed lo, hi }
// ---------------------------------------------------------------------------
constexpr uint32_t kStreamPoints = 0u;
constexpr uint32_t kStreamAndersen = 1u;

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
        const uint32_t lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ uint4 philox_draw(uint64_t seed, uint32_t index, uint32_t stream,
                                             uint32_t draw, uint32_t step, uint32_t sys)
{
    return philox4x32_10(make_uint4(index, (stream << 24) | draw, step, sys),
                         make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
}

// MC rejection sampling on the 2^-24 lattice: accept iff |2v - 2^24|^2 summed
// over the three axes is <= floor(0.99 * 2^48)  (u in the ball |u|^2 < 0.99).
constexpr uint64_t kMcAcceptMax = 278660226943549ULL;
constexpr int kMcMaxAttempts = 64;

// ---------------------------------------------------------------------------
// single-image minimum image (src/cuda_vector_routines.h:28-47):
//   r > h -> r - L ; r < -h -> r + L.  fma(-L, k, r) with k in {-1,0,1} rounds
// exactly like r -+ L.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float min_image1(float r, float L, float h)
{
    const float k = (r > h) ? 1.0f : ((r < -h) ? -1.0f : 0.0f);
    return __builtin_fmaf(-L, k, r);
}

// squared distance in the operation order the oracle reproduces
__device__ __forceinline__ float dist2_f32(float rx, float ry, float rz)
{
    return __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
}

__device__ __forceinline__ float rsqrt_f32(float x) { return __builtin_amdgcn_rsqf(x); }

// wave-uniform "any lane" test
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }

// A read-only input seen through the constant address space.  Its loads at a
// wave-uniform index are scalar loads whatever else the kernel writes: the
// compiler otherwise turns them into per-lane vector loads when it cannot
// prove that no store of the kernel reaches them (as in a kernel that also
// runs the all-pairs blocks, which write their partials).  Only for memory
// that no work-item writes during the launch.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *read_only(const T *p)
{
    return (const __attribute__((address_space(4))) T *)p;
}

}  // namespace isspa
