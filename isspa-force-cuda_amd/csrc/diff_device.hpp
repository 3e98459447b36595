// Analytic per-type functions of the differentiable IS-SPA model (FP32,
// device side).  The FP64 restatement and its derivation live in the oracle;
// the definitions are:
//
//   sigma_ts = (sigma + sigma_s)/2, eps_ts = sqrt(eps eps_s)
//   r_lo = max(x0 - sqrt(g0/alpha), core sigma_ts), r_hi = x0 + sqrt((g0 - 1)/alpha)
//   S(d) = smoothstep((d - r_lo)/delta + 1)
//   g(d) = S(d) max(0, g0 - alpha (d - x0)^2) for d < r_hi, 1 beyond
//   E(d) = c exp(-(d - x0)^2/2) S(d)
//   F(d) = 24 eps_ts/d (2 (sigma_ts/d)^12 - (sigma_ts/d)^6) S(d)
//
// Derivatives are with respect to theta = (sigma, eps, x0, g0, alpha).
#pragma once

#include <hip/hip_runtime.h>

namespace isspa {

// per-type constants, 64 bytes (one scalar load)
struct DiffType {
    float sts, ets;      // sigma_ts, eps_ts
    float x0, g0, al, c; // parabola and E-field amplitude
    float rlo, rhi;      // switch end and parabola end
    float idel;          // 1 / delta
    float eps;           // eps (dF/deps = F / (2 eps))
    float drlo[5];       // d r_lo / d theta
    float pad;
};

struct DiffSwitch {
    float S, dSdz;
};

__device__ __forceinline__ DiffSwitch diff_switch(const DiffType &T, float d)
{
    const float z = (d - T.rlo) * T.idel + 1.0f;
    DiffSwitch o;
    if (z <= 0.0f) {
        o.S = 0.0f;
        o.dSdz = 0.0f;
    } else if (z >= 1.0f) {
        o.S = 1.0f;
        o.dSdz = 0.0f;
    } else {
        o.S = z * z * (3.0f - 2.0f * z);
        o.dSdz = 6.0f * z * (1.0f - z);
    }
    return o;
}

__device__ __forceinline__ float diff_g(const DiffType &T, float d)
{
    if (d >= T.rhi) return 1.0f;
    const float u = d - T.x0;
    const float P = T.g0 - T.al * u * u;
    if (P <= 0.0f) return 0.0f;
    return diff_switch(T, d).S * P;
}

__device__ __forceinline__ float diff_E(const DiffType &T, float d)
{
    const float u = d - T.x0;
    return T.c * expf(-0.5f * u * u) * diff_switch(T, d).S;
}

// the four constants diff_F needs, held in registers by the force pass
struct DiffFParams {
    float sts, ets, rlo, idel;
};

__device__ __forceinline__ DiffFParams diff_f_params(const DiffType &T)
{
    return DiffFParams{T.sts, T.ets, T.rlo, T.idel};
}

__device__ __forceinline__ float diff_F(const DiffFParams &T, float d, float rinv)
{
    const float z = (d - T.rlo) * T.idel + 1.0f;
    const float S = z <= 0.0f ? 0.0f : (z >= 1.0f ? 1.0f : z * z * (3.0f - 2.0f * z));
    const float sr = T.sts * rinv;
    const float sr2 = sr * sr;
    const float sr6 = sr2 * sr2 * sr2;
    return 24.0f * T.ets * rinv * (2.0f * sr6 * sr6 - sr6) * S;
}

__device__ __forceinline__ float diff_F(const DiffType &T, float d, float rinv)
{
    const float sr = T.sts * rinv;
    const float sr2 = sr * sr;
    const float sr6 = sr2 * sr2 * sr2;
    return 24.0f * T.ets * rinv * (2.0f * sr6 * sr6 - sr6) * diff_switch(T, d).S;
}

// values {g, E, F} and derivatives dv[5k + p]
__device__ __forceinline__ void diff_all(const DiffType &T, float d, float rinv, float v[3],
                                         float dv[15])
{
    const DiffSwitch sw = diff_switch(T, d);
    float dS[5];
#pragma unroll
    for (int p = 0; p < 5; ++p) dS[p] = -sw.dSdz * T.idel * T.drlo[p];
#pragma unroll
    for (int k = 0; k < 15; ++k) dv[k] = 0.0f;
    const float u = d - T.x0;
    // g
    if (d >= T.rhi) {
        v[0] = 1.0f;
    } else {
        const float P = T.g0 - T.al * u * u;
        if (P <= 0.0f) {
            v[0] = 0.0f;
        } else {
            v[0] = sw.S * P;
            const float dP[5] = {0.0f, 0.0f, 2.0f * T.al * u, 1.0f, -u * u};
#pragma unroll
            for (int p = 0; p < 5; ++p) dv[p] = dS[p] * P + sw.S * dP[p];
        }
    }
    // E
    const float Er = T.c * expf(-0.5f * u * u);
    v[1] = Er * sw.S;
#pragma unroll
    for (int p = 0; p < 5; ++p) dv[5 + p] = Er * dS[p];
    dv[7] += Er * u * sw.S;
    // F
    const float sr = T.sts * rinv;
    const float sr2 = sr * sr;
    const float sr6 = sr2 * sr2 * sr2;
    const float FLJ = 24.0f * T.ets * rinv * (2.0f * sr6 * sr6 - sr6);
    const float dFds = 24.0f * T.ets * rinv * (24.0f * sr6 * sr6 - 6.0f * sr6) / T.sts;
    v[2] = FLJ * sw.S;
#pragma unroll
    for (int p = 0; p < 5; ++p) dv[10 + p] = FLJ * dS[p];
    dv[10] += 0.5f * dFds * sw.S;
    dv[11] += FLJ / (2.0f * T.eps) * sw.S;
}

}  // namespace isspa
