// Fixed-order per-atom sums shared by the reduction kernels: the IS-SPA chunk
// reduction, the bonded gather, the all-pairs chunk reduction and the MD
// engine's force assembly, which does all three in one launch.  Each kernel
// that sums a term calls the same helper, so a term's per-atom sum has the
// same operations in the same order whichever kernel computes it.
#pragma once

#include "kernels.hpp"

namespace isspa {

// A 256-thread reduction block owns kSumAtoms atoms and kSumSlices slices of
// each atom's list: slice t takes entries t, t + S, t + 2S, ... in order, and
// slice 0 then adds slices 1..S-1 in order.  The slices give the kernel S
// times the waves of one lane per atom, which small systems need to hide the
// load latency (a 44-atom system has 275 IS-SPA chunk partials per atom).
constexpr int kSumAtoms = 32;
constexpr int kSumSlices = 8;

// Every slice's sums into slice 0, added in slice order (slice 0's value,
// then slices 1, 2, ...).  Component k is added up by the work-items of slice
// k mod kSumSlices, so each chain of LDS reads is kSumSlices long rather than
// (kSumSlices - 1) x NC in slice 0 alone; each sum has the same additions in
// the same order, so the same bits.  Every work-item of the block calls it
// (two barriers).
template <int NC>
__device__ __forceinline__ void combine_slices(double (&red)[kSumSlices][NC][kSumAtoms], int al,
                                               int sl, double (&s)[NC])
{
#pragma unroll
    for (int k = 0; k < NC; ++k) red[sl][k][al] = s[k];
    __syncthreads();
    for (int k = sl; k < NC; k += kSumSlices) {
        double acc = red[0][k][al];
#pragma unroll 4
        for (int t = 1; t < kSumSlices; ++t) acc += red[t][k][al];
        red[0][k][al] = acc;
    }
    __syncthreads();
    if (sl == 0)
#pragma unroll
        for (int k = 0; k < NC; ++k) s[k] = red[0][k][al];
}

// Chunks whose partial loads one slice keeps in flight at a time; the
// batch only groups the loads, the adds stay in chunk order.
constexpr int kSumBatch = 6;

// IS-SPA force partials of atom i (COMPS = 6, or 9 in split mode), chunks of
// slice sl, added in chunk order.
// the COMPS components of chunk-partial entry o (stored as double2 pairs)
template <int COMPS>
__device__ __forceinline__ void load_force_entry(const double *part, long long o,
                                                 double (&v)[COMPS])
{
    constexpr int kStride = force_part_stride(COMPS == 9);
    const double2 *e = reinterpret_cast<const double2 *>(part + o * kStride);
#pragma unroll
    for (int h = 0; h < (COMPS + 1) / 2; ++h) {
        const double2 d = e[h];
        v[2 * h] = d.x;
        if (2 * h + 1 < COMPS) v[2 * h + 1] = d.y;
    }
}

template <int COMPS>
__device__ inline void isspa_slice_sum(const ReduceArgs &a, int i, int sl, double (&s)[COMPS])
{
    const WinDesc W = a.wins[a.atom_win[i]];
    const long long base = W.ppo + (i - W.a0);
    int c = sl;
    for (; c + (kSumBatch - 1) * kSumSlices < W.chunks; c += kSumBatch * kSumSlices) {
        double v[kSumBatch][COMPS];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u)
            load_force_entry<COMPS>(a.part, base + (long long)(c + u * kSumSlices) * W.n, v[u]);
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u)
#pragma unroll
            for (int k = 0; k < COMPS; ++k) s[k] += v[u][k];
    }
    for (; c < W.chunks; c += kSumSlices) {
        double v[COMPS];
        load_force_entry<COMPS>(a.part, base + (long long)c * W.n, v);
#pragma unroll
        for (int k = 0; k < COMPS; ++k) s[k] += v[k];
    }
}

// F_i = -q_i vtot_i S_e - vtot_i S_lj from the summed IS-SPA partials
__device__ inline float3 isspa_atom_force(const ReduceArgs &a, int i, const double s[6])
{
    const float4 x = a.xyzq[i];
    const double vt = (double)a.vtot[a.itype[i]];
    const double qv = -(double)x.w * vt;
    return make_float3((float)(qv * s[0] - vt * s[3]), (float)(qv * s[1] - vt * s[4]),
                       (float)(qv * s[2] - vt * s[5]));
}

// bonded contribution slots of atom i, slots ptr[i] + sl, + kSumSlices, ...
__device__ inline void bonded_slice_sum(const float4 *__restrict__ contrib,
                                        const int *__restrict__ ptr, const int *__restrict__ idx,
                                        int i, int sl, double &sx, double &sy, double &sz)
{
    for (int e = ptr[i] + sl; e < ptr[i + 1]; e += kSumSlices) {
        const float4 c = contrib[idx[e]];
        sx += c.x;
        sy += c.y;
        sz += c.z;
    }
}

// all-pairs partials of atom i (COMPS = 3: the total force, entries of the
// total-force layout; 9: the split components, one entry per partner chunk),
// entries of slice sl, added in entry order
template <int COMPS>
__device__ inline void nonbond_slice_sum(const WinDesc *__restrict__ wins,
                                         const int *__restrict__ atom_win,
                                         const double *__restrict__ part, long long B, int i,
                                         int sl, double (&s)[COMPS])
{
    const WinDesc W = wins[atom_win[i]];
    const long long base = (COMPS == 3 ? W.spo : W.npo) + (i - W.a0);
    const int ne = COMPS == 3 ? W.nte : W.nbc;
    int c = sl;
    for (; c + (kSumBatch - 1) * kSumSlices < ne; c += kSumBatch * kSumSlices) {
        double v[kSumBatch][COMPS];
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u) {
            const long long o = base + (long long)(c + u * kSumSlices) * W.n;
#pragma unroll
            for (int k = 0; k < COMPS; ++k) v[u][k] = part[k * B + o];
        }
#pragma unroll
        for (int u = 0; u < kSumBatch; ++u)
#pragma unroll
            for (int k = 0; k < COMPS; ++k) s[k] += v[u][k];
    }
    for (; c < ne; c += kSumSlices) {
        const long long o = base + (long long)c * W.n;
#pragma unroll
        for (int k = 0; k < COMPS; ++k) s[k] += part[k * B + o];
    }
}

}  // namespace isspa
