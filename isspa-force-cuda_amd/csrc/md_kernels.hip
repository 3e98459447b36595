// Nonbond, bonded, integrator, restraint and utility kernels for gfx950.
//
// Semantics follow the reference kernels cited at each function; the
// decomposition is this engine's own: one lane per receiving atom with a
// wave-uniform partner stream for the all-pairs term, per-term contribution
// slots plus a per-atom gather for the bonded terms (no float atomics), and
// single-workgroup FP64 tree reductions for the umbrella centres of mass.
// The all-pairs, integrator and restraint kernels are window-aware (kernels.hpp).
#include "device_common.hpp"
#include "kernels.hpp"
#include "pbc_device.hpp"
#include "reduce_device.hpp"
#include "points_device.hpp"
#include "nonbond_device.hpp"

namespace isspa {

constexpr int kBlk = 256;
static inline int nblocks(long long n) { return (int)((n + kBlk - 1) / kBlk); }

template <bool SPLIT>
__global__ __launch_bounds__(kBlk) void k_nonbond(NonbondArgs a)
{
    extern __shared__ float2 lds_lj[];
    nonbond_block<SPLIT, !SPLIT>(a, blockIdx.x, lds_lj);
}

hipError_t launch_nonbond(const NonbondArgs &a, int nb, bool split, hipStream_t s)
{
    if (nb == 0) return hipSuccess;
    const size_t lds = nonbond_lds_bytes(a.T, a.sym != 0);
    if (split)
        hipLaunchKernelGGL(k_nonbond<true>, dim3(nb), dim3(kBlk), lds, s, a);
    else
        hipLaunchKernelGGL(k_nonbond<false>, dim3(nb), dim3(kBlk), lds, s, a);
    return hipGetLastError();
}

// sum [comps][B] partials over the window's chunks in order; comps == 3:
// f += sum (or, with f null, o0 += sum in FP64); comps == 9: three FP64 outputs
// sum over the partner chunks in the fixed slice order of nonbond_slice_sum
template <int COMPS>
__global__ __launch_bounds__(kBlk) void k_chunk_reduce3(const WinDesc *__restrict__ wins,
                                                        const int *__restrict__ atom_win,
                                                        const double *__restrict__ part,
                                                        long long B, int ntot, float4 *f,
                                                        double *o0, double *o1, double *o2)
{
    __shared__ double red[kSumSlices][COMPS][kSumAtoms];
    const int al = threadIdx.x % kSumAtoms, sl = threadIdx.x / kSumAtoms;
    const int i = blockIdx.x * kSumAtoms + al;
    const bool valid = i < ntot;
    double s[COMPS];
    for (int k = 0; k < COMPS; ++k) s[k] = 0.0;
    if (valid) nonbond_slice_sum<COMPS>(wins, atom_win, part, B, i, sl, s);
    combine_slices<COMPS>(red, al, sl, s);
    if (sl != 0 || !valid) return;
    if constexpr (COMPS == 3) {
        if (f) {
            float4 v = f[i];
            v.x += (float)s[0];
            v.y += (float)s[1];
            v.z += (float)s[2];
            f[i] = v;
        } else {
            for (int k = 0; k < 3; ++k) o0[3 * i + k] += s[k];
        }
    } else {
        for (int k = 0; k < 3; ++k) {
            o0[3 * i + k] = s[k];
            o1[3 * i + k] = s[3 + k];
            o2[3 * i + k] = s[6 + k];
        }
    }
}

hipError_t launch_chunk_reduce3(const WinDesc *wins, const int *atom_win, const double *part,
                                long long B, int ntot, int comps, float4 *f, double *o0,
                                double *o1, double *o2, hipStream_t s)
{
    if (ntot == 0) return hipSuccess;
    const dim3 grid((ntot + kSumAtoms - 1) / kSumAtoms);
    if (comps == 3)
        hipLaunchKernelGGL(k_chunk_reduce3<3>, grid, dim3(kBlk), 0, s, wins, atom_win, part, B,
                           ntot, f, o0, o1, o2);
    else
        hipLaunchKernelGGL(k_chunk_reduce3<9>, grid, dim3(kBlk), 0, s, wins, atom_win, part, B,
                           ntot, f, o0, o1, o2);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bonded terms.  Each term writes its per-atom forces into fixed contribution
// slots; k_bonded_gather sums each atom's slots in FP64 in a fixed order.
// Slot layout: bonds [0, 2nb), angles [2nb, 2nb+3na), dihedrals 6 per term
// (atoms i, j, k, l, then the 1-4 pair's i and l).
// ---------------------------------------------------------------------------
// The bonded terms are evaluated in FP64.  Near equilibrium the reference's
// FP32 forms cancel: 2k (r0/|r| - 1) loses the digits of r0/|r| that agree
// with 1, and the angle and dihedral denominators lose theirs for nearly
// straight angles, so one bond's FP32 error is ~1e-5 of a force of tens of
// kcal/mol/A that mostly cancels in the atom's total.  These terms are O(N)
// and cost microseconds; in FP64 every contribution is accurate before it is
// stored as FP32.
//
// separation a - b, exact in FP64; the image is decided on the FP32
// difference, as every other kernel decides it
__device__ __forceinline__ double3 mi3d(float4 a, float4 b, float L, float h)
{
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const double kx = dx > h ? 1.0 : (dx < -h ? -1.0 : 0.0);
    const double ky = dy > h ? 1.0 : (dy < -h ? -1.0 : 0.0);
    const double kz = dz > h ? 1.0 : (dz < -h ? -1.0 : 0.0);
    const double Ld = (double)L;
    return make_double3((double)a.x - (double)b.x - kx * Ld, (double)a.y - (double)b.y - ky * Ld,
                        (double)a.z - (double)b.z - kz * Ld);
}

__device__ __forceinline__ double dot3(double3 u, double3 v)
{
    return u.x * v.x + u.y * v.y + u.z * v.z;
}

__device__ __forceinline__ float4 f4_of(double s, double3 r)
{
    return make_float4((float)(s * r.x), (float)(s * r.y), (float)(s * r.z), 0.0f);
}

// src/bond_force_cuda.cu:15-53: F = 2k (r0/|r| - 1) r
__device__ void bond_term(const BondedArgs &a, int q)
{
    const int4 at = a.bond_atoms[q];
    const float2 pr = a.bond_params[at.z];
    const double3 r = mi3d(a.xyzq[at.x], a.xyzq[at.y], a.L, a.h);
    const double d = sqrt(dot3(r, r));
    const double fb = (double)pr.x * ((double)pr.y / d - 1.0);
    a.contrib[2 * q] = f4_of(fb, r);
    a.contrib[2 * q + 1] = f4_of(-fb, r);
    const double dd = d - (double)pr.y;
    a.energy[q] = 0.5 * (double)pr.x * dd * dd;
}

// src/angle_force_cuda.cu:12-64: theta = acos(-c12/sqrt(c11 c22)) (clamped),
// fa = 2k (theta - theta0) / sqrt(c11 c22 - c12^2)
__device__ void angle_term(const BondedArgs &a, int q)
{
    const int4 at = a.angle_atoms[q];
    const double3 r1 = mi3d(a.xyzq[at.x], a.xyzq[at.y], a.L, a.h);
    const double3 r2 = mi3d(a.xyzq[at.y], a.xyzq[at.z], a.L, a.h);
    const double c11 = dot3(r1, r1), c22 = dot3(r2, r2), c12 = dot3(r1, r2);
    const double b = -c12 / sqrt(c11 * c22);
    double th;
    if (b >= 1.0) th = 1.0e-16;
    else if (b <= -1.0) th = (double)3.1415926535f;
    else th = acos(b);
    const float2 pr = a.angle_params[at.w];
    const double dth = th - (double)pr.y;
    const double fa = (double)pr.x * dth / sqrt(c11 * c22 - c12 * c12);
    const double a11 = c12 / c11, a22 = c12 / c22;
    const int base = 2 * a.nb + 3 * q;
    a.contrib[base] = make_float4((float)(fa * (a11 * r1.x - r2.x)), (float)(fa * (a11 * r1.y - r2.y)),
                                  (float)(fa * (a11 * r1.z - r2.z)), 0.0f);
    a.contrib[base + 1] =
        make_float4((float)(fa * ((1.0 + a22) * r2.x - (1.0 + a11) * r1.x)),
                    (float)(fa * ((1.0 + a22) * r2.y - (1.0 + a11) * r1.y)),
                    (float)(fa * ((1.0 + a22) * r2.z - (1.0 + a11) * r1.z)), 0.0f);
    a.contrib[base + 2] = make_float4((float)(fa * (r1.x - a22 * r2.x)), (float)(fa * (r1.y - a22 * r2.y)),
                                      (float)(fa * (r1.z - a22 * r2.z)), 0.0f);
    a.energy[a.nb + q] = 0.5 * (double)pr.x * dth * dth;
}

// src/dih_force_cuda.cu:13-136: n V sin(n phi - phase) c22 / (sin phi b) and
// the scaled 1-4 pair
__device__ void dih_term(const BondedArgs &a, int q)
{
    int4 at = a.dih_atoms[q];
    const int ty = a.dih_types[q];
    const float4 p1 = a.xyzq[at.x], p2 = a.xyzq[at.y];
    const float4 p3 = a.xyzq[abs(at.z)], p4 = a.xyzq[abs(at.w)];
    const int base = 2 * a.nb + 3 * a.na + 6 * q;
    double e14c = 0.0, e14v = 0.0;
    if (at.z > 0 && at.w > 0) {  // scaled 1-4 (index 0 never qualifies, quirk Q20 mirrored)
        const float2 lj = a.lj[a.ityp[at.x] * a.T + a.ityp[at.w]];
        const float2 sc = a.sc14[ty];
        const double3 r = mi3d(p1, p4, a.L, a.h);
        const double rm = dot3(r, r);
        const double r6 = 1.0 / (rm * rm * rm);
        const double qq = (double)p1.w * (double)p4.w;
        const double fe = qq / (rm * sqrt(rm) * (double)sc.x);
        const double fv = r6 * (12.0 * (double)lj.x * r6 - 6.0 * (double)lj.y) / ((double)sc.y * rm);
        a.contrib[base + 4] = f4_of(fe + fv, r);
        a.contrib[base + 5] = f4_of(-(fe + fv), r);
        e14c = qq / (sqrt(rm) * (double)sc.x);
        e14v = ((double)lj.x * r6 * r6 - (double)lj.y * r6) / (double)sc.y;
    } else {
        a.contrib[base + 4] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.contrib[base + 5] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const double3 r12 = mi3d(p1, p2, a.L, a.h);
    const double3 r23 = mi3d(p2, p3, a.L, a.h);
    const double3 r34 = mi3d(p3, p4, a.L, a.h);
    const double c11 = dot3(r12, r12), c22 = dot3(r23, r23), c12 = dot3(r12, r23);
    const double c33 = dot3(r34, r34), c23 = dot3(r23, r34), c13 = dot3(r12, r34);
    const double t1 = c13 * c22 - c12 * c23;
    const double t2 = c11 * c23 - c12 * c13;
    const double t3 = c12 * c12 - c11 * c22;
    const double t4 = c22 * c33 - c23 * c23;
    const double t5 = c13 * c23 - c12 * c33;
    const double t6 = -t1;
    const double b = sqrt(-t3 * t4);
    const double aa = t6 / b;
    const float4 pr = a.dih_params[ty];
    const double pn = (double)pr.x, V = (double)pr.y, ph = (double)pr.z;
    double fd = 0.0, phi;
    if (aa <= -1.0) {
        phi = (double)3.1415926535f;
    } else if (aa >= 1.0) {
        phi = 0.0;
    } else {
        phi = acos(aa);
        fd = pn * V * sin(pn * phi - ph) * c22 / (sin(phi) * b);
    }
    const double g1 = fd / t3, g4 = -fd / t4;
    const double3 f1 = make_double3(g1 * (t1 * r12.x + t2 * r23.x + t3 * r34.x),
                                    g1 * (t1 * r12.y + t2 * r23.y + t3 * r34.y),
                                    g1 * (t1 * r12.z + t2 * r23.z + t3 * r34.z));
    const double3 f4 = make_double3(g4 * (t4 * r12.x + t5 * r23.x + t6 * r34.x),
                                    g4 * (t4 * r12.y + t5 * r23.y + t6 * r34.y),
                                    g4 * (t4 * r12.z + t5 * r23.z + t6 * r34.z));
    const double a12 = c12 / c22, a23 = c23 / c22;
    a.contrib[base] = f4_of(1.0, f1);
    a.contrib[base + 1] = make_float4((float)(-(1.0 + a12) * f1.x + a23 * f4.x),
                                      (float)(-(1.0 + a12) * f1.y + a23 * f4.y),
                                      (float)(-(1.0 + a12) * f1.z + a23 * f4.z), 0.f);
    a.contrib[base + 2] = make_float4((float)(a12 * f1.x - (1.0 + a23) * f4.x),
                                      (float)(a12 * f1.y - (1.0 + a23) * f4.y),
                                      (float)(a12 * f1.z - (1.0 + a23) * f4.z), 0.f);
    a.contrib[base + 3] = f4_of(1.0, f4);
    const size_t eb = (size_t)a.nb + a.na + 3 * (size_t)q;
    a.energy[eb] = V * (1.0 + cos(pn * phi - ph));
    a.energy[eb + 1] = e14c;
    a.energy[eb + 2] = e14v;
}

__device__ __forceinline__ void bonded_one(const BondedArgs &a, int q)
{
    if (q < a.nb) {
        if (a.terms & 1) {
            bond_term(a, q);
        } else {
            a.contrib[2 * q] = a.contrib[2 * q + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.energy[q] = 0.0;
        }
    } else if (q < a.nb + a.na) {
        const int k = q - a.nb;
        if (a.terms & 2) {
            angle_term(a, k);
        } else {
            const int base = 2 * a.nb + 3 * k;
            for (int s = 0; s < 3; ++s) a.contrib[base + s] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.energy[a.nb + k] = 0.0;
        }
    } else if (q < a.nb + a.na + a.nd) {
        const int k = q - a.nb - a.na;
        if (a.terms & 4) {
            dih_term(a, k);
        } else {
            const int base = 2 * a.nb + 3 * a.na + 6 * k;
            for (int s = 0; s < 6; ++s) a.contrib[base + s] = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int s = 0; s < 3; ++s) a.energy[(size_t)a.nb + a.na + 3 * (size_t)k + s] = 0.0;
        }
    }
}

__global__ __launch_bounds__(kBlk) void k_bonded_terms(BondedArgs a)
{
    bonded_one(a, blockIdx.x * kBlk + threadIdx.x);
}

hipError_t launch_bonded_terms(const BondedArgs &a, hipStream_t s)
{
    const long n = (long)a.nb + a.na + a.nd;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bonded_terms, dim3(nblocks(n)), dim3(kBlk), 0, s, a);
    return hipGetLastError();
}

// A block owns kSumAtoms atoms and kSumSlices slot slices: slice t sums slots
// t, t + S, ... of its atom in order, then the slices are added in order (a
// fixed order independent of scheduling, with S times the waves of one lane
// per atom).
__global__ __launch_bounds__(kBlk) void k_bonded_gather(const float4 *__restrict__ contrib,
                                                        const int *__restrict__ ptr,
                                                        const int *__restrict__ idx, int n,
                                                        float4 *f)
{
    __shared__ double red[kSumSlices][3][kSumAtoms];
    const int al = threadIdx.x % kSumAtoms, sl = threadIdx.x / kSumAtoms;
    const int i = blockIdx.x * kSumAtoms + al;
    const bool valid = i < n;
    double s[3] = {0.0, 0.0, 0.0};
    if (valid) bonded_slice_sum(contrib, ptr, idx, i, sl, s[0], s[1], s[2]);
    combine_slices<3>(red, al, sl, s);
    if (sl != 0 || !valid) return;
    float4 v = f[i];
    v.x += (float)s[0];
    v.y += (float)s[1];
    v.z += (float)s[2];
    f[i] = v;
}

hipError_t launch_bonded_gather(const float4 *contrib, const int *ptr, const int *idx, int n,
                                float4 *f, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bonded_gather, dim3((n + kSumAtoms - 1) / kSumAtoms), dim3(kBlk), 0,
                       s, contrib, ptr, idx, n, f);
    return hipGetLastError();
}

// energies: bond, angle, torsion, 1-4 Coulomb, 1-4 LJ (one workgroup, fixed tree)
__global__ __launch_bounds__(kBlk) void k_bonded_energy(const double *__restrict__ e, int nb,
                                                        int na, int nd, double *out)
{
    __shared__ double red[5][kBlk];
    double acc[5] = {0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < nb; q += kBlk) acc[0] += e[q];
    for (int q = threadIdx.x; q < na; q += kBlk) acc[1] += e[nb + q];
    for (int q = threadIdx.x; q < nd; q += kBlk)
        for (int s = 0; s < 3; ++s) acc[2 + s] += e[(size_t)nb + na + 3 * (size_t)q + s];
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) out[k] = red[k][0];
}

hipError_t launch_bonded_energy(const double *eterm, int nb, int na, int nd, double *out5,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_bonded_energy, dim3(1), dim3(kBlk), 0, s, eterm, nb, na, nd, out5);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// A.15 leap-frog + Andersen (src/leapfrog_cuda.cu:33-67).  On a collision the
// velocity is redrawn from Maxwell-Boltzmann, then half-kicked (quirk Q21).
// Atom i of window w draws with counter {i - a0_w, draw, step, sys_w}.
// ---------------------------------------------------------------------------
// advance = {step, blocks done}, called by every thread of every block after
// its last read of the step: the last block to get here moves the step on and
// clears the count.  Every block read the step before its count, so no block
// of this launch sees the new value.
__device__ __forceinline__ void advance_step(uint32_t *advance, uint32_t step)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(advance + 1, 1u);
        if (done == gridDim.x - 1) {
            advance[0] = step + 1u;
            atomicExch(advance + 1, 0u);
            __threadfence();
        }
    }
}

// atom i of window W
__device__ __forceinline__ void leapfrog_atom(const WinDesc W, int i, float4 *xyzq, float4 *vel,
                                              const float4 fi, float kT, float dt, uint32_t thr,
                                              uint64_t seed, uint32_t step)
{
    const uint32_t il = (uint32_t)(i - W.a0);
    const uint4 o = philox_draw(seed, il, kStreamAndersen, 0u, step, W.sys);
    float4 v = vel[i];
    float4 x = xyzq[i];
    const float m = v.w;
    if (o.x < thr) {
        const uint4 z = philox_draw(seed, il, kStreamAndersen, 1u, step, W.sys);
        const float s24 = 5.9604644775390625e-08f;
        const float u1a = (float)((z.x >> 8) + 1u) * s24, u2a = (float)(z.y >> 8) * s24;
        const float u1b = (float)((z.z >> 8) + 1u) * s24, u2b = (float)(z.w >> 8) * s24;
        const float ra = sqrtf(-2.0f * logf(u1a)), rb = sqrtf(-2.0f * logf(u1b));
        const float twopi = 6.283185307179586f;
        const float sc = sqrtf(kT / m);
        const float hk = dt / (m * 2.0f);
        v.x = ra * cosf(twopi * u2a) * sc + fi.x * hk;
        v.y = ra * sinf(twopi * u2a) * sc + fi.y * hk;
        v.z = rb * cosf(twopi * u2b) * sc + fi.z * hk;
    } else {
        const float kk = dt / m;
        v.x += fi.x * kk;
        v.y += fi.y * kk;
        v.z += fi.z * kk;
    }
    x.x += v.x * dt;
    x.y += v.y * dt;
    x.z += v.z * dt;
    // wrap (src/cuda_vector_routines.h:49-68)
    const float L = W.L;
    x.x = x.x > L ? x.x - L : (x.x < 0.0f ? x.x + L : x.x);
    x.y = x.y > L ? x.y - L : (x.y < 0.0f ? x.y + L : x.y);
    x.z = x.z > L ? x.z - L : (x.z < 0.0f ? x.z + L : x.z);
    vel[i] = v;
    xyzq[i] = x;
}

__global__ __launch_bounds__(kBlk) void k_leapfrog(const WinDesc *__restrict__ wins,
                                                   const int *__restrict__ atom_win, WinDesc one,
                                                   int ntot,
                                                   float4 *xyzq, float4 *vel,
                                                   const float4 *__restrict__ f, float kT,
                                                   float dt, uint32_t thr, uint64_t seed,
                                                   uint32_t step_arg,
                                                   const uint32_t *step_d, uint32_t *advance)
{
    const int i = blockIdx.x * kBlk + threadIdx.x;
    const uint32_t step = step_d ? *step_d : step_arg;
    if (i < ntot)
        leapfrog_atom(atom_win ? wins[atom_win[i]] : one, i, xyzq, vel, f[i], kT, dt, thr, seed,
                      step);
    if (advance) advance_step(advance, step);
}

hipError_t launch_leapfrog(const WinDesc *wins, const int *atom_win, const WinDesc &one,
                           int ntot, float4 *xyzq, float4 *vel, const float4 *f, float kT,
                           float dt, uint32_t thr, uint64_t seed, uint32_t step,
                           const uint32_t *step_d, hipStream_t s, uint32_t *advance)
{
    if (ntot == 0) return advance ? launch_step_advance(advance, s) : hipSuccess;
    hipLaunchKernelGGL(k_leapfrog, dim3(nblocks(ntot)), dim3(kBlk), 0, s, wins, atom_win, one, ntot,
                       xyzq, vel, f, kT, dt, thr, seed, step, step_d, advance);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The MD engine's per-step force: f = ((bonded + IS-SPA) + all-pairs) per atom,
// plus the restraint force of restrained atoms.  Each term is summed by the
// same helper and in the same order as its own reduction kernel and added in
// the order the separate calls add them into a zeroed f, so the result equals
// zero + gather + IS-SPA reduction + chunk reduction + restraint bit for bit,
// in one launch.  With
// lf.xyzq the same thread then integrates its atom (leap-frog + Andersen, as
// k_leapfrog) and the last block advances the step, so a whole step ends with
// this launch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlk) void k_assemble_forces(AssembleArgs a)
{
    constexpr int nc = 12;  // bonded xyz, IS-SPA 6, all-pairs xyz
    __shared__ double red[kSumSlices][nc][kSumAtoms];
    const int al = threadIdx.x % kSumAtoms, sl = threadIdx.x / kSumAtoms;
    const int i = blockIdx.x * kSumAtoms + al;
    const bool valid = i < a.n;
    const uint32_t step = a.lf.step_d ? *a.lf.step_d : 0u;
    double s[nc];
    for (int k = 0; k < nc; ++k) s[k] = 0.0;
    if (valid) {
        if (a.contrib) {
            double bx = 0.0, by = 0.0, bz = 0.0;
            bonded_slice_sum(a.contrib, a.gptr, a.gidx, i, sl, bx, by, bz);
            s[0] = bx;
            s[1] = by;
            s[2] = bz;
        }
        if (a.isspa.part) {
            double t[6] = {0, 0, 0, 0, 0, 0};
            isspa_slice_sum<6>(a.isspa, i, sl, t);
            for (int k = 0; k < 6; ++k) s[3 + k] = t[k];
        }
        if (a.nb_part) {
            double t[3] = {0, 0, 0};
            nonbond_slice_sum<3>(a.nb_wins, a.nb_atom_win, a.nb_part, a.nb_B, i, sl, t);
            for (int k = 0; k < 3; ++k) s[9 + k] = t[k];
        }
    }
    combine_slices<nc>(red, al, sl, s);
    if (sl == 0 && valid) {
        float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (a.contrib) {
            f.x += (float)s[0];
            f.y += (float)s[1];
            f.z += (float)s[2];
        }
        if (a.isspa.part) {
            const double si[6] = {s[3], s[4], s[5], s[6], s[7], s[8]};
            const float3 fi = isspa_atom_force(a.isspa, i, si);
            f.x += fi.x;
            f.y += fi.y;
            f.z += fi.z;
        }
        if (a.nb_part) {
            f.x += (float)s[9];
            f.y += (float)s[10];
            f.z += (float)s[11];
        }
        if (a.fus) {
            const float4 u = a.fus[i];
            if (u.w != 0.0f) {
                f.x += u.x;
                f.y += u.y;
                f.z += u.z;
            }
        }
        a.f[i] = f;
        if (a.lf.xyzq)
            leapfrog_atom(a.lf.wins[a.lf.atom_win[i]], i, a.lf.xyzq, a.lf.vel, f, a.lf.kT, a.lf.dt,
                          a.lf.thr, a.lf.seed, step);
    }
    if (a.lf.advance) advance_step(a.lf.advance, step);
}

hipError_t launch_assemble_forces(const AssembleArgs &a, hipStream_t s)
{
    if (a.n == 0) return a.lf.advance ? launch_step_advance(a.lf.advance, s) : hipSuccess;
    hipLaunchKernelGGL(k_assemble_forces, dim3((a.n + kSumAtoms - 1) / kSumAtoms), dim3(kBlk), 0,
                       s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// A.16 umbrella restraint (src/us_force_cuda.cu:14-47).  One workgroup per
// restraint: the two centres of mass by an FP64 tree reduction (the
// reference synchronises only within a block, quirk Q12), then each bias
// atom's force.  Atoms in a restraint are unique, so the updates do not race.
// ---------------------------------------------------------------------------
// restraint r by one block; the bias force of each restrained atom is added
// into f, or with `store` written to f (w = 1 marking the atom)
__device__ __forceinline__ void us_block(int r, const int *__restrict__ boff,
                                         const int2 *__restrict__ atoms,
                                         const float *__restrict__ mass,
                                         const float2 *__restrict__ kx,
                                         const float *__restrict__ Ls,
                                         const float4 *__restrict__ xyzq, float4 *f, double *com,
                                         bool store)
{
    __shared__ double red[6][kBlk];
    const int b0 = boff[r], b1 = boff[r + 1];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int b = b0 + threadIdx.x; b < b1; b += kBlk) {
        const int2 at = atoms[b];
        const float4 x = xyzq[at.x];
        const double m = mass[b];
        acc[3 * at.y + 0] += m * x.x;
        acc[3 * at.y + 1] += m * x.y;
        acc[3 * at.y + 2] += m * x.z;
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    const double L = Ls[r];
    const double h = 0.5 * L;
    double dv[3];
    for (int k = 0; k < 3; ++k) {
        double d = red[k][0] - red[3 + k][0];
        if (d > h) d -= L;
        else if (d < -h) d += L;
        dv[k] = d;
    }
    const double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    const float2 p = kx[r];  // {k, x0}: group 0 pulls with -k, group 1 with +k
    for (int b = b0 + threadIdx.x; b < b1; b += kBlk) {
        const int2 at = atoms[b];
        const double kk = at.y == 0 ? -(double)p.x : (double)p.x;
        const double pre = kk * (d - (double)p.y) * (double)mass[b] / d;
        const float fx = (float)(pre * dv[0]), fy = (float)(pre * dv[1]),
                    fz = (float)(pre * dv[2]);
        if (store) {
            f[at.x] = make_float4(fx, fy, fz, 1.0f);
        } else {
            float4 v = f[at.x];
            v.x += fx;
            v.y += fy;
            v.z += fz;
            f[at.x] = v;
        }
    }
    if (com && threadIdx.x < 6) com[6 * r + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlk) void k_us(const int *__restrict__ boff,
                                             const int2 *__restrict__ atoms,
                                             const float *__restrict__ mass,
                                             const float2 *__restrict__ kx,
                                             const float *__restrict__ Ls,
                                             const float4 *__restrict__ xyzq, float4 *f,
                                             double *com)
{
    us_block(blockIdx.x, boff, atoms, mass, kx, Ls, xyzq, f, com, false);
}

hipError_t launch_us(int nres, const int *boff, const int2 *atoms, const float *mass,
                     const float2 *kx, const float *L, const float4 *xyzq, float4 *f, double *com,
                     hipStream_t s)
{
    if (nres == 0) return hipSuccess;
    hipLaunchKernelGGL(k_us, dim3(nres), dim3(kBlk), 0, s, boff, atoms, mass, kx, L, xyzq, f, com);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// MD engine step prologue: the launches of a step that depend only on the
// coordinates, as block ranges of one grid -- minimum-image flags (one block
// per window), bonded terms (256 terms a block), restraints (one block each,
// bias forces stored for the force assembly) and the MC points.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlk) void k_step_prologue(PrologueArgs a)
{
    int b = blockIdx.x;
    if (b < a.nflag) {
        const WinDesc W = a.pts.wins[b];
        pbc_flag_block(W, a.pts.xyzq, a.rmax_all, a.flag + b);
        return;
    }
    b -= a.nflag;
    if (b < a.bonded_blocks) {
        bonded_one(a.bonded, b * kBlk + (int)threadIdx.x);
        return;
    }
    b -= a.bonded_blocks;
    if (b < a.nres) {
        us_block(b, a.boff, a.us_atoms, a.us_mass, a.kx, a.Ls, a.pts.xyzq, a.fus, a.com, true);
        return;
    }
    b -= a.nres;
    const long long gp = (long long)b * kBlk + threadIdx.x;
    if (gp < (long long)a.pts.ntot * a.pts.nmc) mc_point(a.pts, gp);
}

hipError_t launch_step_prologue(PrologueArgs a, hipStream_t s)
{
    const long nterms = (long)a.bonded.nb + a.bonded.na + a.bonded.nd;
    a.bonded_blocks = a.bonded.contrib ? nblocks(nterms) : 0;
    const long long P = (long long)a.pts.ntot * a.pts.nmc;
    const long long grid = (long long)a.nflag + a.bonded_blocks + a.nres + (P + kBlk - 1) / kBlk;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_step_prologue, dim3((unsigned)grid), dim3(kBlk), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// utilities
// ---------------------------------------------------------------------------
// flag[w] for every window w: one block per window
__global__ __launch_bounds__(256) void k_pbc_flag(const WinDesc *__restrict__ wins,
                                                  const float4 *__restrict__ xyzq,
                                                  float rmax_all, int *flag)
{
    const WinDesc W = wins[blockIdx.x];
    pbc_flag_block(W, xyzq, rmax_all, flag + blockIdx.x);
}

hipError_t launch_pbc_flag(const WinDesc *wins, int W, const float4 *xyzq, float rmax_all,
                           int *flag, hipStream_t s)
{
    if (W == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pbc_flag, dim3(W), dim3(256), 0, s, wins, xyzq, rmax_all, flag);
    return hipGetLastError();
}

__global__ void k_zero16(uint4 *p, size_t n16)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n16;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = make_uint4(0u, 0u, 0u, 0u);
}
__global__ void k_zero1(unsigned char *p, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0;
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return hipSuccess;
    if (((uintptr_t)p & 15u) == 0 && (bytes & 15u) == 0) {
        const size_t n16 = bytes / 16;
        const int g = (int)std::min<size_t>(4096, (n16 + kBlk - 1) / kBlk);
        hipLaunchKernelGGL(k_zero16, dim3(g), dim3(kBlk), 0, s, (uint4 *)p, n16);
    } else {
        const int g = (int)std::min<size_t>(4096, (bytes + kBlk - 1) / kBlk);
        hipLaunchKernelGGL(k_zero1, dim3(g), dim3(kBlk), 0, s, (unsigned char *)p, bytes);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlk) void k_add4(float4 *dst, const float4 *__restrict__ src, int n)
{
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    float4 d = dst[i];
    const float4 s = src[i];
    d.x += s.x;
    d.y += s.y;
    d.z += s.z;
    dst[i] = d;
}

hipError_t launch_add4(float4 *dst, const float4 *src, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_add4, dim3(nblocks(n)), dim3(kBlk), 0, s, dst, src, n);
    return hipGetLastError();
}

__global__ void k_step_advance(uint32_t *step) { *step += 1u; }
__global__ void k_set_word(uint32_t *dst, uint32_t v) { *dst = v; }

hipError_t launch_step_advance(uint32_t *step_d, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, s, step_d);
    return hipGetLastError();
}

hipError_t launch_set_word(uint32_t *dst, uint32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, s, dst, v);
    return hipGetLastError();
}

}  // namespace isspa
