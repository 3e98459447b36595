// Launch interface between the C-ABI host code (capi.hip) and the gfx950
// kernels.  Every launcher enqueues on `stream` and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace isspa {

// Per-type tables on the device, each type padded with one trailing zero row
// so that a lerp at the last row reads 0 (the reference reads the next type's
// first row there, 0 in every shipped table; quirk Q4).
struct TableView {
    const float *g, *e, *f;
    int n_g, n_e, n_f;     // logical rows
    int ld_g, ld_e, ld_f;  // padded rows per type
    float g_r0, g_dr, g_inv;
    float e_r0, e_dr, e_inv;
    float f_r0, f_dr, f_inv;
};

// Per-point record written by the field epilogue and streamed by the force pass.
// 16 floats = one 64-byte scalar load per point.
//   [0..2]  p      MC point position
//   [3..5]  Ehat   unit mean-field direction
//   [6..8]  V      A1*Ehat + cp*E0/3  (E-hat-like terms of dipole + CDD)
//   [9..11] E0c    cp*E0              (CDD r-coefficient vector)
//   [12]    A1 = p0 <P1> w   [13] A2 = q0 <P2> w   [14] A3 = o0 <P3> w   [15] w = g/cnt
constexpr int kRecFloats = 16;

struct FieldArgs {
    const float4 *pts;
    const float4 *xyzq;
    const int *itype;
    TableView tab;
    int n, P, atoms_per_split;
    float L, h, e0;
    const int *pbc_flag;
    float *part_g;      // [split][P]
    int *part_cnt;      // [split][P]
    double *part_s0;    // [split][3][P]  sum_out q r / d^3
    double *part_ein;   // [split][3][P]  sum_in (-e0 q / d^3 + Etab / d) r
};

struct EpilogueArgs {
    const float4 *pts;
    int P, splits;
    const float *part_g;
    const int *part_cnt;
    const double *part_s0, *part_ein;
    float e0, p0, q0, o0;
    float *rec;         // [P][16]
    double *g_out;      // [P]      (may be null)
    int *cnt_out;       // [P]
    double *E_out;      // [P][3]
    double *E0_out;     // [P][3]
};

struct ForceArgs {
    const float4 *rec4;  // [P][4] float4 view of the records
    const float4 *xyzq;
    const int *itype;
    const float *rmax2;  // per type
    TableView tab;
    int n, P, pts_per_chunk;
    float L, h;
    const int *pbc_flag;
    double *part;        // [chunk][6][N] (S_e, S_lj) or [chunk][9][N] (pol, cdd, lj)
};

struct ReduceArgs {
    const double *part;
    int n, chunks;
    const float4 *xyzq;
    const int *itype;
    const float *vtot;   // per type
    float4 *f;           // += total (may be null)
    double *f_lj, *f_pol, *f_cdd;  // component outputs (split mode)
};

hipError_t launch_mc_points(const float4 *xyzq, const int *itype, const float *rmax,
                            const float *rmax2, int n, int nmc, uint64_t seed, uint32_t sys,
                            uint32_t step, const uint32_t *step_d, float4 *pts, hipStream_t s);
hipError_t launch_field(const FieldArgs &a, int splits, hipStream_t s);
hipError_t launch_epilogue(const EpilogueArgs &a, hipStream_t s);
hipError_t launch_force(const ForceArgs &a, int chunks, bool split, hipStream_t s);
hipError_t launch_force_reduce(const ReduceArgs &a, bool split, hipStream_t s);

// ---- nonbond ----
struct NonbondArgs {
    const float4 *xyzq;
    const float4 *aux;     // per atom j: {4 R_j^2, cA_j, cB_j, LJ type (int bits)}
    const float2 *lj;      // [T*T] {12A, 6B}
    int T, n, atoms_per_chunk;
    const uint32_t *excl;  // [N][4] bit window j - i + 64 in [0,128)
    const int *excl_ptr;   // CSR fallback (null when the windows are complete)
    const int *excl_idx;
    float L, h, cfar;
    const int *pbc_flag;
    double *part;          // [chunk][3][N] or [chunk][9][N] (lj, coul, pair)
};
hipError_t launch_nonbond(const NonbondArgs &a, int chunks, bool split, hipStream_t s);
hipError_t launch_chunk_reduce3(const double *part, int n, int chunks, int comps, float4 *f,
                                double *out0, double *out1, double *out2, hipStream_t s);

// ---- bonded ----
struct BondedArgs {
    const float4 *xyzq;
    float L, h;
    int nb, na, nd;
    const int4 *bond_atoms;
    const float2 *bond_params;
    const int4 *angle_atoms;
    const float2 *angle_params;
    const int4 *dih_atoms;
    const int *dih_types;
    const float4 *dih_params;
    const float2 *sc14;
    const int *ityp;
    const float2 *lj;      // [T*T] {A, B}
    int T;
    float4 *contrib;       // [2 nb + 3 na + 6 nd]
    double *energy;        // [nb + na + nd] x {E} : per-term energies (torsion, 1-4 c, 1-4 lj for dih)
    int terms;
};
hipError_t launch_bonded_terms(const BondedArgs &a, hipStream_t s);
hipError_t launch_bonded_gather(const float4 *contrib, const int *ptr, const int *idx, int n,
                                float4 *f, hipStream_t s);
hipError_t launch_bonded_energy(const double *eterm, int nb, int na, int nd, double *out5,
                                hipStream_t s);

// ---- integrator / restraint / utilities ----
hipError_t launch_leapfrog(int n, float4 *xyzq, float4 *vel, const float4 *f, float kT, float dt,
                           uint32_t thr, float L, uint64_t seed, uint32_t sys, uint32_t step,
                           const uint32_t *step_d, hipStream_t s);
hipError_t launch_us(int nbias, const int2 *atoms, const float *mass, float k0, float k1,
                     float x0, float L, const float4 *xyzq, float4 *f, double *com,
                     hipStream_t s);
hipError_t launch_pbc_flag(const float4 *xyzq, int n, float rmax_all, float L, int *flag,
                           hipStream_t s);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_add4(float4 *dst, const float4 *src, int n, hipStream_t s);
hipError_t launch_step_advance(uint32_t *step_d, hipStream_t s);
hipError_t launch_set_word(uint32_t *dst, uint32_t v, hipStream_t s);

}  // namespace isspa
