// Launch interface between the C-ABI host code (capi_*.hip) and the gfx950
// kernels.  Every launcher enqueues on `stream` and returns the launch status.
//
// All kernels are window-aware: a launch covers W independent systems
// ("windows", e.g. umbrella-sampling windows of uneven size) whose atoms are
// concatenated.  Window w owns atoms [a0, a0 + n) and MC points
// [a0 nMC, (a0 + n) nMC).  A single system is the case W = 1.  Grids are 1-D
// over a per-block work table (int4 {window, block, split/chunk, -}), so
// uneven windows share one launch without idle blocks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "diff_device.hpp"

namespace isspa {

struct WinDesc {
    int a0, n;          // atoms of the window
    int splits, aps;    // field pass: atom splits, atoms per split
    int chunks, ppc;    // force pass: point chunks, points per chunk
    int nbc, apc;       // nonbond: partner chunks, atoms per chunk
    long long fpo;      // field partial offset:  sum over earlier windows of splits * P
    long long ppo;      // force partial offset:  sum of chunks * n
    long long npo;      // nonbond partial offset: sum of nbc * n
    uint32_t sys;       // Philox counter word 3
    float L, h;         // box length and half length
    int nte;            // nonbond total force: entries per atom (nbc, or the
                        // number of 256-atom blocks when pairs are taken once)
    int fg;             // field pass: splits per block (1, 2 or 4; divides splits);
                        // the block's waves take fg splits of 256 / fg points
                        // and leave one partial per group of fg splits
    int fgpad;
    long long spo;      // nonbond total-force partial offset: sum of nte * n
};

// splits per field block for a window of `splits` splits
inline int field_group(int splits) { return splits % 4 == 0 ? 4 : (splits % 2 == 0 ? 2 : 1); }

// Per-type tables on the device, each type padded with one trailing zero row
// so that a lerp at the last row reads 0 (the reference reads the next type's
// first row there, 0 in every shipped table; quirk Q4).  All windows share the
// grid; their types are concatenated.
struct TableView {
    const float *g, *e, *f;
    int n_g, n_e, n_f;     // logical rows
    int ld_g, ld_e, ld_f;  // padded rows per type
    int ntypes;
    float g_r0, g_dr, g_inv;
    float e_r0, e_dr, e_inv;
    float f_r0, f_dr, f_inv;
};

// Per-point record written by the field epilogue and streamed by the force pass.
// 16 floats = one 64-byte scalar load per point.
//   [0..2]  p      MC point position
//   [3..5]  Ehat   unit mean-field direction
//   [6..8]  V      A1*Ehat + cp*E0/3  (E-hat-like terms of dipole + CDD)
//   [9..11] D      3 A1 Ehat + cp*E0  (r-coefficient vector of dipole + CDD)
//   [12] A2 = q0 <P2> w   [13] 3 A2   [14] A3 = o0 <P3> w   [15] w = g/cnt
// with A1 = p0 <P1> w.  A second array of P float4 {cp*E0, A1} (read only by
// the in-sphere branch, the component split and the adjoint) follows the
// records in the same allocation.
constexpr int kRecFloats = 16;
constexpr int kAuxFloats = 4;

// all-pairs blocks: 256 receivers; kinds of the block table's .w
constexpr int kNbBlock = 256;
constexpr int kNbKindChunk = 0;  // receivers x a partner chunk, every pair in both orders
constexpr int kNbKindNear = 1;   // receiver block x itself or a neighbouring block, ordered
constexpr int kNbKindSym = 2;    // super-tile: receiver block x partner block, each pair once
constexpr int kSymMinAtoms = 4096;            // windows that take each pair once
constexpr int kSymLdsDoubles = 3 * kNbBlock;  // a super-tile's partner sums in LDS

struct NonbondArgs {
    const WinDesc *wins;
    const int4 *blocks;    // {w, receiver block, partner chunk or block, kind}
    const float4 *xyzq;
    const float4 *aux;     // per atom j: {4 R_j^2, cA_j, cB_j, LJ type (int bits)}
    const float2 *lj;      // [T*T] {12A, 6B}
    int T;
    const uint32_t *excl;  // [N][4] bit window j - i + 64 in [0,128)
    const int *excl_ptr;   // CSR fallback (null when the windows are complete)
    const int *excl_idx;
    float cfar;
    const int *pbc_flag;
    double *part;          // [comps][B], in the layout of the launch's mode
    long long B;
    int sym;               // the block table holds super-tiles
};

// dynamic LDS of an all-pairs block: the T x T LJ table, then (super-tiles)
// the partner sums on an 8-byte boundary
inline size_t nonbond_lds_bytes(int T, bool sym)
{
    return sym ? sizeof(float) * 2 * (((size_t)T * T + 1) & ~(size_t)1) + sizeof(double) * kSymLdsDoubles
               : sizeof(float) * 2 * (size_t)T * T;
}

struct FieldArgs {
    const WinDesc *wins;
    const int4 *blocks;   // {w, point block, split, -}
    const float4 *pts;
    const float4 *xyzq;
    const int *itype;
    TableView tab;
    const DiffType *dtypes;  // non-null: analytic functions instead of tables
    int nmc;
    float e0;
    const int *pbc_flag;  // per window
    double *part_g;       // [F] product of the in-sphere g factors
    int *part_cnt;        // [F]
    double *part_s0;      // [3][F]  sum_out q r / d^3
    double *part_ein;     // [3][F]  sum_in (-e0 q / d^3 + Etab / d) r
    long long F;          // total field partial length
    // all-pairs partials computed by nb_blocks extra blocks of the same launch
    // (tables only, nb_blocks = 0: none); nfield is set by the launcher
    NonbondArgs nb;
    int nb_blocks, nfield;
    int max_fg;  // the largest W.fg of the launch (its LDS holds the split exchange if > 1)
};

struct EpilogueArgs {
    const WinDesc *wins;
    int nwin;
    const int *atom_win;
    const float4 *pts;
    long long Ptot;
    int nmc;
    const double *part_g;
    const int *part_cnt;
    const double *part_s0, *part_ein;
    long long F;
    float e0, p0, q0, o0;
    float *rec;         // [P][16]
    float4 *aux;        // [P] {cp E0, A1}
    double *g_out;      // [P]      inspection outputs, each may be null
    int *cnt_out;       // [P]
    double *E_out;      // [P][3]
    double *E0_out;     // [P][3]
};

struct ForceArgs {
    const WinDesc *wins;
    const int4 *blocks;   // {w, receiver block, chunk, -}
    const float4 *rec4;   // [P][4] float4 view of the records
    const float4 *aux4;   // [P] {cp E0, A1}
    const float4 *xyzq;
    const int *itype;
    const float *rmax2;   // per type
    TableView tab;
    const DiffType *dtypes;  // non-null: analytic F instead of the table
    int nmc;
    const int *pbc_flag;
    double *part;         // [C][stride]: (S_e, S_lj), or (pol, lj, cdd, pad) split
    long long C;
    int xcd_runs;         // 1: each XCD takes a contiguous run of the table (equal work items)
    int replicas;         // most receiver replicas of any window (force_replicas of the smallest)
    int red_off;          // float offset of the replica-sum area in the dynamic LDS (set at launch)
};

// Windows of at most 128 atoms fill only one or two of a 256-thread force
// block's four waves with receivers.  The other waves then take a share of
// the chunk's points for the same receivers (4 replicas of the receivers at
// <= 64 atoms, 2 at <= 128), and the block adds the replicas' sums in replica
// order before it writes the chunk's partial: the same parallelism with a
// quarter (or half) of the chunks, so a quarter of the partials to reduce.
__host__ __device__ constexpr int force_replicas(int n) { return n <= 64 ? 4 : (n <= 128 ? 2 : 1); }
// doubles of the replica-sum area: replicas 1..r-1 of 64 * (4 / r) receivers
__host__ __device__ constexpr int force_replica_doubles(bool split)
{
    return 3 * 64 * (split ? 9 : 6);
}

// Workgroup b of a launch goes to XCD b % 8 (round-robin).  With xcd_runs the
// table entry of workgroup b is the (b / 8)-th of XCD b % 8's contiguous run,
// so the receiver blocks of one chunk share an L2 instead of being dealt to
// all eight.  A bijection on [0, nb) for any nb.
__device__ __forceinline__ int xcd_run_entry(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// The force pass's chunk partials are stored entry by entry, entry
// (chunk, atom) holding its components together: 6 doubles (S_e, S_lj), or
// 10 in split mode (pol, lj, cdd and a pad, so every entry stays 16-byte
// aligned), read and written as double2.
__host__ __device__ constexpr int force_part_stride(bool split) { return split ? 10 : 6; }

struct ReduceArgs {
    const WinDesc *wins;
    const int *atom_win;
    const double *part;
    long long C;
    int ntot;
    const float4 *xyzq;
    const int *itype;
    const float *vtot;    // per type
    float4 *f;            // += total (may be null)
    float4 *f_alone;      // = total, the IS-SPA force kept apart (may be null)
    double *f_lj, *f_pol, *f_cdd;  // component outputs (split mode)
};

// MC points: one per (atom, draw) of every window
struct PointsArgs {
    const WinDesc *wins;
    const int *atom_win;
    const float4 *xyzq;
    const int *itype;
    const float *rmax, *rmax2;
    int ntot, nmc;
    uint64_t seed;
    uint32_t step;            // used when step_d is null
    const uint32_t *step_d;
    float4 *pts;
};

hipError_t launch_mc_points(const WinDesc *wins, const int *atom_win, const float4 *xyzq,
                            const int *itype, const float *rmax, const float *rmax2, int ntot,
                            int nmc, uint64_t seed, uint32_t step, const uint32_t *step_d,
                            float4 *pts, hipStream_t s, int nwin = 0, float rmax_all = 0.0f,
                            int *flag = nullptr);
hipError_t launch_field(const FieldArgs &a, int nblocks, hipStream_t s);
hipError_t launch_epilogue(const EpilogueArgs &a, hipStream_t s);
// the epilogue and the force pass in one launch, for models whose windows all
// fit one receiver block, whose chunks hold at most 256 points and whose force
// grid has at most 256 blocks (table model, total force only)
bool force_epi_fits(const TableView &t, int max_window_atoms, int max_chunk_points, int nblocks);
hipError_t launch_force_epi(const ForceArgs &a, const EpilogueArgs &ea, int nblocks, hipStream_t s);
hipError_t launch_force(const ForceArgs &a, int nblocks, bool split, hipStream_t s);
hipError_t launch_force_reduce(const ReduceArgs &a, bool split, hipStream_t s);

// One step's force of the MD engine in one launch: f = bonded gather + IS-SPA
// reduction + all-pairs chunk reduction (each optional: null contrib, null
// isspa.part, null nb_part), overwriting f (w = 0).
struct AssembleArgs {
    int n;
    float4 *f;
    const float4 *contrib;  // bonded contribution slots
    const int *gptr, *gidx;
    ReduceArgs isspa;       // IS-SPA chunk partials (isspa.f unused)
    const WinDesc *nb_wins; // all-pairs chunk partials
    const int *nb_atom_win;
    const double *nb_part;
    long long nb_B;
    const float4 *fus;      // restraint bias forces (w = 1 marks a restrained atom), added last
    struct {                // leap-frog + Andersen of the assembled atoms (xyzq null: none)
        const WinDesc *wins;    // integrator windows, both required with xyzq
        const int *atom_win;
        float4 *xyzq, *vel;
        float kT, dt;
        uint32_t thr;
        uint64_t seed;
        const uint32_t *step_d;
        uint32_t *advance;  // {step, blocks done}: the last block moves the step on
    } lf;
};
hipError_t launch_assemble_forces(const AssembleArgs &a, hipStream_t s);

// ---- nonbond ----
hipError_t launch_nonbond(const NonbondArgs &a, int nblocks, bool split, hipStream_t s);
hipError_t launch_chunk_reduce3(const WinDesc *wins, const int *atom_win, const double *part,
                                long long B, int ntot, int comps, float4 *f, double *o0,
                                double *o1, double *o2, hipStream_t s);

// ---- bonded (window-agnostic: terms carry global atom indices) ----
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
    double *energy;        // per-term energies (3 per dihedral: torsion, 1-4 Coulomb, 1-4 LJ)
    int terms;
};
hipError_t launch_bonded_terms(const BondedArgs &a, hipStream_t s);
hipError_t launch_bonded_gather(const float4 *contrib, const int *ptr, const int *idx, int n,
                                float4 *f, hipStream_t s);
// The MD engine's step prologue: every launch of a step that depends only on
// the coordinates, in one grid.  Each part is skipped when empty (nflag 0,
// bonded.contrib null, nres 0, pts.ntot 0).
struct PrologueArgs {
    int nflag;              // minimum-image flags of the IS-SPA windows
    float rmax_all;
    int *flag;
    BondedArgs bonded;      // bonded terms into their contribution slots
    int bonded_blocks;      // set by the launcher
    int nres;               // restraints: bias forces stored into fus (w = 1)
    const int *boff;
    const int2 *us_atoms;
    const float *us_mass;
    const float2 *kx;
    const float *Ls;
    float4 *fus;
    double *com;
    PointsArgs pts;         // MC points (pts.wins also serves the flags)
};
hipError_t launch_step_prologue(PrologueArgs a, hipStream_t s);

hipError_t launch_bonded_energy(const double *eterm, int nb, int na, int nd, double *out5,
                                hipStream_t s);

// ---- differentiable model: parameter preparation, forward reduction, adjoint ----
struct DiffPrepArgs {
    int T;
    const double *theta;   // [T][5] sigma, eps, x0, g0, alpha
    const double *camp;    // [T]
    double sig_s, eps_s, core, delta;
    DiffType *types;       // out [T]
    float2 *nb_lj;         // out [T*T] {12A, 6B} for the nonbond kernel
    float4 *nb_ld;         // out [T*T] {A, B, dA/dsigma_ij, dB/dsigma_ij}
};
struct DiffAtomArgs {
    int n;
    const float4 *xyz;     // positions (.w ignored)
    const double *q;       // charges
    const double *lam;     // [N][3] cotangent, or null (forward)
    const int *itype;
    const float *vtot, *rmax2;
    float4 *xq;            // out: positions with FP32 charges
    float4 *arec;          // out [N][4]: {x, R2}, {-q vtot lam, type}, {-vtot lam, q}, {lam, 0}
};
struct DiffReduceArgs {
    int n;
    const WinDesc *wins;
    const int *atom_win;
    const double *part;    // force partials [6][C] (S_e, S_lj)
    long long C;
    const int *itype;
    const float *vtot;
    const double *q;
    double *se;            // out [N][3]
    double *f;             // out [N][3] (overwritten)
};
struct VjpArgs {
    const WinDesc *wins;
    const int *atom_win;
    const int4 *fblk, *pblk;   // point-lane and atom-lane block tables
    int nmc;
    long long Ptot;
    const float4 *pts, *rec4, *aux4, *arec, *xq;
    const int *itype;
    const float *rmax2;
    const DiffType *dtypes;
    int ntypes;
    const int *pbc_flag;
    const double *g_out, *E_out;
    const int *cnt_out;
    float e0, p0, q0, o0;
    double *ppart;             // [10][F] point-record bars
    long long F;
    float4 *prec;              // [P][4] {p, R2}, {dL/dE, g dL/dg}, {dL/dE0, w}, -
    double *apart;             // [6][C] atom bars: q, theta(5)
    long long C;
};
struct VjpNbArgs {
    const WinDesc *wins;
    const int4 *blocks;
    const float4 *xq, *aux, *arec;
    const uint32_t *excl;
    const int *excl_ptr, *excl_idx;
    float cfar;
    const int *pbc_flag;
    const float4 *ld;          // [T*T] {A, B, dA, dB}
    int T;
    double *part;              // [3][B] q, sigma, eps bars
    long long B;
};
struct VjpFinishArgs {
    int n;
    const WinDesc *wins, *nb_wins;
    const int *atom_win;
    const double *apart;
    long long C;
    const double *nbpart;
    long long B;
    const int *itype;
    const float *vtot;
    const double *lam, *se, *theta;
    double *gq;                // out [N]
    double *gatom;             // out [N][5]
};
hipError_t launch_diff_prep_types(const DiffPrepArgs &a, hipStream_t s);
hipError_t launch_diff_prep_atoms(const DiffAtomArgs &a, hipStream_t s);
hipError_t launch_diff_reduce(const DiffReduceArgs &a, hipStream_t s);
hipError_t launch_vjp_points(const VjpArgs &a, int nblocks, hipStream_t s);
hipError_t launch_vjp_epilogue(const VjpArgs &a, hipStream_t s);
hipError_t launch_vjp_atoms(const VjpArgs &a, int nblocks, hipStream_t s);
hipError_t launch_vjp_nonbond(const VjpNbArgs &a, int nblocks, hipStream_t s);
hipError_t launch_vjp_finish(const VjpFinishArgs &a, hipStream_t s);
hipError_t launch_type_reduce(const double *gatom, const int *itype, int n, int T,
                              double *gtheta, hipStream_t s);

// ---- integrator / restraint / utilities ----
// atom_win == null: every atom belongs to the single window `one`
hipError_t launch_leapfrog(const WinDesc *wins, const int *atom_win, const WinDesc &one,
                           int ntot, float4 *xyzq, float4 *vel, const float4 *f, float kT,
                           float dt, uint32_t thr, uint64_t seed, uint32_t step,
                           const uint32_t *step_d, hipStream_t s, uint32_t *advance = nullptr);
// one workgroup per restraint; bias entries [boff[r], boff[r+1]) {atom, group}
hipError_t launch_us(int nres, const int *boff, const int2 *atoms, const float *mass,
                     const float2 *kx /* {k, x0} */, const float *L, const float4 *xyzq,
                     float4 *f, double *com, hipStream_t s);
hipError_t launch_pbc_flag(const WinDesc *wins, int W, const float4 *xyzq, float rmax_all,
                           int *flag, hipStream_t s);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_add4(float4 *dst, const float4 *src, int n, hipStream_t s);
hipError_t launch_step_advance(uint32_t *step_d, hipStream_t s);
hipError_t launch_set_word(uint32_t *dst, uint32_t v, hipStream_t s);

}  // namespace isspa
