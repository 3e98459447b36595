// Device-side state of the C-ABI models, shared by the C-ABI translation
// units (capi_models.hip, capi_diff.hip).  Not part of the public ABI.
#pragma once

#include <vector>

#include "capi_common.hpp"

using namespace isspa;

// 256 CUs x 4 SIMDs x 8 resident waves = 8192 wave slots; the IS-SPA kernels
// use few enough VGPRs for full residency.  Work items are equal, so a grid
// runs in whole rounds of 8192 waves: targets are multiples of a round and
// the split never overshoots (a grid of 4.05 rounds takes 5).
static constexpr long kWaveSlots = 8192;
// Without a target set, one wave per kPointsPerWave MC points, within
// [kMinAutoWaves, kMaxAutoWaves]: measured on MI355X, more pieces than that
// cost more in partial sums (the epilogue and the chunk reduction read every
// piece of a point or atom in turn) than they gain in parallelism (1,100 atoms,
// nMC = 20: 116 us per step at 4,096 waves against 179 us at 32,768), and
// the largest systems tune to about 80,000-100,000 waves.
static constexpr long kPointsPerWave = 5;
static constexpr long kMinAutoWaves = 2048;
static constexpr long kMaxAutoWaves = 131072;

struct Windows {
    int W = 0, ntot = 0;
    std::vector<int> a0, n;
    std::vector<uint32_t> sys;
    std::vector<float> L;
    std::vector<int> atom_win;  // host copy
};

struct isspa_model {
    isspa_model_desc d{};
    Windows lay;
    long long P = 0;
    float rmax_all = 0.f;
    int *itype = nullptr, *atom_win = nullptr;
    float *rmax = nullptr, *rmax2 = nullptr, *vtot = nullptr;
    float *g = nullptr, *e = nullptr, *f = nullptr;
    float4 *pts = nullptr;
    float *rec = nullptr;  // [P][kRecFloats] records, then [P] float4 aux
    double *g_out = nullptr, *E_out = nullptr, *E0_out = nullptr;
    int *cnt_out = nullptr;
    int *pbc_flag = nullptr;
    // tiling
    int user_splits = 0, user_chunks = 0;
    long target_waves = 0;  // 0: automatic (kPointsPerWave)
    std::vector<WinDesc> wins_h;
    WinDesc *wins = nullptr;
    // field blocks: gblk (fg splits a block, the field pass) and fblk (one
    // split a block, the adjoint's point pass)
    int4 *fblk = nullptr, *pblk = nullptr, *gblk = nullptr;
    int nfblk = 0, npblk = 0, ngblk = 0;
    int max_fg = 1;  // the largest splits-per-field-block of the windows
    int max_win_atoms = 0, max_chunk_pts = 0;  // over the windows, for the fused force launch
    int min_win_atoms = 0;                     // the smallest window: receiver replicas
    double *part_g = nullptr;
    int *part_cnt = nullptr;
    double *part_s0 = nullptr, *part_ein = nullptr;
    double *part_f = nullptr;
    long long F = 0, C = 0, cap_field = 0, cap_force = 0;
    float4 *prof_f = nullptr;  // force sink of isspa_model_profile
    TableView tv{};
    // bumped whenever the launch tables are rebuilt; a graph captured under
    // another generation holds freed pointers and must not be replayed
    unsigned generation = 0;
};

struct nonbond_model {
    Windows lay;
    int n = 0, T = 0;
    float cfar = 0.f;
    float4 *aux = nullptr;
    float2 *lj = nullptr;
    uint32_t *excl = nullptr;
    int *excl_ptr = nullptr, *excl_idx = nullptr;
    int *pbc_flag = nullptr, *atom_win = nullptr;
    std::vector<WinDesc> wins_h;
    WinDesc *wins = nullptr;
    // split components: every pair in both orders, one entry per partner chunk
    int4 *blk = nullptr;
    int nblk = 0;
    long long B = 0;
    // total force: windows of kSymMinAtoms or more take each unordered pair
    // once (super-tiles); the others keep their partner chunks
    int4 *blk_tot = nullptr;
    int nblk_tot = 0;
    long long B_tot = 0;
    bool any_sym = false;
    double *part = nullptr;  // max(9 B, 3 B_tot) doubles, shared by both layouts
};

// the {cp E0, A1} array stored after the point records
inline float4 *rec_aux(isspa_model *m)
{
    return reinterpret_cast<float4 *>(m->rec + (size_t)m->P * kRecFloats);
}

FieldArgs field_args(isspa_model *m, const float4 *x);
EpilogueArgs epilogue_args(isspa_model *m, bool inspect);
ForceArgs force_args(isspa_model *m, const float4 *x);
ReduceArgs reduce_args(isspa_model *m, const float4 *x, float4 *f, double *flj, double *fpol,
                       double *fcdd);
// the all-pairs term for positions x: split -> three FP64 component outputs;
// otherwise f (float4, +=) or, with f null, o0 (double[N*3], +=)
int nonbond_run(nonbond_model *m, const float *xyzq4_d, bool split, float4 *f, double *o0,
                double *o1, double *o2, hipStream_t s);
// the same in two halves: per-chunk partials, then their fixed-order sum
// pbc_flag: per-window minimum-image flags already computed on stream s with a
// margin at least as large as the all-pairs one (the IS-SPA model's), or null
// to compute them here
int nonbond_partials(nonbond_model *m, const float *xyzq4_d, bool split, hipStream_t s,
                     const int *pbc_flag = nullptr);
int nonbond_reduce(nonbond_model *m, bool split, float4 *f, double *o0, double *o1, double *o2,
                   hipStream_t s);
