// The reference's wrapper functions over the C ABI (see force_calls.h).
// Device models are built on first use from the host classes' arrays; one
// call in 16 is timed with events on the null stream (LagTimer below).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/isspa_amd.h"
#include "force_calls.h"

void isspa_check(int status, const char *what)
{
    if (status != ISSPA_OK) {
        fprintf(stderr, "%s failed (status %d): %s\n", what, status, isspa_last_error());
        exit(3);
    }
}

namespace {

// Per-wrapper timing without a synchronisation per call.  A timed event pair
// around a launch makes the device drain before and after it (about 4 us each
// on MI355X, more than a small system's kernels take), so only one call in
// kPeriod is bracketed by events; every call returns the duration of the
// latest bracketed call that has been read back.  That one is read at the next
// bracketed call, kPeriod calls later, when the device has long passed its
// end mark.  Summed over a run this estimates the reference's per-call timing
// (src/total_force_cuda.cpp:109-153 adds each call's milliseconds) from one
// call in kPeriod, lagged by kPeriod calls.
struct LagTimer {
    static constexpr unsigned kPeriod = 16;
    void *pair = nullptr;
    bool pending = false, marked = false;
    unsigned n = 0;
    float last = 0.0f;
    void begin()
    {
        marked = n++ % kPeriod == 0;
        if (!marked) return;
        if (!pair) isspa_check(isspa_events_create(&pair), "isspa_events_create");
        if (pending) {
            isspa_check(isspa_events_elapsed(pair, &last), "isspa_events_elapsed");
            pending = false;
        }
        isspa_check(isspa_events_mark_begin(pair, nullptr), "isspa_events_mark_begin");
    }
    float end()
    {
        if (marked) {
            isspa_check(isspa_events_mark_end(pair, nullptr), "isspa_events_mark_end");
            pending = true;
        }
        return last;
    }
};

void split_lj(const atom &a, std::vector<float> &A, std::vector<float> &B)
{
    const int ntri = a.nTypes * (a.nTypes + 1) / 2;
    A.resize(ntri);
    B.resize(ntri);
    for (int k = 0; k < ntri; ++k) {
        A[k] = a.lj_h[k].x;
        B[k] = a.lj_h[k].y;
    }
}

}  // namespace

// ------------------------------------------------------------ IS-SPA ------
void isspa_grid_block(int nAtoms, int /*nPairs*/, float lbox, isspa &is)
{
    if (is.model) return;
    isspa_model_desc d{};
    d.n_atoms = nAtoms;
    d.n_types = is.nTypes;
    d.n_mc = is.nMC;
    d.n_g = is.nGRs;
    d.n_e = is.nERs;
    d.n_f = is.nRs;
    d.g_r0 = is.gRparams.x; d.g_dr = is.gRparams.y;
    d.e_r0 = is.eRparams.x; d.e_dr = is.eRparams.y;
    d.f_r0 = is.forceRparams.x; d.f_dr = is.forceRparams.y;
    d.lbox = lbox;
    isspa_default_constants(&d.e0, &d.p0, &d.q0, &d.o0, &d.ep);
    d.seed = (uint64_t)(uint32_t)is.seed;
    d.system_id = 0;
    isspa_check(isspa_model_create(&d, is.isspaTypes_h, is.rmax_h, is.isspaGTable_h,
                                   is.isspaETable_h, is.isspaForceTable_h, &is.model),
                "isspa_model_create");
    printf("IS-SPA model: %d atoms, %d types, nMC=%d\n", nAtoms, is.nTypes, is.nMC);
}

float isspa_force_cuda(float4 *xyz_d, float4 *f_d, float4 *isspaf_d, isspa &is, int nAtoms)
{
    if (!is.model) {
        fprintf(stderr, "isspa_force_cuda: isspa_grid_block was not called\n");
        exit(3);
    }
    (void)nAtoms;
    static LagTimer t;
    t.begin();
    if (isspaf_d) {
        // the IS-SPA force alone is also kept in isspaf (the reference zeroes
        // it and never fills it, quirk Q17); one reduction writes both
        isspa_check(isspa_model_force_keep(is.model, (const float *)xyz_d, (float *)f_d,
                                           (float *)isspaf_d, is.step, nullptr, nullptr),
                    "isspa_model_force_keep");
    } else {
        isspa_check(isspa_model_force(is.model, (const float *)xyz_d, (float *)f_d, is.step,
                                      nullptr, nullptr), "isspa_model_force");
    }
    ++is.step;
    return t.end();
}

// ------------------------------------------------------------ nonbond -----
extern "C" void nonbond_force_cuda_grid_block(atom &atoms, float /*rCut2*/, float lbox)
{
    atoms.lbox = lbox;
    atoms.nThreads = ((atoms.nAtoms + 63) / 64) * 64;
    atoms.blockSize = 256;
    atoms.gridSize = (atoms.nAtoms + 255) / 256;
}

float nonbond_force_cuda(atom &atoms, isspa &is, int nAtoms)
{
    if (!atoms.nonbond) {
        std::vector<float> A, B;
        split_lj(atoms, A, B);
        float ep = 0.0f;
        isspa_default_constants(nullptr, nullptr, nullptr, nullptr, &ep);
        isspa_check(nonbond_model_create(nAtoms, atoms.nTypes, atoms.ityp_h,
                                         atoms.nonBondedParmIndex_h, A.data(), B.data(),
                                         atoms.nExcludedAtoms_h, atoms.excludedAtomsList_h,
                                         atoms.excludedAtomsListLength, is.isspaTypes_h,
                                         is.nTypes, is.rmax_h, ep, atoms.lbox, &atoms.nonbond),
                    "nonbond_model_create");
    }
    static LagTimer t;
    t.begin();
    isspa_check(nonbond_force(atoms.nonbond, (const float *)atoms.pos_d, (float *)atoms.for_d,
                              nullptr), "nonbond_force");
    return t.end();
}

// ------------------------------------------------------------ bonded ------
static void grid_block(int n, int *gridSize, int *blockSize, int *minGridSize)
{
    *blockSize = 256;
    *gridSize = (n + 255) / 256;
    *minGridSize = 1;
}
extern "C" void bond_force_cuda_grid_block(int n, int *g, int *b, int *m) { grid_block(n, g, b, m); }
extern "C" void angle_force_cuda_grid_block(int n, int *g, int *b, int *m) { grid_block(n, g, b, m); }
extern "C" void dih_force_cuda_grid_block(int n, int *g, int *b, int *m) { grid_block(n, g, b, m); }
extern "C" void leapfrog_cuda_grid_block(int n, int *g, int *b, int *m) { grid_block(n, g, b, m); }

float bond_force_cuda(float4 *xyz_d, float4 *f_d, int nAtoms, float lbox, bond &bonds)
{
    if (!bonds.model)
        isspa_check(bonded_model_create(nAtoms, bonds.nBonds, (const int *)bonds.bondAtoms_h,
                                        bonds.nTypes, (const float *)bonds.bondParams_h, 0, nullptr,
                                        0, nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr, 0,
                                        nullptr, nullptr, nullptr, nullptr, lbox, &bonds.model),
                    "bonded_model_create(bonds)");
    static LagTimer t;
    t.begin();
    isspa_check(bonded_force(bonds.model, ISSPA_TERM_BOND, (const float *)xyz_d, (float *)f_d,
                             nullptr, nullptr), "bonded_force(bonds)");
    return t.end();
}

float angle_force_cuda(float4 *xyz_d, float4 *f_d, int nAtoms, float lbox, angle &angles)
{
    if (!angles.model)
        isspa_check(bonded_model_create(nAtoms, 0, nullptr, 0, nullptr, angles.nAngles,
                                        (const int *)angles.angleAtoms_h, angles.nTypes,
                                        (const float *)angles.angleParams_h, 0, nullptr, nullptr,
                                        0, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr,
                                        lbox, &angles.model),
                    "bonded_model_create(angles)");
    static LagTimer t;
    t.begin();
    isspa_check(bonded_force(angles.model, ISSPA_TERM_ANGLE, (const float *)xyz_d, (float *)f_d,
                             nullptr, nullptr), "bonded_force(angles)");
    return t.end();
}

float dih_force_cuda(atom &atoms, dih &dihs, float lbox)
{
    if (!dihs.model) {
        std::vector<float> A, B;
        split_lj(atoms, A, B);
        isspa_check(bonded_model_create(atoms.nAtoms, 0, nullptr, 0, nullptr, 0, nullptr, 0,
                                        nullptr, dihs.nDihs, (const int *)dihs.dihAtoms_h,
                                        dihs.dihTypes_h, dihs.nTypes,
                                        (const float *)dihs.dihParams_h,
                                        (const float *)dihs.scaled14Factors_h, atoms.nTypes,
                                        atoms.ityp_h, atoms.nonBondedParmIndex_h, A.data(),
                                        B.data(), lbox, &dihs.model),
                    "bonded_model_create(dihedrals)");
    }
    static LagTimer t;
    t.begin();
    isspa_check(bonded_force(dihs.model, ISSPA_TERM_DIH, (const float *)atoms.pos_d,
                             (float *)atoms.for_d, nullptr, nullptr), "bonded_force(dihedrals)");
    return t.end();
}

// ------------------------------------------------------------ integrate ---
float leapfrog_cuda(atom &atoms, config &c)
{
    static LagTimer t;
    t.begin();
    isspa_check(leapfrog_andersen(atoms.nAtoms, (float *)atoms.pos_d, (float *)atoms.vel_d,
                                  (const float *)atoms.for_d, c.T, c.dt, c.pnu, c.lbox,
                                  (uint64_t)(uint32_t)atoms.seed, 0u, atoms.step, nullptr, nullptr),
                "leapfrog_andersen");
    ++atoms.step;
    return t.end();
}

// ------------------------------------------------------------ umbrella ----
void us_grid_block(us &bias)
{
    bias.usBlockSize = 256;
    bias.usGridSize = 1;
}

float us_force_cuda(float4 *xyz_d, float4 *f_d, us &bias, float lbox, int nAtoms)
{
    if (!bias.model)
        isspa_check(us_model_create(nAtoms, bias.totalBiasAtoms, (const int *)bias.atomList_h,
                                    bias.mass_h, bias.kumb_h, bias.x0, lbox, &bias.model),
                    "us_model_create");
    static LagTimer t;
    t.begin();
    isspa_check(us_force(bias.model, (const float *)xyz_d, (float *)f_d, bias.groupComPos_d,
                         nullptr), "us_force");
    return t.end();
}

extern "C" void init_rand_states(void *, int, int)
{
    // Philox draws are keyed by (seed, step, index): there is no state to set
}
