// C-ABI models (include/isspa_amd.h): IS-SPA, nonbond, bonded, umbrella,
// integrator and the hipGraph-captured MD engine.  Every model is built over
// a window layout: W independent systems with concatenated atoms; a single
// system is W = 1.  Launch geometry (per-window splits/chunks and the 1-D
// block work tables) is decided here on the host.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "models.hpp"

using namespace isspa;

#define fail isspa_fail

// ---------------------------------------------------------------------------
// window layout
// ---------------------------------------------------------------------------

static int build_windows(int ntot, const isspa_windows *win, uint32_t sys0, float L0, Windows &out)
{
    out = Windows{};
    out.ntot = ntot;
    if (!win) {
        out.W = 1;
        out.a0 = {0};
        out.n = {ntot};
        out.sys = {sys0};
        out.L = {L0};
    } else {
        if (win->n_windows < 1 || !win->window_atoms)
            return fail(ISSPA_ERR_ARG, "windows: n_windows=%d, window_atoms %s", win->n_windows,
                        win->window_atoms ? "set" : "NULL");
        out.W = win->n_windows;
        int acc = 0;
        for (int w = 0; w < out.W; ++w) {
            const int nw = win->window_atoms[w];
            if (nw < 1) return fail(ISSPA_ERR_ARG, "windows: window %d has %d atoms", w, nw);
            out.a0.push_back(acc);
            out.n.push_back(nw);
            out.sys.push_back(win->system_ids ? win->system_ids[w] : (uint32_t)w);
            const float L = win->lbox ? win->lbox[w] : L0;
            if (!(L > 0)) return fail(ISSPA_ERR_ARG, "windows: lbox of window %d is %g", w, L);
            out.L.push_back(L);
            acc += nw;
        }
        if (acc != ntot)
            return fail(ISSPA_ERR_ARG, "windows: window atoms sum to %d, model has %d", acc, ntot);
    }
    out.atom_win.resize(ntot);
    for (int w = 0; w < out.W; ++w)
        for (int i = 0; i < out.n[w]; ++i) out.atom_win[out.a0[w] + i] = w;
    return ISSPA_OK;
}

static WinDesc base_desc(const Windows &L, int w)
{
    WinDesc d{};
    d.a0 = L.a0[w];
    d.n = L.n[w];
    d.sys = L.sys[w];
    d.L = L.L[w];
    d.h = 0.5f * L.L[w];
    return d;
}


// pieces for `lanes` lanes doing `work` items each, at most `target` waves in
// total, with at least `min_work` items per piece
// Smallest automatic pieces: a field split keeps at least kMinSplitAtoms atoms
// and a force chunk at least kMinChunkPoints points.  Below these the FP64
// partial traffic and the reduction outgrow the gain in parallelism; small
// windows of a batch need pieces this fine to fill the device.
constexpr long long kMinSplitAtoms = 8;
constexpr long long kMinChunkPoints = 16;
// and an all-pairs partner chunk at least this many partners
constexpr long long kMinNonbondPartners = 16;

static int auto_pieces(long long lanes, long long work, long long min_work, double target)
{
    const long long lane_blocks = std::max(1LL, (lanes + 255) / 256);
    const long long waves = 4 * lane_blocks;
    long long s = (long long)std::floor(target / (double)waves);
    s = std::min(s, std::max(1LL, work / min_work));
    return (int)std::max(1LL, s);
}

// ---------------------------------------------------------------------------
// IS-SPA model
// ---------------------------------------------------------------------------

static void isspa_free(isspa_model *m)
{
    dfree(m->itype); dfree(m->atom_win); dfree(m->rmax); dfree(m->rmax2); dfree(m->vtot);
    dfree(m->g); dfree(m->e); dfree(m->f);
    dfree(m->pts); dfree(m->rec); dfree(m->g_out); dfree(m->E_out); dfree(m->E0_out);
    dfree(m->cnt_out); dfree(m->pbc_flag); dfree(m->wins); dfree(m->fblk); dfree(m->gblk);
    dfree(m->pblk);
    dfree(m->part_g); dfree(m->part_cnt); dfree(m->part_s0); dfree(m->part_ein); dfree(m->part_f);
    dfree(m->prof_f);
}

// per-window splits/chunks, partial offsets and the block work tables; blocks
// are ordered by decreasing work so the longest ones start first
static int isspa_tile(isspa_model *m)
{
    const Windows &L = m->lay;
    const int nmc = m->d.n_mc;
    double work_tot = 0;
    for (int w = 0; w < L.W; ++w) work_tot += (double)L.n[w] * L.n[w] * nmc;
    const double target =
        m->target_waves > 0
            ? (double)m->target_waves
            : (double)std::min(kMaxAutoWaves,
                               std::max(kMinAutoWaves, (long)(m->P / kPointsPerWave)));
    m->wins_h.assign(L.W, WinDesc{});
    m->max_win_atoms = m->max_chunk_pts = 0;
    m->max_fg = 1;
    m->min_win_atoms = 1 << 30;
    long long F = 0, C = 0;
    std::vector<std::pair<long long, int4>> fb, gb, pb;
    for (int w = 0; w < L.W; ++w) {
        WinDesc d = base_desc(L, w);
        const long long n = L.n[w], P = n * nmc;
        const double tgt = std::max(1.0, target * ((double)n * P / work_tot));
        int s = m->user_splits > 0 ? m->user_splits : auto_pieces(P, n, kMinSplitAtoms, tgt);
        s = (int)std::max(1LL, std::min((long long)s, n));
        // automatic splits from 4 up come in multiples of 4, so that each
        // field block sums four splits and the epilogue reads a quarter of them
        if (m->user_splits <= 0 && s >= 4) s -= s % 4;
        d.aps = (int)((n + s - 1) / s);
        d.splits = (int)((n + d.aps - 1) / d.aps);
        // a chunk keeps kMinChunkPoints points for each receiver replica
        int c = m->user_chunks > 0
                    ? m->user_chunks
                    : auto_pieces(n, P, kMinChunkPoints * force_replicas((int)n), tgt);
        c = (int)std::max(1LL, std::min((long long)c, P));
        d.ppc = (int)((P + c - 1) / c);
        d.chunks = (int)((P + d.ppc - 1) / d.ppc);
        m->max_win_atoms = std::max(m->max_win_atoms, (int)n);
        m->min_win_atoms = std::min(m->min_win_atoms, (int)n);
        m->max_chunk_pts = std::max(m->max_chunk_pts, d.ppc);
        d.fpo = F;
        d.ppo = C;
        F += (long long)d.splits * P;
        C += (long long)d.chunks * n;
        const int pblocks = (int)((P + 255) / 256), rblocks = (int)((n + 255) / 256);
        for (int b = 0; b < pblocks; ++b)
            for (int sp = 0; sp < d.splits; ++sp) fb.push_back({(long long)d.aps, make_int4(w, b, sp, 0)});
        // the field pass: fg splits of the same 256 / fg points a block
        d.fg = field_group(d.splits);
        m->max_fg = std::max(m->max_fg, d.fg);
        const int gpts = 256 / d.fg, gblocks = (int)((P + gpts - 1) / gpts);
        for (int b = 0; b < gblocks; ++b)
            for (int gr = 0; gr < d.splits / d.fg; ++gr)
                gb.push_back({(long long)d.aps, make_int4(w, b, gr, 0)});
        // chunk-major: the receiver blocks of one point chunk are consecutive,
        // so the round-robin dispatch over the 8 XCDs gives each XCD's L2 a
        // share of them that reuses the chunk's records
        for (int ch = 0; ch < d.chunks; ++ch)
            for (int b = 0; b < rblocks; ++b) pb.push_back({(long long)d.ppc, make_int4(w, b, ch, 0)});
        m->wins_h[w] = d;
    }
    auto by_work = [](const std::pair<long long, int4> &a, const std::pair<long long, int4> &b) {
        return a.first > b.first;
    };
    std::stable_sort(fb.begin(), fb.end(), by_work);
    std::stable_sort(gb.begin(), gb.end(), by_work);
    std::stable_sort(pb.begin(), pb.end(), by_work);
    std::vector<int4> fbv, gbv, pbv;
    for (auto &x : fb) fbv.push_back(x.second);
    for (auto &x : gb) gbv.push_back(x.second);
    for (auto &x : pb) pbv.push_back(x.second);
    dfree(m->wins); dfree(m->fblk); dfree(m->gblk); dfree(m->pblk);
    CHK(upload(&m->wins, m->wins_h.data(), m->wins_h.size()));
    CHK(upload(&m->fblk, fbv.data(), fbv.size()));
    CHK(upload(&m->gblk, gbv.data(), gbv.size()));
    CHK(upload(&m->pblk, pbv.data(), pbv.size()));
    m->nfblk = (int)fbv.size();
    m->ngblk = (int)gbv.size();
    m->npblk = (int)pbv.size();
    m->F = F;
    m->C = C;
    m->generation++;
    if (F > m->cap_field) {
        dfree(m->part_g); dfree(m->part_cnt); dfree(m->part_s0); dfree(m->part_ein);
        CHK(dalloc(&m->part_g, F));
        CHK(dalloc(&m->part_cnt, F));
        CHK(dalloc(&m->part_s0, 3 * F));
        CHK(dalloc(&m->part_ein, 3 * F));
        m->cap_field = F;
    }
    if (C > m->cap_force) {
        dfree(m->part_f);
        CHK(dalloc(&m->part_f, (size_t)force_part_stride(true) * C));
        m->cap_force = C;
    }
    return ISSPA_OK;
}

static void pad_table(const float *src, int T, int n, std::vector<float> &out)
{
    out.assign((size_t)T * (n + 1), 0.0f);
    for (int t = 0; t < T; ++t)
        for (int b = 0; b < n; ++b) out[(size_t)t * (n + 1) + b] = src[(size_t)t * n + b];
}

extern "C" int isspa_model_create_batch(const isspa_model_desc *d, const isspa_windows *win,
                                        const int *isspa_type, const float *rmax, const float *g,
                                        const float *e, const float *f, isspa_model **out)
{
    if (!d || !out || !isspa_type || !rmax || !g || !e || !f)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: null argument");
    *out = nullptr;
    if (d->n_atoms < 1 || d->n_types < 1 || d->n_mc < 1)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: n_atoms=%d n_types=%d n_mc=%d must be >= 1",
                    d->n_atoms, d->n_types, d->n_mc);
    // the force pass indexes 4 float4 per point with 32-bit integers
    if ((long long)d->n_atoms * d->n_mc >= (1LL << 29))
        return fail(ISSPA_ERR_ARG, "isspa_model_create: n_atoms*n_mc must be below 2^29");
    if (d->n_g < 2 || d->n_e < d->n_g || d->n_f < 2)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: table rows n_g=%d n_e=%d n_f=%d (need n_g>=2, n_e>=n_g, n_f>=2)",
                    d->n_g, d->n_e, d->n_f);
    if (!(d->g_dr > 0 && d->e_dr > 0 && d->f_dr > 0 && d->lbox > 0))
        return fail(ISSPA_ERR_ARG, "isspa_model_create: grid spacings and lbox must be > 0");
    for (int i = 0; i < d->n_atoms; ++i)
        if (isspa_type[i] < 0 || isspa_type[i] >= d->n_types)
            return fail(ISSPA_ERR_ARG, "isspa_model_create: isspa_type[%d]=%d out of range [0,%d)",
                        i, isspa_type[i], d->n_types);
    const int T = d->n_types;
    for (int t = 0; t < T; ++t)
        if (!(rmax[t] > 0)) return fail(ISSPA_ERR_ARG, "isspa_model_create: rmax[%d]=%g must be > 0", t, rmax[t]);
    isspa_model *m = new (std::nothrow) isspa_model;
    if (!m) return fail(ISSPA_ERR_ALLOC, "isspa_model_create: host allocation");
    m->d = *d;
    int rc = build_windows(d->n_atoms, win, d->system_id, d->lbox, m->lay);
    if (rc != ISSPA_OK) {
        delete m;
        return rc;
    }
    m->P = (long long)d->n_atoms * d->n_mc;
    std::vector<float> r2(T), vt(T);
    const float pi = 3.1415926535f;  // src/constants.h:6
    for (int t = 0; t < T; ++t) {
        r2[t] = rmax[t] * rmax[t];
        // src/isspa_class.cpp:110-112, double arithmetic stored as float
        vt[t] = (float)(4.0 / 3.0 * (double)pi * rmax[t] * rmax[t] * rmax[t] * 0.0074 /
                        (double)(float)d->n_mc);
        m->rmax_all = std::max(m->rmax_all, rmax[t]);
    }
    std::vector<float> gp, ep, fp;
    pad_table(g, T, d->n_g, gp);
    pad_table(e, T, d->n_e, ep);
    pad_table(f, T, d->n_f, fp);
#define TRY(x)                      \
    if ((rc = (x)) != ISSPA_OK) {   \
        isspa_free(m);              \
        delete m;                   \
        return rc;                  \
    }
    TRY(upload(&m->itype, isspa_type, d->n_atoms));
    TRY(upload(&m->atom_win, m->lay.atom_win.data(), m->lay.atom_win.size()));
    TRY(upload(&m->rmax, rmax, T));
    TRY(upload(&m->rmax2, r2.data(), T));
    TRY(upload(&m->vtot, vt.data(), T));
    TRY(upload(&m->g, gp.data(), gp.size()));
    TRY(upload(&m->e, ep.data(), ep.size()));
    TRY(upload(&m->f, fp.data(), fp.size()));
    TRY(dalloc(&m->pts, m->P));
    TRY(dalloc(&m->rec, (size_t)m->P * (kRecFloats + kAuxFloats)));  // records, then aux
    TRY(dalloc(&m->g_out, m->P));
    TRY(dalloc(&m->cnt_out, m->P));
    TRY(dalloc(&m->E_out, 3 * (size_t)m->P));
    TRY(dalloc(&m->E0_out, 3 * (size_t)m->P));
    TRY(dalloc(&m->pbc_flag, m->lay.W));
    TRY(isspa_tile(m));
#undef TRY
    TableView &tv = m->tv;
    tv.g = m->g; tv.e = m->e; tv.f = m->f;
    tv.n_g = d->n_g; tv.n_e = d->n_e; tv.n_f = d->n_f;
    tv.ld_g = d->n_g + 1; tv.ld_e = d->n_e + 1; tv.ld_f = d->n_f + 1;
    tv.ntypes = d->n_types;
    tv.g_r0 = d->g_r0; tv.g_dr = d->g_dr; tv.g_inv = 1.0f / d->g_dr;
    tv.e_r0 = d->e_r0; tv.e_dr = d->e_dr; tv.e_inv = 1.0f / d->e_dr;
    tv.f_r0 = d->f_r0; tv.f_dr = d->f_dr; tv.f_inv = 1.0f / d->f_dr;
    *out = m;
    return ISSPA_OK;
}

extern "C" int isspa_model_create(const isspa_model_desc *d, const int *isspa_type,
                                  const float *rmax, const float *g, const float *e,
                                  const float *f, isspa_model **out)
{
    return isspa_model_create_batch(d, nullptr, isspa_type, rmax, g, e, f, out);
}

extern "C" int isspa_model_destroy(isspa_model *m)
{
    if (!m) return ISSPA_OK;
    isspa_free(m);
    delete m;
    return ISSPA_OK;
}

extern "C" int isspa_model_set_tiling(isspa_model *m, int field_splits, int force_chunks)
{
    if (!m) return fail(ISSPA_ERR_ARG, "null model");
    if (field_splits < 0 || force_chunks < 0) return fail(ISSPA_ERR_ARG, "negative tiling");
    m->user_splits = field_splits;
    m->user_chunks = force_chunks;
    return isspa_tile(m);
}

extern "C" int isspa_model_set_target_waves(isspa_model *m, int waves)
{
    if (!m || waves < 0) return fail(ISSPA_ERR_ARG, "isspa_model_set_target_waves: bad argument");
    m->target_waves = waves;
    return isspa_tile(m);
}

extern "C" int isspa_model_get_tiling(const isspa_model *m, int *field_splits, int *force_chunks)
{
    if (!m) return fail(ISSPA_ERR_ARG, "null model");
    // the largest window's tiling (the only one for a single system)
    int best = 0;
    for (int w = 1; w < m->lay.W; ++w)
        if (m->lay.n[w] > m->lay.n[best]) best = w;
    if (field_splits) *field_splits = m->wins_h[best].splits;
    if (force_chunks) *force_chunks = m->wins_h[best].chunks;
    return ISSPA_OK;
}

FieldArgs field_args(isspa_model *m, const float4 *x)
{
    FieldArgs fa{};
    fa.wins = m->wins; fa.blocks = m->gblk; fa.max_fg = m->max_fg; fa.pts = m->pts; fa.xyzq = x; fa.itype = m->itype;
    fa.tab = m->tv; fa.nmc = m->d.n_mc; fa.e0 = m->d.e0; fa.pbc_flag = m->pbc_flag;
    fa.part_g = m->part_g; fa.part_cnt = m->part_cnt; fa.part_s0 = m->part_s0;
    fa.part_ein = m->part_ein; fa.F = m->F;
    return fa;
}

// inspect: also write the per-point g, count, E and E0 that
// isspa_model_get_field returns and the adjoint reads (the component entry
// point and the differentiable forward)
EpilogueArgs epilogue_args(isspa_model *m, bool inspect)
{
    EpilogueArgs ea{};
    ea.wins = m->wins; ea.nwin = m->lay.W; ea.atom_win = m->atom_win; ea.pts = m->pts;
    ea.Ptot = m->P; ea.nmc = m->d.n_mc; ea.part_g = m->part_g; ea.part_cnt = m->part_cnt;
    ea.part_s0 = m->part_s0; ea.part_ein = m->part_ein; ea.F = m->F;
    ea.e0 = m->d.e0; ea.p0 = m->d.p0; ea.q0 = m->d.q0; ea.o0 = m->d.o0;
    ea.rec = m->rec; ea.aux = rec_aux(m);
    if (inspect) {
        ea.g_out = m->g_out; ea.cnt_out = m->cnt_out; ea.E_out = m->E_out;
        ea.E0_out = m->E0_out;
    }
    return ea;
}

ForceArgs force_args(isspa_model *m, const float4 *x)
{
    ForceArgs fa{};
    fa.wins = m->wins; fa.blocks = m->pblk; fa.rec4 = reinterpret_cast<const float4 *>(m->rec);
    fa.aux4 = rec_aux(m);
    fa.xyzq = x; fa.itype = m->itype; fa.rmax2 = m->rmax2; fa.tab = m->tv; fa.nmc = m->d.n_mc;
    fa.pbc_flag = m->pbc_flag; fa.part = m->part_f; fa.C = m->C;
    // one system: every block has the same work, so each XCD may take a
    // contiguous run of chunks (batches keep their longest-first order)
    fa.xcd_runs = m->lay.W == 1;
    fa.replicas = force_replicas(m->min_win_atoms);
    return fa;
}

ReduceArgs reduce_args(isspa_model *m, const float4 *x, float4 *f, double *flj, double *fpol,
                       double *fcdd)
{
    ReduceArgs ra{};
    ra.wins = m->wins; ra.atom_win = m->atom_win; ra.part = m->part_f; ra.C = m->C;
    ra.ntot = m->d.n_atoms; ra.xyzq = x; ra.itype = m->itype; ra.vtot = m->vtot; ra.f = f;
    ra.f_lj = flj; ra.f_pol = fpol; ra.f_cdd = fcdd;
    return ra;
}

// points + field + epilogue (shared by the total and component entry points)
static int isspa_field_pass(isspa_model *m, const float *xyzq4_d, uint32_t step,
                            const uint32_t *step_d, hipStream_t s, bool inspect = false)
{
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    HIPCHK(launch_mc_points(m->wins, m->atom_win, x, m->itype, m->rmax, m->rmax2, m->d.n_atoms,
                            m->d.n_mc, m->d.seed, step, step_d, m->pts, s, m->lay.W, m->rmax_all,
                            m->pbc_flag));
    HIPCHK(launch_field(field_args(m, x), m->ngblk, s));
    HIPCHK(launch_epilogue(epilogue_args(m, inspect), s));
    return ISSPA_OK;
}

static int isspa_force_pass(isspa_model *m, const float *xyzq4_d, bool split, float4 *f,
                            double *flj, double *fpol, double *fcdd, hipStream_t s,
                            float4 *f_alone = nullptr)
{
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    HIPCHK(launch_force(force_args(m, x), m->npblk, split, s));
    ReduceArgs ra = reduce_args(m, x, f, flj, fpol, fcdd);
    ra.f_alone = f_alone;
    HIPCHK(launch_force_reduce(ra, split, s));
    return ISSPA_OK;
}

extern "C" int isspa_model_force(isspa_model *m, const float *xyzq4_d, float *f4_d,
                                 uint32_t step, const uint32_t *step_d, void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "isspa_model_force: null argument");
    CHK(isspa_field_pass(m, xyzq4_d, step, step_d, S(stream)));
    return isspa_force_pass(m, xyzq4_d, false, reinterpret_cast<float4 *>(f4_d), nullptr, nullptr,
                            nullptr, S(stream));
}

extern "C" int isspa_model_force_keep(isspa_model *m, const float *xyzq4_d, float *f4_d,
                                      float *f_isspa4_d, uint32_t step, const uint32_t *step_d,
                                      void *stream)
{
    if (!m || !xyzq4_d || !f4_d || !f_isspa4_d)
        return fail(ISSPA_ERR_ARG, "isspa_model_force_keep: null argument");
    CHK(isspa_field_pass(m, xyzq4_d, step, step_d, S(stream)));
    return isspa_force_pass(m, xyzq4_d, false, reinterpret_cast<float4 *>(f4_d), nullptr, nullptr,
                            nullptr, S(stream), reinterpret_cast<float4 *>(f_isspa4_d));
}

extern "C" int isspa_model_force_components(isspa_model *m, const float *xyzq4_d,
                                            double *f_lj_d, double *f_pol_d, double *f_cdd_d,
                                            uint32_t step, void *stream)
{
    if (!m || !xyzq4_d || !f_lj_d || !f_pol_d || !f_cdd_d)
        return fail(ISSPA_ERR_ARG, "isspa_model_force_components: null argument");
    CHK(isspa_field_pass(m, xyzq4_d, step, nullptr, S(stream), true));
    return isspa_force_pass(m, xyzq4_d, true, nullptr, f_lj_d, f_pol_d, f_cdd_d, S(stream));
}

extern "C" int isspa_model_profile(isspa_model *m, const float *xyzq4_d, int reps, float *ms,
                                   void *stream)
{
    if (!m || !xyzq4_d || !ms || reps < 1) return fail(ISSPA_ERR_ARG, "isspa_model_profile: bad argument");
    hipStream_t s = S(stream);
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    if (!m->prof_f) CHK(dalloc(&m->prof_f, (size_t)m->d.n_atoms));
    hipEvent_t ev[2];
    HIPCHK(hipEventCreate(&ev[0]));
    HIPCHK(hipEventCreate(&ev[1]));
    int rc = ISSPA_OK;
    hipError_t e = launch_pbc_flag(m->wins, m->lay.W, x, m->rmax_all, m->pbc_flag, s);
    for (int stage = 0; stage < 5 && rc == ISSPA_OK && e == hipSuccess; ++stage) {
        // one untimed launch, then `reps` timed back-to-back launches
        for (int r = 0; r <= reps && e == hipSuccess; ++r) {
            if (r == 1) e = hipEventRecord(ev[0], s);
            if (e != hipSuccess) break;
            switch (stage) {
            case 0:
                e = launch_mc_points(m->wins, m->atom_win, x, m->itype, m->rmax, m->rmax2,
                                     m->d.n_atoms, m->d.n_mc, m->d.seed, 0u, nullptr, m->pts, s);
                break;
            case 1: e = launch_field(field_args(m, x), m->ngblk, s); break;
            case 2: e = launch_epilogue(epilogue_args(m, false), s); break;
            case 3: e = launch_force(force_args(m, x), m->npblk, false, s); break;
            default:
                e = launch_force_reduce(reduce_args(m, x, m->prof_f, nullptr, nullptr, nullptr),
                                        false, s);
            }
        }
        if (e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
        ms[stage] = t / (float)reps;
    }
    if (e != hipSuccess) rc = fail(ISSPA_ERR_HIP, "isspa_model_profile: %s", hipGetErrorString(e));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return rc;
}

extern "C" int isspa_model_get_field(isspa_model *m, float *pts4_d, double *g_d, int *cnt_d,
                                     double *E_d, double *E0_d, void *stream)
{
    if (!m) return fail(ISSPA_ERR_ARG, "null model");
    const size_t P = m->P;
    hipStream_t s = S(stream);
    if (pts4_d) HIPCHK(hipMemcpyAsync(pts4_d, m->pts, P * sizeof(float4), hipMemcpyDeviceToDevice, s));
    if (g_d) HIPCHK(hipMemcpyAsync(g_d, m->g_out, P * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (cnt_d) HIPCHK(hipMemcpyAsync(cnt_d, m->cnt_out, P * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (E_d) HIPCHK(hipMemcpyAsync(E_d, m->E_out, 3 * P * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (E0_d) HIPCHK(hipMemcpyAsync(E0_d, m->E0_out, 3 * P * sizeof(double), hipMemcpyDeviceToDevice, s));
    return ISSPA_OK;
}

// ---------------------------------------------------------------------------
// nonbond model
// ---------------------------------------------------------------------------

static void nonbond_free(nonbond_model *m)
{
    dfree(m->aux); dfree(m->lj); dfree(m->excl); dfree(m->excl_ptr); dfree(m->excl_idx);
    dfree(m->pbc_flag); dfree(m->atom_win); dfree(m->wins); dfree(m->blk); dfree(m->blk_tot);
    dfree(m->part);
}

extern "C" int nonbond_model_create_batch(int n, int T, const int *ityp, const int *nbparm,
                                          const float *lj_a, const float *lj_b, const int *pre,
                                          const int *list, int L, const int *isspa_type, int Ti,
                                          const float *rmax, float ep, float lbox,
                                          const isspa_windows *win, nonbond_model **out)
{
    if (!out || !ityp || !nbparm || !lj_a || !lj_b || !pre || (!list && L > 0) || !isspa_type || !rmax)
        return fail(ISSPA_ERR_ARG, "nonbond_model_create: null argument");
    *out = nullptr;
    if (n < 1 || T < 1 || Ti < 1 || L < 0 || !(lbox > 0))
        return fail(ISSPA_ERR_ARG, "nonbond_model_create: bad sizes n=%d T=%d Ti=%d L=%d", n, T, Ti, L);
    if ((size_t)T * T * sizeof(float2) > 65536)
        return fail(ISSPA_ERR_NOTSUP, "nonbond_model_create: %d LJ types exceed the 64 KB LDS pair table", T);
    const int ntri = T * (T + 1) / 2;
    for (int i = 0; i < n; ++i) {
        if (ityp[i] < 0 || ityp[i] >= T) return fail(ISSPA_ERR_ARG, "ityp[%d]=%d out of range", i, ityp[i]);
        if (isspa_type[i] < 0 || isspa_type[i] >= Ti)
            return fail(ISSPA_ERR_ARG, "isspa_type[%d]=%d out of range", i, isspa_type[i]);
        if (pre[i] < (i ? pre[i - 1] : 0) || pre[i] > L)
            return fail(ISSPA_ERR_ARG, "exclusion prefix sum not monotone / exceeds list at %d", i);
    }
    for (int k = 0; k < T * T; ++k)
        if (nbparm[k] < 0 || nbparm[k] >= ntri)
            return fail(ISSPA_ERR_ARG, "nbparm[%d]=%d out of range [0,%d)", k, nbparm[k], ntri);
    nonbond_model *m = new (std::nothrow) nonbond_model;
    if (!m) return fail(ISSPA_ERR_ALLOC, "host allocation");
    int rc = build_windows(n, win, 0u, lbox, m->lay);
    if (rc != ISSPA_OK) {
        delete m;
        return rc;
    }
    m->n = n;
    m->T = T;
    const double fep = 1.0 - 1.0 / (double)ep;
    m->cfar = (float)(-2.0 * fep / 3.0);
    std::vector<float4> aux(n);
    for (int j = 0; j < n; ++j) {
        const double R = rmax[isspa_type[j]];
        aux[j].x = (float)(4.0 * R * R);
        aux[j].y = (float)(-fep / (3.0 * R * R * R));
        aux[j].z = (float)(fep / (8.0 * R * R * R * R));
        aux[j].w = __builtin_bit_cast(float, ityp[j]);
    }
    std::vector<float2> lj((size_t)T * T);
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) {
            const int p = nbparm[a * T + b];
            lj[(size_t)a * T + b] = make_float2(12.0f * lj_a[p], 6.0f * lj_b[p]);
        }
    // exclusions: decode with the reference's scan (src/nonbond_force_cuda.cu:71-108)
    // into a symmetric per-atom set, then into 128-bit windows j - i + 64.
    std::vector<std::vector<int>> ex(n);
    for (int lo = 0; lo < n; ++lo) {
        const int start = lo == 0 ? 0 : pre[lo - 1], stop = pre[lo];
        for (int e = start; e < stop; ++e) {
            const int hi = list[e];
            if (hi <= lo || hi >= n) continue;
            bool found = false;
            for (int k = start; k < stop; ++k) {
                if (list[k] == hi) { found = true; break; }
                if (list[k] > hi) break;
            }
            if (!found) continue;
            if (m->lay.atom_win[lo] != m->lay.atom_win[hi]) {
                nonbond_free(m);
                delete m;
                return fail(ISSPA_ERR_ARG, "exclusion %d-%d crosses windows", lo, hi);
            }
            ex[lo].push_back(hi);
            ex[hi].push_back(lo);
        }
    }
    std::vector<uint32_t> winmask((size_t)4 * n, 0u);
    std::vector<int> fptr(n + 1, 0), fidx;
    bool fallback = false;
    for (int i = 0; i < n; ++i) {
        std::sort(ex[i].begin(), ex[i].end());
        ex[i].erase(std::unique(ex[i].begin(), ex[i].end()), ex[i].end());
        for (int j : ex[i]) {
            const int w = j - i + 64;
            if (w >= 0 && w < 128) winmask[4 * (size_t)i + (w >> 5)] |= 1u << (w & 31);
            else fallback = true;
        }
    }
    if (fallback)
        for (int i = 0; i < n; ++i) {
            for (int j : ex[i])
                if (j - i + 64 < 0 || j - i + 64 >= 128) fidx.push_back(j);
            fptr[i + 1] = (int)fidx.size();
        }
    // partner chunks per window, work-balanced over the launch
    const Windows &Ly = m->lay;
    double work_tot = 0;
    for (int w = 0; w < Ly.W; ++w) work_tot += (double)Ly.n[w] * Ly.n[w];
    m->wins_h.assign(Ly.W, WinDesc{});
    std::vector<std::pair<long long, int4>> bl;
    long long B = 0;
    for (int w = 0; w < Ly.W; ++w) {
        WinDesc d = base_desc(Ly, w);
        const long long nw = Ly.n[w];
        const double tgt = std::max(1.0, kWaveSlots * ((double)nw * nw / work_tot));
        const int c = auto_pieces(nw, nw, kMinNonbondPartners, tgt);
        d.apc = (int)((nw + c - 1) / c);
        d.nbc = (int)((nw + d.apc - 1) / d.apc);
        d.npo = B;
        B += (long long)d.nbc * nw;
        for (int b = 0; b < (int)((nw + 255) / 256); ++b)
            for (int ch = 0; ch < d.nbc; ++ch) bl.push_back({(long long)d.apc, make_int4(w, b, ch, 0)});
        m->wins_h[w] = d;
    }
    // total force: a window of kSymMinAtoms or more takes each unordered pair
    // once in super-tiles of 256 x 256 atoms (and the pairs of a block with
    // itself and its neighbours in both orders, one ordered block per
    // receiver block and partner block) when no pair of two blocks at least
    // two apart can be excluded (no exclusion beyond the 128-wide windows)
    // and the LJ pair table is symmetric; its atoms then have one partial
    // entry per 256-atom block.  Other windows keep their partner chunks.
    // The weights order the blocks longest first: partners per lane, a
    // super-tile's pair step costing about 1.25 of an ordered pair.
    bool lj_sym = true;
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < a; ++b) {
            const float2 u = lj[(size_t)a * T + b], v = lj[(size_t)b * T + a];
            lj_sym = lj_sym && u.x == v.x && u.y == v.y;
        }
    // the LJ table and a super-tile's partner sums share the block's 64 KB of LDS
    const bool sym_lds_ok = nonbond_lds_bytes(T, true) <= 65536;
    std::vector<std::pair<long long, int4>> blt;
    long long Bt = 0;
    for (int w = 0; w < Ly.W; ++w) {
        WinDesc &d = m->wins_h[w];
        const int nw = Ly.n[w];
        const bool sym = !fallback && lj_sym && sym_lds_ok && nw >= kSymMinAtoms;
        const int nsb = (nw + kNbBlock - 1) / kNbBlock;
        d.nte = sym ? nsb : d.nbc;
        d.spo = Bt;
        Bt += (long long)d.nte * nw;
        if (!sym) {
            for (int b = 0; b < nsb; ++b)
                for (int ch = 0; ch < d.nbc; ++ch)
                    blt.push_back({(long long)d.apc, make_int4(w, b, ch, kNbKindChunk)});
            continue;
        }
        m->any_sym = true;
        for (int rb = 0; rb < nsb; ++rb) {
            for (int q = std::max(0, rb - 1); q <= std::min(nsb - 1, rb + 1); ++q)
                blt.push_back({(long long)std::min(kNbBlock, nw - q * kNbBlock),
                               make_int4(w, rb, q, kNbKindNear)});
            for (int pb = rb + 2; pb < nsb; ++pb)
                blt.push_back({(long long)(1.25 * kNbBlock), make_int4(w, rb, pb, kNbKindSym)});
        }
    }
    auto by_weight = [](std::vector<std::pair<long long, int4>> &v) {
        std::stable_sort(v.begin(), v.end(),
                         [](const std::pair<long long, int4> &a, const std::pair<long long, int4> &b) {
                             return a.first > b.first;
                         });
        std::vector<int4> out;
        for (auto &x : v) out.push_back(x.second);
        return out;
    };
    std::vector<int4> blv = by_weight(bl), bltv = by_weight(blt);
    m->nblk = (int)blv.size();
    m->B = B;
    m->nblk_tot = (int)bltv.size();
    m->B_tot = Bt;
#define TRY(x)                    \
    if ((rc = (x)) != ISSPA_OK) { \
        nonbond_free(m);          \
        delete m;                 \
        return rc;                \
    }
    TRY(upload(&m->aux, aux.data(), aux.size()));
    TRY(upload(&m->lj, lj.data(), lj.size()));
    TRY(upload(&m->excl, winmask.data(), winmask.size()));
    if (fallback) {
        TRY(upload(&m->excl_ptr, fptr.data(), fptr.size()));
        TRY(upload(&m->excl_idx, fidx.data(), std::max<size_t>(1, fidx.size())));
    }
    TRY(dalloc(&m->pbc_flag, Ly.W));
    TRY(upload(&m->atom_win, Ly.atom_win.data(), Ly.atom_win.size()));
    TRY(upload(&m->wins, m->wins_h.data(), m->wins_h.size()));
    TRY(upload(&m->blk, blv.data(), blv.size()));
    TRY(upload(&m->blk_tot, bltv.data(), bltv.size()));
    TRY(dalloc(&m->part, (size_t)std::max(9 * B, 3 * Bt)));
#undef TRY
    *out = m;
    return ISSPA_OK;
}

extern "C" int nonbond_model_create(int n, int T, const int *ityp, const int *nbparm,
                                    const float *lj_a, const float *lj_b, const int *pre,
                                    const int *list, int L, const int *isspa_type, int Ti,
                                    const float *rmax, float ep, float lbox, nonbond_model **out)
{
    return nonbond_model_create_batch(n, T, ityp, nbparm, lj_a, lj_b, pre, list, L, isspa_type,
                                      Ti, rmax, ep, lbox, nullptr, out);
}

extern "C" int nonbond_model_destroy(nonbond_model *m)
{
    if (!m) return ISSPA_OK;
    nonbond_free(m);
    delete m;
    return ISSPA_OK;
}

// split: the components' block table and partial layout; otherwise the total force's
static NonbondArgs nonbond_args(nonbond_model *m, const float4 *x, const int *pbc_flag, bool split)
{
    NonbondArgs a{};
    a.wins = m->wins; a.blocks = split ? m->blk : m->blk_tot; a.xyzq = x; a.aux = m->aux;
    a.lj = m->lj; a.T = m->T;
    a.excl = m->excl; a.excl_ptr = m->excl_ptr; a.excl_idx = m->excl_idx; a.cfar = m->cfar;
    a.pbc_flag = pbc_flag; a.part = m->part; a.B = split ? m->B : m->B_tot;
    a.sym = !split && m->any_sym;
    return a;
}

int nonbond_partials(nonbond_model *m, const float *xyzq4_d, bool split, hipStream_t s,
                     const int *pbc_flag)
{
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    // A flag set with a larger margin only sends a window through the
    // minimum-image loop when no image is needed, where that loop's shift
    // (fma(-L, 0, r)) returns r unchanged: the result is the same bit for bit.
    if (!pbc_flag) {
        HIPCHK(launch_pbc_flag(m->wins, m->lay.W, x, 0.0f, m->pbc_flag, s));
        pbc_flag = m->pbc_flag;
    }
    HIPCHK(launch_nonbond(nonbond_args(m, x, pbc_flag, split), split ? m->nblk : m->nblk_tot,
                          split, s));
    return ISSPA_OK;
}

int nonbond_reduce(nonbond_model *m, bool split, float4 *f, double *o0, double *o1, double *o2,
                   hipStream_t s)
{
    HIPCHK(launch_chunk_reduce3(m->wins, m->atom_win, m->part, split ? m->B : m->B_tot, m->n,
                                split ? 9 : 3, f, o0, o1, o2, s));
    return ISSPA_OK;
}

int nonbond_run(nonbond_model *m, const float *xyzq4_d, bool split, float4 *f, double *o0,
                       double *o1, double *o2, hipStream_t s)
{
    CHK(nonbond_partials(m, xyzq4_d, split, s));
    return nonbond_reduce(m, split, f, o0, o1, o2, s);
}

extern "C" int nonbond_force(nonbond_model *m, const float *xyzq4_d, float *f4_d, void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "nonbond_force: null argument");
    return nonbond_run(m, xyzq4_d, false, reinterpret_cast<float4 *>(f4_d), nullptr, nullptr,
                       nullptr, S(stream));
}

extern "C" int nonbond_force_components(nonbond_model *m, const float *xyzq4_d, double *f_lj_d,
                                        double *f_coul_d, double *f_pair_d, void *stream)
{
    if (!m || !xyzq4_d || !f_lj_d || !f_coul_d || !f_pair_d)
        return fail(ISSPA_ERR_ARG, "nonbond_force_components: null argument");
    return nonbond_run(m, xyzq4_d, true, nullptr, f_lj_d, f_coul_d, f_pair_d, S(stream));
}

// ---------------------------------------------------------------------------
// bonded model (window-agnostic: terms carry global atom indices)
// ---------------------------------------------------------------------------
struct bonded_model {
    int n = 0, nb = 0, na = 0, nd = 0, T = 0;
    float lbox = 0.f;
    int4 *bond_atoms = nullptr, *angle_atoms = nullptr, *dih_atoms = nullptr;
    float2 *bond_params = nullptr, *angle_params = nullptr, *sc14 = nullptr, *lj = nullptr;
    float4 *dih_params = nullptr;
    int *dih_types = nullptr, *ityp = nullptr;
    float4 *contrib = nullptr;
    double *energy = nullptr;
    int *gptr = nullptr, *gidx = nullptr;
};

static void bonded_free(bonded_model *m)
{
    dfree(m->bond_atoms); dfree(m->angle_atoms); dfree(m->dih_atoms); dfree(m->bond_params);
    dfree(m->angle_params); dfree(m->sc14); dfree(m->lj); dfree(m->dih_params);
    dfree(m->dih_types); dfree(m->ityp); dfree(m->contrib); dfree(m->energy);
    dfree(m->gptr); dfree(m->gidx);
}

extern "C" int bonded_model_create(int n, int nb, const int *bond_atoms, int nbt,
                                   const float *bond_params, int na, const int *angle_atoms,
                                   int nat, const float *angle_params, int nd,
                                   const int *dih_atoms, const int *dih_types, int ndt,
                                   const float *dih_params, const float *sc14, int T,
                                   const int *ityp, const int *nbparm, const float *lj_a,
                                   const float *lj_b, float lbox, bonded_model **out)
{
    if (!out) return fail(ISSPA_ERR_ARG, "bonded_model_create: null out");
    *out = nullptr;
    if (n < 1 || nb < 0 || na < 0 || nd < 0 || !(lbox > 0))
        return fail(ISSPA_ERR_ARG, "bonded_model_create: bad sizes");
    if ((nb && (!bond_atoms || !bond_params)) || (na && (!angle_atoms || !angle_params)) ||
        (nd && (!dih_atoms || !dih_types || !dih_params || !sc14 || !ityp || !nbparm || !lj_a || !lj_b)))
        return fail(ISSPA_ERR_ARG, "bonded_model_create: null array for a non-empty term list");
    auto bad_atom = [n](int a) { return a < 0 || a >= n; };
    for (int q = 0; q < nb; ++q)
        if (bad_atom(bond_atoms[4 * q]) || bad_atom(bond_atoms[4 * q + 1]) ||
            bond_atoms[4 * q + 2] < 0 || bond_atoms[4 * q + 2] >= nbt)
            return fail(ISSPA_ERR_ARG, "bond %d has an atom or type out of range", q);
    for (int q = 0; q < na; ++q)
        if (bad_atom(angle_atoms[4 * q]) || bad_atom(angle_atoms[4 * q + 1]) ||
            bad_atom(angle_atoms[4 * q + 2]) || angle_atoms[4 * q + 3] < 0 ||
            angle_atoms[4 * q + 3] >= nat)
            return fail(ISSPA_ERR_ARG, "angle %d has an atom or type out of range", q);
    for (int q = 0; q < nd; ++q)
        if (bad_atom(dih_atoms[4 * q]) || bad_atom(dih_atoms[4 * q + 1]) ||
            bad_atom(std::abs(dih_atoms[4 * q + 2])) || bad_atom(std::abs(dih_atoms[4 * q + 3])) ||
            dih_types[q] < 0 || dih_types[q] >= ndt)
            return fail(ISSPA_ERR_ARG, "dihedral %d has an atom or type out of range", q);
    if (nd) {
        if (T < 1) return fail(ISSPA_ERR_ARG, "bonded_model_create: dihedrals need LJ types");
        for (int i = 0; i < n; ++i)
            if (ityp[i] < 0 || ityp[i] >= T) return fail(ISSPA_ERR_ARG, "ityp[%d] out of range", i);
        for (int k = 0; k < T * T; ++k)
            if (nbparm[k] < 0 || nbparm[k] >= T * (T + 1) / 2)
                return fail(ISSPA_ERR_ARG, "nbparm[%d] out of range", k);
    }
    bonded_model *m = new (std::nothrow) bonded_model;
    if (!m) return fail(ISSPA_ERR_ALLOC, "host allocation");
    m->n = n; m->nb = nb; m->na = na; m->nd = nd; m->T = T; m->lbox = lbox;
    // per-atom gather lists over the contribution slots, in slot order
    std::vector<std::vector<int>> slots(n);
    for (int q = 0; q < nb; ++q) {
        slots[bond_atoms[4 * q]].push_back(2 * q);
        slots[bond_atoms[4 * q + 1]].push_back(2 * q + 1);
    }
    for (int q = 0; q < na; ++q)
        for (int s = 0; s < 3; ++s) slots[angle_atoms[4 * q + s]].push_back(2 * nb + 3 * q + s);
    for (int q = 0; q < nd; ++q) {
        const int base = 2 * nb + 3 * na + 6 * q;
        const int at[4] = {dih_atoms[4 * q], dih_atoms[4 * q + 1], std::abs(dih_atoms[4 * q + 2]),
                           std::abs(dih_atoms[4 * q + 3])};
        for (int s = 0; s < 4; ++s) slots[at[s]].push_back(base + s);
        slots[at[0]].push_back(base + 4);
        slots[at[3]].push_back(base + 5);
    }
    std::vector<int> gptr(n + 1, 0), gidx;
    for (int i = 0; i < n; ++i) {
        std::sort(slots[i].begin(), slots[i].end());
        gidx.insert(gidx.end(), slots[i].begin(), slots[i].end());
        gptr[i + 1] = (int)gidx.size();
    }
    if (gidx.empty()) gidx.push_back(0);  // no terms: one unread entry to upload
    std::vector<float2> lj;
    if (nd && T > 0) {
        lj.resize((size_t)T * T);
        for (int a = 0; a < T; ++a)
            for (int b = 0; b < T; ++b) {
                const int p = nbparm[a * T + b];
                lj[(size_t)a * T + b] = make_float2(lj_a[p], lj_b[p]);
            }
    }
    int rc;
#define TRY(x)                    \
    if ((rc = (x)) != ISSPA_OK) { \
        bonded_free(m);           \
        delete m;                 \
        return rc;                \
    }
    if (nb) {
        TRY(upload(&m->bond_atoms, reinterpret_cast<const int4 *>(bond_atoms), nb));
        TRY(upload(&m->bond_params, reinterpret_cast<const float2 *>(bond_params), nbt));
    }
    if (na) {
        TRY(upload(&m->angle_atoms, reinterpret_cast<const int4 *>(angle_atoms), na));
        TRY(upload(&m->angle_params, reinterpret_cast<const float2 *>(angle_params), nat));
    }
    if (nd) {
        TRY(upload(&m->dih_atoms, reinterpret_cast<const int4 *>(dih_atoms), nd));
        TRY(upload(&m->dih_types, dih_types, nd));
        TRY(upload(&m->dih_params, reinterpret_cast<const float4 *>(dih_params), ndt));
        TRY(upload(&m->sc14, reinterpret_cast<const float2 *>(sc14), ndt));
        TRY(upload(&m->ityp, ityp, n));
        TRY(upload(&m->lj, lj.data(), lj.size()));
    }
    TRY(dalloc(&m->contrib, (size_t)2 * nb + 3 * na + 6 * nd));
    TRY(dalloc(&m->energy, (size_t)nb + na + 3 * nd));
    TRY(upload(&m->gptr, gptr.data(), gptr.size()));
    TRY(upload(&m->gidx, gidx.data(), gidx.size()));
#undef TRY
    *out = m;
    return ISSPA_OK;
}

extern "C" int bonded_model_destroy(bonded_model *m)
{
    if (!m) return ISSPA_OK;
    bonded_free(m);
    delete m;
    return ISSPA_OK;
}

static BondedArgs bonded_args(bonded_model *m, int terms, const float *xyzq4_d)
{
    BondedArgs a{};
    a.xyzq = reinterpret_cast<const float4 *>(xyzq4_d);
    a.L = m->lbox; a.h = 0.5f * m->lbox;
    a.nb = m->nb; a.na = m->na; a.nd = m->nd;
    a.bond_atoms = m->bond_atoms; a.bond_params = m->bond_params;
    a.angle_atoms = m->angle_atoms; a.angle_params = m->angle_params;
    a.dih_atoms = m->dih_atoms; a.dih_types = m->dih_types; a.dih_params = m->dih_params;
    a.sc14 = m->sc14; a.ityp = m->ityp; a.lj = m->lj; a.T = m->T;
    a.contrib = m->contrib; a.energy = m->energy; a.terms = terms;
    return a;
}

// the per-term kernel: every term's forces into its contribution slots
static int bonded_terms(bonded_model *m, int terms, const float *xyzq4_d, hipStream_t s)
{
    HIPCHK(launch_bonded_terms(bonded_args(m, terms, xyzq4_d), s));
    return ISSPA_OK;
}

extern "C" int bonded_force(bonded_model *m, int terms, const float *xyzq4_d, float *f4_d,
                            double *energies_d, void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "bonded_force: null argument");
    hipStream_t s = S(stream);
    CHK(bonded_terms(m, terms, xyzq4_d, s));
    HIPCHK(launch_bonded_gather(m->contrib, m->gptr, m->gidx, m->n,
                                reinterpret_cast<float4 *>(f4_d), s));
    if (energies_d) HIPCHK(launch_bonded_energy(m->energy, m->nb, m->na, m->nd, energies_d, s));
    return ISSPA_OK;
}

// ---------------------------------------------------------------------------
// integrator
// ---------------------------------------------------------------------------
static uint32_t andersen_threshold(float pnu)
{
    const double t = (double)pnu * 4294967296.0;
    if (t <= 0.0) return 0u;
    if (t >= 4294967295.0) return 4294967295u;
    return (uint32_t)t;
}

extern "C" int leapfrog_andersen(int n, float *xyzq4_d, float *vel4_d, const float *f4_d,
                                 float kT, float dt, float pnu, float lbox, uint64_t seed,
                                 uint32_t system_id, uint32_t step, const uint32_t *step_d,
                                 void *stream)
{
    if (n < 0 || !xyzq4_d || !vel4_d || !f4_d || !(lbox > 0))
        return fail(ISSPA_ERR_ARG, "leapfrog_andersen: bad argument");
    if (!(pnu >= 0.0f && pnu <= 1.0f)) return fail(ISSPA_ERR_ARG, "leapfrog_andersen: pnu=%g not in [0,1]", pnu);
    WinDesc one{};
    one.n = n;
    one.sys = system_id;
    one.L = lbox;
    one.h = 0.5f * lbox;
    HIPCHK(launch_leapfrog(nullptr, nullptr, one, n, reinterpret_cast<float4 *>(xyzq4_d),
                           reinterpret_cast<float4 *>(vel4_d), reinterpret_cast<const float4 *>(f4_d),
                           kT, dt, andersen_threshold(pnu), seed, step, step_d, S(stream)));
    return ISSPA_OK;
}

// ---------------------------------------------------------------------------
// umbrella restraints
// ---------------------------------------------------------------------------
struct us_model {
    int nres = 0, nbias = 0;
    int *boff = nullptr;
    int2 *atoms = nullptr;
    float *mass = nullptr;
    float2 *kx = nullptr;
    float *L = nullptr;
};

static void us_free(us_model *m)
{
    dfree(m->boff); dfree(m->atoms); dfree(m->mass); dfree(m->kx); dfree(m->L);
}

extern "C" int us_model_create_batch(int n_atoms, int nres, const int *boff, const int *atom_list,
                                     const float *mass, const float *k, const float *x0,
                                     const float *lbox, us_model **out)
{
    if (!out || nres < 1 || !boff || !atom_list || !mass || !k || !x0 || !lbox)
        return fail(ISSPA_ERR_ARG, "us_model_create: bad argument");
    *out = nullptr;
    if (boff[0] != 0) return fail(ISSPA_ERR_ARG, "us_model_create: bias offsets must start at 0");
    for (int r = 0; r < nres; ++r) {
        if (boff[r + 1] <= boff[r]) return fail(ISSPA_ERR_ARG, "us_model_create: restraint %d is empty", r);
        if (!(lbox[r] > 0)) return fail(ISSPA_ERR_ARG, "us_model_create: lbox of restraint %d", r);
        std::vector<int> seen;
        int ng[2] = {0, 0};
        for (int b = boff[r]; b < boff[r + 1]; ++b) {
            const int a = atom_list[2 * b], g = atom_list[2 * b + 1];
            if (a < 0 || a >= n_atoms || (g != 0 && g != 1))
                return fail(ISSPA_ERR_ARG, "us_model_create: entry %d = {%d,%d} invalid", b, a, g);
            seen.push_back(a);
            ng[g]++;
        }
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            return fail(ISSPA_ERR_ARG, "us_model_create: an atom appears twice in restraint %d", r);
        if (!ng[0] || !ng[1]) return fail(ISSPA_ERR_ARG, "us_model_create: restraint %d needs both groups", r);
    }
    us_model *m = new (std::nothrow) us_model;
    if (!m) return fail(ISSPA_ERR_ALLOC, "host allocation");
    m->nres = nres;
    m->nbias = boff[nres];
    std::vector<float2> kx(nres);
    for (int r = 0; r < nres; ++r) kx[r] = make_float2(k[r], x0[r]);
    int rc;
    if ((rc = upload(&m->boff, boff, (size_t)nres + 1)) != ISSPA_OK ||
        (rc = upload(&m->atoms, reinterpret_cast<const int2 *>(atom_list), m->nbias)) != ISSPA_OK ||
        (rc = upload(&m->mass, mass, m->nbias)) != ISSPA_OK ||
        (rc = upload(&m->kx, kx.data(), kx.size())) != ISSPA_OK ||
        (rc = upload(&m->L, lbox, nres)) != ISSPA_OK) {
        us_free(m);
        delete m;
        return rc;
    }
    *out = m;
    return ISSPA_OK;
}

extern "C" int us_model_create(int n_atoms, int nbias, const int *atom_list, const float *mass,
                               const float *kumb, float x0, float lbox, us_model **out)
{
    if (!kumb || !out) return fail(ISSPA_ERR_ARG, "us_model_create: bad argument");
    if (kumb[0] != -kumb[1])
        return fail(ISSPA_ERR_NOTSUP, "us_model_create: kumb must be {-k, +k} (us::populate_mass)");
    const int boff[2] = {0, nbias};
    return us_model_create_batch(n_atoms, 1, boff, atom_list, mass, &kumb[1], &x0, &lbox, out);
}

extern "C" int us_model_destroy(us_model *m)
{
    if (!m) return ISSPA_OK;
    us_free(m);
    delete m;
    return ISSPA_OK;
}

extern "C" int us_force(us_model *m, const float *xyzq4_d, float *f4_d, double *com_d,
                        void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "us_force: null argument");
    HIPCHK(launch_us(m->nres, m->boff, m->atoms, m->mass, m->kx, m->L,
                     reinterpret_cast<const float4 *>(xyzq4_d), reinterpret_cast<float4 *>(f4_d),
                     com_d, S(stream)));
    return ISSPA_OK;
}

// ---------------------------------------------------------------------------
// MD engine: one step = zero f, bonded, IS-SPA, nonbond, US, leap-frog, step++
// (order of src/total_force_cuda.cpp:109-153), as five launches: prologue,
// field (+ all-pairs below kForkMinAtoms), epilogue, force, assembly (which
// also integrates and advances the step); see engine_forces.
// ---------------------------------------------------------------------------
struct md_engine {
    isspa_model *isspa = nullptr;
    nonbond_model *nb = nullptr;
    bonded_model *bonded = nullptr;
    us_model *us = nullptr;
    int n = 0;
    float *x = nullptr, *v = nullptr, *f = nullptr;
    float kT = 0.f, dt = 0.f, pnu = 0.f, lbox = 0.f;
    uint64_t seed = 0;
    uint32_t sys = 0;
    WinDesc *wins = nullptr;   // integrator windows (owned)
    int *atom_win = nullptr;
    uint32_t *step_d = nullptr;  // [2]: step, leap-frog blocks done
    double *com_d = nullptr;
    float4 *fus = nullptr;       // restraint bias forces of the step (w = 1: restrained)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned captured_generation = 0;
    // recorded after every replay: a recapture or destroy waits on it before
    // freeing the executable graph, which may still be running
    hipEvent_t ev_replayed = nullptr;
    bool replay_pending = false;
    // the all-pairs kernel runs on a side stream beside the IS-SPA passes
    // (default priority; fork and join by events, which a capture records as
    // graph edges); its chunk reduction runs after the join
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

constexpr int kForkMinAtoms = 4096;

// Destroys the executable graph and its graph once the last replay has ended
// (a replay still in flight reads the executable's kernel arguments).
static int engine_drop_graph(md_engine *e)
{
    hipError_t he = hipSuccess;
    if (e->replay_pending && e->ev_replayed) he = hipEventSynchronize(e->ev_replayed);
    e->replay_pending = false;
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    if (he != hipSuccess) return fail(ISSPA_ERR_HIP, "waiting for the last replay: %s", hipGetErrorString(he));
    return ISSPA_OK;
}

static void engine_free(md_engine *e)
{
    (void)engine_drop_graph(e);
    if (e->ev_replayed) (void)hipEventDestroy(e->ev_replayed);
    e->ev_replayed = nullptr;
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    if (e->ev_join) (void)hipEventDestroy(e->ev_join);
    e->side = nullptr;
    e->ev_fork = e->ev_join = nullptr;
    dfree(e->step_d); dfree(e->com_d); dfree(e->fus); dfree(e->wins); dfree(e->atom_win);
}

extern "C" int md_engine_create(isspa_model *isspa, nonbond_model *nb, bonded_model *bonded,
                                us_model *us, int n, float *x, float *v, float *f, float kT,
                                float dt, float pnu, float lbox, uint64_t seed, uint32_t sys,
                                md_engine **out)
{
    if (!out || n < 1 || !x || !v || !f) return fail(ISSPA_ERR_ARG, "md_engine_create: bad argument");
    if (isspa && isspa->d.n_atoms != n) return fail(ISSPA_ERR_ARG, "md_engine_create: isspa model has %d atoms, engine %d", isspa->d.n_atoms, n);
    if (nb && nb->n != n) return fail(ISSPA_ERR_ARG, "md_engine_create: nonbond model atom count mismatch");
    if (bonded && bonded->n != n) return fail(ISSPA_ERR_ARG, "md_engine_create: bonded model atom count mismatch");
    if (!(pnu >= 0.0f && pnu <= 1.0f)) return fail(ISSPA_ERR_ARG, "md_engine_create: pnu not in [0,1]");
    if (isspa && nb && isspa->lay.W != nb->lay.W)
        return fail(ISSPA_ERR_ARG, "md_engine_create: IS-SPA and nonbond window counts differ");
    md_engine *e = new (std::nothrow) md_engine;
    if (!e) return fail(ISSPA_ERR_ALLOC, "host allocation");
    e->isspa = isspa; e->nb = nb; e->bonded = bonded; e->us = us; e->n = n;
    e->x = x; e->v = v; e->f = f; e->kT = kT; e->dt = dt; e->pnu = pnu; e->lbox = lbox;
    e->seed = seed; e->sys = sys;
    // integrator windows: those of the IS-SPA (else nonbond) model, else one
    Windows lay;
    int rc = ISSPA_OK;
    if (isspa) lay = isspa->lay;
    else if (nb) lay = nb->lay;
    else rc = build_windows(n, nullptr, sys, lbox, lay);
    std::vector<WinDesc> wd;
    for (int w = 0; rc == ISSPA_OK && w < lay.W; ++w) wd.push_back(base_desc(lay, w));
    if (rc == ISSPA_OK) rc = upload(&e->wins, wd.data(), wd.size());
    if (rc == ISSPA_OK) rc = upload(&e->atom_win, lay.atom_win.data(), lay.atom_win.size());
    // {step, leap-frog blocks done}: the leap-frog launch advances the step
    if (rc == ISSPA_OK) rc = dalloc(&e->step_d, 2);
    if (rc == ISSPA_OK) rc = dalloc(&e->com_d, 6 * (size_t)(us ? us->nres : 1));
    if (rc == ISSPA_OK && us) {
        // entries of atoms no restraint holds stay zero (w = 0) for good
        rc = dalloc(&e->fus, (size_t)n);
        if (rc == ISSPA_OK) {
            hipError_t he = hipMemset(e->fus, 0, (size_t)n * sizeof(float4));
            if (he != hipSuccess) rc = fail(ISSPA_ERR_HIP, "hipMemset: %s", hipGetErrorString(he));
        }
    }
    if (rc == ISSPA_OK) {
        hipError_t he = hipMemset(e->step_d, 0, 2 * sizeof(uint32_t));
        if (he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_replayed, hipEventDisableTiming);
        if (he != hipSuccess) rc = fail(ISSPA_ERR_HIP, "md_engine_create: %s", hipGetErrorString(he));
    }
    // below a few thousand atoms the all-pairs kernel is a few microseconds and
    // the fork and join would cost more than the overlap saves
    if (rc == ISSPA_OK && isspa && nb && n >= kForkMinAtoms) {
        hipError_t he = hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming);
        if (he != hipSuccess) rc = fail(ISSPA_ERR_HIP, "side stream: %s", hipGetErrorString(he));
    }
    if (rc != ISSPA_OK) {
        engine_free(e);
        delete e;
        return rc;
    }
    *out = e;
    return ISSPA_OK;
}

extern "C" int md_engine_destroy(md_engine *e)
{
    if (!e) return ISSPA_OK;
    engine_free(e);
    delete e;
    return ISSPA_OK;
}

extern "C" int md_engine_set_step(md_engine *e, uint32_t step, void *stream)
{
    if (!e) return fail(ISSPA_ERR_ARG, "null engine");
    HIPCHK(launch_set_word(e->step_d, step, S(stream)));
    return ISSPA_OK;
}

// One prologue launch computes everything that depends only on the
// coordinates (IS-SPA minimum-image flags, bonded terms, restraint forces and
// centres of mass, MC points); the IS-SPA passes and the all-pairs partials
// follow, and one assembly launch forms f = bonded + IS-SPA + all-pairs +
// restraint -- bit for bit what zeroing f and calling bonded_force,
// isspa_model_force, nonbond_force and us_force in that order give.
static int engine_forces(md_engine *e, hipStream_t s, bool integrate)
{
    const float4 *x4 = reinterpret_cast<const float4 *>(e->x);
    const bool fork = e->side != nullptr;
    if (fork) {
        // all-pairs partials beside the IS-SPA passes: they read x and write
        // only the nonbond model's own buffers
        HIPCHK(hipEventRecord(e->ev_fork, s));
        HIPCHK(hipStreamWaitEvent(e->side, e->ev_fork, 0));
        CHK(nonbond_partials(e->nb, e->x, false, e->side));
        HIPCHK(hipEventRecord(e->ev_join, e->side));
    }
    PrologueArgs p{};
    AssembleArgs a{};
    a.n = e->n;
    a.f = reinterpret_cast<float4 *>(e->f);
    isspa_model *m = e->isspa;
    if (m) {
        p.nflag = m->lay.W;
        p.rmax_all = m->rmax_all;
        p.flag = m->pbc_flag;
        p.pts = PointsArgs{m->wins, m->atom_win, x4, m->itype, m->rmax, m->rmax2, m->d.n_atoms,
                           m->d.n_mc, m->d.seed, 0u, e->step_d, m->pts};
    }
    if (e->bonded) {
        p.bonded = bonded_args(e->bonded, ISSPA_TERM_ALL, e->x);
        a.contrib = e->bonded->contrib;
        a.gptr = e->bonded->gptr;
        a.gidx = e->bonded->gidx;
    }
    if (e->us) {
        p.nres = e->us->nres;
        p.boff = e->us->boff;
        p.us_atoms = e->us->atoms;
        p.us_mass = e->us->mass;
        p.kx = e->us->kx;
        p.Ls = e->us->L;
        p.fus = e->fus;
        p.com = e->com_d;
        a.fus = e->fus;
    }
    HIPCHK(launch_step_prologue(p, s));
    // without the side stream the all-pairs blocks ride on the field launch,
    // reusing the prologue's IS-SPA minimum-image flags (the field kernel
    // carries partner-chunk blocks only: windows large enough for super-tiles
    // always have the side stream)
    const bool nb_in_field = m && e->nb && !fork && !e->nb->any_sym;
    if (m) {
        FieldArgs fa = field_args(m, x4);
        if (nb_in_field) {
            fa.nb = nonbond_args(e->nb, x4, m->pbc_flag, false);
            fa.nb_blocks = e->nb->nblk_tot;
        }
        HIPCHK(launch_field(fa, m->ngblk, s));
        const ForceArgs fo = force_args(m, x4);
        if (!fo.dtypes && force_epi_fits(fo.tab, m->max_win_atoms, m->max_chunk_pts, m->npblk)) {
            // small windows: the epilogue inside the force launch
            HIPCHK(launch_force_epi(fo, epilogue_args(m, false), m->npblk, s));
        } else {
            HIPCHK(launch_epilogue(epilogue_args(m, false), s));
            HIPCHK(launch_force(fo, m->npblk, false, s));
        }
        a.isspa = reduce_args(m, x4, nullptr, nullptr, nullptr, nullptr);
    }
    if (fork)
        HIPCHK(hipStreamWaitEvent(s, e->ev_join, 0));
    else if (e->nb && !nb_in_field)
        CHK(nonbond_partials(e->nb, e->x, false, s));
    if (e->nb) {
        a.nb_wins = e->nb->wins;
        a.nb_atom_win = e->nb->atom_win;
        a.nb_part = e->nb->part;
        a.nb_B = e->nb->B_tot;
    }
    if (integrate) {
        // the assembly launch also integrates each atom and advances the step
        a.lf.wins = e->wins;
        a.lf.atom_win = e->atom_win;
        a.lf.xyzq = reinterpret_cast<float4 *>(e->x);
        a.lf.vel = reinterpret_cast<float4 *>(e->v);
        a.lf.kT = e->kT;
        a.lf.dt = e->dt;
        a.lf.thr = andersen_threshold(e->pnu);
        a.lf.seed = e->seed;
        a.lf.step_d = e->step_d;
        a.lf.advance = e->step_d;
    }
    HIPCHK(launch_assemble_forces(a, s));
    return ISSPA_OK;
}

static int engine_step(md_engine *e, hipStream_t s) { return engine_forces(e, s, true); }

extern "C" int md_engine_forces(md_engine *e, void *stream)
{
    if (!e) return fail(ISSPA_ERR_ARG, "null engine");
    return engine_forces(e, S(stream), false);
}

extern "C" int md_engine_step(md_engine *e, void *stream)
{
    if (!e) return fail(ISSPA_ERR_ARG, "null engine");
    return engine_step(e, S(stream));
}

extern "C" int md_engine_get_com(md_engine *e, double *com_d, void *stream)
{
    if (!e || !com_d) return fail(ISSPA_ERR_ARG, "md_engine_get_com: bad argument");
    const size_t k = 6 * (size_t)(e->us ? e->us->nres : 1);
    HIPCHK(hipMemcpyAsync(com_d, e->com_d, k * sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
    return ISSPA_OK;
}

extern "C" int md_engine_capture(md_engine *e, int steps_per_graph, void *stream)
{
    if (!e || steps_per_graph < 1) return fail(ISSPA_ERR_ARG, "md_engine_capture: bad argument");
    if (!stream) return fail(ISSPA_ERR_ARG, "md_engine_capture: needs a non-null stream");
    hipStream_t s = S(stream);
    CHK(engine_drop_graph(e));
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = ISSPA_OK;
    for (int k = 0; k < steps_per_graph && rc == ISSPA_OK; ++k) rc = engine_step(e, s);
    hipGraph_t g = nullptr;
    hipError_t he = hipStreamEndCapture(s, &g);
    if (rc != ISSPA_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (he != hipSuccess) return fail(ISSPA_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(he));
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    e->captured_generation = e->isspa ? e->isspa->generation : 0;
    return ISSPA_OK;
}

extern "C" int md_engine_replay(md_engine *e, int replays, void *stream)
{
    if (!e || replays < 0) return fail(ISSPA_ERR_ARG, "md_engine_replay: bad argument");
    if (!e->exec) return fail(ISSPA_ERR_STATE, "md_engine_replay: no captured graph");
    if (e->isspa && e->isspa->generation != e->captured_generation)
        return fail(ISSPA_ERR_STATE, "md_engine_replay: the IS-SPA tiling changed after capture; capture again");
    for (int r = 0; r < replays; ++r) HIPCHK(hipGraphLaunch(e->exec, S(stream)));
    if (replays > 0) {
        HIPCHK(hipEventRecord(e->ev_replayed, S(stream)));
        e->replay_pending = true;
    }
    return ISSPA_OK;
}
