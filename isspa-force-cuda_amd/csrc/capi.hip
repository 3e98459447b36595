// C-ABI implementation (include/isspa_amd.h): model objects, device buffers,
// launch geometry, and the hipGraph-captured MD step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/isspa_amd.h"
#include "kernels.hpp"

using namespace isspa;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(ISSPA_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                            \
    } while (0)

#define CHK(expr)                 \
    do {                          \
        int r_ = (expr);          \
        if (r_ != ISSPA_OK) return r_; \
    } while (0)

extern "C" const char *isspa_last_error(void) { return g_err.c_str(); }
extern "C" int isspa_abi_version(void) { return 1; }

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

template <class T>
static int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess)
        return fail(ISSPA_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T),
                    hipGetErrorString(e));
    return ISSPA_OK;
}

template <class T>
static int upload(T **p, const T *h, size_t count)
{
    CHK(dalloc(p, count));
    if (count) HIPCHK(hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice));
    return ISSPA_OK;
}

template <class T>
static void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// ---------------------------------------------------------------------------
// device memory / streams
// ---------------------------------------------------------------------------
extern "C" int isspa_dev_malloc(void **ptr_d, size_t bytes)
{
    if (!ptr_d) return fail(ISSPA_ERR_ARG, "isspa_dev_malloc: null out pointer");
    return dalloc((unsigned char **)ptr_d, bytes);
}
extern "C" int isspa_dev_free(void *ptr_d)
{
    if (ptr_d) HIPCHK(hipFree(ptr_d));
    return ISSPA_OK;
}
extern "C" int isspa_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_zero(void *dst, size_t bytes, void *stream)
{
    HIPCHK(launch_zero(dst, bytes, S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_add4(float *dst4_d, const float *src4_d, int n, void *stream)
{
    if (n < 0 || (n > 0 && (!dst4_d || !src4_d))) return fail(ISSPA_ERR_ARG, "isspa_add4: bad argument");
    HIPCHK(launch_add4(reinterpret_cast<float4 *>(dst4_d), reinterpret_cast<const float4 *>(src4_d),
                       n, S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_stream_create(void **stream)
{
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return ISSPA_OK;
}
extern "C" int isspa_stream_destroy(void *stream)
{
    if (stream) HIPCHK(hipStreamDestroy(S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_stream_sync(void *stream)
{
    HIPCHK(hipStreamSynchronize(S(stream)));
    return ISSPA_OK;
}

struct timer_pair {
    hipEvent_t a, b;
};
extern "C" int isspa_timer_begin(void **timer, void *stream)
{
    timer_pair *t = new (std::nothrow) timer_pair;
    if (!t) return fail(ISSPA_ERR_ALLOC, "timer allocation");
    HIPCHK(hipEventCreate(&t->a));
    HIPCHK(hipEventCreate(&t->b));
    HIPCHK(hipEventRecord(t->a, S(stream)));
    *timer = t;
    return ISSPA_OK;
}
extern "C" int isspa_timer_end(void *timer, void *stream, float *ms)
{
    timer_pair *t = (timer_pair *)timer;
    if (!t) return fail(ISSPA_ERR_ARG, "null timer");
    HIPCHK(hipEventRecord(t->b, S(stream)));
    HIPCHK(hipEventSynchronize(t->b));
    HIPCHK(hipEventElapsedTime(ms, t->a, t->b));
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
    return ISSPA_OK;
}

extern "C" int isspa_default_constants(float *e0, float *p0, float *q0, float *o0, float *ep)
{
    // src/constants.h:6-13, evaluated in FP32 (volatile keeps the compiler from
    // folding in higher precision)
    volatile float pi = 3.1415926535f, rho = 0.0075f, epv = 2.3473f;
    volatile float p0v = 4.173705894203977f;
    const float e0v = 3.0f / (4.0f * pi * rho * p0v) * (1.0f - 1.0f / epv);
    if (e0) *e0 = e0v;
    if (p0) *p0 = p0v;
    if (q0) *q0 = 1.6899960102882601f;
    if (o0) *o0 = 9.3658298730779350f;
    if (ep) *ep = epv;
    return ISSPA_OK;
}

// launch geometry: enough waves to cover the 256 CUs several times over
static constexpr long kTargetWaves = 8192;

static int auto_splits(long lanes, long work, long min_work)
{
    const long waves = std::max(1L, (lanes + 63) / 64);
    long s = (kTargetWaves + waves - 1) / waves;
    s = std::min(s, std::max(1L, work / min_work));
    return (int)std::max(1L, s);
}

// ---------------------------------------------------------------------------
// IS-SPA model
// ---------------------------------------------------------------------------
struct isspa_model {
    isspa_model_desc d{};
    int P = 0;
    float rmax_all = 0.f;
    int *itype = nullptr;
    float *rmax = nullptr, *rmax2 = nullptr, *vtot = nullptr;
    float *g = nullptr, *e = nullptr, *f = nullptr;
    float4 *pts = nullptr;
    float *rec = nullptr;
    double *g_out = nullptr, *E_out = nullptr, *E0_out = nullptr;
    int *cnt_out = nullptr;
    int *pbc_flag = nullptr;
    // tiling and partial buffers
    int splits = 0, chunks = 0, atoms_per_split = 0, pts_per_chunk = 0;
    int user_splits = 0, user_chunks = 0;
    float *part_g = nullptr;
    int *part_cnt = nullptr;
    double *part_s0 = nullptr, *part_ein = nullptr;
    double *part_f = nullptr;
    size_t cap_field = 0, cap_force = 0;
    float4 *prof_f = nullptr;  // force sink of isspa_model_profile
    TableView tv{};
};

static int isspa_tile(isspa_model *m)
{
    const int n = m->d.n_atoms, P = m->P;
    int s = m->user_splits > 0 ? m->user_splits : auto_splits(P, n, 64);
    s = std::max(1, std::min(s, std::max(1, n)));
    m->atoms_per_split = (n + s - 1) / s;
    m->splits = (n + m->atoms_per_split - 1) / m->atoms_per_split;
    int c = m->user_chunks > 0 ? m->user_chunks : auto_splits(n, P, 128);
    c = std::max(1, std::min(c, std::max(1, P)));
    m->pts_per_chunk = (P + c - 1) / c;
    m->chunks = (P + m->pts_per_chunk - 1) / m->pts_per_chunk;
    const size_t need_field = (size_t)m->splits * P;
    if (need_field > m->cap_field) {
        dfree(m->part_g);
        dfree(m->part_cnt);
        dfree(m->part_s0);
        dfree(m->part_ein);
        CHK(dalloc(&m->part_g, need_field));
        CHK(dalloc(&m->part_cnt, need_field));
        CHK(dalloc(&m->part_s0, 3 * need_field));
        CHK(dalloc(&m->part_ein, 3 * need_field));
        m->cap_field = need_field;
    }
    const size_t need_force = (size_t)m->chunks * 9 * n;
    if (need_force > m->cap_force) {
        dfree(m->part_f);
        CHK(dalloc(&m->part_f, need_force));
        m->cap_force = need_force;
    }
    return ISSPA_OK;
}

static void isspa_free(isspa_model *m)
{
    dfree(m->itype); dfree(m->rmax); dfree(m->rmax2); dfree(m->vtot);
    dfree(m->g); dfree(m->e); dfree(m->f);
    dfree(m->pts); dfree(m->rec); dfree(m->g_out); dfree(m->E_out); dfree(m->E0_out);
    dfree(m->cnt_out); dfree(m->pbc_flag);
    dfree(m->part_g); dfree(m->part_cnt); dfree(m->part_s0); dfree(m->part_ein); dfree(m->part_f);
    dfree(m->prof_f);
}

static int pad_table(const float *src, int T, int n, std::vector<float> &out)
{
    out.assign((size_t)T * (n + 1), 0.0f);
    for (int t = 0; t < T; ++t)
        for (int b = 0; b < n; ++b) out[(size_t)t * (n + 1) + b] = src[(size_t)t * n + b];
    return n + 1;
}

extern "C" int isspa_model_create(const isspa_model_desc *d, const int *isspa_type,
                                  const float *rmax, const float *g, const float *e,
                                  const float *f, isspa_model **out)
{
    if (!d || !out || !isspa_type || !rmax || !g || !e || !f)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: null argument");
    *out = nullptr;
    if (d->n_atoms < 1 || d->n_types < 1 || d->n_mc < 1)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: n_atoms=%d n_types=%d n_mc=%d must be >= 1",
                    d->n_atoms, d->n_types, d->n_mc);
    if ((long)d->n_atoms * d->n_mc > 0x7fffffffL)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: n_atoms*n_mc exceeds 2^31");
    if (d->n_g < 2 || d->n_e < d->n_g || d->n_f < 2)
        return fail(ISSPA_ERR_ARG, "isspa_model_create: table rows n_g=%d n_e=%d n_f=%d (need n_g>=2, n_e>=n_g, n_f>=2)",
                    d->n_g, d->n_e, d->n_f);
    if (!(d->g_dr > 0 && d->e_dr > 0 && d->f_dr > 0 && d->lbox > 0))
        return fail(ISSPA_ERR_ARG, "isspa_model_create: grid spacings and lbox must be > 0");
    for (int i = 0; i < d->n_atoms; ++i)
        if (isspa_type[i] < 0 || isspa_type[i] >= d->n_types)
            return fail(ISSPA_ERR_ARG, "isspa_model_create: isspa_type[%d]=%d out of range [0,%d)",
                        i, isspa_type[i], d->n_types);
    isspa_model *m = new (std::nothrow) isspa_model;
    if (!m) return fail(ISSPA_ERR_ALLOC, "isspa_model_create: host allocation");
    m->d = *d;
    m->P = d->n_atoms * d->n_mc;
    const int T = d->n_types;
    std::vector<float> r2(T), vt(T);
    const float pi = 3.1415926535f;  // src/constants.h:6
    for (int t = 0; t < T; ++t) {
        if (!(rmax[t] > 0)) {
            delete m;
            return fail(ISSPA_ERR_ARG, "isspa_model_create: rmax[%d]=%g must be > 0", t, rmax[t]);
        }
        r2[t] = rmax[t] * rmax[t];
        // src/isspa_class.cpp:110-112, double arithmetic stored as float
        vt[t] = (float)(4.0 / 3.0 * (double)pi * rmax[t] * rmax[t] * rmax[t] * 0.0074 /
                        (double)(float)d->n_mc);
        m->rmax_all = std::max(m->rmax_all, rmax[t]);
    }
    std::vector<float> gp, ep, fp;
    int ldg = pad_table(g, T, d->n_g, gp);
    int lde = pad_table(e, T, d->n_e, ep);
    int ldf = pad_table(f, T, d->n_f, fp);
    int rc = ISSPA_OK;
#define TRY(x)                      \
    if ((rc = (x)) != ISSPA_OK) {   \
        isspa_free(m);              \
        delete m;                   \
        return rc;                  \
    }
    TRY(upload(&m->itype, isspa_type, d->n_atoms));
    TRY(upload(&m->rmax, rmax, T));
    TRY(upload(&m->rmax2, r2.data(), T));
    TRY(upload(&m->vtot, vt.data(), T));
    TRY(upload(&m->g, gp.data(), gp.size()));
    TRY(upload(&m->e, ep.data(), ep.size()));
    TRY(upload(&m->f, fp.data(), fp.size()));
    TRY(dalloc(&m->pts, m->P));
    TRY(dalloc(&m->rec, (size_t)m->P * kRecFloats));
    TRY(dalloc(&m->g_out, m->P));
    TRY(dalloc(&m->cnt_out, m->P));
    TRY(dalloc(&m->E_out, 3 * (size_t)m->P));
    TRY(dalloc(&m->E0_out, 3 * (size_t)m->P));
    TRY(dalloc(&m->pbc_flag, 1));
    TRY(isspa_tile(m));
#undef TRY
    TableView &tv = m->tv;
    tv.g = m->g; tv.e = m->e; tv.f = m->f;
    tv.n_g = d->n_g; tv.n_e = d->n_e; tv.n_f = d->n_f;
    tv.ld_g = ldg; tv.ld_e = lde; tv.ld_f = ldf;
    tv.g_r0 = d->g_r0; tv.g_dr = d->g_dr; tv.g_inv = 1.0f / d->g_dr;
    tv.e_r0 = d->e_r0; tv.e_dr = d->e_dr; tv.e_inv = 1.0f / d->e_dr;
    tv.f_r0 = d->f_r0; tv.f_dr = d->f_dr; tv.f_inv = 1.0f / d->f_dr;
    *out = m;
    return ISSPA_OK;
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

extern "C" int isspa_model_get_tiling(const isspa_model *m, int *field_splits, int *force_chunks)
{
    if (!m) return fail(ISSPA_ERR_ARG, "null model");
    if (field_splits) *field_splits = m->splits;
    if (force_chunks) *force_chunks = m->chunks;
    return ISSPA_OK;
}

// points + field + epilogue (shared by the total and component entry points)
static int isspa_field_pass(isspa_model *m, const float *xyzq4_d, uint32_t step,
                            const uint32_t *step_d, hipStream_t s)
{
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    const isspa_model_desc &d = m->d;
    HIPCHK(launch_pbc_flag(x, d.n_atoms, m->rmax_all, d.lbox, m->pbc_flag, s));
    HIPCHK(launch_mc_points(x, m->itype, m->rmax, m->rmax2, d.n_atoms, d.n_mc, d.seed,
                            d.system_id, step, step_d, m->pts, s));
    FieldArgs fa;
    fa.pts = m->pts; fa.xyzq = x; fa.itype = m->itype; fa.tab = m->tv;
    fa.n = d.n_atoms; fa.P = m->P; fa.atoms_per_split = m->atoms_per_split;
    fa.L = d.lbox; fa.h = 0.5f * d.lbox; fa.e0 = d.e0; fa.pbc_flag = m->pbc_flag;
    fa.part_g = m->part_g; fa.part_cnt = m->part_cnt; fa.part_s0 = m->part_s0;
    fa.part_ein = m->part_ein;
    HIPCHK(launch_field(fa, m->splits, s));
    EpilogueArgs ea;
    ea.pts = m->pts; ea.P = m->P; ea.splits = m->splits;
    ea.part_g = m->part_g; ea.part_cnt = m->part_cnt; ea.part_s0 = m->part_s0;
    ea.part_ein = m->part_ein;
    ea.e0 = d.e0; ea.p0 = d.p0; ea.q0 = d.q0; ea.o0 = d.o0;
    ea.rec = m->rec; ea.g_out = m->g_out; ea.cnt_out = m->cnt_out; ea.E_out = m->E_out;
    ea.E0_out = m->E0_out;
    HIPCHK(launch_epilogue(ea, s));
    return ISSPA_OK;
}

static int isspa_force_pass(isspa_model *m, const float *xyzq4_d, bool split, float4 *f,
                            double *flj, double *fpol, double *fcdd, hipStream_t s)
{
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    const isspa_model_desc &d = m->d;
    ForceArgs fa;
    fa.rec4 = reinterpret_cast<const float4 *>(m->rec);
    fa.xyzq = x; fa.itype = m->itype; fa.rmax2 = m->rmax2; fa.tab = m->tv;
    fa.n = d.n_atoms; fa.P = m->P; fa.pts_per_chunk = m->pts_per_chunk;
    fa.L = d.lbox; fa.h = 0.5f * d.lbox; fa.pbc_flag = m->pbc_flag; fa.part = m->part_f;
    HIPCHK(launch_force(fa, m->chunks, split, s));
    ReduceArgs ra;
    ra.part = m->part_f; ra.n = d.n_atoms; ra.chunks = m->chunks; ra.xyzq = x;
    ra.itype = m->itype; ra.vtot = m->vtot; ra.f = f; ra.f_lj = flj; ra.f_pol = fpol;
    ra.f_cdd = fcdd;
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

extern "C" int isspa_model_force_components(isspa_model *m, const float *xyzq4_d,
                                            double *f_lj_d, double *f_pol_d, double *f_cdd_d,
                                            uint32_t step, void *stream)
{
    if (!m || !xyzq4_d || !f_lj_d || !f_pol_d || !f_cdd_d)
        return fail(ISSPA_ERR_ARG, "isspa_model_force_components: null argument");
    CHK(isspa_field_pass(m, xyzq4_d, step, nullptr, S(stream)));
    return isspa_force_pass(m, xyzq4_d, true, nullptr, f_lj_d, f_pol_d, f_cdd_d, S(stream));
}

extern "C" int isspa_model_profile(isspa_model *m, const float *xyzq4_d, int reps, float *ms,
                                   void *stream)
{
    if (!m || !xyzq4_d || !ms || reps < 1) return fail(ISSPA_ERR_ARG, "isspa_model_profile: bad argument");
    hipStream_t s = S(stream);
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    const isspa_model_desc &d = m->d;
    if (!m->prof_f) CHK(dalloc(&m->prof_f, (size_t)d.n_atoms));
    hipEvent_t ev[2];
    HIPCHK(hipEventCreate(&ev[0]));
    HIPCHK(hipEventCreate(&ev[1]));
    HIPCHK(launch_pbc_flag(x, d.n_atoms, m->rmax_all, d.lbox, m->pbc_flag, s));
    int rc = ISSPA_OK;
    for (int stage = 0; stage < 5 && rc == ISSPA_OK; ++stage) {
        // one untimed launch, then `reps` timed back-to-back launches
        for (int r = 0; r <= reps && rc == ISSPA_OK; ++r) {
            if (r == 1) HIPCHK(hipEventRecord(ev[0], s));
            hipError_t e = hipSuccess;
            if (stage == 0) {
                e = launch_mc_points(x, m->itype, m->rmax, m->rmax2, d.n_atoms, d.n_mc, d.seed,
                                     d.system_id, 0u, nullptr, m->pts, s);
            } else if (stage == 1 || stage == 2) {
                FieldArgs fa;
                fa.pts = m->pts; fa.xyzq = x; fa.itype = m->itype; fa.tab = m->tv;
                fa.n = d.n_atoms; fa.P = m->P; fa.atoms_per_split = m->atoms_per_split;
                fa.L = d.lbox; fa.h = 0.5f * d.lbox; fa.e0 = d.e0; fa.pbc_flag = m->pbc_flag;
                fa.part_g = m->part_g; fa.part_cnt = m->part_cnt; fa.part_s0 = m->part_s0;
                fa.part_ein = m->part_ein;
                EpilogueArgs ea;
                ea.pts = m->pts; ea.P = m->P; ea.splits = m->splits;
                ea.part_g = m->part_g; ea.part_cnt = m->part_cnt; ea.part_s0 = m->part_s0;
                ea.part_ein = m->part_ein;
                ea.e0 = d.e0; ea.p0 = d.p0; ea.q0 = d.q0; ea.o0 = d.o0;
                ea.rec = m->rec; ea.g_out = m->g_out; ea.cnt_out = m->cnt_out;
                ea.E_out = m->E_out; ea.E0_out = m->E0_out;
                e = stage == 1 ? launch_field(fa, m->splits, s) : launch_epilogue(ea, s);
            } else {
                ForceArgs fa;
                fa.rec4 = reinterpret_cast<const float4 *>(m->rec);
                fa.xyzq = x; fa.itype = m->itype; fa.rmax2 = m->rmax2; fa.tab = m->tv;
                fa.n = d.n_atoms; fa.P = m->P; fa.pts_per_chunk = m->pts_per_chunk;
                fa.L = d.lbox; fa.h = 0.5f * d.lbox; fa.pbc_flag = m->pbc_flag;
                fa.part = m->part_f;
                ReduceArgs ra;
                ra.part = m->part_f; ra.n = d.n_atoms; ra.chunks = m->chunks; ra.xyzq = x;
                ra.itype = m->itype; ra.vtot = m->vtot; ra.f = m->prof_f; ra.f_lj = nullptr;
                ra.f_pol = nullptr; ra.f_cdd = nullptr;
                e = stage == 3 ? launch_force(fa, m->chunks, false, s)
                               : launch_force_reduce(ra, false, s);
            }
            if (e != hipSuccess)
                rc = fail(ISSPA_ERR_HIP, "isspa_model_profile stage %d: %s", stage, hipGetErrorString(e));
        }
        if (rc != ISSPA_OK) break;
        HIPCHK(hipEventRecord(ev[1], s));
        HIPCHK(hipEventSynchronize(ev[1]));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, ev[0], ev[1]));
        ms[stage] = t / (float)reps;
    }
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
struct nonbond_model {
    int n = 0, T = 0;
    float lbox = 0.f, cfar = 0.f;
    float4 *aux = nullptr;
    float2 *lj = nullptr;
    uint32_t *excl = nullptr;
    int *excl_ptr = nullptr, *excl_idx = nullptr;
    int *pbc_flag = nullptr;
    int chunks = 0, atoms_per_chunk = 0;
    double *part = nullptr;
};

static void nonbond_free(nonbond_model *m)
{
    dfree(m->aux); dfree(m->lj); dfree(m->excl); dfree(m->excl_ptr); dfree(m->excl_idx);
    dfree(m->pbc_flag); dfree(m->part);
}

extern "C" int nonbond_model_create(int n, int T, const int *ityp, const int *nbparm,
                                    const float *lj_a, const float *lj_b, const int *pre,
                                    const int *list, int L, const int *isspa_type, int Ti,
                                    const float *rmax, float ep, float lbox, nonbond_model **out)
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
    m->n = n; m->T = T; m->lbox = lbox;
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
        // every hi > lo that the scan would find: entries in ascending walk
        // order up to the first entry larger than the partner
        for (int e = start; e < stop; ++e) {
            const int hi = list[e];
            if (hi <= lo || hi >= n) continue;
            bool found = false;
            for (int k = start; k < stop; ++k) {
                if (list[k] == hi) { found = true; break; }
                if (list[k] > hi) break;
            }
            if (found) {
                ex[lo].push_back(hi);
                ex[hi].push_back(lo);
            }
        }
    }
    std::vector<uint32_t> win((size_t)4 * n, 0u);
    std::vector<int> fptr(n + 1, 0), fidx;
    bool fallback = false;
    for (int i = 0; i < n; ++i) {
        std::sort(ex[i].begin(), ex[i].end());
        ex[i].erase(std::unique(ex[i].begin(), ex[i].end()), ex[i].end());
        for (int j : ex[i]) {
            const int w = j - i + 64;
            if (w >= 0 && w < 128) win[4 * (size_t)i + (w >> 5)] |= 1u << (w & 31);
            else fallback = true;
        }
    }
    if (fallback)
        for (int i = 0; i < n; ++i) {
            for (int j : ex[i])
                if (std::abs(j - i) >= 64) fidx.push_back(j);
            fptr[i + 1] = (int)fidx.size();
        }
    int rc;
#define TRY(x)                    \
    if ((rc = (x)) != ISSPA_OK) { \
        nonbond_free(m);          \
        delete m;                 \
        return rc;                \
    }
    TRY(upload(&m->aux, aux.data(), aux.size()));
    TRY(upload(&m->lj, lj.data(), lj.size()));
    TRY(upload(&m->excl, win.data(), win.size()));
    if (fallback) {
        TRY(upload(&m->excl_ptr, fptr.data(), fptr.size()));
        TRY(upload(&m->excl_idx, fidx.data(), std::max<size_t>(1, fidx.size())));
    }
    TRY(dalloc(&m->pbc_flag, 1));
    int c = auto_splits(n, n, 64);
    m->atoms_per_chunk = (n + c - 1) / c;
    m->chunks = (n + m->atoms_per_chunk - 1) / m->atoms_per_chunk;
    TRY(dalloc(&m->part, (size_t)m->chunks * 9 * n));
#undef TRY
    *out = m;
    return ISSPA_OK;
}

extern "C" int nonbond_model_destroy(nonbond_model *m)
{
    if (!m) return ISSPA_OK;
    nonbond_free(m);
    delete m;
    return ISSPA_OK;
}

static int nonbond_run(nonbond_model *m, const float *xyzq4_d, bool split, float4 *f, double *o0,
                       double *o1, double *o2, hipStream_t s)
{
    const float4 *x = reinterpret_cast<const float4 *>(xyzq4_d);
    HIPCHK(launch_pbc_flag(x, m->n, 0.0f, m->lbox, m->pbc_flag, s));
    NonbondArgs a;
    a.xyzq = x; a.aux = m->aux; a.lj = m->lj; a.T = m->T; a.n = m->n;
    a.atoms_per_chunk = m->atoms_per_chunk; a.excl = m->excl; a.excl_ptr = m->excl_ptr;
    a.excl_idx = m->excl_idx; a.L = m->lbox; a.h = 0.5f * m->lbox; a.cfar = m->cfar;
    a.pbc_flag = m->pbc_flag; a.part = m->part;
    HIPCHK(launch_nonbond(a, m->chunks, split, s));
    HIPCHK(launch_chunk_reduce3(m->part, m->n, m->chunks, split ? 9 : 3, f, o0, o1, o2, s));
    return ISSPA_OK;
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
// bonded model
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
    TRY(upload(&m->gidx, gidx.data(), std::max<size_t>(1, gidx.size())));
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

extern "C" int bonded_force(bonded_model *m, int terms, const float *xyzq4_d, float *f4_d,
                            double *energies_d, void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "bonded_force: null argument");
    hipStream_t s = S(stream);
    BondedArgs a;
    a.xyzq = reinterpret_cast<const float4 *>(xyzq4_d);
    a.L = m->lbox; a.h = 0.5f * m->lbox;
    a.nb = m->nb; a.na = m->na; a.nd = m->nd;
    a.bond_atoms = m->bond_atoms; a.bond_params = m->bond_params;
    a.angle_atoms = m->angle_atoms; a.angle_params = m->angle_params;
    a.dih_atoms = m->dih_atoms; a.dih_types = m->dih_types; a.dih_params = m->dih_params;
    a.sc14 = m->sc14; a.ityp = m->ityp; a.lj = m->lj; a.T = m->T;
    a.contrib = m->contrib; a.energy = m->energy; a.terms = terms;
    HIPCHK(launch_bonded_terms(a, s));
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
    if (n < 0 || !xyzq4_d || !vel4_d || !f4_d) return fail(ISSPA_ERR_ARG, "leapfrog_andersen: bad argument");
    if (!(pnu >= 0.0f && pnu <= 1.0f)) return fail(ISSPA_ERR_ARG, "leapfrog_andersen: pnu=%g not in [0,1]", pnu);
    HIPCHK(launch_leapfrog(n, reinterpret_cast<float4 *>(xyzq4_d), reinterpret_cast<float4 *>(vel4_d),
                           reinterpret_cast<const float4 *>(f4_d), kT, dt, andersen_threshold(pnu),
                           lbox, seed, system_id, step, step_d, S(stream)));
    return ISSPA_OK;
}

// ---------------------------------------------------------------------------
// umbrella restraint
// ---------------------------------------------------------------------------
struct us_model {
    int nbias = 0;
    int2 *atoms = nullptr;
    float *mass = nullptr;
    float k0 = 0.f, k1 = 0.f, x0 = 0.f, lbox = 0.f;
};

extern "C" int us_model_create(int n_atoms, int nbias, const int *atom_list, const float *mass,
                               const float *kumb, float x0, float lbox, us_model **out)
{
    if (!out || nbias < 1 || !atom_list || !mass || !kumb || !(lbox > 0))
        return fail(ISSPA_ERR_ARG, "us_model_create: bad argument");
    *out = nullptr;
    std::vector<int> seen;
    int ng[2] = {0, 0};
    for (int b = 0; b < nbias; ++b) {
        const int a = atom_list[2 * b], g = atom_list[2 * b + 1];
        if (a < 0 || a >= n_atoms || (g != 0 && g != 1))
            return fail(ISSPA_ERR_ARG, "us_model_create: entry %d = {%d,%d} invalid", b, a, g);
        seen.push_back(a);
        ng[g]++;
    }
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
        return fail(ISSPA_ERR_ARG, "us_model_create: an atom appears twice in the bias groups");
    if (!ng[0] || !ng[1]) return fail(ISSPA_ERR_ARG, "us_model_create: both groups need atoms");
    us_model *m = new (std::nothrow) us_model;
    if (!m) return fail(ISSPA_ERR_ALLOC, "host allocation");
    m->nbias = nbias; m->k0 = kumb[0]; m->k1 = kumb[1]; m->x0 = x0; m->lbox = lbox;
    int rc;
    if ((rc = upload(&m->atoms, reinterpret_cast<const int2 *>(atom_list), nbias)) != ISSPA_OK ||
        (rc = upload(&m->mass, mass, nbias)) != ISSPA_OK) {
        dfree(m->atoms); dfree(m->mass);
        delete m;
        return rc;
    }
    *out = m;
    return ISSPA_OK;
}

extern "C" int us_model_destroy(us_model *m)
{
    if (!m) return ISSPA_OK;
    dfree(m->atoms);
    dfree(m->mass);
    delete m;
    return ISSPA_OK;
}

extern "C" int us_force(us_model *m, const float *xyzq4_d, float *f4_d, double *com_d,
                        void *stream)
{
    if (!m || !xyzq4_d || !f4_d) return fail(ISSPA_ERR_ARG, "us_force: null argument");
    HIPCHK(launch_us(m->nbias, m->atoms, m->mass, m->k0, m->k1, m->x0, m->lbox,
                     reinterpret_cast<const float4 *>(xyzq4_d), reinterpret_cast<float4 *>(f4_d),
                     com_d, S(stream)));
    return ISSPA_OK;
}

// ---------------------------------------------------------------------------
// MD engine: one step = zero f, bonded, IS-SPA, nonbond, US, leap-frog, step++
// (order of src/total_force_cuda.cpp:109-153)
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
    uint32_t *step_d = nullptr;
    double *com_d = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

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
    md_engine *e = new (std::nothrow) md_engine;
    if (!e) return fail(ISSPA_ERR_ALLOC, "host allocation");
    e->isspa = isspa; e->nb = nb; e->bonded = bonded; e->us = us; e->n = n;
    e->x = x; e->v = v; e->f = f; e->kT = kT; e->dt = dt; e->pnu = pnu; e->lbox = lbox;
    e->seed = seed; e->sys = sys;
    int rc;
    if ((rc = dalloc(&e->step_d, 1)) != ISSPA_OK || (rc = dalloc(&e->com_d, 6)) != ISSPA_OK) {
        dfree(e->step_d);
        delete e;
        return rc;
    }
    hipError_t he = hipMemset(e->step_d, 0, sizeof(uint32_t));
    if (he != hipSuccess) {
        dfree(e->step_d); dfree(e->com_d);
        delete e;
        return fail(ISSPA_ERR_HIP, "hipMemset: %s", hipGetErrorString(he));
    }
    *out = e;
    return ISSPA_OK;
}

extern "C" int md_engine_destroy(md_engine *e)
{
    if (!e) return ISSPA_OK;
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    dfree(e->step_d);
    dfree(e->com_d);
    delete e;
    return ISSPA_OK;
}

extern "C" int md_engine_set_step(md_engine *e, uint32_t step, void *stream)
{
    if (!e) return fail(ISSPA_ERR_ARG, "null engine");
    HIPCHK(launch_set_word(e->step_d, step, S(stream)));
    return ISSPA_OK;
}

static int engine_forces(md_engine *e, hipStream_t s)
{
    HIPCHK(launch_zero(e->f, (size_t)e->n * sizeof(float4), s));
    if (e->bonded) CHK(bonded_force(e->bonded, ISSPA_TERM_ALL, e->x, e->f, nullptr, s));
    if (e->isspa) {
        CHK(isspa_field_pass(e->isspa, e->x, 0u, e->step_d, s));
        CHK(isspa_force_pass(e->isspa, e->x, false, reinterpret_cast<float4 *>(e->f), nullptr,
                             nullptr, nullptr, s));
    }
    if (e->nb) CHK(nonbond_force(e->nb, e->x, e->f, s));
    if (e->us) CHK(us_force(e->us, e->x, e->f, e->com_d, s));
    return ISSPA_OK;
}

static int engine_step(md_engine *e, hipStream_t s)
{
    CHK(engine_forces(e, s));
    HIPCHK(launch_leapfrog(e->n, reinterpret_cast<float4 *>(e->x), reinterpret_cast<float4 *>(e->v),
                           reinterpret_cast<const float4 *>(e->f), e->kT, e->dt,
                           andersen_threshold(e->pnu), e->lbox, e->seed, e->sys, 0u, e->step_d, s));
    HIPCHK(launch_step_advance(e->step_d, s));
    return ISSPA_OK;
}

extern "C" int md_engine_forces(md_engine *e, void *stream)
{
    if (!e) return fail(ISSPA_ERR_ARG, "null engine");
    return engine_forces(e, S(stream));
}

extern "C" int md_engine_step(md_engine *e, void *stream)
{
    if (!e) return fail(ISSPA_ERR_ARG, "null engine");
    return engine_step(e, S(stream));
}

extern "C" int md_engine_capture(md_engine *e, int steps_per_graph, void *stream)
{
    if (!e || steps_per_graph < 1) return fail(ISSPA_ERR_ARG, "md_engine_capture: bad argument");
    if (!stream) return fail(ISSPA_ERR_ARG, "md_engine_capture: needs a non-null stream");
    hipStream_t s = S(stream);
    if (e->exec) {
        (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    if (e->graph) {
        (void)hipGraphDestroy(e->graph);
        e->graph = nullptr;
    }
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
    return ISSPA_OK;
}

extern "C" int md_engine_replay(md_engine *e, int replays, void *stream)
{
    if (!e || replays < 0) return fail(ISSPA_ERR_ARG, "md_engine_replay: bad argument");
    if (!e->exec) return fail(ISSPA_ERR_STATE, "md_engine_replay: no captured graph");
    for (int r = 0; r < replays; ++r) HIPCHK(hipGraphLaunch(e->exec, S(stream)));
    return ISSPA_OK;
}
