// C-ABI utilities (include/isspa_amd.h): error string, device memory,
// streams, timers, model constants.
#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>

#include "capi_common.hpp"

using namespace isspa;

static thread_local std::string g_err;

int isspa_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *isspa_last_error(void) { return g_err.c_str(); }
extern "C" int isspa_abi_version(void) { return 2; }

// ---------------------------------------------------------------------------
// device memory / streams
// ---------------------------------------------------------------------------
extern "C" int isspa_dev_malloc(void **ptr_d, size_t bytes)
{
    if (!ptr_d) return isspa_fail(ISSPA_ERR_ARG, "isspa_dev_malloc: null out pointer");
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
    if (n < 0 || (n > 0 && (!dst4_d || !src4_d))) return isspa_fail(ISSPA_ERR_ARG, "isspa_add4: bad argument");
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
    if (!t) return isspa_fail(ISSPA_ERR_ALLOC, "timer allocation");
    HIPCHK(hipEventCreate(&t->a));
    HIPCHK(hipEventCreate(&t->b));
    HIPCHK(hipEventRecord(t->a, S(stream)));
    *timer = t;
    return ISSPA_OK;
}
extern "C" int isspa_timer_end(void *timer, void *stream, float *ms)
{
    timer_pair *t = (timer_pair *)timer;
    if (!t) return isspa_fail(ISSPA_ERR_ARG, "null timer");
    HIPCHK(hipEventRecord(t->b, S(stream)));
    HIPCHK(hipEventSynchronize(t->b));
    HIPCHK(hipEventElapsedTime(ms, t->a, t->b));
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
    return ISSPA_OK;
}

extern "C" int isspa_events_create(void **pair)
{
    if (!pair) return isspa_fail(ISSPA_ERR_ARG, "isspa_events_create: null argument");
    timer_pair *t = new (std::nothrow) timer_pair;
    if (!t) return isspa_fail(ISSPA_ERR_ALLOC, "event pair allocation");
    hipError_t e = hipEventCreate(&t->a);
    if (e == hipSuccess) e = hipEventCreate(&t->b);
    if (e != hipSuccess) {
        delete t;
        return isspa_fail(ISSPA_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
    }
    *pair = t;
    return ISSPA_OK;
}
extern "C" int isspa_events_mark_begin(void *pair, void *stream)
{
    if (!pair) return isspa_fail(ISSPA_ERR_ARG, "null event pair");
    HIPCHK(hipEventRecord(((timer_pair *)pair)->a, S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_events_mark_end(void *pair, void *stream)
{
    if (!pair) return isspa_fail(ISSPA_ERR_ARG, "null event pair");
    HIPCHK(hipEventRecord(((timer_pair *)pair)->b, S(stream)));
    return ISSPA_OK;
}
extern "C" int isspa_events_elapsed(void *pair, float *ms)
{
    if (!pair || !ms) return isspa_fail(ISSPA_ERR_ARG, "isspa_events_elapsed: null argument");
    timer_pair *t = (timer_pair *)pair;
    HIPCHK(hipEventSynchronize(t->b));
    HIPCHK(hipEventElapsedTime(ms, t->a, t->b));
    return ISSPA_OK;
}
extern "C" int isspa_events_destroy(void *pair)
{
    if (!pair) return ISSPA_OK;
    timer_pair *t = (timer_pair *)pair;
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

