// Helpers shared by the C-ABI translation units: error reporting, checked
// device allocation, stream cast.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/isspa_amd.h"
#include "kernels.hpp"

int isspa_fail(int code, const char *fmt, ...);

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return isspa_fail(ISSPA_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                                           \
    } while (0)

#define CHK(expr)                      \
    do {                               \
        int r_ = (expr);               \
        if (r_ != ISSPA_OK) return r_; \
    } while (0)

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

template <class T>
static int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess)
        return isspa_fail(ISSPA_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T),
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
