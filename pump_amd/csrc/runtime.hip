// runtime.hip -- thin device-memory, stream and timer helpers so that host
// programs (the Python tests, the benchmark) can drive the C ABI without a
// second GPU runtime in the process.
#include <cstdio>
#include <cstring>

#include "../../include/pump_host.h"
#include "pump_internal.hpp"

using namespace pump;

extern "C" void *pump_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        note_hip_error(e, __FILE__, __LINE__);
        return nullptr;
    }
    return p;
}

extern "C" int pump_dev_free(void *p)
{
    if (p)
        PUMP_HIP(hipFree(p));
    return PUMP_OK;
}

extern "C" int pump_dev_upload(void *dst, const void *src, size_t bytes)
{
    PUMP_REQUIRE(dst && (src || bytes == 0));
    if (bytes)
        PUMP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PUMP_OK;
}

extern "C" int pump_dev_download(void *dst, const void *src, size_t bytes)
{
    PUMP_REQUIRE(src && (dst || bytes == 0));
    if (bytes)
        PUMP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PUMP_OK;
}

extern "C" int pump_dev_copy(void *dst, const void *src, size_t bytes)
{
    PUMP_REQUIRE(dst && src);
    if (bytes)
        PUMP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return PUMP_OK;
}

extern "C" int pump_dev_memset(void *dst, int value, size_t bytes)
{
    PUMP_REQUIRE(dst);
    if (bytes)
        PUMP_HIP(hipMemset(dst, value, bytes));
    return PUMP_OK;
}

extern "C" int pump_dev_sync(void)
{
    PUMP_HIP(hipDeviceSynchronize());
    return PUMP_OK;
}

extern "C" int pump_dev_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

extern "C" int pump_dev_name(char *buf, int len)
{
    PUMP_REQUIRE(buf && len > 0);
    hipDeviceProp_t p;
    int dev = 0;
    PUMP_HIP(hipGetDevice(&dev));
    PUMP_HIP(hipGetDeviceProperties(&p, dev));
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return PUMP_OK;
}

extern "C" int pump_stream_create(pump_stream_t *s)
{
    PUMP_REQUIRE(s);
    hipStream_t h = nullptr;
    PUMP_HIP(hipStreamCreate(&h));
    *s = h;
    return PUMP_OK;
}

extern "C" int pump_stream_destroy(pump_stream_t s)
{
    if (s)
        PUMP_HIP(hipStreamDestroy(as_stream(s)));
    return PUMP_OK;
}

extern "C" int pump_stream_sync(pump_stream_t s)
{
    PUMP_HIP(hipStreamSynchronize(as_stream(s)));
    return PUMP_OK;
}

struct pump_timer {
    hipEvent_t a, b;
    hipStream_t s;
};

extern "C" pump_timer *pump_timer_create(pump_stream_t s)
{
    pump_timer *t = new pump_timer();
    t->s = as_stream(s);
    if (hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
        delete t;
        return nullptr;
    }
    return t;
}

extern "C" int pump_timer_start(pump_timer *t)
{
    PUMP_REQUIRE(t);
    PUMP_HIP(hipEventRecord(t->a, t->s));
    return PUMP_OK;
}

extern "C" float pump_timer_stop(pump_timer *t)
{
    if (!t)
        return -1.0f;
    float ms = -1.0f;
    if (hipEventRecord(t->b, t->s) != hipSuccess || hipEventSynchronize(t->b) != hipSuccess ||
        hipEventElapsedTime(&ms, t->a, t->b) != hipSuccess)
        return -1.0f;
    return ms;
}

extern "C" void pump_timer_destroy(pump_timer *t)
{
    if (!t)
        return;
    hipEventDestroy(t->a);
    hipEventDestroy(t->b);
    delete t;
}
