// pf_runtime.hip -- thin C-ABI wrappers over the HIP runtime (see
// include/pf_runtime.h).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/pf_runtime.h"

static int pf_ok(hipError_t e) { return e == hipSuccess ? 0 : -1; }

extern "C" int pf_device_count(void)
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}

extern "C" const char *pf_device_name(void)
{
    static char name[256];
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return "";
    strncpy(name, prop.name, sizeof name - 1);
    name[sizeof name - 1] = 0;
    return name;
}

extern "C" int pf_device_arch(char *buf, size_t cap)
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (!buf || cap == 0) return -1;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return -1;
    strncpy(buf, prop.gcnArchName, cap - 1);
    buf[cap - 1] = 0;
    return 0;
}

extern "C" int pf_device_cu_count(void)
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    return v;
}

extern "C" void *pf_malloc(size_t bytes)
{
    void *p = nullptr;
    return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : nullptr;
}

extern "C" int pf_free(void *dptr) { return pf_ok(hipFree(dptr)); }

extern "C" int pf_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    return pf_ok(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

extern "C" int pf_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    return pf_ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

extern "C" int pf_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    return pf_ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}

extern "C" int pf_memset(void *dst, int value, size_t bytes)
{
    return pf_ok(hipMemset(dst, value, bytes));
}

extern "C" int pf_sync(void) { return pf_ok(hipDeviceSynchronize()); }

// A plain streaming copy (16-byte vectors, grid-stride): the reference point
// for HBM byte counters, and a device-to-device copy usable on any stream.
__global__ __launch_bounds__(256) void pf_copy_kernel(uint4 *__restrict__ dst,
                                                     const uint4 *__restrict__ src, size_t n16)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

extern "C" int pf_device_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    if ((bytes & 15u) || ((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u)) return -1;
    const size_t n16 = bytes / 16u;
    if (n16 == 0) return 0;
    size_t blocks = (n16 + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    pf_copy_kernel<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(
        (uint4 *)dst, (const uint4 *)src, n16);
    return pf_ok(hipGetLastError());
}

extern "C" void *pf_stream_create(void)
{
    hipStream_t s = nullptr;
    return hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess ? (void *)s : nullptr;
}

extern "C" int pf_stream_destroy(void *stream)
{
    return pf_ok(hipStreamDestroy((hipStream_t)stream));
}

extern "C" int pf_stream_sync(void *stream)
{
    return pf_ok(hipStreamSynchronize((hipStream_t)stream));
}

extern "C" void *pf_event_create(void)
{
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? (void *)e : nullptr;
}

extern "C" int pf_event_destroy(void *ev) { return pf_ok(hipEventDestroy((hipEvent_t)ev)); }

extern "C" int pf_event_record(void *ev, void *stream)
{
    return pf_ok(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
}

extern "C" int pf_event_sync(void *ev) { return pf_ok(hipEventSynchronize((hipEvent_t)ev)); }

extern "C" float pf_event_elapsed_ms(void *start, void *stop)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop) != hipSuccess) return -1.f;
    return ms;
}
