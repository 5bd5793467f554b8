// ppl_runtime.hip -- device selection, memory, streams and error reporting of
// the C ABI.  Replaces the GPU_NVidia / GPU_Param plumbing of pelican-lofar
// (src/lib/src/GPU_NVidia.cpp, src/lib/src/GPU_Param.cpp) with plain HIP.
#include "ppl_internal.h"

#include <hip/hip_ext.h>

#include <cstring>

namespace ppl {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

}  // namespace ppl

using namespace ppl;

extern "C" const char* ppl_last_error(void) { return g_err; }

extern "C" int ppl_version(void) { return 1; }

extern "C" int ppl_init(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device available");
        return PPL_E_NODEV;
    }
    PPL_CHECK_ARG(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
    PPL_HIP(hipSetDevice(device));
    return detect_noalloc_setup();
}

extern "C" int ppl_shutdown(void)
{
    PPL_HIP(hipDeviceSynchronize());
    return PPL_OK;
}

extern "C" int ppl_device_alloc(void** ptr, size_t bytes)
{
    PPL_CHECK_ARG(ptr, "null pointer");
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        set_error("hipMalloc(%zu) failed", bytes);
        return PPL_E_NOMEM;
    }
    return PPL_OK;
}

extern "C" int ppl_device_free(void* ptr)
{
    if (ptr) PPL_HIP(hipFree(ptr));
    return PPL_OK;
}

extern "C" int ppl_host_alloc(void** ptr, size_t bytes)
{
    PPL_CHECK_ARG(ptr, "null pointer");
    if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc(%zu) failed", bytes);
        return PPL_E_NOMEM;
    }
    return PPL_OK;
}

extern "C" int ppl_host_free(void* ptr)
{
    if (ptr) PPL_HIP(hipHostFree(ptr));
    return PPL_OK;
}

extern "C" int ppl_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    PPL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(stream)));
    return PPL_OK;
}

extern "C" int ppl_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    PPL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(stream)));
    return PPL_OK;
}

extern "C" int ppl_memset_async(void* dst, int value, size_t bytes, void* stream)
{
    PPL_HIP(hipMemsetAsync(dst, value, bytes, stream_of(stream)));
    return PPL_OK;
}

extern "C" int ppl_stream_create(void** stream)
{
    PPL_CHECK_ARG(stream, "null pointer");
    hipStream_t s;
    PPL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return PPL_OK;
}

// A stream whose kernels run only on CU slots [first, first + count) of
// every XCD.  A CU-mask bit i selects CU i / nXcd of XCD i % nXcd, and a
// launch's workgroups go to the XCDs round robin whatever the mask, so the
// range is the same on every XCD.
extern "C" int ppl_stream_create_cu_range(int first, int count, void** stream)
{
    PPL_CHECK_ARG(stream, "null pointer");
    int dev = 0;
    PPL_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PPL_HIP(hipGetDeviceProperties(&prop, dev));
    constexpr int kXcd = 8;
    const int nCu = prop.multiProcessorCount;
    PPL_CHECK_ARG(nCu % kXcd == 0 && nCu <= 1024, "%d CUs is not a whole number of %d XCDs", nCu, kXcd);
    const int perXcd = nCu / kXcd;
    PPL_CHECK_ARG(first >= 0 && count >= 1 && first + count <= perXcd, "CU range [%d, %d) outside 0..%d per XCD",
                  first, first + count, perXcd);
    uint32_t mask[32] = {0};
    for (int u = first; u < first + count; ++u)
        for (int x = 0; x < kXcd; ++x) {
            const int bit = u * kXcd + x;
            mask[bit / 32] |= 1u << (bit % 32);
        }
    hipStream_t s;
    PPL_HIP(hipExtStreamCreateWithCUMask(&s, (uint32_t)((nCu + 31) / 32), mask));
    *stream = s;
    return PPL_OK;
}

extern "C" int ppl_stream_destroy(void* stream)
{
    if (stream) PPL_HIP(hipStreamDestroy(stream_of(stream)));
    return PPL_OK;
}

extern "C" int ppl_stream_sync(void* stream)
{
    PPL_HIP(hipStreamSynchronize(stream_of(stream)));
    return PPL_OK;
}

namespace {
struct Cb {
    void (*fn)(void*, int);
    void* user;
};
void host_trampoline(void* p)
{
    Cb* cb = static_cast<Cb*>(p);
    cb->fn(cb->user, PPL_OK);
    delete cb;
}
}  // namespace

extern "C" int ppl_stream_callback(void* stream, void (*fn)(void* user, int status), void* user)
{
    PPL_CHECK_ARG(fn, "null callback");
    Cb* cb = new Cb{fn, user};
    hipError_t e = hipLaunchHostFunc(stream_of(stream), host_trampoline, cb);
    if (e != hipSuccess) {
        delete cb;
        set_error("hipLaunchHostFunc: %s", hipGetErrorString(e));
        return PPL_E_HIP;
    }
    return PPL_OK;
}
