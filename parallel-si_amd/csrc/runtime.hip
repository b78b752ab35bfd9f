// Library runtime: last-error string, the library stream, device memory and
// timing helpers exported through the C-ABI.  Plays the role of the
// reference's gpu_util.{h,cpp} (gpu_util.cpp:10-48) but reports failures as
// status codes instead of printing and calling exit(1).
#include "common.h"

#include <mutex>
#include <string>

namespace {
thread_local std::string t_err;
hipStream_t g_stream = nullptr;
std::once_flag g_stream_once;
hipError_t g_stream_err = hipSuccess;
}

extern "C" void psi_set_error_(const char* msg) { t_err = msg ? msg : ""; }

extern "C" const char* psi_last_error(void) { return t_err.c_str(); }

namespace psi {

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
             file, line);
    t_err = buf;
    return PSI_ERR_HIP;
}

std::recursive_mutex& setup_mutex()
{
    static std::recursive_mutex m;
    return m;
}

hipStream_t stream()
{
    std::call_once(g_stream_once, [] {
        g_stream_err = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
    });
    return g_stream;
}

int ensure(void** buf, size_t* cap, size_t bytes)
{
    if (*cap >= bytes && *buf) return PSI_OK;
    if (*buf) {
        PSI_HIP(hipDeviceSynchronize());           // queued work may still use it
        PSI_HIP(hipFree(*buf));
    }
    *buf = nullptr;
    *cap = 0;
    size_t b = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(buf, b);
    if (e != hipSuccess) {
        *buf = nullptr;
        return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    *cap = b;
    return PSI_OK;
}

} // namespace psi

extern "C" int psi_device_count(int* count)
{
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return psi::hip_fail(e, "hipGetDeviceCount", __FILE__, __LINE__);
    }
    return PSI_OK;
}

extern "C" int psi_init(void)
{
    int c = 0;
    int rc = psi_device_count(&c);
    if (rc) return rc;
    if (c < 1) return psi::fail(PSI_ERR_NO_DEVICE, "no HIP device");
    psi::stream();
    if (g_stream_err != hipSuccess) return psi::hip_fail(g_stream_err, "hipStreamCreate", __FILE__, __LINE__);
    // the fixed list-ranking workspace of the current device, resolved here
    // so psi_list_rank_noalloc makes no set-up call of its own
    return psi::lr_scratch_address(nullptr);
}

extern "C" int psi_synchronize(void)
{
    PSI_HIP(hipStreamSynchronize(psi::stream()));
    PSI_HIP(hipDeviceSynchronize());
    return PSI_OK;
}

extern "C" int psi_malloc(void** ptr, size_t bytes)
{
    PSI_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return PSI_OK;
}

extern "C" int psi_free(void* ptr)
{
    if (ptr) PSI_HIP(hipFree(ptr));
    return PSI_OK;
}

extern "C" int psi_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    PSI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, psi::stream()));
    PSI_HIP(hipStreamSynchronize(psi::stream()));
    return PSI_OK;
}

extern "C" int psi_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    PSI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, psi::stream()));
    PSI_HIP(hipStreamSynchronize(psi::stream()));
    return PSI_OK;
}

extern "C" int psi_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    PSI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, psi::stream()));
    return PSI_OK;
}

// Events on the library stream, for timing regions of device work.
extern "C" int psi_event_create(void** ev)
{
    hipEvent_t e;
    PSI_HIP(hipEventCreate(&e));
    *ev = (void*)e;
    return PSI_OK;
}

extern "C" int psi_event_destroy(void* ev)
{
    PSI_HIP(hipEventDestroy((hipEvent_t)ev));
    return PSI_OK;
}

extern "C" int psi_event_record(void* ev)
{
    PSI_HIP(hipEventRecord((hipEvent_t)ev, psi::stream()));
    return PSI_OK;
}

extern "C" int psi_event_elapsed_ms(void* start, void* end, float* ms)
{
    PSI_HIP(hipEventSynchronize((hipEvent_t)end));
    PSI_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return PSI_OK;
}
