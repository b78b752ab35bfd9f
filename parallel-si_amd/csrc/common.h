// Shared helpers for the HIP sources: error state, checked HIP calls, launch
// sizing and small device utilities.  gfx950 only (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>

#include "../../include/psi.h"

extern "C" void psi_set_error_(const char* msg);

namespace psi {

constexpr int kWave = 64;

// Records a HIP failure as the library's last error and returns PSI_ERR_HIP.
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define PSI_HIP(call)                                                         \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return ::psi::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define PSI_LAUNCH_CHECK()                                                    \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess) return ::psi::hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
    } while (0)

inline int fail(int code, const std::string& msg)
{
    psi_set_error_(msg.c_str());
    return code;
}

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Library stream (default priority, non-blocking w.r.t. the null stream).
hipStream_t stream();
// Device address of the current GPU's fixed list-ranking workspace (the
// array psi_list_rank_noalloc uses), resolved once per device.
int lr_scratch_address(void** out);

// Grows *buf to at least `bytes`; contents are not preserved.
int ensure(void** buf, size_t* cap, size_t bytes);
// Held while a stream is captured into a graph, and around the creation and
// destruction of solver handles and pools (allocation, frees, streams,
// synchronous copies): HIP invalidates a capture that another thread's such
// call overlaps.
std::recursive_mutex& setup_mutex();

// f(lo, hi) over [0, n) split into one range per worker of a persistent host
// pool (up to 16 threads); calls do not nest.
void parallel_for(long long n, const std::function<void(long long, long long)>& f);
// Large transfers through two pinned chunks, host side on the pool.
struct HostPart {
    const void* src;
    size_t bytes;
};
// The concatenated parts to device dst, ordered on stream s (the sources
// may be released on return).
int staged_upload(hipStream_t s, void* dst, const HostPart* parts, int nparts);
// Device src to consume(chunk, byte offset, length), chunk by chunk in order.
int staged_download(hipStream_t s, const void* src, size_t total,
                    const std::function<void(const char*, size_t, size_t)>& consume);

__host__ __device__ inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline int ceil_log2(long long x)
{
    int r = 0;
    while ((1LL << r) < x) r++;
    return r;
}

__device__ inline int ceil_log2_dev(int x)
{
    int r = 0;
    while ((1 << r) < x) r++;
    return r;
}

} // namespace psi
