// Host-side helpers for large host<->device transfers: a persistent worker
// pool, and copies of pageable ranges through two pinned chunks.  A pageable
// copy bounces through a driver buffer on one thread (and faults fresh pages
// in on it); here host threads fill or drain one pinned chunk while the DMA
// of the other is in flight, so a transfer runs at about the pinned rate.
#include "common.h"

#include <condition_variable>
#include <thread>
#include <vector>

namespace psi {
namespace {

class WorkerPool {
  public:
    explicit WorkerPool(int T) : T_(T)
    {
        for (int t = 1; t < T_; t++) th_.emplace_back([this, t] { loop(t); });
    }
    int size() const { return T_; }
    // f(t) for t in [0, T): the caller runs t = 0.
    void run(const std::function<void(int)>& f)
    {
        std::lock_guard<std::mutex> one(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            pending_ = T_ - 1;
            gen_++;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

  private:
    void loop(int t)
    {
        unsigned seen = 0;
        for (;;) {
            const std::function<void(int)>* job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                job = job_;
            }
            (*job)(t);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    int T_;
    std::vector<std::thread> th_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int pending_ = 0;
    unsigned gen_ = 0;
};

WorkerPool& pool()
{
    // never destroyed: the workers wait on the pool until the process ends
    static WorkerPool* p = new WorkerPool(
        (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency())));
    return *p;
}

constexpr size_t kChunk = 64ull << 20;

struct Stage {
    std::mutex mu;
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};
};
Stage g_stage;

int stage_ready()
{
    if (g_stage.buf[0]) return PSI_OK;
    std::lock_guard<std::recursive_mutex> lk(setup_mutex());
    for (int k = 0; k < 2; k++) {
        PSI_HIP(hipHostMalloc((void**)&g_stage.buf[k], kChunk));
        PSI_HIP(hipEventCreateWithFlags(&g_stage.ev[k], hipEventDisableTiming));
    }
    return PSI_OK;
}

} // namespace

void parallel_for(long long n, const std::function<void(long long, long long)>& f)
{
    if (n <= 0) return;
    WorkerPool& p = pool();
    const int T = (int)std::min<long long>(p.size(), n);
    if (T <= 1) {
        f(0, n);
        return;
    }
    p.run([&](int t) {
        if (t < T) f(n * t / T, n * (t + 1) / T);
    });
}

int staged_upload(hipStream_t s, void* dst, const HostPart* parts, int nparts)
{
    std::lock_guard<std::mutex> lk(g_stage.mu);
    if (int rc = stage_ready()) return rc;
    // chunk offsets of the parts, concatenated
    std::vector<size_t> at(nparts + 1, 0);
    for (int i = 0; i < nparts; i++) at[i + 1] = at[i] + parts[i].bytes;
    const size_t total = at[nparts];
    for (size_t off = 0, k = 0; off < total; off += kChunk, k++) {
        const size_t len = std::min(kChunk, total - off);
        const int b = (int)(k & 1);
        if (g_stage.busy[b]) PSI_HIP(hipEventSynchronize(g_stage.ev[b]));
        char* buf = g_stage.buf[b];
        parallel_for((long long)len, [&](long long lo, long long hi) {
            // bytes [off+lo, off+hi) of the concatenation
            size_t x = off + (size_t)lo, end = off + (size_t)hi;
            int i = (int)(std::upper_bound(at.begin(), at.end(), x) - at.begin()) - 1;
            while (x < end) {
                const size_t take = std::min(end, at[i + 1]) - x;
                if (take) memcpy(buf + (x - off), (const char*)parts[i].src + (x - at[i]), take);
                x += take;
                i++;
            }
        });
        PSI_HIP(hipMemcpyAsync((char*)dst + off, buf, len, hipMemcpyHostToDevice, s));
        PSI_HIP(hipEventRecord(g_stage.ev[b], s));
        g_stage.busy[b] = true;
    }
    return PSI_OK;
}

int staged_download(hipStream_t s, const void* src, size_t total,
                    const std::function<void(const char*, size_t, size_t)>& consume)
{
    std::lock_guard<std::mutex> lk(g_stage.mu);
    if (int rc = stage_ready()) return rc;
    const size_t nchunks = (total + kChunk - 1) / kChunk;
    auto issue = [&](size_t k) -> int {
        const int b = (int)(k & 1);
        if (g_stage.busy[b]) PSI_HIP(hipEventSynchronize(g_stage.ev[b]));
        const size_t off = k * kChunk, len = std::min(kChunk, total - off);
        PSI_HIP(hipMemcpyAsync(g_stage.buf[b], (const char*)src + off, len, hipMemcpyDeviceToHost, s));
        PSI_HIP(hipEventRecord(g_stage.ev[b], s));
        g_stage.busy[b] = true;
        return PSI_OK;
    };
    if (nchunks && issue(0)) return PSI_ERR_HIP;
    for (size_t k = 0; k < nchunks; k++) {
        if (k + 1 < nchunks)
            if (int rc = issue(k + 1)) return rc;   // in flight while chunk k is consumed
        const int b = (int)(k & 1);
        PSI_HIP(hipEventSynchronize(g_stage.ev[b]));
        const size_t off = k * kChunk;
        consume(g_stage.buf[b], off, std::min(kChunk, total - off));
    }
    return PSI_OK;
}

} // namespace psi
