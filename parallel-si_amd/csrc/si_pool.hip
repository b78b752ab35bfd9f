// C ABI of the serving pool: a request loop for the GPU solve (solve_si,
// gpumain.cpp:18-53, run per request) with continuous batching.
//
// The pool owns fixed-capacity device arrays sized from an HBM budget.  A
// submitted game is admitted as soon as a control slot, a vertex range and an
// edge range are free (first fit, FIFO with bounded backfill); its arrays are
// uploaded and initialised on an admission stream while the SI steps of the
// games already admitted run on the step stream, and it joins the steps when
// its control block is written on the step stream.  Each psi_pool_step call
// waits for the previous step batch, launches the next one, and then, while
// it runs, copies out the games that finished, frees their ranges and admits
// queued requests.  A game's result depends only on that game: every kernel
// reads and writes a game's vertices through its own control block, so the
// result equals the game's solve alone.
#include "si_core.h"

#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

using namespace psi;

namespace {

// First-fit allocator of ranges in [0, cap).
struct Ranges {
    std::map<long long, long long> free_;   // start -> length
    long long cap = 0, used = 0;
    void init(long long c)
    {
        cap = c;
        used = 0;
        free_.clear();
        free_[0] = c;
    }
    long long alloc(long long len)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            long long at = it->first, rest = it->second - len;
            free_.erase(it);
            if (rest) free_[at + len] = rest;
            used += len;
            return at;
        }
        return -1;
    }
    void release(long long at, long long len)
    {
        used -= len;
        auto next = free_.lower_bound(at);
        if (next != free_.end() && next->first == at + len) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == at) {
                prev->second += len;
                return;
            }
        }
        free_[at] = len;
    }
    long long largest() const
    {
        long long b = 0;
        for (auto& kv : free_) b = std::max(b, kv.second);
        return b;
    }
};

struct Request {
    int64_t ticket;
    GameView gv;
    int reset, max_rounds;
};

struct Result {
    int iter = 0, br = 0, capped = 0, n = 0;
    std::vector<int> strat, lab;
};

// Admissions that may pass a waiting request before it is admitted in order.
constexpr int kMaxBypass = 16;

} // namespace

struct psi_pool {
    Core core;
    hipStream_t s1 = nullptr;        // step stream (core.s)
    hipStream_t s2 = nullptr;        // admission and harvest stream
    Ranges va, ea;
    int slots = 0, max_n = 0;
    std::vector<int> free_slots;
    std::vector<int64_t> slot_ticket;     // -1: free
    std::vector<long long> slot_ebase;
    GameCtl* hctl = nullptr;              // pinned: control blocks being activated
    GameCtl* hread = nullptr;             // pinned: control blocks read back
    int* herr = nullptr;                  // pinned: the jump rounds' error flag
    int* hseq = nullptr;                  // pinned: admissions written (step program)
    bool program = false;                 // steps as the step program (fixed at creation)
    int admitted = 0;
    std::deque<Request> queue;
    int head_bypassed = 0;
    std::unordered_map<int64_t, int> state;      // 0 queued, 1 running, 2 finished
    std::unordered_map<int64_t, Result> results;
    int64_t next_ticket = 1;
    int running = 0;
    long long bytes_per_vertex = 0, budget = 0;
};

static int pool_admit(psi_pool* p, const Request& r, bool* admitted)
{
    *admitted = false;
    if (p->free_slots.empty()) return PSI_OK;
    const long long vb = p->va.alloc((long long)r.gv.n + 1);
    if (vb < 0) return PSI_OK;
    const long long eb = p->ea.alloc(r.gv.m);
    if (eb < 0) {
        p->va.release(vb, (long long)r.gv.n + 1);
        return PSI_OK;
    }
    const int slot = p->free_slots.back();
    p->free_slots.pop_back();
    if (int rc = core_pool_admit(p->core, slot, vb, eb, r.gv, p->s2)) {
        p->va.release(vb, (long long)r.gv.n + 1);
        p->ea.release(eb, r.gv.m);
        p->free_slots.push_back(slot);
        return rc;
    }
    // The caller's arrays are read by the copies above: wait for them.
    PSI_HIP(hipStreamSynchronize(p->s2));
    GameCtl& g = p->hctl[slot];
    std::memset(&g, 0, sizeof g);
    g.vbase = (int)vb;
    g.n = r.gv.n;
    g.reset = r.reset;
    g.max_rounds = r.max_rounds;
    // Ordered after the step batch in flight: the game joins the next one
    // (the step program activates it at its next step boundary, since a
    // batch of its launches may end inside a step).
    if (p->program) {
        g.done = 1;
        g.activate = 1;
    }
    PSI_HIP(hipMemcpyAsync(p->core.gctl + slot, &g, sizeof g, hipMemcpyHostToDevice, p->core.s));
    if (p->program) {
        *p->hseq = ++p->admitted;
        PSI_HIP(hipMemcpyAsync(&p->core.prog->host_seq, p->hseq, sizeof(int), hipMemcpyHostToDevice,
                               p->core.s));
    }
    p->slot_ticket[slot] = r.ticket;
    p->slot_ebase[slot] = eb;
    p->state[r.ticket] = 1;
    p->running++;
    *admitted = true;
    return PSI_OK;
}

// FIFO; a request that does not fit may be passed by later ones that do, at
// most kMaxBypass times, after which admission waits for it.
static int pool_admit_queued(psi_pool* p)
{
    while (!p->queue.empty()) {
        bool ok = false;
        if (int rc = pool_admit(p, p->queue.front(), &ok)) {
            p->state.erase(p->queue.front().ticket);
            p->queue.pop_front();
            return rc;
        }
        if (ok) {
            p->queue.pop_front();
            p->head_bypassed = 0;
            continue;
        }
        if (p->head_bypassed >= kMaxBypass) return PSI_OK;
        for (auto it = p->queue.begin() + 1; it != p->queue.end() && p->head_bypassed < kMaxBypass;) {
            if (int rc = pool_admit(p, *it, &ok)) {
                p->state.erase(it->ticket);
                p->queue.erase(it);
                return rc;
            }
            if (ok) {
                it = p->queue.erase(it);
                p->head_bypassed++;
            } else {
                ++it;
            }
        }
        return PSI_OK;
    }
    return PSI_OK;
}

extern "C" int psi_pool_bytes_per_vertex(int max_n, int64_t* bytes)
{
    if (max_n <= 0) return fail(PSI_ERR_INVALID, "max_n must be positive");
    *bytes = core_pool_bytes_per_vertex(max_n);
    return PSI_OK;
}

extern "C" int psi_pool_create(int64_t hbm_budget, double edges_per_vertex, int max_games, int max_n,
                               psi_pool** out)
{
    *out = nullptr;
    if (hbm_budget <= 0 || edges_per_vertex <= 0 || max_games <= 0 || max_n <= 0)
        return fail(PSI_ERR_INVALID, "pool: budget, edges per vertex, games and max_n must be positive");
    const long long bpv = core_pool_bytes_per_vertex(max_n);
    const double per = (double)bpv + 4.0 * edges_per_vertex;
    long long vcap = (long long)((double)hbm_budget / per);
    long long ecap = (long long)(edges_per_vertex * (double)vcap);
    vcap = std::min(vcap, (1LL << 31) - 2);
    ecap = std::min(ecap, (1LL << 31) - 2);
    if (vcap < (long long)max_n + 1)
        return fail(PSI_ERR_BUDGET, "pool: the budget does not hold one game of max_n vertices");
    psi_pool* p = new psi_pool();
    p->bytes_per_vertex = bpv;
    p->budget = hbm_budget;
    p->slots = max_games;
    p->max_n = max_n;
    int rc = psi_init();
    std::unique_lock<std::recursive_mutex> lk(setup_mutex());
    if (rc == PSI_OK) {
        hipError_t e = hipStreamCreateWithFlags(&p->s1, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->s2, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipHostMalloc((void**)&p->hctl, sizeof(GameCtl) * (max_games + 1));
        if (e == hipSuccess) e = hipHostMalloc((void**)&p->hread, sizeof(GameCtl) * (max_games + 1));
        if (e == hipSuccess) e = hipHostMalloc((void**)&p->herr, sizeof(int));
        if (e == hipSuccess) e = hipHostMalloc((void**)&p->hseq, sizeof(int));
        p->program = g_si_tune.program != 0;
        if (e != hipSuccess) rc = hip_fail(e, "pool streams and pinned buffers", __FILE__, __LINE__);
    }
    if (rc == PSI_OK) rc = core_create_pool(p->core, vcap, ecap, max_games, max_n, p->s1);
    lk.unlock();
    if (rc) {
        psi_pool_destroy(p);
        return rc;
    }
    p->va.init(vcap);
    p->ea.init(ecap);
    p->slot_ticket.assign(max_games, -1);
    p->slot_ebase.assign(max_games, 0);
    for (int k = max_games - 1; k >= 0; k--) p->free_slots.push_back(k);
    *out = p;
    return PSI_OK;
}

extern "C" int psi_pool_destroy(psi_pool* p)
{
    if (!p) return PSI_OK;
    if (p->s1) (void)hipStreamSynchronize(p->s1);
    if (p->s2) (void)hipStreamSynchronize(p->s2);
    std::lock_guard<std::recursive_mutex> lk(setup_mutex());
    core_destroy(p->core);
    if (p->s1) (void)hipStreamDestroy(p->s1);
    if (p->s2) (void)hipStreamDestroy(p->s2);
    if (p->hctl) (void)hipHostFree(p->hctl);
    if (p->hread) (void)hipHostFree(p->hread);
    if (p->herr) (void)hipHostFree(p->herr);
    if (p->hseq) (void)hipHostFree(p->hseq);
    delete p;
    return PSI_OK;
}

extern "C" int psi_pool_submit(psi_pool* p, int n, int64_t m, const int* owner, const int* priority,
                               const int* offsets, const int* targets, int reset, int max_rounds,
                               int64_t* ticket)
{
    *ticket = -1;
    if (n <= 0 || m <= 0) return fail(PSI_ERR_INVALID, "game with no vertices");
    if (n > p->max_n || (long long)n + 1 > p->va.cap || m > p->ea.cap)
        return fail(PSI_ERR_BUDGET, "game larger than the pool admits");
    GameView gv{n, (long long)m, owner, priority, offsets, targets};
    int mp = 0;
    // validated here, so that a queued request cannot fail at admission
    if (const char* e = core_validate_game(gv, &mp)) return fail(PSI_ERR_INVALID, e);
    Request r{p->next_ticket++, gv, reset ? 1 : 0, std::max(max_rounds, 0)};
    p->state[r.ticket] = 0;
    p->queue.push_back(r);
    if (p->queue.size() == 1) {
        if (int rc = pool_admit_queued(p)) return rc;
    }
    *ticket = r.ticket;
    return PSI_OK;
}

extern "C" int psi_pool_step(psi_pool* p, int steps, int64_t* done_tickets, int cap, int* n_done)
{
    *n_done = 0;
    Core& c = p->core;
    // 1. wait for the batch in flight and read every control block
    PSI_HIP(hipMemcpyAsync(p->hread, c.gctl, sizeof(GameCtl) * c.G, hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipMemcpyAsync(p->herr, (char*)c.jc + offsetof(JumpCnt, err), sizeof(int),
                           hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    if (*p->herr) return fail(PSI_ERR_INVALID, "internal: jump rounds exhausted");
    std::vector<int> fin;
    for (int k = 0; k < p->slots; k++)
        if (p->slot_ticket[k] >= 0 && p->hread[k].done && !p->hread[k].activate) fin.push_back(k);
    // 2. next batch, if any admitted game is unfinished
    if (p->running > (int)fin.size()) {
        if (int rc = p->program ? core_pool_launch_program(c, steps) : core_launch_steps(c, steps))
            return rc;
    }
    // 3. while it runs: copy out finished games and free their ranges
    for (int k : fin) {
        const int64_t t = p->slot_ticket[k];
        const int n = c.h_n[k];
        const long long vb = c.h_vbase[k];
        Result& r = p->results[t];
        r.n = n;
        r.iter = p->hread[k].iter;
        r.br = p->hread[k].br;
        r.capped = p->hread[k].capped;
        r.strat.resize(n);
        std::vector<int8_t> lab(n);
        if (int rc = core_pool_localize(c, k, p->s2)) return rc;
        PSI_HIP(hipMemcpyAsync(r.strat.data(), c.strat + vb, (size_t)n * 4, hipMemcpyDeviceToHost, p->s2));
        PSI_HIP(hipMemcpyAsync(lab.data(), c.lab + vb, (size_t)n, hipMemcpyDeviceToHost, p->s2));
        PSI_HIP(hipStreamSynchronize(p->s2));
        r.lab.assign(lab.begin(), lab.end());
        const long long m = c.h_m[k];
        if (int rc = core_pool_release(c, k, p->s2)) return rc;
        p->va.release(vb, (long long)n + 1);
        p->ea.release(p->slot_ebase[k], m);
        p->slot_ticket[k] = -1;
        p->free_slots.push_back(k);
        p->running--;
        p->state[t] = 2;
        if (*n_done < cap && done_tickets) done_tickets[*n_done] = t;
        (*n_done)++;
    }
    // 4. admit queued requests into the freed space
    return pool_admit_queued(p);
}

extern "C" int psi_pool_poll(psi_pool* p, int64_t ticket, int* state)
{
    auto it = p->state.find(ticket);
    *state = it == p->state.end() ? -1 : it->second;
    return PSI_OK;
}

extern "C" int psi_pool_result(psi_pool* p, int64_t ticket, int* iter, int* br, int* capped,
                               int* strategy, int* labels)
{
    auto it = p->results.find(ticket);
    if (it == p->results.end()) return fail(PSI_ERR_INVALID, "no finished request with this ticket");
    Result& r = it->second;
    if (iter) *iter = r.iter;
    if (br) *br = r.br;
    if (capped) *capped = r.capped;
    if (strategy) std::memcpy(strategy, r.strat.data(), (size_t)r.n * 4);
    if (labels) std::memcpy(labels, r.lab.data(), (size_t)r.n * 4);
    p->results.erase(it);
    p->state.erase(ticket);
    return PSI_OK;
}

extern "C" int psi_pool_info(psi_pool* p, int* running, int* queued, int64_t* vertex_cap,
                             int64_t* edge_cap, int64_t* vertices_free, int64_t* largest_free)
{
    if (running) *running = p->running;
    if (queued) *queued = (int)p->queue.size();
    if (vertex_cap) *vertex_cap = p->va.cap;
    if (edge_cap) *edge_cap = p->ea.cap;
    if (vertices_free) *vertices_free = p->va.cap - p->va.used;
    if (largest_free) *largest_free = p->va.largest();
    return PSI_OK;
}
