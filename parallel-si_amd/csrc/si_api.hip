// C ABI of the solver: games, single-game and batched solvers, the
// step-wise StrategyImprovement operations (si.h:6-16) and the
// device-resident solve (solve_si, gpumain.cpp:18-53).
#include "si_core.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace psi;

struct psi_game {
    int n = 0, m = 0, max_pri = 0;
    std::vector<int> owner, priority, offsets, targets;
    // Device copy of the CSR, made by psi_game_device_csr on first request.
    int* d_owner = nullptr;
    int* d_priority = nullptr;
    int* d_offsets = nullptr;
    int* d_targets = nullptr;
    bool uploaded = false;
};

struct psi_solver {
    Core core;
    bool trace = false;
    hipStream_t s = nullptr;   // the handle's own stream: handles solve concurrently
    // A one-game handle of a game with no vertices has no device state.  The
    // reference solves such a game at once: solve_si runs one valuation and
    // one outer step (gpumain.cpp:18-53: iter 1, br 1, also under CPUBV and
    // CPUList), and solve_bf's first best response counts no round
    // (main.cpp:57-77, cpu_bf.cpp:25-47: br 0).
    bool empty = false;
    int e_iter = 0, e_br = 0;
};

extern "C" int psi_game_create(int n, int m, const int* owner, const int* priority,
                               const int* offsets, const int* targets, psi_game** out)
{
    *out = nullptr;
    // n = 0 is the reference's game with no vertices (a header line only)
    if (n < 0 || m < 0) return fail(PSI_ERR_INVALID, "psi_game_create: negative size");
    if (offsets[0] != 0 || offsets[n] != m)
        return fail(PSI_ERR_INVALID, "psi_game_create: offsets must run from 0 to m");
    auto* g = new psi_game;
    g->n = n;
    g->m = m;
    g->owner.assign(owner, owner + n);
    g->priority.assign(priority, priority + n);
    g->offsets.assign(offsets, offsets + n + 1);
    g->targets.assign(targets, targets + m);
    for (int v = 0; v < n; v++) {
        if (offsets[v + 1] <= offsets[v]) { delete g; return fail(PSI_ERR_INVALID, "sinks are not allowed!"); }
        if (owner[v] != 0 && owner[v] != 1) { delete g; return fail(PSI_ERR_INVALID, "invalid player"); }
        if (priority[v] < 0 || priority[v] > 0x3FFFFFFF) {
            delete g;
            return fail(PSI_ERR_INVALID, "priority out of range [0, 2^30)");
        }
        g->max_pri = std::max(g->max_pri, priority[v]);
    }
    for (int e = 0; e < m; e++)
        if (targets[e] < 0 || targets[e] >= n) { delete g; return fail(PSI_ERR_INVALID, "successor out of range"); }
    *out = g;
    return PSI_OK;
}

extern "C" int psi_game_destroy(psi_game* g)
{
    if (!g) return PSI_OK;
    if (g->uploaded) {
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        for (int* p : {g->d_owner, g->d_priority, g->d_offsets, g->d_targets})
            if (p) (void)hipFree(p);
    }
    delete g;
    return PSI_OK;
}

// GPUGame's device arrays (gpu_game.cpp:11-41: player, priority, vertices =
// offsets, edges = targets, int32, uploaded synchronously).  Owned by the game
// handle and freed with it (the reference's destructor frees only player and
// priority, gpu_game.cpp:43-47).
extern "C" int psi_game_device_csr(psi_game* g, int** player, int** priority, int** vertices,
                                   int** edges)
{
    if (!g) return fail(PSI_ERR_INVALID, "psi_game_device_csr: null game");
    if (!g->uploaded) {
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        int* d[4] = {nullptr, nullptr, nullptr, nullptr};
        const std::vector<int>* h[4] = {&g->owner, &g->priority, &g->offsets, &g->targets};
        for (int i = 0; i < 4; i++) {
            size_t bytes = h[i]->size() * sizeof(int);
            if (bytes == 0) continue;              // a game with no vertices: null arrays
            hipError_t e = hipMalloc((void**)&d[i], bytes);
            if (e == hipSuccess) e = hipMemcpy(d[i], h[i]->data(), bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                for (int k = 0; k <= i; k++) (void)hipFree(d[k]);
                return hip_fail(e, "psi_game_device_csr", __FILE__, __LINE__);
            }
        }
        g->d_owner = d[0];
        g->d_priority = d[1];
        g->d_offsets = d[2];
        g->d_targets = d[3];
        g->uploaded = true;
    }
    if (player) *player = g->d_owner;
    if (priority) *priority = g->d_priority;
    if (vertices) *vertices = g->d_offsets;
    if (edges) *edges = g->d_targets;
    return PSI_OK;
}

extern "C" int psi_game_info(const psi_game* g, int* n, int* m, int* max_pri)
{
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_pri) *max_pri = g->max_pri;
    return PSI_OK;
}

static int create_from_views(int n_games, const GameView* views, int flags, int trace_cap,
                             psi_solver** out)
{
    *out = nullptr;
    if (n_games <= 0) return fail(PSI_ERR_INVALID, "no games");
    if (int rc = psi_init()) return rc;
    std::lock_guard<std::recursive_mutex> lk(setup_mutex());
    auto* s = new psi_solver;
    s->trace = (flags & PSI_SOLVER_TRACE) != 0;
    // one game with no vertices: a handle without device state (a batch
    // holds games of at least one vertex, as core_create checks)
    s->empty = n_games == 1 && views[0].n == 0 && views[0].m == 0;
    hipError_t e = hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking);
    int rc = e == hipSuccess ? PSI_OK : hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    if (!rc && s->empty) {
        *out = s;
        return PSI_OK;
    }
    if (!rc) rc = core_create(s->core, n_games, views, s->trace, trace_cap, s->s);
    if (!rc) rc = core_init_strat(s->core);
    if (rc) {
        if (s->s) (void)hipStreamSynchronize(s->s);
        core_destroy(s->core);
        if (s->s) (void)hipStreamDestroy(s->s);
        delete s;
        return rc;
    }
    *out = s;
    return PSI_OK;
}

extern "C" int psi_solver_create_batch(int n_games, const int* n_per, const int64_t* m_per,
                                       const int* owner, const int* priority, const int* offsets,
                                       const int* targets, int flags, int trace_cap,
                                       psi_solver** out)
{
    std::vector<GameView> v(std::max(n_games, 0));
    long long vb = 0, eb = 0;
    for (int g = 0; g < n_games; g++) {
        v[g] = GameView{n_per[g], (long long)m_per[g], owner + vb, priority + vb, offsets + vb + g,
                        targets + eb};
        vb += n_per[g];
        eb += m_per[g];
    }
    return create_from_views(n_games, v.data(), flags, trace_cap, out);
}

extern "C" int psi_solver_create_games(int n_games, const int* n_per, const int64_t* m_per,
                                       const int* const* owner, const int* const* priority,
                                       const int* const* offsets, const int* const* targets,
                                       int flags, int trace_cap, psi_solver** out)
{
    std::vector<GameView> v(std::max(n_games, 0));
    for (int g = 0; g < n_games; g++)
        v[g] = GameView{n_per[g], (long long)m_per[g], owner[g], priority[g], offsets[g], targets[g]};
    return create_from_views(n_games, v.data(), flags, trace_cap, out);
}

extern "C" int psi_solver_create(const psi_game* g, int flags, int trace_cap, psi_solver** out)
{
    int64_t m = g->m;
    return psi_solver_create_batch(1, &g->n, &m, g->owner.data(), g->priority.data(),
                                   g->offsets.data(), g->targets.data(), flags, trace_cap, out);
}

extern "C" int psi_solver_destroy(psi_solver* s)
{
    if (!s) return PSI_OK;
    (void)hipStreamSynchronize(s->s);
    std::lock_guard<std::recursive_mutex> lk(setup_mutex());
    core_destroy(s->core);
    (void)hipStreamDestroy(s->s);
    delete s;
    return PSI_OK;
}

extern "C" int psi_solver_info(psi_solver* s, int* n_games, int64_t* n_total, int64_t* m_total)
{
    if (n_games) *n_games = s->empty ? 1 : s->core.G;
    if (n_total) *n_total = s->core.N;
    if (m_total) *m_total = s->core.M;
    return PSI_OK;
}

extern "C" int psi_init_strat(psi_solver* s) { return s->empty ? PSI_OK : core_init_strat(s->core); }

extern "C" int psi_compute_valuation(psi_solver* s)
{
    return s->empty ? PSI_OK : core_valuation(s->core, false);
}

extern "C" int psi_mark_solved(psi_solver* s, int player)
{
    if (player != 0 && player != 1) return fail(PSI_ERR_INVALID, "player must be 0 or 1");
    return s->empty ? PSI_OK : core_mark_solved(s->core, player);
}

extern "C" int psi_switch_strategy(psi_solver* s, int player, int* switched)
{
    if (player != 0 && player != 1) return fail(PSI_ERR_INVALID, "player must be 0 or 1");
    if (s->empty) {
        *switched = 0;
        return PSI_OK;
    }
    return core_switch(s->core, player, switched);
}

extern "C" int psi_reset_opponent_strategy(psi_solver* s)
{
    return s->empty ? PSI_OK : core_reset_opp(s->core);
}

// The solves of a game with no vertices.
static int empty_solve(psi_solver* s, int br)
{
    s->e_iter = 1;
    s->e_br = br;
    return PSI_OK;
}

static int check_game(psi_solver* s, int game);

extern "C" int psi_si_tune(int param, int value)
{
    if (param == PSI_SIT_PROGRAM) {
        if (value != 0 && value != 1) return fail(PSI_ERR_INVALID, "psi_si_tune: program is 0 or 1");
        g_si_tune.program = value;
        return PSI_OK;
    }
    if (param == PSI_SIT_TWO_LEVELS) {
        if (value != 0 && value != 1) return fail(PSI_ERR_INVALID, "psi_si_tune: two levels is 0 or 1");
        g_si_tune.two_levels = value;
        return PSI_OK;
    }
    if (param == PSI_SIT_PUSH_MIN_N) {
        if (value < 0) return fail(PSI_ERR_INVALID, "psi_si_tune: push_min_n is 0 or positive");
        g_si_tune.push_min_n = value;
        return PSI_OK;
    }
    if (param == PSI_SIT_PUSH) {
        if (value < 0) return fail(PSI_ERR_INVALID, "psi_si_tune: push is 0 or positive");
        g_si_tune.push = value;
        return PSI_OK;
    }
    if (value < 64 || value > (1 << 20)) return fail(PSI_ERR_INVALID, "psi_si_tune: value out of range");
    switch (param) {
    case PSI_SIT_PHASE_GRID: g_si_tune.phase_grid = (unsigned)value; return PSI_OK;
    case PSI_SIT_JUMP_GRID: g_si_tune.jump_grid = (unsigned)value; return PSI_OK;
    case PSI_SIT_STEP_GRID: g_si_tune.step_grid = (unsigned)value; return PSI_OK;
    case PSI_SIT_POOL_GRID: g_si_tune.pool_grid = (unsigned)value; return PSI_OK;
    default: return fail(PSI_ERR_INVALID, "psi_si_tune: bad parameter");
    }
}

extern "C" int psi_solve_bf(psi_solver* s, int max_rounds, int steps_per_sync)
{
    if (s->empty) return empty_solve(s, 0);
    Core& c = s->core;
    if (int rc = core_init_strat(c)) return rc;
    if (int rc = core_set_params(c, 0, max_rounds, 1)) return rc;
    if (steps_per_sync <= 0) steps_per_sync = 4;
    if (int rc = core_solve_bf(c, steps_per_sync)) return rc;
    std::vector<GameCtl> gc(c.G);
    PSI_HIP(hipMemcpyAsync(gc.data(), c.gctl, c.G * sizeof(GameCtl), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    for (int g = 0; g < c.G; g++)
        if (gc[g].err)
            return fail(PSI_ERR_INVALID, "game " + std::to_string(g) +
                                             ": Game contains an odd cycle that uses only odd nodes");
    return PSI_OK;
}

// solve_si with CPUBV semantics (the cpubv mode, main.cpp:163-167): the
// canonical solve, in which a game stops at the first valuation that finds
// a strategy cycle whose largest priority is odd (cpu_bv.cpp:73-84 prints
// the message and exits).
extern "C" int psi_solve_bv(psi_solver* s, int max_rounds, int steps_per_sync)
{
    if (s->empty) return empty_solve(s, 1);
    Core& c = s->core;
    if (int rc = core_init_strat(c)) return rc;
    if (int rc = core_set_params(c, 0, max_rounds, 2)) return rc;
    if (steps_per_sync <= 0) steps_per_sync = 4;
    if (int rc = core_solve(c, steps_per_sync)) return rc;
    std::vector<GameCtl> gc(c.G);
    PSI_HIP(hipMemcpyAsync(gc.data(), c.gctl, c.G * sizeof(GameCtl), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    for (int g = 0; g < c.G; g++)
        if (gc[g].err)
            return fail(PSI_ERR_INVALID, "game " + std::to_string(g) +
                                             ": Game contains an odd cycle that uses only odd nodes");
    return PSI_OK;
}

// solve_si with CPUList's semantics (the cpulist mode, main.cpp:178-191;
// CPUList::compute_valuation, cpu_list_ranking.cpp:242-255): a vertex whose
// strategy reaches the sink is finite, every other vertex +inf, and solved
// flags are recorded but never read.  Such runs need not end (Odd can close
// an odd cycle and reopen it): max_rounds caps them (PSI_ERR_CAP).
extern "C" int psi_solve_list(psi_solver* s, int max_rounds, int steps_per_sync)
{
    if (s->empty) return empty_solve(s, 1);
    Core& c = s->core;
    if (int rc = core_init_strat(c)) return rc;
    if (int rc = core_set_params(c, 0, max_rounds, 3)) return rc;
    if (steps_per_sync <= 0) steps_per_sync = 4;
    if (int rc = core_solve(c, steps_per_sync)) return rc;
    std::vector<GameCtl> gc(c.G);
    PSI_HIP(hipMemcpyAsync(gc.data(), c.gctl, c.G * sizeof(GameCtl), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    for (int g = 0; g < c.G; g++)
        if (gc[g].capped) return fail(PSI_ERR_CAP, "game " + std::to_string(g) + ": round cap reached");
    return PSI_OK;
}

extern "C" int psi_get_bf_status(psi_solver* s, int game, int* status)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) {
        *status = 0;
        return PSI_OK;
    }
    Core& c = s->core;
    GameCtl gc;
    PSI_HIP(hipMemcpyAsync(&gc, c.gctl + game, sizeof gc, hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    *status = gc.err ? 1 : 0;
    return PSI_OK;
}

extern "C" int psi_solve(psi_solver* s, int reset, int max_rounds, int steps_per_sync)
{
    if (s->empty) return empty_solve(s, 1);
    Core& c = s->core;
    if (int rc = core_init_strat(c)) return rc;
    if (int rc = core_set_params(c, reset, max_rounds)) return rc;
    if (steps_per_sync <= 0) steps_per_sync = 4;
    return core_solve(c, steps_per_sync);
}

extern "C" int psi_solve_profiled(psi_solver* s, int reset, int max_rounds, double* times_ms,
                                  int64_t* counts)
{
    if (s->empty) {
        std::fill(times_ms, times_ms + PSI_K_COUNT, 0.0);
        std::fill(counts, counts + PSI_K_COUNT, (int64_t)0);
        return empty_solve(s, 1);
    }
    Core& c = s->core;
    if (int rc = core_init_strat(c)) return rc;
    if (int rc = core_set_params(c, reset, max_rounds)) return rc;
    std::vector<long long> cn(PSI_K_COUNT);
    int rc = core_solve_profiled(c, times_ms, cn.data());
    for (int k = 0; k < PSI_K_COUNT; k++) counts[k] = cn[k];
    return rc;
}

extern "C" int psi_get_profile_rounds(psi_solver* s, int* out, int cap, int* count)
{
    const std::vector<int>& r = s->core.prof_rounds;
    *count = (int)r.size();
    for (int k = 0; k < cap && k < (int)r.size(); k++) out[k] = r[k];
    return PSI_OK;
}

extern "C" int psi_solve_program_profiled(psi_solver* s, int reset, int max_rounds, double* times_ms,
                                          int64_t* counts)
{
    if (s->empty) {
        std::fill(times_ms, times_ms + PSI_K_COUNT, 0.0);
        std::fill(counts, counts + PSI_K_COUNT, (int64_t)0);
        return empty_solve(s, 1);
    }
    Core& c = s->core;
    if (int rc = core_init_strat(c)) return rc;
    if (int rc = core_set_params(c, reset, max_rounds)) return rc;
    std::vector<long long> cn(PSI_K_COUNT);
    int rc = core_solve_program_profiled(c, times_ms, cn.data());
    for (int k = 0; k < PSI_K_COUNT; k++) counts[k] = cn[k];
    return rc;
}

// Launch timeline of the last psi_solve_program_profiled: the phase word
// each launch ran (low 7 bits: the phase) and its kernel duration in us.
extern "C" int psi_get_program_launches(psi_solver* s, int* phase, float* us, int cap, int* count)
{
    const Core& c = s->core;
    *count = (int)c.prof_phase.size();
    for (int k = 0; k < cap && k < *count; k++) {
        if (phase) phase[k] = c.prof_phase[k];
        if (us) us[k] = c.prof_us[k];
    }
    return PSI_OK;
}

static int check_game(psi_solver* s, int game)
{
    if (game < 0 || game >= (s->empty ? 1 : s->core.G))
        return fail(PSI_ERR_INVALID, "game index out of range");
    return PSI_OK;
}

template <class T> static int copy_game(psi_solver* s, int game, const T* dev, T* out)
{
    Core& c = s->core;
    int vb = c.h_vbase[game], n = c.h_n[game];
    PSI_HIP(hipMemcpyAsync(out, dev + vb, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    return PSI_OK;
}

extern "C" int psi_get_stats(psi_solver* s, int game, int* iter, int* br, int* capped)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) {
        if (iter) *iter = s->e_iter;
        if (br) *br = s->e_br;
        if (capped) *capped = 0;
        return PSI_OK;
    }
    Core& c = s->core;
    GameCtl gc;
    PSI_HIP(hipMemcpyAsync(&gc, c.gctl + game, sizeof gc, hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    if (iter) *iter = gc.iter;
    if (br) *br = gc.br;
    if (capped) *capped = gc.capped;
    return PSI_OK;
}

// Strategy in the game's local vertex ids (-1 = sink).
extern "C" int psi_get_strategy(psi_solver* s, int game, int* out)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) return PSI_OK;
    if (int rc = copy_game(s, game, s->core.strat, out)) return rc;
    int vb = s->core.h_vbase[game], n = s->core.h_n[game];
    for (int v = 0; v < n; v++)
        if (out[v] >= 0) out[v] -= vb;
    return PSI_OK;
}

// infinite[] labels: +1, 0, -1 (cpu_parity.h:15).
extern "C" int psi_get_labels(psi_solver* s, int game, int* out)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) return PSI_OK;
    int n = s->core.h_n[game];
    std::vector<int8_t> t(n);
    if (int rc = copy_game(s, game, s->core.lab, t.data())) return rc;
    for (int v = 0; v < n; v++) out[v] = t[v];
    return PSI_OK;
}

extern "C" int psi_get_solved(psi_solver* s, int game, int* out)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) return PSI_OK;
    int n = s->core.h_n[game];
    std::vector<int8_t> t(n);
    if (int rc = copy_game(s, game, s->core.solved, t.data())) return rc;
    for (int v = 0; v < n; v++) out[v] = t[v];
    return PSI_OK;
}

// Path length to the sink of finite vertices after the last valuation (0 otherwise).
extern "C" int psi_get_depths(psi_solver* s, int game, int* out)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) return PSI_OK;
    if (int rc = copy_game(s, game, s->core.dep, out)) return rc;
    int n = s->core.h_n[game];
    std::vector<int8_t> t(n);
    if (int rc = copy_game(s, game, s->core.lab, t.data())) return rc;
    for (int v = 0; v < n; v++)
        if (t[v] != 0) out[v] = 0;
    return PSI_OK;
}

// Valuation digests sum_p count_v[p] * splitmix64(p) of finite vertices (trace mode).
extern "C" int psi_get_digests(psi_solver* s, int game, uint64_t* out)
{
    if (int rc = check_game(s, game)) return rc;
    if (!s->trace) return fail(PSI_ERR_INVALID, "digests need a solver created with PSI_SOLVER_TRACE");
    if (s->empty) return PSI_OK;
    return copy_game(s, game, (const uint64_t*)s->core.dig, out);
}

// Per-round hashes (strategy valued, solved, labels, digests), 4 per round.
extern "C" int psi_get_trace(psi_solver* s, int game, uint64_t* hashes, int cap, int* count)
{
    if (int rc = check_game(s, game)) return rc;
    Core& c = s->core;
    if (!s->trace) return fail(PSI_ERR_INVALID, "trace needs a solver created with PSI_SOLVER_TRACE");
    if (s->empty) {
        // every round values empty arrays: each hash is the empty sum
        const int k = std::max(0, std::min(s->e_br, cap));
        std::fill(hashes, hashes + (size_t)k * 4, (uint64_t)0);
        *count = k;
        return PSI_OK;
    }
    GameCtl gc;
    PSI_HIP(hipMemcpyAsync(&gc, c.gctl + game, sizeof gc, hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    int k = std::min(std::min(gc.br, c.trace_cap), cap);
    if (k > 0) {
        PSI_HIP(hipMemcpyAsync(hashes, c.tr + (size_t)game * c.trace_cap * 4, (size_t)k * 4 * 8,
                               hipMemcpyDeviceToHost, c.s));
        PSI_HIP(hipStreamSynchronize(c.s));
    }
    *count = k;
    return PSI_OK;
}

// get_winning(out, player) (cpu_parity.cpp:288-297, gpu_parity.cu:340-354): player 0 gets the +inf
// vertices, player 1 every other vertex, ascending local ids.
extern "C" int psi_get_winning(psi_solver* s, int game, int player, int* out, int* len)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) {
        *len = 0;
        return PSI_OK;
    }
    int n = s->core.h_n[game];
    std::vector<int> lab(n);
    if (int rc = psi_get_labels(s, game, lab.data())) return rc;
    int k = 0;
    for (int v = 0; v < n; v++)
        if ((player == 0) == (lab[v] == 1)) out[k++] = v;
    *len = k;
    return PSI_OK;
}

// Loads a state (local ids) into game `game`: strategy, solved flags, labels.
extern "C" int psi_set_state(psi_solver* s, int game, const int* strat, const int* solved,
                             const int* labels)
{
    if (int rc = check_game(s, game)) return rc;
    if (s->empty) return PSI_OK;
    Core& c = s->core;
    int vb = c.h_vbase[game], n = c.h_n[game];
    std::vector<int> st(n);
    std::vector<int8_t> so(n), la(n);
    for (int v = 0; v < n; v++) {
        if (strat[v] < -1 || strat[v] >= n) return fail(PSI_ERR_INVALID, "strategy out of range");
        st[v] = strat[v] >= 0 ? strat[v] + vb : -1;
        so[v] = solved ? (int8_t)(solved[v] != 0) : 0;
        la[v] = labels ? (int8_t)labels[v] : 0;
    }
    PSI_HIP(hipMemcpyAsync(c.strat + vb, st.data(), n * 4, hipMemcpyHostToDevice, c.s));
    PSI_HIP(hipMemcpyAsync(c.solved + vb, so.data(), n, hipMemcpyHostToDevice, c.s));
    PSI_HIP(hipMemcpyAsync(c.lab + vb, la.data(), n, hipMemcpyHostToDevice, c.s));
    // a loaded strategy need not be a successor, which the switch's change
    // marks assume of the current choice
    c.push_ok = false;
    if (int rc = core_state_loaded(c, game)) return rc;
    PSI_HIP(hipStreamSynchronize(c.s));
    return PSI_OK;
}

extern "C" int psi_solver_synchronize(psi_solver* s)
{
    PSI_HIP(hipStreamSynchronize(s->s));
    return PSI_OK;
}

// An event (psi_event_create) recorded on the handle's own stream, the one
// its kernels run on.
extern "C" int psi_solver_event_record(psi_solver* s, void* ev)
{
    PSI_HIP(hipEventRecord((hipEvent_t)ev, s->s));
    return PSI_OK;
}

// Jump-round counters of the last valuation: out[2r] = vertices resolved in
// round r, out[2r+1] = vertices still jumping after round r (r < rounds);
// *last = last round that ran, *err = 1 if the launched rounds ran out.
extern "C" int psi_get_jump_stats(psi_solver* s, int* out, int rounds, int* last, int* err)
{
    if (s->empty) {
        std::fill(out, out + 2 * std::max(0, std::min(rounds, kMaxRounds)), 0);
        *last = -1;
        *err = 0;
        return PSI_OK;
    }
    Core& c = s->core;
    JumpCnt jc;
    PSI_HIP(hipMemcpyAsync(&jc, c.jc, sizeof jc, hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    for (int r = 0; r < rounds && r < kMaxRounds; r++) {
        unsigned long long res = 0, act = 0;
        for (int k = 0; k < kSlots; k++) {
            res += jc.slot[r][k] >> 32;
            act += jc.slot[r][k] & 0xFFFFFFFFu;
        }
        out[2 * r] = (int)res;
        out[2 * r + 1] = (int)act;
    }
    *last = jc.last;
    *err = jc.err;
    return PSI_OK;
}

// All games at once, concatenated in game order: strategies (local ids, -1
// sink), labels, and per game (iter, br, capped).  Any pointer may be null.
extern "C" int psi_get_results(psi_solver* s, int* strat_all, int* labels_all, int* stats_all)
{
    if (s->empty) {
        if (stats_all) {
            stats_all[0] = s->e_iter;
            stats_all[1] = s->e_br;
            stats_all[2] = 0;
        }
        return PSI_OK;
    }
    Core& c = s->core;
    if (c.N * 4 >= (32LL << 20)) {
        // large: through pinned chunks, localised and widened on the host pool
        if (strat_all) {
            int rc = staged_download(c.s, c.strat, (size_t)c.N * 4, [&](const char* buf, size_t off, size_t len) {
                const long long v0 = (long long)(off / 4), nv = (long long)(len / 4);
                const int* in = (const int*)buf;
                parallel_for(nv, [&](long long lo, long long hi) {
                    long long v = v0 + lo;
                    int g = (int)(std::upper_bound(c.h_vbase.begin(), c.h_vbase.end(), (int)v) - c.h_vbase.begin()) - 1;
                    for (long long k = lo; k < hi; k++, v++) {
                        while (g + 1 < c.G && v >= c.h_vbase[g + 1]) g++;
                        const int x = in[k];
                        strat_all[v] = x >= 0 ? x - c.h_vbase[g] : x;
                    }
                });
            });
            if (rc) return rc;
        }
        if (labels_all) {
            int rc = staged_download(c.s, c.lab, (size_t)c.N, [&](const char* buf, size_t off, size_t len) {
                const int8_t* in = (const int8_t*)buf;
                parallel_for((long long)len, [&](long long lo, long long hi) {
                    for (long long k = lo; k < hi; k++) labels_all[off + k] = in[k];
                });
            });
            if (rc) return rc;
        }
        if (stats_all) {
            std::vector<GameCtl> gc(c.G);
            PSI_HIP(hipMemcpyAsync(gc.data(), c.gctl, c.G * sizeof(GameCtl), hipMemcpyDeviceToHost, c.s));
            PSI_HIP(hipStreamSynchronize(c.s));
            for (int g = 0; g < c.G; g++) {
                stats_all[3 * g] = gc[g].iter;
                stats_all[3 * g + 1] = gc[g].br;
                stats_all[3 * g + 2] = gc[g].capped;
            }
        }
        return PSI_OK;
    }
    if (strat_all) {
        PSI_HIP(hipMemcpyAsync(strat_all, c.strat, (size_t)c.N * 4, hipMemcpyDeviceToHost, c.s));
    }
    std::vector<int8_t> lab;
    if (labels_all) {
        lab.resize(c.N);
        PSI_HIP(hipMemcpyAsync(lab.data(), c.lab, (size_t)c.N, hipMemcpyDeviceToHost, c.s));
    }
    std::vector<GameCtl> gc;
    if (stats_all) {
        gc.resize(c.G);
        PSI_HIP(hipMemcpyAsync(gc.data(), c.gctl, c.G * sizeof(GameCtl), hipMemcpyDeviceToHost, c.s));
    }
    PSI_HIP(hipStreamSynchronize(c.s));
    parallel_games(c.G, [&](int g) {
        long long vb = c.h_vbase[g], n = c.h_n[g];
        if (strat_all)
            for (long long v = vb; v < vb + n; v++)
                if (strat_all[v] >= 0) strat_all[v] -= (int)vb;
        if (labels_all)
            for (long long v = vb; v < vb + n; v++) labels_all[v] = lab[v];
        if (stats_all) {
            stats_all[3 * g] = gc[g].iter;
            stats_all[3 * g + 1] = gc[g].br;
            stats_all[3 * g + 2] = gc[g].capped;
        }
    });
    return PSI_OK;
}

// CSR of a game (the arrays given to psi_game_create); sizes from psi_game_info.
extern "C" int psi_game_csr(const psi_game* g, int* owner, int* priority, int* offsets, int* targets)
{
    std::copy(g->owner.begin(), g->owner.end(), owner);
    std::copy(g->priority.begin(), g->priority.end(), priority);
    std::copy(g->offsets.begin(), g->offsets.end(), offsets);
    std::copy(g->targets.begin(), g->targets.end(), targets);
    return PSI_OK;
}
