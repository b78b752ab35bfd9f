// Device-resident strategy-improvement engine shared by the single-game
// solver (GPUParity / GPUList roles) and the batched and serving forms.
// Many games are solved together as one disjoint union; every game keeps its
// own control block, so the SI control flow (inner loop, outer loop,
// convergence) of each game is exactly that of solve_si (gpumain.cpp:18-53)
// run on that game alone.
#pragma once

#include "common.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace psi {

// Runs f(g) for g in [0, G) on up to 16 host threads.
template <class F> void parallel_games(int G, F f)
{
    int T = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    T = std::min(T, G);
    if (T <= 1) {
        for (int g = 0; g < G; g++) f(g);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([&] {
            for (int g; (g = next.fetch_add(1)) < G;) f(g);
        });
    for (auto& x : th) x.join();
}

// Per-game control block (device memory).
struct GameCtl {
    int done;       // 1 once switch_strategy(0) switched nothing (or capped)
    int br;         // valuation rounds (gpumain.cpp:34)
    int iter;       // outer iterations (gpumain.cpp:45)
    int sw1;        // a player-1 switch happened this step
    int sw0;        // a player-0 switch happened this step
    int outer;      // this step runs mark_solved(0) + switch_strategy(0)
    int reset_pending;
    int capped;
    int max_rounds; // 0 = no cap
    int reset;      // cpu-reset mode (main.cpp:25-26)
    int vbase;      // first global vertex id of the game
    int n;          // vertices of the game
    int clear0;     // the last step was an outer step: restart the switch(0) change record
    int bf;         // 1: solve_bf (main.cpp:57-77), later best responses by relaxation;
                    // 2: CPUBV's solve_si (main.cpp:163-167), every valuation stops on an
                    // odd strategy cycle; 3: CPUList's semantics (cpulist): vertices not
                    // reaching the sink +inf, solved flags not read by the valuation
    int bf_init;    // solve_bf: the first best response (an SI inner loop) is done
    int err;        // odd strategy cycle: the game stopped there (cpu_bv.cpp:73-84)
    int activate;   // serving pool, step program: admitted (done until the next step boundary)
};

// Jump-round bookkeeping (device memory), rewritten by every valuation.
// Counts are spread over kSlots addresses per round (one atomic per block into
// slot blockIdx % kSlots) so that no single address takes more than a few
// hundred atomics per round.
constexpr int kSlots = 64;
constexpr int kCntStride = 32;   // 128-byte lines
constexpr int kMaxRounds = 40;
struct JumpCnt {
    unsigned long long slot[kMaxRounds][kSlots]; // (resolved << 32) | still-jumping
    unsigned long long summary[2];               // by round parity: (stop + 1) << 32 | cycle-bound count
    int last;       // last round that ran
    int err;        // 1: ran out of launched rounds
};

// Device state of the step program (the solve loop as a sequence of identical
// launches, each running the next phase of the loop), indexed by launch number
// i so that launch i+1 finds launch i's words without a dependent load:
// st[i & 1] = the phase launch i ran (phase | round << 8); psum[i & 1] = its
// jump-round decision word; pslot[i & 3] = its jump round's counters and
// pmis[i & 3] its cycle-label mismatch flags (both zeroed by launch i-2).  A
// graph of these launches holds a multiple of 4.
struct StepProg {
    int st[2];
    int any_outer;  // some game takes the outer step (set by the control phase)
    int live;       // games not done
    int host_seq;   // serving pool: admissions written by the host (between graphs)
    int dev_seq[2]; // by launch parity: admissions activated as of that launch
    unsigned long long psum[2];
    int pushed;     // the last switch(1) ran on change marks (few valuations changed)
    int bf0;        // game 0's mode (bf of its control block) as the program started
    unsigned long long pslot[4][kSlots];
    unsigned pmis[4][kSlots];
    unsigned ccnt[kSlots * kCntStride]; // vertices whose valuation changed at the last
                                        // valuation, over kSlots slots a cache line apart
    int full0;      // the next switch(0) scans every entry (a valuation since its last pass
                    // set no marks)
};

struct Core {
    int G = 0;
    long long N = 0, M = 0;
    int lmax = 0;          // jump rounds launched per valuation
    bool trace = false;
    int trace_cap = 0;     // rounds of trace per game
    int max_pri = 0;
    // game (device)
    int* vinfo = nullptr;  // priority | owner << 30
    int* off = nullptr;    // N+1 global CSR offsets
    int* tgt = nullptr;    // M global target ids
    int* gid = nullptr;    // N game of each vertex
    // state (device)
    int* strat = nullptr;
    int* par = nullptr;    // link at the last valuation: successor, -1 sink, -2/-3 solved +inf/-inf
    int* dep = nullptr;    // path length to the sink of finite vertices
    int* key = nullptr;    // (label+1) << 30 | max priority on the path
    int8_t* solved = nullptr;
    int8_t* lab = nullptr; // infinite[]: -1, 0, +1
    uint8_t* blive = nullptr; // per 256-vertex block: holds a vertex of an unfinished game
    uint8_t* bact = nullptr;  // 2 x blocks: block holds a vertex still jumping (by round parity)
    uint8_t* blast = nullptr; // per block: last jump level written this valuation
    int2* Sl = nullptr;    // jump records: level j at Sl + level_slot(j)*N
    unsigned long long* chg = nullptr;  // bit per vertex: valuation changed at the last valuation,
                                        // then (at chg + chg_words) since the last switch(0) pass
    long long chg_words = 0;
    int levels = 0;        // buffers allocated
    int* own[2] = {nullptr, nullptr};   // vertices owned by each player
    long long own_n[2] = {0, 0};
    // One game (not a pool): in-edge lists, each in-edge as its source's
    // owner-list entry (index << 1 | owner), and per owner-list entry a mark
    // set when a successor's valuation changed (the step program's switch
    // visits only marked entries when few valuations changed).
    int* rin_off = nullptr;            // N+1
    int* rin_src = nullptr;            // M
    uint8_t* dmark[2] = {nullptr, nullptr};
    // Step program, one game: d1 marks (a byte per vertex, the valuation's
    // tag) the vertices whose link changed for the coming valuation and
    // their link-children; l12ok
    // (per block) says the block's level-1 and level-2 jump records were
    // written by the last valuation's first jump launch.  A first jump launch
    // after a switch on change marks takes a vertex's level-1 and level-2
    // records from the last valuation unless the vertex or its 2nd successor
    // is marked (then the first four vertices of its path are the same).
    uint8_t* d1 = nullptr;
    uint8_t* l12ok = nullptr;
    // bchg (per block, step program, one game): some record of the block
    // ended this valuation with the change bit (set by the INIT for its
    // terminals, by the jump launches for what they resolve); the label
    // launch only writes the change words of a block without it and without
    // a vertex still jumping after the last round.
    uint8_t* bchg = nullptr;
    unsigned char* rin_mem = nullptr;  // the allocation these live in
    bool push_ok = false;              // marks usable (cleared when a state is loaded)
    unsigned long long* Zl = nullptr;  // trace: digests per jump level (as Sl)
    unsigned long long* dig = nullptr;
    GameCtl* gctl = nullptr;
    JumpCnt* jc = nullptr;
    unsigned long long* tr = nullptr;  // G * trace_cap * 4
    int* flag = nullptr;               // [0] step-wise switch flag, [1] live count,
                                       // [2] one past the last occupied vertex
    bool pool = false;                 // serving pool (games come and go)
    int maxn_cap = 0;                  // pool: longest admissible game
    int* cnt_scratch = nullptr;        // pool: owner-list block counts of one admission
    int* succ = nullptr;               // solve_bf: CPU_BellmanFord's succ (allocated on first use)
    hipStream_t s = nullptr;
    // host copies of per-game sizes
    std::vector<int> h_n, h_vbase;
    std::vector<long long> h_m;
    unsigned long long* prof_ties = nullptr;  // profiled solves: compares taking the exact path
    std::vector<int> prof_rounds;             // last profiled solve: jump rounds run per valuation
    std::vector<int> prof_phase;              // last program-profiled solve: phase of each launch
    std::vector<float> prof_us;               // ... and its duration (kernel start to end)
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    StepProg* prog = nullptr;          // step program state (not in a serving pool)
    void* pargs = nullptr;             // step program arguments (device copy)
    int pargs_two = -1;                // two_levels in that copy
    long long pargs_push = -1;         // and its mark threshold
    hipGraphExec_t pgraph = nullptr;   // captured step-program launches
    int pgraph_launches = 0;
    unsigned pgraph_grid = 0;          // workgroups per launch in that graph
};

// Launch shapes of the SI step kernels (grid-stride launches).
struct SiTuning {
    unsigned jump_grid = 2048;   // blocks of a pointer-jumping round
    unsigned step_grid = 16384;  // blocks of val_init / label / switch / mark
    unsigned pool_grid = 4096;   // the same in a serving pool
    unsigned phase_grid = 2048;  // blocks of a step-program launch
    int program = 1;             // solve loop as a step program (else fixed launches per step)
    int two_levels = 1;          // step program: jump launches advance two levels
    int push = 16;               // step program, one game: marks when at most N / push
                                 // valuations changed (0: never)
    int push_min_n = 1 << 20;    // ... on games of at least this many vertices (the in-edge
                                 // lists are built at creation for those only)
};
extern SiTuning g_si_tune;

// One game's host CSR arrays (local vertex ids).
struct GameView {
    int n;
    long long m;
    const int *owner, *pri, *off, *tgt;
};

// nullptr, or the reason the game is refused (the reference's messages where
// it has one); *max_pri = largest priority.
const char* core_validate_game(const GameView& gv, int* max_pri);
// Builds a core from G games (validated; packed on host threads).
int core_create(Core& c, int G, const GameView* games, bool trace, int trace_cap, hipStream_t s);
void core_destroy(Core& c);
int core_init_strat(Core& c);
int core_set_params(Core& c, int reset, int max_rounds, int bf = 0);
// One valuation (val_init, jump rounds, labels); fuse_mark1 also applies
// mark_solved(1) and counts the round (the device-resident loop's form).
int core_valuation(Core& c, bool loop_form);
int core_mark_solved(Core& c, int player);
// After strategies, solved flags and labels of one game were loaded from the host.
int core_state_loaded(Core& c, int game);
int core_reset_opp(Core& c);
int core_switch(Core& c, int player, int* switched_host);   // step-wise, syncs
// Enqueues `steps` SI steps (device-resident loop body).
int core_enqueue_steps(Core& c, int steps);
// Runs the loop until every game is done; graph-captured step batches.
int core_solve(Core& c, int steps_per_check);
// Launches one graph of `steps` SI steps (captured on first use) on c.s.
int core_launch_steps(Core& c, int steps);
int core_all_done(Core& c, int* done);
// Serving pool: fixed capacity, games admitted and released between steps.
// Slot G-1 is a permanently finished sentinel that free vertices belong to.
int core_create_pool(Core& c, long long vcap, long long ecap, int slots, int maxn, hipStream_t s);
long long core_pool_bytes_per_vertex(int maxn);
// Places a validated game at vertices [vbase, vbase+n] (the last closes its
// offsets) and edges [ebase, ebase+m) on stream s2; it joins the steps once
// its control block is written on the step stream.
int core_pool_admit(Core& c, int slot, long long vbase, long long ebase, const GameView& gv,
                    hipStream_t s2);
int core_pool_localize(Core& c, int slot, hipStream_t s2);
int core_pool_release(Core& c, int slot, hipStream_t s2);
// Serving pool under the step program: launches one graph of 16 x steps
// program launches (admitted games join at the next step boundary).
int core_pool_launch_program(Core& c, int steps);
int core_solve_profiled(Core& c, double* times, long long* counts);
int core_solve_program_profiled(Core& c, double* times, long long* counts);
// solve_bf for every game of the core (after core_init_strat and
// core_set_params(..., bf = 1)): the device loop, then the Odd strategy of
// games with at least one relaxation replaced by succ (cpu_bf.cpp:156-165).
int core_solve_bf(Core& c, int steps_per_check);

} // namespace psi
