/*
 * psi.h - C ABI of the MI355X (gfx950) strategy-improvement parity-game
 * solver.  Plain C types only: int32 vertex ids, host or device pointers as
 * stated per function, int status codes (0 = ok, < 0 = error, message in
 * psi_last_error()).  No function prints or exits.
 *
 * The reference exposes this path only as C++ classes (no C ABI).  Each entry
 * point below names the reference interface it replaces, file:line in
 * jfearnley/parallel-si.
 *
 * Threading: a handle is used by one host thread at a time; distinct handles
 * may be used concurrently.  The standalone list-ranking entry points share
 * one library workspace: concurrent calls of psi_list_rank take turns, and
 * concurrent calls of psi_list_rank_noalloc race (its own note).
 */
#ifndef PSI_H
#define PSI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PSI_OK 0
#define PSI_ERR_INVALID -1   /* bad argument or malformed input               */
#define PSI_ERR_HIP -2       /* HIP runtime error (message has the details)   */
#define PSI_ERR_NO_DEVICE -3 /* no GPU visible                                */
#define PSI_ERR_CAP -4       /* round cap reached before convergence          */
#define PSI_ERR_BUDGET -5    /* device-memory budget exceeded (admission)     */

/* Last error message of the calling thread (replaces gpu_assert's
 * print-and-exit, gpu_util.cpp:10-17). */
const char* psi_last_error(void);

/* ---- runtime (gpu_util.h:8-16 roles) ------------------------------------ */
int psi_init(void);                       /* checks a device, creates the stream */
int psi_device_count(int* count);
int psi_synchronize(void);                /* waits for all library work          */
int psi_malloc(void** ptr, size_t bytes); /* init_zero / init_from_vector roles,  */
int psi_free(void* ptr);                  /* gpu_util.cpp:24-37                   */
int psi_memcpy_h2d(void* dst, const void* src, size_t bytes);
int psi_memcpy_d2h(void* dst, const void* src, size_t bytes);
int psi_memcpy_d2d(void* dst, const void* src, size_t bytes);
int psi_event_create(void** ev);
int psi_event_destroy(void* ev);
int psi_event_record(void* ev);           /* on the library stream               */
int psi_event_elapsed_ms(void* start, void* end, float* ms);

/* ---- list ranking ------------------------------------------------------- */
/* Replaces the list-ranking core of GPUList::compute_valuation
 * (gpu_list_ranking.cu:279-326: d_pick_splitters :96, d_traverse_sublists
 * :153, process_reduced :197, d_broadcast_val :244) and of CPUList
 * (cpu_list_ranking.cpp:108-239) as a standalone entry with unit weights.
 * succ[i] in [-1, n) (-1 ends a list); heads[] holds the first node of every
 * list.  rank[i] = distance of i from its head, -1 if no head reaches i.  */
#define PSI_LR_AUTO 0      /* by length: Wyllie (heads as given) up to PSI_LRT_WYLLIE_MAX
                              nodes, else WALK if the mean list length n/n_heads is
                              short, else HJ                                          */
#define PSI_LR_HJ 1        /* Helman-JaJa: hashed block splitters + head slots     */
#define PSI_LR_WYLLIE 2    /* pointer jumping; heads = nodes without predecessor   */
#define PSI_LR_WALK 3      /* head slots only (one walk per list, hop-capped)      */
#define PSI_LR_CHECKED 0x100 /* flag: pre-fill rank with -1 so unreached nodes read -1 */

/* Device pointers; asynchronous on the library stream. */
int psi_list_rank(const int* d_succ, int n, const int* d_heads, int n_heads, int* d_rank,
                  int method);
/* Host pointers; synchronous; always CHECKED; validates ranges. */
int psi_list_rank_host(const int* succ, int n, const int* heads, int n_heads, int* rank,
                       int method);
/* psi_list_rank with no allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * methods and results as psi_list_rank; its workspace and control words are
 * a fixed array of 64 MiB per GPU, statically allocated in the library and
 * held until the process ends, instead of psi_list_rank's grown workspace.
 * Inputs whose workspace would exceed it (e.g. PSI_LR_WYLLIE beyond ~2.7M
 * nodes; the 16,777,216-node list by HJ and the 1,048,576-list forest by
 * AUTO fit with the default densities)
 * are rejected with PSI_ERR_INVALID and a message naming the bytes needed.
 * Every call on a GPU uses that GPU's one array: calls that overlap on one
 * GPU (other host threads, other streams) race.  Calls no hipMalloc/hipFree;
 * psi_init resolves the array's device address (else the first call). */
int psi_list_rank_noalloc(const int* d_succ, int n, const int* d_heads, int n_heads, int* d_rank,
                          int method);
/* Device error bits of the last call of either entry point (1: index out of
 * range, 2: a walk did not end) and the number of overflow sublists the
 * level-0 walk created (Helman-JaJa; head walks write ranks directly and
 * create none). */
int psi_list_rank_status(int* error_bits, int* overflow_used);

/* Euler-tour path sums over a forest of links: the valuation engine of
 * GPUList (gpu_list_ranking.cu:52-92 d_init_list / d_build_list: the tour of
 * the sink tree, one atomic exchange per vertex into its parent's list;
 * :96-277 splitters, sublist walks, reduced list, broadcast: here psi_list_rank
 * on the 2n+2-node tour and a scan in tour order), with a 64-bit weight per
 * vertex instead of a dense count vector.  link[v] in [-1, n) or below -1:
 * -1 = child of the root (the sink), below -1 = detached.  sum[v] = sum of w
 * over v and its ancestors (mod 2^64), reach[v] = 1, for every v whose links
 * reach the root; sum 0, reach 0 otherwise (cycles, detached chains).  Device
 * pointers; method PSI_LR_AUTO, HJ or WALK; reach may be null. */
int psi_tour_path_sums(const int* d_link, int n, const uint64_t* d_w, uint64_t* d_sum, int* d_reach,
                       int method);

/* Tuning knobs (the autotuner sets these). */
#define PSI_LRT_LOG_D0 1        /* level-0 density 2^-v (2^24+ nodes; denser below) */
#define PSI_LRT_LOCAL_BITS 2    /* level-0 hop cap 2^v                           */
#define PSI_LRT_LOG_DK 3        /* reduced-level splitter density 2^-v           */
#define PSI_LRT_HOP_CAP_K 4     /* reduced-level hop cap                         */
#define PSI_LRT_WALK_MEAN_MAX 5 /* AUTO threshold on mean list length            */
#define PSI_LRT_WYLLIE_MAX 6    /* AUTO: Wyllie for at most this many nodes      */
int psi_list_rank_tune(int param, int value);

/* ---- games --------------------------------------------------------------- */
/* Replaces Game (game.h:9-34) as input and GPUGame (gpu_game.h:6-21,
 * gpu_game.cpp:11-41) as the int32 CSR layout: offsets[n+1] (offsets[n] = m),
 * targets[m] (file order = the switch tie-break order), owner[n] in {0,1},
 * priority[n] in [0, 2^30), n < 2^28.  Unlike the reference, successor ids are range
 * checked and negative priorities are rejected (both are undefined behaviour
 * there).  The arrays are copied.  n = 0 (m = 0, offsets = {0}) is the game
 * of a file with a header line only: a one-game solver of it has no device
 * state and solves as the reference does (iter 1, br 1; solve_bf br 0).  A
 * batch or a serving pool holds games of at least one vertex. */
typedef struct psi_game psi_game;
int psi_game_create(int n, int m, const int* owner, const int* priority, const int* offsets,
                    const int* targets, psi_game** out);
int psi_game_destroy(psi_game* g);
int psi_game_info(const psi_game* g, int* n, int* m, int* max_pri);
int psi_game_csr(const psi_game* g, int* owner, int* priority, int* offsets, int* targets);
/* GPUGame's public device arrays (gpu_game.h:11-16, uploaded as
 * gpu_game.cpp:11-41 does): player[n], priority[n], vertices[n+1] (offsets),
 * edges[m] (targets), int32 device pointers owned by the game handle (made on
 * the first call, freed by psi_game_destroy).  Pointers may be null. */
int psi_game_device_csr(psi_game* g, int** player, int** priority, int** vertices, int** edges);
/* Parses PGSolver text (Game::parse_pgsolver, game.cpp:51-107: header line
 * skipped; "id priority owner s1,s2,...[ label];" per vertex, ids in order
 * from 0).  Errors return PSI_ERR_INVALID with the reference's message. */
int psi_game_from_gm(const char* text, size_t len, psi_game** out);

/* ---- solvers --------------------------------------------------------------
 * Replaces GPUParity(Game&, GPUGame&) (gpu_parity.h:8-38) and GPUList
 * (gpu_list_ranking.h:11-41): allocates device state and uploads the CSR.
 * Semantics are CPUParity's (cpu_parity.cpp), the reference's canonical CPU
 * path, for every game: results are bit-exact with it.
 * A batch solver holds n_games games back to back (per-game n, m; owner and
 * priority concatenated; each game's n+1 offsets concatenated; targets in
 * local ids concatenated) and solves each as if alone (per-game control). */
#define PSI_SOLVER_TRACE 1   /* keep per-round hashes and valuation digests */
typedef struct psi_solver psi_solver;
int psi_solver_create(const psi_game* g, int flags, int trace_cap, psi_solver** out);
int psi_solver_create_batch(int n_games, const int* n_per, const int64_t* m_per, const int* owner,
                            const int* priority, const int* offsets, const int* targets, int flags,
                            int trace_cap, psi_solver** out);
/* The same with each game in its own arrays (no concatenation by the caller). */
int psi_solver_create_games(int n_games, const int* n_per, const int64_t* m_per,
                            const int* const* owner, const int* const* priority,
                            const int* const* offsets, const int* const* targets, int flags,
                            int trace_cap, psi_solver** out);
int psi_solver_destroy(psi_solver* s);
int psi_solver_info(psi_solver* s, int* n_games, int64_t* n_total, int64_t* m_total);
int psi_solver_synchronize(psi_solver* s);
/* Records an event (psi_event_create) on the handle's own stream, where its
 * kernels run (timing; no reference counterpart). */
int psi_solver_event_record(psi_solver* s, void* ev);

/* StrategyImprovement (si.h:6-16), one call each, all games of the handle:
 *   init_strat              gpu_parity.cu:51-56, cpu_parity.cpp:276-286
 *   compute_valuation       gpu_parity.cu:151-187, cpu_parity.cpp:116-133
 *                           (compute_first_val is the same)
 *   mark_solved(player)     cpu_parity.cpp:265-274 (GPUParity's is a no-op,
 *                           gpu_parity.cu:333-336)
 *   switch_strategy(player) gpu_parity.cu:303-317: *switched = 1 if any vertex
 *                           switched (the reference returns this 0/1 flag)
 *   reset_opponent_strategy cpu_parity.cpp:300-307 */
int psi_init_strat(psi_solver* s);
int psi_compute_valuation(psi_solver* s);
int psi_mark_solved(psi_solver* s, int player);
int psi_switch_strategy(psi_solver* s, int player, int* switched);
int psi_reset_opponent_strategy(psi_solver* s);

/* The whole solve_si loop (gpumain.cpp:18-53) on the device, every game with
 * its own convergence: reset = the cpu-reset mode (main.cpp:25-26);
 * max_rounds > 0 caps valuation rounds per game (then psi_get_stats reports
 * capped); steps_per_sync SI steps run per host check (a captured hipGraph). */
int psi_solve(psi_solver* s, int reset, int max_rounds, int steps_per_sync);
/* solve_bf (main.cpp:57-77) with the CPU_BellmanFord best response
 * (cpu_bf.cpp:25-166), every game of the handle on the device: the first best
 * response is an SI inner loop (CPUBV semantics: an odd strategy cycle stops
 * the game, as cpu_bv.cpp:73-84 exits; PSI_ERR_INVALID with that message and
 * psi_get_bf_status = 1); later ones are relaxations to Odd's best response,
 * in which a vertex that cannot reach the sink stays +inf.  Odd's final
 * strategy is succ (the last finite choice of each Odd vertex, 0 before any),
 * as the reference leaves it.  The relaxation runs in parallel rounds
 * (Jacobi, on exact path valuations) where the reference updates in place
 * (Gauss-Seidel): values, labels, solved sets, Even's strategies, iteration
 * counts and winning regions equal the reference's on every game whose
 * relaxations converge; Odd's succ equals it where the minimum successor is
 * unique; br counts device valuation rounds, not the reference's sweeps. */
int psi_solve_bf(psi_solver* s, int max_rounds, int steps_per_sync);
/* solve_si with CPUBV<ArrayVal> (the CPU command line's cpubv mode,
 * main.cpp:163-167; CPUBV::compute_valuation, cpu_bv.cpp:49-106): the
 * canonical solve, except that a game stops at the first valuation that finds
 * a strategy cycle whose largest priority is odd, where the reference prints
 * "Game contains an odd cycle that uses only odd nodes" and exits
 * (cpu_bv.cpp:73-84).  Such games report psi_get_bf_status = 1 and the call
 * returns PSI_ERR_INVALID with that message; iter/br then count what was
 * completed before the stop.  Other games end with CPUParity's result. */
int psi_solve_bv(psi_solver* s, int max_rounds, int steps_per_sync);
/* solve_si with CPUList<ArrayVal>'s semantics (the cpulist mode,
 * main.cpp:178-191; the valuation of cpu_list_ranking.cpp:242-255 and
 * gpu_list_ranking.cu:279-326): a vertex whose strategy reaches the sink is
 * finite, every other vertex +inf (broadcast_val, cpu_list_ranking.cpp:
 * 208-239); solved flags are recorded (CPUBV::mark_solved) but the valuation
 * never reads them.  Bit-exact with the reference's cpulist (counts,
 * strategies, labels, per-round hashes).  Such a run need not end (on test2.gm
 * Odd's vertices switch back and forth): max_rounds > 0 caps it, and a capped
 * game returns PSI_ERR_CAP (its state is the reference's after that round). */
int psi_solve_list(psi_solver* s, int max_rounds, int steps_per_sync);
int psi_get_bf_status(psi_solver* s, int game, int* status);   /* 1 = odd cycle */
int psi_get_stats(psi_solver* s, int game, int* iter, int* br, int* capped);
/* Profiling (no reference counterpart; the reference times solve_si as a
 * whole, gpumain.cpp:56-71).  The same solve launched kernel by kernel with
 * HIP events between launches (no graph): device milliseconds and launch
 * counts per kernel class.  Jump rounds are timed by their kernels' own start
 * and end (events attached to the dispatch, as a profiler reads them); the
 * other classes by the events around each launch. */
#define PSI_K_VAL_INIT 0
#define PSI_K_JUMP 1      /* pointer-jumping rounds that ran               */
#define PSI_K_JUMP_IDLE 2 /* rounds launched after convergence (no-ops)    */
#define PSI_K_LABEL 3
#define PSI_K_SWITCH1 4
#define PSI_K_SWITCH0 5
#define PSI_K_CTL 6
#define PSI_K_TOTAL 7
#define PSI_K_GATHERS 8   /* counts only: successor records gathered by jump rounds */
#define PSI_K_TIES 9      /* counts only: compares that took the exact (equal path max) path */
#define PSI_K_SCANS 10    /* counts only: exact compares that walked the private segments */
#define PSI_K_SCAN_STEPS 11 /* counts only: vertices visited by those walks */
#define PSI_K_PUSH 12     /* step program: change marks for the next switch passes */
#define PSI_K_COUNT 13
int psi_solve_profiled(psi_solver* s, int reset, int max_rounds, double* times_ms,
                       int64_t* counts);  /* both arrays have PSI_K_COUNT entries */
/* Jump rounds that ran in each valuation of the last psi_solve_profiled, in
 * order (its launches per valuation: lmax rounds, these first). */
int psi_get_profile_rounds(psi_solver* s, int* out, int cap, int* count);
/* The step program (PSI_SIT_PROGRAM) launched one launch at a time with HIP
 * events around each: the same classes by the phase each launch ran (jump
 * rounds after convergence are never launched; mark_solved(0) counts with
 * switch(0); TOTAL counts steps).  Not on a serving pool's solver. */
int psi_solve_program_profiled(psi_solver* s, int reset, int max_rounds, double* times_ms,
                               int64_t* counts);
/* Its launch timeline: the phase word of each launch (low 7 bits: phase) and
 * the launch's kernel duration in microseconds (profiling; no reference
 * counterpart). */
int psi_get_program_launches(psi_solver* s, int* phase, float* us, int cap, int* count);

/* Results of one game, local vertex ids.  get_winning: gpu_parity.cu:340-354,
 * cpu_parity.cpp:288-297. */
int psi_get_winning(psi_solver* s, int game, int player, int* out, int* len);
int psi_get_strategy(psi_solver* s, int game, int* out);   /* -1 = sink */
int psi_get_labels(psi_solver* s, int game, int* out);     /* infinite[]: -1, 0, +1 */
int psi_get_solved(psi_solver* s, int game, int* out);
int psi_get_depths(psi_solver* s, int game, int* out);     /* path length to sink */
/* Every game at once, concatenated in game order (strategies in local ids,
 * labels) and stats_all[3g..3g+2] = (iter, br, capped).  Pointers may be null. */
int psi_get_results(psi_solver* s, int* strat_all, int* labels_all, int* stats_all);
/* Trace mode only.  Digest of a finite valuation: sum_p count[p]*splitmix64(p)
 * mod 2^64 (0 for infinite vertices).  Trace: per valuation round, hashes of
 * the strategy valued, solved flags, labels and digests (4 x uint64). */
int psi_get_digests(psi_solver* s, int game, uint64_t* out);
int psi_get_trace(psi_solver* s, int game, uint64_t* hashes, int cap, int* count);
/* Pointer-jumping statistics of the last valuation: out[2r] vertices that
 * reached a terminal in round r, out[2r+1] still jumping after it; *last =
 * last round that ran; *err = 1 if the rounds were not enough.  The counts are
 * those of valuations launched round by round (psi_compute_valuation,
 * psi_solve_profiled, psi_solve with PSI_SIT_PROGRAM 0); the step program
 * keeps only *last and *err. */
int psi_get_jump_stats(psi_solver* s, int* out, int rounds, int* last, int* err);
/* Loads strategy (-1 sink), solved flags and labels of solved vertices.  The
 * game's run restarts from that state: its round and iteration counts return
 * to zero and a finished game takes part in the next valuation again. */
int psi_set_state(psi_solver* s, int game, const int* strat, const int* solved,
                  const int* labels);

/* ---- serving pool -------------------------------------------------------
 * No class of the reference has this role: it serves the GPU solve
 * (solve_si, gpumain.cpp:18-53, with GPUParity, gpu_parity.cu) to a stream of
 * requests.  Continuous batching: device arrays of fixed capacity, sized from
 * an HBM budget; a request is admitted when a control slot and vertex and
 * edge ranges are free, and joins the SI steps of the games already running;
 * finished games are copied out and their ranges reused.  Each result equals
 * the game's solve alone (psi_solver_create + psi_solve).
 *
 * Capacity: vertices = budget / (psi_pool_bytes_per_vertex + 4 *
 * edges_per_vertex), edges = edges_per_vertex * vertices; a game takes n + 1
 * vertices and m edges.  max_n bounds the admissible game size. */
typedef struct psi_pool psi_pool;
int psi_pool_bytes_per_vertex(int max_n, int64_t* bytes);
int psi_pool_create(int64_t hbm_budget, double edges_per_vertex, int max_games, int max_n,
                    psi_pool** out);
int psi_pool_destroy(psi_pool* p);
/* Queues one game (host CSR arrays, as psi_game_create) and admits it at
 * once if it fits.  The arrays are read when the request is admitted: they
 * must stay valid until psi_pool_poll reports state >= 1.  PSI_ERR_INVALID
 * with the reference's message for a malformed game; PSI_ERR_BUDGET if it can
 * never fit (n > max_n or beyond the capacities).  reset/max_rounds as
 * psi_solve.  *ticket names the request. */
int psi_pool_submit(psi_pool* p, int n, int64_t m, const int* owner, const int* priority,
                    const int* offsets, const int* targets, int reset, int max_rounds,
                    int64_t* ticket);
/* Waits for the step batch in flight, launches the next `steps` SI steps of
 * every admitted unfinished game, and while they run copies out the games
 * that finished, frees their ranges and admits queued requests.  The tickets
 * finished in this call go to done_tickets[0..min(*n_done, cap)). */
int psi_pool_step(psi_pool* p, int steps, int64_t* done_tickets, int cap, int* n_done);
/* state: 0 queued, 1 running, 2 finished (result ready), -1 unknown. */
int psi_pool_poll(psi_pool* p, int64_t ticket, int* state);
/* Result of a finished request (strategy in the game's own ids, -1 sink;
 * labels -1/0/+1); frees it.  Pointers may be null. */
int psi_pool_result(psi_pool* p, int64_t ticket, int* iter, int* br, int* capped, int* strategy,
                    int* labels);
int psi_pool_info(psi_pool* p, int* running, int* queued, int64_t* vertex_cap, int64_t* edge_cap,
                  int64_t* vertices_free, int64_t* largest_free);

/* ---- launch-shape tuning (the autotuner sets these) -----------------------
 * Grid sizes (blocks of 256 threads) of the SI step kernels, which all run
 * grid-stride loops; they apply to step graphs captured afterwards.  No
 * reference counterpart (the reference hard-codes its launch shapes,
 * gpu_util.h:8, gpu_util.cpp:19-22, gpu_list_ranking.cu:19). */
#define PSI_SIT_JUMP_GRID 1  /* pointer-jumping rounds (default 2048)          */
#define PSI_SIT_STEP_GRID 2  /* val_init / label / switch / mark (16384)       */
#define PSI_SIT_POOL_GRID 3  /* the same kernels in a serving pool (4096)      */
#define PSI_SIT_PHASE_GRID 4 /* step-program launches (2048)                   */
/* PSI_SIT_PROGRAM: 1 (default) runs psi_solve's loop as a step program (each
 * launch runs the next phase of the loop: no launch is spent on a jump round
 * after convergence); 0 launches every phase of every step. */
#define PSI_SIT_PROGRAM 5
/* PSI_SIT_TWO_LEVELS: 1 (default) the step program's pointer-jumping
 * launches advance two levels each (three dependent gathers per vertex),
 * 0 one level; results are identical. */
#define PSI_SIT_TWO_LEVELS 6
/* PSI_SIT_PUSH: k > 0 (default 16): in the step program on a one-game
 * handle, after a valuation in which at most n / k vertices changed, the
 * owner-list entries of their in-neighbours are marked and the next switch
 * passes visit only marked entries (else every entry); 0 never marks.
 * Results are identical. */
#define PSI_SIT_PUSH 7
/* PSI_SIT_PUSH_MIN_N: the marks of PSI_SIT_PUSH on games of at least this
 * many vertices (default 1048576; their in-edge lists are built when the
 * handle is created, so this applies to handles created afterwards). */
#define PSI_SIT_PUSH_MIN_N 8
int psi_si_tune(int param, int value);

/* ---- seeded generators (host only, no device calls) --------------------- */
/* Random game: owner Bernoulli(1/2), priority U[0,pri_bound), out-degree
 * U[deg_lo,deg_hi] distinct uniform successors; CSR with offsets[n+1]. */
int psi_gen_random_game(uint64_t seed, int n, int pri_bound, int deg_lo, int deg_hi, int* owner,
                        int* priority, int* offsets, int* targets, int64_t targets_cap,
                        int64_t* m_out);
/* Chain game: priority i, owner i%2, edge i->i+1 (last -> 0), plus one random
 * edge with probability 1/10. */
int psi_gen_chain_game(uint64_t seed, int n, int* owner, int* priority, int* offsets,
                       int* targets, int64_t targets_cap, int64_t* m_out);
/* count sizes log-uniform in [lo, hi], rounded. */
int psi_gen_sizes(uint64_t seed, int count, double lo, double hi, int* out);
int psi_gen_permutation(uint64_t seed, int n, int* perm);
/* One list over a random permutation of 0..n-1. */
int psi_gen_list(uint64_t seed, int n, int* succ, int* head);
/* n_lists lists, n_nodes nodes, log-uniform raw lengths in [1,len_hi]. */
int psi_gen_forest(uint64_t seed, int n_nodes, int n_lists, int len_hi, int* succ, int* heads,
                   int* lengths_out);

#ifdef __cplusplus
}
#endif

#endif /* PSI_H */
