"""ctypes binding of libpsi.so (the C ABI in include/psi.h).

Host glue for tests and the benchmark: numpy arrays in, numpy arrays out.
The library must be built (``make -C parallel-si_amd``); importing this module
without it raises, so no caller can silently fall back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsi.so")

PSI_OK = 0
PSI_ERR_INVALID = -1
PSI_ERR_HIP = -2
PSI_ERR_NO_DEVICE = -3
PSI_ERR_CAP = -4
PSI_ERR_BUDGET = -5

LR_AUTO, LR_HJ, LR_WYLLIE, LR_WALK = 0, 1, 2, 3
LR_CHECKED = 0x100
LRT_LOG_D0, LRT_LOCAL_BITS, LRT_LOG_DK, LRT_HOP_CAP_K, LRT_WALK_MEAN_MAX = 1, 2, 3, 4, 5
LRT_WYLLIE_MAX = 6
LRT_WYLLIE_MAX_DEFAULT = 8192

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (run make -C parallel-si_amd)")
lib = C.CDLL(LIB_PATH)

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("psi_last_error", C.c_char_p)
_sig("psi_init", C.c_int)
_sig("psi_device_count", C.c_int, C.POINTER(C.c_int))
_sig("psi_synchronize", C.c_int)
_sig("psi_malloc", C.c_int, C.POINTER(_vp), C.c_size_t)
_sig("psi_free", C.c_int, _vp)
_sig("psi_memcpy_h2d", C.c_int, _vp, _vp, C.c_size_t)
_sig("psi_memcpy_d2h", C.c_int, _vp, _vp, C.c_size_t)
_sig("psi_memcpy_d2d", C.c_int, _vp, _vp, C.c_size_t)
_sig("psi_event_create", C.c_int, C.POINTER(_vp))
_sig("psi_event_destroy", C.c_int, _vp)
_sig("psi_event_record", C.c_int, _vp)
_sig("psi_solver_event_record", C.c_int, _vp, _vp)
_sig("psi_event_elapsed_ms", C.c_int, _vp, _vp, C.POINTER(C.c_float))
_sig("psi_list_rank", C.c_int, _vp, C.c_int, _vp, C.c_int, _vp, C.c_int)
_sig("psi_list_rank_host", C.c_int, _i32p, C.c_int, _i32p, C.c_int, _i32p, C.c_int)
_sig("psi_list_rank_status", C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("psi_list_rank_tune", C.c_int, C.c_int, C.c_int)
_sig("psi_gen_random_game", C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
     _i32p, _i32p, _i32p, _i32p, C.c_int64, C.POINTER(C.c_int64))
_sig("psi_gen_chain_game", C.c_int, C.c_uint64, C.c_int, _i32p, _i32p, _i32p, _i32p,
     C.c_int64, C.POINTER(C.c_int64))
_sig("psi_gen_sizes", C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_double, _i32p)
_sig("psi_gen_permutation", C.c_int, C.c_uint64, C.c_int, _i32p)
_sig("psi_gen_list", C.c_int, C.c_uint64, C.c_int, _i32p, C.POINTER(C.c_int))
_sig("psi_gen_forest", C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _vp)


class PsiError(RuntimeError):
    def __init__(self, rc, where):
        msg = lib.psi_last_error().decode(errors="replace")
        super().__init__(f"{where}: rc={rc}: {msg}")
        self.rc = rc


def check(rc, where):
    if rc != PSI_OK:
        raise PsiError(rc, where)
    return rc


def init():
    check(lib.psi_init(), "psi_init")


def synchronize():
    check(lib.psi_synchronize(), "psi_synchronize")


class Game:
    """A game as int32 CSR arrays (the GPUGame layout, gpu_game.h:6-21)."""

    def __init__(self, owner, priority, offsets, targets):
        self.owner = np.ascontiguousarray(owner, dtype=np.int32)
        self.priority = np.ascontiguousarray(priority, dtype=np.int32)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.targets = np.ascontiguousarray(targets, dtype=np.int32)

    @property
    def n(self):
        return int(self.owner.shape[0])

    @property
    def m(self):
        return int(self.targets.shape[0])

    @property
    def max_pri(self):
        return int(self.priority.max()) if self.n else 0

    def to_gm(self):
        """PGSolver text (the format game.cpp:51-107 parses)."""
        lines = [f"parity {self.n - 1};"]
        for v in range(self.n):
            succ = ",".join(str(int(t)) for t in self.targets[self.offsets[v]:self.offsets[v + 1]])
            lines.append(f"{v} {int(self.priority[v])} {int(self.owner[v])} {succ};")
        return "\n".join(lines) + "\n"


def gen_random_game(seed, n, pri_bound, deg_lo, deg_hi):
    owner = np.empty(n, np.int32)
    pri = np.empty(n, np.int32)
    off = np.empty(n + 1, np.int32)
    cap = n * min(deg_hi, n, 64)
    tgt = np.empty(cap, np.int32)
    m = C.c_int64()
    check(lib.psi_gen_random_game(seed, n, pri_bound, deg_lo, deg_hi, owner, pri, off, tgt, cap,
                                  C.byref(m)), "psi_gen_random_game")
    return Game(owner, pri, off, tgt[: m.value].copy())


def gen_chain_game(seed, n):
    owner = np.empty(n, np.int32)
    pri = np.empty(n, np.int32)
    off = np.empty(n + 1, np.int32)
    cap = 2 * n
    tgt = np.empty(cap, np.int32)
    m = C.c_int64()
    check(lib.psi_gen_chain_game(seed, n, owner, pri, off, tgt, cap, C.byref(m)),
          "psi_gen_chain_game")
    return Game(owner, pri, off, tgt[: m.value].copy())


def _splitmix64(x):
    """splitmix64 finaliser over a uint64 numpy array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def hash_i32(a):
    """Order-independent hash of an int32 array, the one the fixtures and the
    trace hashes use: sum over i of splitmix64(i << 32 | (uint32) a[i]),
    mod 2^64."""
    a = np.ascontiguousarray(a, np.int32)
    i = np.arange(a.shape[0], dtype=np.uint64)
    x = (i << np.uint64(32)) | a.view(np.uint32).astype(np.uint64)
    return int(_splitmix64(x).sum(dtype=np.uint64))


def gen_sizes(seed, count, lo, hi):
    out = np.empty(count, np.int32)
    check(lib.psi_gen_sizes(seed, count, float(lo), float(hi), out), "psi_gen_sizes")
    return out


def gen_list(seed, n):
    succ = np.empty(n, np.int32)
    head = C.c_int()
    check(lib.psi_gen_list(seed, n, succ, C.byref(head)), "psi_gen_list")
    return succ, np.array([head.value], np.int32)


def gen_forest(seed, n_nodes, n_lists, len_hi=4096):
    succ = np.empty(n_nodes, np.int32)
    heads = np.empty(n_lists, np.int32)
    lens = np.empty(n_lists, np.int32)
    check(lib.psi_gen_forest(seed, n_nodes, n_lists, len_hi, succ, heads,
                             lens.ctypes.data_as(_vp)), "psi_gen_forest")
    return succ, heads, lens


class DeviceArray:
    """A device allocation with numpy-typed copies in and out."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _vp()
        check(lib.psi_malloc(C.byref(self.ptr), self.nbytes), "psi_malloc")

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(max(arr.nbytes, 1))
        if arr.nbytes:
            check(lib.psi_memcpy_h2d(d.ptr, arr.ctypes.data_as(_vp), arr.nbytes), "h2d")
        return d

    def to_host(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            check(lib.psi_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib.psi_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Timer:
    """HIP events on the library stream, or on a solver handle's own stream
    (the stream its kernels run on) when one is given."""

    def __init__(self, solver=None):
        self.a, self.b = _vp(), _vp()
        self.solver = solver
        check(lib.psi_event_create(C.byref(self.a)), "event")
        check(lib.psi_event_create(C.byref(self.b)), "event")

    def _record(self, ev):
        if self.solver is not None:
            check(lib.psi_solver_event_record(self.solver.h, ev), "record")
        else:
            check(lib.psi_event_record(ev), "record")

    def start(self):
        self._record(self.a)

    def stop(self):
        self._record(self.b)

    def ms(self):
        f = C.c_float()
        check(lib.psi_event_elapsed_ms(self.a, self.b, C.byref(f)), "elapsed")
        return f.value


def list_rank_host(succ, heads, method=LR_AUTO):
    succ = np.ascontiguousarray(succ, np.int32)
    heads = np.ascontiguousarray(heads, np.int32)
    rank = np.empty_like(succ)
    check(lib.psi_list_rank_host(succ, succ.shape[0], heads, heads.shape[0], rank, method),
          "psi_list_rank_host")
    return rank


def list_rank_device(d_succ, n, d_heads, n_heads, d_rank, method=LR_AUTO):
    check(lib.psi_list_rank(d_succ.ptr, n, d_heads.ptr, n_heads, d_rank.ptr, method),
          "psi_list_rank")


_sig("psi_tour_path_sums", C.c_int, _vp, C.c_int, _vp, _vp, _vp, C.c_int)


def tour_path_sums(link, w, method=LR_AUTO):
    """Host wrapper of psi_tour_path_sums: (sum uint64[n], reach int32[n])."""
    link = np.ascontiguousarray(link, np.int32)
    w = np.ascontiguousarray(w, np.uint64)
    n = link.shape[0]
    d_link, d_w = DeviceArray.from_host(link), DeviceArray.from_host(w)
    d_sum, d_reach = DeviceArray(8 * max(n, 1)), DeviceArray(4 * max(n, 1))
    check(lib.psi_tour_path_sums(d_link.ptr, n, d_w.ptr, d_sum.ptr, d_reach.ptr, method),
          "psi_tour_path_sums")
    return d_sum.to_host(np.uint64, n), d_reach.to_host(np.int32, n)


def list_rank_status():
    bits, over = C.c_int(), C.c_int()
    check(lib.psi_list_rank_status(C.byref(bits), C.byref(over)), "status")
    return bits.value, over.value


SIT_JUMP_GRID, SIT_STEP_GRID, SIT_POOL_GRID, SIT_PHASE_GRID, SIT_PROGRAM = 1, 2, 3, 4, 5
SIT_TWO_LEVELS = 6
SIT_PUSH, SIT_PUSH_DEFAULT = 7, 16
SIT_PUSH_MIN_N, SIT_PUSH_MIN_N_DEFAULT = 8, 1 << 20
_sig("psi_si_tune", C.c_int, C.c_int, C.c_int)


def si_tune(param, value):
    check(lib.psi_si_tune(param, value), "psi_si_tune")


def list_rank_tune(param, value):
    check(lib.psi_list_rank_tune(param, value), "psi_list_rank_tune")


# ---- solvers ---------------------------------------------------------------
SOLVER_TRACE = 1
_pp = C.POINTER(_vp)
_sig("psi_game_create", C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, _pp)
_sig("psi_game_destroy", C.c_int, _vp)
_sig("psi_game_info", C.c_int, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("psi_solver_create", C.c_int, _vp, C.c_int, C.c_int, _pp)
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_sig("psi_solver_create_batch", C.c_int, C.c_int, _i32p, _i64p, _i32p, _i32p, _i32p, _i32p,
     C.c_int, C.c_int, _pp)
_sig("psi_solver_create_games", C.c_int, C.c_int, _i32p, _i64p, C.POINTER(_vp), C.POINTER(_vp),
     C.POINTER(_vp), C.POINTER(_vp), C.c_int, C.c_int, _pp)
_sig("psi_solver_destroy", C.c_int, _vp)
_sig("psi_solver_info", C.c_int, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int64),
     C.POINTER(C.c_int64))
_sig("psi_solver_synchronize", C.c_int, _vp)
_sig("psi_init_strat", C.c_int, _vp)
_sig("psi_compute_valuation", C.c_int, _vp)
_sig("psi_mark_solved", C.c_int, _vp, C.c_int)
_sig("psi_switch_strategy", C.c_int, _vp, C.c_int, C.POINTER(C.c_int))
_sig("psi_reset_opponent_strategy", C.c_int, _vp)
_sig("psi_solve", C.c_int, _vp, C.c_int, C.c_int, C.c_int)
_sig("psi_solve_bf", C.c_int, _vp, C.c_int, C.c_int)
_sig("psi_solve_bv", C.c_int, _vp, C.c_int, C.c_int)
_sig("psi_solve_list", C.c_int, _vp, C.c_int, C.c_int)
_sig("psi_get_bf_status", C.c_int, _vp, C.c_int, C.POINTER(C.c_int))
_sig("psi_get_stats", C.c_int, _vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_int))
_sig("psi_get_winning", C.c_int, _vp, C.c_int, C.c_int, _i32p, C.POINTER(C.c_int))
for _f in ("psi_get_strategy", "psi_get_labels", "psi_get_solved", "psi_get_depths"):
    _sig(_f, C.c_int, _vp, C.c_int, _i32p)
_sig("psi_get_digests", C.c_int, _vp, C.c_int, _u64p)
_sig("psi_get_trace", C.c_int, _vp, C.c_int, _u64p, C.c_int, C.POINTER(C.c_int))
_sig("psi_set_state", C.c_int, _vp, C.c_int, _i32p, _i32p, _i32p)
_sig("psi_get_results", C.c_int, _vp, _i32p, _i32p, _i32p)


class Solver:
    """One psi_solver handle: a single game, or a batch of games solved each
    as if alone."""

    def __init__(self, games, trace=False, trace_cap=4096):
        if isinstance(games, Game):
            games = [games]
        self.games = list(games)
        n_per = np.array([g.n for g in self.games], np.int32)
        m_per = np.array([g.m for g in self.games], np.int64)
        G = len(self.games)
        ptrs = [(_vp * G)(*[getattr(g, a).ctypes.data for g in self.games])
                for a in ("owner", "priority", "offsets", "targets")]
        self.h = _vp()
        self.trace = trace
        self.trace_cap = trace_cap
        check(lib.psi_solver_create_games(G, n_per, m_per, *ptrs, SOLVER_TRACE if trace else 0,
                                          trace_cap, C.byref(self.h)), "psi_solver_create_games")

    def close(self):
        if self.h:
            lib.psi_solver_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, reset=False, max_rounds=0, steps_per_sync=4):
        check(lib.psi_solve(self.h, int(reset), max_rounds, steps_per_sync), "psi_solve")

    def solve_bf(self, max_rounds=0, steps_per_sync=4):
        """solve_bf with the Bellman-Ford best response (cpu_bf role)."""
        check(lib.psi_solve_bf(self.h, max_rounds, steps_per_sync), "psi_solve_bf")

    def solve_bv(self, max_rounds=0, steps_per_sync=4):
        """solve_si with CPUBV semantics (the cpubv role): returns False if a
        game stopped on an odd strategy cycle (bf_status(game) == 1)."""
        rc = lib.psi_solve_bv(self.h, max_rounds, steps_per_sync)
        if rc == PSI_ERR_INVALID and b"odd cycle" in lib.psi_last_error():
            return False
        check(rc, "psi_solve_bv")
        return True

    def solve_list(self, max_rounds=0, steps_per_sync=4):
        """solve_si with CPUList semantics (the cpulist role): returns False
        if a game reached max_rounds (such runs need not end)."""
        rc = lib.psi_solve_list(self.h, max_rounds, steps_per_sync)
        if rc == PSI_ERR_CAP:
            return False
        check(rc, "psi_solve_list")
        return True

    def bf_status(self, game=0):
        s = C.c_int()
        check(lib.psi_get_bf_status(self.h, game, C.byref(s)), "psi_get_bf_status")
        return s.value

    def sync(self):
        check(lib.psi_solver_synchronize(self.h), "sync")

    # step-wise si.h operations
    def init_strat(self):
        check(lib.psi_init_strat(self.h), "init_strat")

    def compute_valuation(self):
        check(lib.psi_compute_valuation(self.h), "compute_valuation")

    def mark_solved(self, player):
        check(lib.psi_mark_solved(self.h, player), "mark_solved")

    def switch_strategy(self, player):
        s = C.c_int()
        check(lib.psi_switch_strategy(self.h, player, C.byref(s)), "switch_strategy")
        return s.value

    def reset_opponent_strategy(self):
        check(lib.psi_reset_opponent_strategy(self.h), "reset_opponent_strategy")

    def set_state(self, strat, solved=None, labels=None, game=0):
        n = self.games[game].n
        strat = np.ascontiguousarray(strat, np.int32)
        solved = np.ascontiguousarray(solved if solved is not None else np.zeros(n), np.int32)
        labels = np.ascontiguousarray(labels if labels is not None else np.zeros(n), np.int32)
        check(lib.psi_set_state(self.h, game, strat, solved, labels), "set_state")

    def stats(self, game=0):
        it, br, cap = C.c_int(), C.c_int(), C.c_int()
        check(lib.psi_get_stats(self.h, game, C.byref(it), C.byref(br), C.byref(cap)), "stats")
        return it.value, br.value, bool(cap.value)

    def _get(self, fn, game, dtype=np.int32):
        out = np.empty(self.games[game].n, dtype)
        check(getattr(lib, fn)(self.h, game, out), fn)
        return out

    def strategy(self, game=0):
        return self._get("psi_get_strategy", game)

    def labels(self, game=0):
        return self._get("psi_get_labels", game)

    def solved(self, game=0):
        return self._get("psi_get_solved", game)

    def depths(self, game=0):
        return self._get("psi_get_depths", game)

    def digests(self, game=0):
        return self._get("psi_get_digests", game, np.uint64)

    def winning(self, player, game=0):
        out = np.empty(self.games[game].n, np.int32)
        k = C.c_int()
        check(lib.psi_get_winning(self.h, game, player, out, C.byref(k)), "get_winning")
        return out[: k.value]

    def results(self):
        """Per game: ((iter, br, capped), strategy, labels), one copy each."""
        N = sum(g.n for g in self.games)
        st = np.empty(N, np.int32)
        lb = np.empty(N, np.int32)
        stats = np.empty(3 * len(self.games), np.int32)
        check(lib.psi_get_results(self.h, st, lb, stats), "psi_get_results")
        out, o = [], 0
        for k, g in enumerate(self.games):
            out.append(((int(stats[3 * k]), int(stats[3 * k + 1]), bool(stats[3 * k + 2])),
                        st[o:o + g.n], lb[o:o + g.n]))
            o += g.n
        return out

    def trace_hashes(self, game=0):
        buf = np.zeros(4 * self.trace_cap, np.uint64)
        k = C.c_int()
        check(lib.psi_get_trace(self.h, game, buf, self.trace_cap, C.byref(k)), "get_trace")
        return buf[: 4 * k.value].reshape(-1, 4)

K_NAMES = ["val_init", "jump", "jump_idle", "label", "switch1", "switch0", "ctl", "total",
           "gathers", "ties", "scans", "scan_steps", "push"]
_sig("psi_solve_profiled", C.c_int, _vp, C.c_int, C.c_int,
     np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"), _i64p)


def solve_profiled(solver, reset=False, max_rounds=0):
    t = np.zeros(len(K_NAMES), np.float64)
    c = np.zeros(len(K_NAMES), np.int64)
    check(lib.psi_solve_profiled(solver.h, int(reset), max_rounds, t, c), "psi_solve_profiled")
    return {k: (float(t[i]), int(c[i])) for i, k in enumerate(K_NAMES)}

_sig("psi_get_profile_rounds", C.c_int, _vp, _i32p, C.c_int, C.POINTER(C.c_int))


def profile_rounds(solver, cap=1 << 16):
    """Jump rounds run per valuation in the last solve_profiled."""
    out = np.zeros(cap, np.int32)
    k = C.c_int()
    check(lib.psi_get_profile_rounds(solver.h, out, cap, C.byref(k)), "psi_get_profile_rounds")
    return out[: min(k.value, cap)].tolist()

_sig("psi_solve_program_profiled", C.c_int, _vp, C.c_int, C.c_int,
     np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"), _i64p)


def solve_program_profiled(solver, reset=False, max_rounds=0):
    """The step program launch by launch: {class: (ms, launches)} as
    solve_profiled (ties/scans not counted)."""
    t = np.zeros(len(K_NAMES), np.float64)
    c = np.zeros(len(K_NAMES), np.int64)
    check(lib.psi_solve_program_profiled(solver.h, int(reset), max_rounds, t, c),
          "psi_solve_program_profiled")
    return {k: (float(t[i]), int(c[i])) for i, k in enumerate(K_NAMES)}

_sig("psi_get_program_launches", C.c_int, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_int))


def program_launches(solver, cap=1 << 16):
    """Launch timeline of the last solve_program_profiled: list of
    (phase, round, us); phases as PH_NAMES."""
    ph = np.zeros(cap, np.int32)
    us = np.zeros(cap, np.float32)
    k = C.c_int()
    check(lib.psi_get_program_launches(solver.h, ph.ctypes.data, us.ctypes.data, cap, C.byref(k)),
          "psi_get_program_launches")
    k = min(k.value, cap)
    return [(PH_NAMES[int(p) & 0x7f] if (int(p) & 0x7f) < len(PH_NAMES) else str(int(p) & 0x7f),
             int(p) >> 8, float(u)) for p, u in zip(ph[:k], us[:k])]


PH_NAMES = ["start", "init", "jump", "label", "switch1", "ctl1", "mark0", "switch0", "ctl0", "done",
            "act", "push"]

_sig("psi_get_jump_stats", C.c_int, _vp, _i32p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int))


def jump_stats(solver, rounds=40):
    out = np.zeros(2 * rounds, np.int32)
    last, err = C.c_int(), C.c_int()
    check(lib.psi_get_jump_stats(solver.h, out, rounds, C.byref(last), C.byref(err)), "jump_stats")
    return out.reshape(-1, 2), last.value, err.value

_sig("psi_game_from_gm", C.c_int, C.c_char_p, C.c_size_t, _pp)
_sig("psi_game_csr", C.c_int, _vp, _i32p, _i32p, _i32p, _i32p)


def game_from_gm(text):
    """Parses PGSolver text with the library's loader (the reference grammar).
    Raises PsiError with the reference's message on malformed input."""
    b = text.encode() if isinstance(text, str) else text
    h = _vp()
    check(lib.psi_game_from_gm(b, len(b), C.byref(h)), "psi_game_from_gm")
    n, m, mp = C.c_int(), C.c_int(), C.c_int()
    lib.psi_game_info(h, C.byref(n), C.byref(m), C.byref(mp))
    owner = np.empty(n.value, np.int32)
    pri = np.empty(n.value, np.int32)
    off = np.empty(n.value + 1, np.int32)
    tgt = np.empty(m.value, np.int32)
    lib.psi_game_csr(h, owner, pri, off, tgt)
    lib.psi_game_destroy(h)
    return Game(owner, pri, off, tgt)


# ---- serving pool ----------------------------------------------------------
_sig("psi_pool_bytes_per_vertex", C.c_int, C.c_int, C.POINTER(C.c_int64))
_sig("psi_pool_create", C.c_int, C.c_int64, C.c_double, C.c_int, C.c_int, _pp)
_sig("psi_pool_destroy", C.c_int, _vp)
_sig("psi_pool_submit", C.c_int, _vp, C.c_int, C.c_int64, _i32p, _i32p, _i32p, _i32p, C.c_int,
     C.c_int, C.POINTER(C.c_int64))
_sig("psi_pool_step", C.c_int, _vp, C.c_int, _i64p, C.c_int, C.POINTER(C.c_int))
_sig("psi_pool_poll", C.c_int, _vp, C.c_int64, C.POINTER(C.c_int))
_sig("psi_pool_result", C.c_int, _vp, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_int), _vp, _vp)
_sig("psi_pool_info", C.c_int, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
     C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64))

QUEUED, RUNNING, FINISHED = 0, 1, 2


def pool_bytes_per_vertex(max_n):
    b = C.c_int64()
    check(lib.psi_pool_bytes_per_vertex(max_n, C.byref(b)), "psi_pool_bytes_per_vertex")
    return b.value


class Pool:
    """Serving pool (psi_pool_*): submit games, step, collect results.  The
    pool keeps each submitted Game referenced until it is admitted (the C
    side reads its arrays then)."""

    def __init__(self, hbm_budget, edges_per_vertex=5.0, max_games=256, max_n=1 << 22):
        self.h = _vp()
        self.max_games = max_games
        check(lib.psi_pool_create(int(hbm_budget), float(edges_per_vertex), max_games, max_n,
                                  C.byref(self.h)), "psi_pool_create")
        self._held = {}
        self._n = {}
        self._done = np.zeros(max_games, np.int64)

    def close(self):
        if self.h:
            lib.psi_pool_destroy(self.h)
            self.h = _vp()
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, g, reset=False, max_rounds=0):
        t = C.c_int64()
        check(lib.psi_pool_submit(self.h, g.n, g.m, g.owner, g.priority, g.offsets, g.targets,
                                  int(reset), max_rounds, C.byref(t)), "psi_pool_submit")
        self._held[t.value] = g
        self._n[t.value] = g.n
        return t.value

    def step(self, steps=4):
        """Returns the tickets that finished in this call."""
        k = C.c_int()
        try:
            check(lib.psi_pool_step(self.h, steps, self._done, self.max_games, C.byref(k)),
                  "psi_pool_step")
        finally:
            # a request that was admitted (state >= RUNNING) or whose admission
            # failed (its ticket is gone, state -1) no longer needs its arrays
            for t in list(self._held):
                st = self.poll(t)
                if st >= RUNNING or st < 0:
                    del self._held[t]
        return [int(x) for x in self._done[: k.value]]

    def poll(self, ticket):
        s = C.c_int()
        check(lib.psi_pool_poll(self.h, ticket, C.byref(s)), "psi_pool_poll")
        return s.value

    def result(self, ticket):
        """((iter, br, capped), strategy, labels); frees the stored result."""
        n = self._n.pop(ticket)
        st = np.empty(n, np.int32)
        lb = np.empty(n, np.int32)
        it, br, cap = C.c_int(), C.c_int(), C.c_int()
        check(lib.psi_pool_result(self.h, ticket, C.byref(it), C.byref(br), C.byref(cap),
                                  st.ctypes.data, lb.ctypes.data), "psi_pool_result")
        return (it.value, br.value, bool(cap.value)), st, lb

    def info(self):
        r, q = C.c_int(), C.c_int()
        vc, ec, vf, lf = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(lib.psi_pool_info(self.h, C.byref(r), C.byref(q), C.byref(vc), C.byref(ec),
                                C.byref(vf), C.byref(lf)), "psi_pool_info")
        return {"running": r.value, "queued": q.value, "vertex_cap": vc.value,
                "edge_cap": ec.value, "vertices_free": vf.value, "largest_free": lf.value}

    def drain(self, steps=4):
        """Steps until nothing is queued or running; returns finished tickets."""
        out = []
        while True:
            out += self.step(steps)
            i = self.info()
            if i["running"] == 0 and i["queued"] == 0:
                return out
