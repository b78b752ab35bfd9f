"""ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference path that the
tests, the smoke check and the benchmark's CPU-baseline leg use as the checker.
The product path never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (run make -C oracle)")
lib = C.CDLL(LIB_PATH)

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

lib.orc_last_error.restype = C.c_char_p
lib.orc_splitmix64.restype = C.c_uint64
lib.orc_splitmix64.argtypes = [C.c_uint64]
lib.orc_hash_i32.restype = C.c_uint64
lib.orc_hash_i32.argtypes = [_i32p, C.c_int]
lib.orc_hash_u64.restype = C.c_uint64
lib.orc_hash_u64.argtypes = [_u64p, C.c_int]
lib.orc_parse.restype = C.c_int
lib.orc_parse.argtypes = [C.c_char_p, C.c_long, C.POINTER(C.c_int), C.POINTER(C.c_int),
                          C.POINTER(C.c_int)]
lib.orc_parsed_arrays.restype = None
lib.orc_parsed_arrays.argtypes = [_i32p, _i32p, _i32p, _i32p]
lib.orc_solve.restype = C.c_int
lib.orc_solve.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int,
                          _vp, _vp, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                          _vp, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_int),
                          C.POINTER(C.c_double)]
lib.orc_valuate.restype = C.c_int
lib.orc_valuate.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p,
                            _i32p, _u64p, _vp]
lib.orc_switch.restype = C.c_int
lib.orc_switch.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, _i32p, _i32p, _i32p,
                           C.c_int, _i32p]
lib.orc_compare_pairs.restype = C.c_int
lib.orc_compare_pairs.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, _i32p, _i32p,
                                  _i32p, C.c_int, _i32p]
lib.orc_list_rank.restype = C.c_int
lib.orc_list_rank.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i32p]
lib.orc_list_scan.restype = C.c_int
lib.orc_list_scan.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i64p, _i64p]
lib.orc_zielonka.restype = C.c_int
lib.orc_zielonka.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]

DENSE, SPARSE = 0, 1


class OracleError(RuntimeError):
    pass


def _err(where):
    return OracleError(f"{where}: {lib.orc_last_error().decode()}")


def splitmix64(x):
    return int(lib.orc_splitmix64(int(x) & 0xFFFFFFFFFFFFFFFF))


def hash_i32(a):
    a = np.ascontiguousarray(a, np.int32)
    return int(lib.orc_hash_i32(a, a.shape[0]))


def hash_u64(a):
    a = np.ascontiguousarray(a, np.uint64)
    return int(lib.orc_hash_u64(a, a.shape[0]))


def parse_gm(text):
    """Returns (owner, priority, offsets, targets, max_pri) or raises OracleError
    with the reference's message (game.cpp:51-107)."""
    b = text.encode() if isinstance(text, str) else text
    n, m, mp = C.c_int(), C.c_int(), C.c_int()
    if lib.orc_parse(b, len(b), C.byref(n), C.byref(m), C.byref(mp)) != 0:
        raise _err("parse")
    owner = np.empty(n.value, np.int32)
    pri = np.empty(n.value, np.int32)
    off = np.empty(n.value + 1, np.int32)
    tgt = np.empty(m.value, np.int32)
    lib.orc_parsed_arrays(owner, pri, off, tgt)
    return owner, pri, off, tgt, mp.value


class SolveResult:
    pass


def solve(game, rep=SPARSE, reset=False, max_rounds=0, trace=False, full_trace_rounds=0):
    """CPUParity solve_si.  Returns SolveResult with strat, inf, solved, iter, br,
    seconds, capped, and (if trace) hashes[br,4] = (strategy valued, solved,
    labels, valuation digests) per round; with full_trace_rounds, per-round
    arrays strat_rounds/inf_rounds/dig_rounds."""
    g = game
    n = g.n
    strat = np.empty(n, np.int32)
    inf = np.empty(n, np.int32)
    solved = np.empty(n, np.int32)
    it, br, nr = C.c_int(), C.c_int(), C.c_int()
    secs = C.c_double()
    cap = full_trace_rounds if full_trace_rounds else (max_rounds if max_rounds else 4096)
    hashes = np.zeros(4 * cap, np.uint64) if (trace or full_trace_rounds) else None
    sr = np.zeros(cap * n, np.int32) if full_trace_rounds else None
    ir = np.zeros(cap * n, np.int32) if full_trace_rounds else None
    dr = np.zeros(cap * n, np.uint64) if full_trace_rounds else None

    def p(a):
        return a.ctypes.data_as(_vp) if a is not None else None

    rc = lib.orc_solve(n, g.offsets, g.targets, g.owner, g.priority, rep, int(reset), max_rounds,
                       p(strat), p(inf), p(solved), C.byref(it), C.byref(br), p(hashes), p(sr),
                       p(ir), p(dr), cap, C.byref(nr), C.byref(secs))
    if rc < 0:
        raise _err("solve")
    r = SolveResult()
    r.strat, r.inf, r.solved = strat, inf, solved
    r.iter, r.br, r.seconds, r.capped = it.value, br.value, secs.value, rc == 1
    k = min(nr.value, cap)
    r.hashes = hashes.reshape(-1, 4)[:k] if hashes is not None else None
    if full_trace_rounds:
        r.strat_rounds = sr.reshape(cap, n)[:k]
        r.inf_rounds = ir.reshape(cap, n)[:k]
        r.dig_rounds = dr.reshape(cap, n)[:k]
    return r


def valuate(game, strat, solved, inf_in):
    g = game
    n = g.n
    inf = np.empty(n, np.int32)
    dig = np.empty(n, np.uint64)
    depth = np.empty(n, np.int32)
    if lib.orc_valuate(n, g.offsets, g.targets, g.owner, g.priority,
                       np.ascontiguousarray(strat, np.int32), np.ascontiguousarray(solved, np.int32),
                       np.ascontiguousarray(inf_in, np.int32), inf, dig,
                       depth.ctypes.data_as(_vp)) != 0:
        raise _err("valuate")
    return inf, dig, depth


def switch(game, strat, solved, inf_in, player, rep=SPARSE):
    g = game
    out = np.empty(g.n, np.int32)
    c = lib.orc_switch(g.n, g.offsets, g.targets, g.owner, g.priority, rep,
                       np.ascontiguousarray(strat, np.int32), np.ascontiguousarray(solved, np.int32),
                       np.ascontiguousarray(inf_in, np.int32), player, out)
    if c < 0:
        raise _err("switch")
    return c, out


def compare_pairs(game, strat, us, ws, rep=SPARSE):
    g = game
    us = np.ascontiguousarray(us, np.int32)
    ws = np.ascontiguousarray(ws, np.int32)
    out = np.empty(us.shape[0], np.int32)
    if lib.orc_compare_pairs(g.n, g.offsets, g.targets, g.owner, g.priority, rep,
                             np.ascontiguousarray(strat, np.int32), us, ws, us.shape[0], out) != 0:
        raise _err("compare")
    return out


def list_rank(succ, heads):
    succ = np.ascontiguousarray(succ, np.int32)
    heads = np.ascontiguousarray(heads, np.int32)
    rank = np.empty_like(succ)
    if lib.orc_list_rank(succ, succ.shape[0], heads, heads.shape[0], rank) != 0:
        raise _err("list_rank")
    return rank


def list_scan(succ, heads, w):
    succ = np.ascontiguousarray(succ, np.int32)
    heads = np.ascontiguousarray(heads, np.int32)
    w = np.ascontiguousarray(w, np.int64)
    out = np.empty(succ.shape[0], np.int64)
    if lib.orc_list_scan(succ, succ.shape[0], heads, heads.shape[0], w, out) != 0:
        raise _err("list_scan")
    return out


def zielonka(game):
    g = game
    win = np.empty(g.n, np.int32)
    if lib.orc_zielonka(g.n, g.offsets, g.targets, g.owner, g.priority, win) != 0:
        raise _err("zielonka")
    return win


lib.orc_solve_list.restype = C.c_int
lib.orc_solve_list.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, _i32p,
                               _i32p, _i32p, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_double)]


def solve_list(game, threads=8, max_rounds=0, trace=False):
    """solve_si with the CPUList restatement (the cpulist mode: Euler tour +
    Helman-JaJa valuation on `threads` threads; vertices not reaching the
    sink are +inf).  Returns SolveResult with strat, inf, solved, iter, br,
    seconds, capped and (if trace) hashes[br, 4]."""
    g = game
    n = g.n
    strat = np.empty(n, np.int32)
    inf = np.empty(n, np.int32)
    solved = np.empty(n, np.int32)
    it, br, nr = C.c_int(), C.c_int(), C.c_int()
    secs = C.c_double()
    cap = max_rounds if max_rounds else 4096
    hashes = np.zeros(4 * cap, np.uint64) if trace else None
    rc = lib.orc_solve_list(n, g.offsets, g.targets, g.owner, g.priority, threads, max_rounds, strat,
                            inf, solved, C.byref(it), C.byref(br),
                            hashes.ctypes.data_as(_vp) if trace else None, cap, C.byref(nr),
                            C.byref(secs))
    if rc < 0:
        raise _err("solve_list")
    r = SolveResult()
    r.strat, r.inf, r.solved = strat, inf, solved
    r.iter, r.br, r.seconds, r.capped = it.value, br.value, secs.value, rc == 1
    r.hashes = hashes.reshape(-1, 4)[:min(nr.value, cap)] if trace else None
    return r


lib.orc_solve_bv.restype = C.c_int
lib.orc_solve_bv.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, _i32p, _i32p,
                             _i32p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
BV_OK, BV_ODD_CYCLE = 0, 1


def solve_bv(game, rep=SPARSE, max_rounds=0):
    """solve_si with CPUBV<ArrayVal> (the cpubv mode): CPUParity's solve that
    stops at the first valuation with an odd strategy cycle.  Returns an
    object with strat, inf, solved, iter, br, status (BV_OK / BV_ODD_CYCLE),
    capped."""
    g = game
    n = g.n
    strat = np.empty(n, np.int32)
    inf = np.empty(n, np.int32)
    solved = np.empty(n, np.int32)
    it, br, st = C.c_int(), C.c_int(), C.c_int()
    rc = lib.orc_solve_bv(n, g.offsets, g.targets, g.owner, g.priority, rep, max_rounds, strat, inf,
                          solved, C.byref(it), C.byref(br), C.byref(st))
    if rc < 0:
        raise _err("solve_bv")

    class R:
        pass
    r = R()
    r.strat, r.inf, r.solved = strat, inf, solved
    r.iter, r.br, r.status, r.capped = it.value, br.value, st.value, rc == 1
    return r


lib.orc_solve_bf.restype = C.c_int
lib.orc_solve_bf.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, _i32p, _i32p, _i32p,
                             _u64p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
BF_OK, BF_ODD_CYCLE, BF_CAPPED, BF_NO_CONVERGENCE = 0, 1, 2, 3


def solve_bf(game, max_iters=0):
    """solve_bf with CPU_BellmanFord<ArrayVal> (dense rows).  Returns an object
    with strat, inf, solved, dig (final valuation digests), iter, br, status."""
    g = game
    n = g.n
    strat = np.empty(n, np.int32)
    inf = np.empty(n, np.int32)
    solved = np.empty(n, np.int32)
    dig = np.empty(n, np.uint64)
    it, br, st = C.c_int(), C.c_int(), C.c_int()
    rc = lib.orc_solve_bf(n, g.offsets, g.targets, g.owner, g.priority, max_iters, strat, inf, solved,
                          dig, C.byref(it), C.byref(br), C.byref(st))
    if rc < 0:
        raise RuntimeError(lib.orc_last_error().decode())

    class R:
        pass
    r = R()
    r.strat, r.inf, r.solved, r.dig = strat, inf, solved, dig
    r.iter, r.br, r.status = it.value, br.value, st.value
    return r
