"""List ranking on the GPU (Helman-JaJa, head walks, Wyllie) against the
oracle's brute-force walk: bit-exact ranks on every shape, including the
16,777,216-node list and the 1,048,576-list forest of the benchmark."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

METHODS = ["AUTO", "HJ", "WALK", "WYLLIE"]


def _m(psi, name):
    return getattr(psi, "LR_" + name)


def _check(psi, succ, heads, method, expect=None):
    succ = np.ascontiguousarray(succ, np.int32)
    heads = np.ascontiguousarray(heads, np.int32)
    got = psi.list_rank_host(succ, heads, _m(psi, method))
    if expect is None:
        expect = oracle.list_rank(succ, heads)
    bad = np.nonzero(got != expect)[0]
    assert bad.size == 0, (method, succ.size, heads.size, bad[:10], got[bad[:10]], expect[bad[:10]])


def perm_list(rng, n):
    perm = rng.permutation(n).astype(np.int32)
    succ = np.full(n, -1, np.int32)
    succ[perm[:-1]] = perm[1:]
    return succ, perm[:1].copy()


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 1000, 4097, 100000, 300001])
def test_random_list(gpu, method, n):
    print(f"list n={n} method={method}", flush=True)
    succ, heads = perm_list(np.random.default_rng(n), n)
    _check(gpu, succ, heads, method)


@pytest.mark.parametrize("method", METHODS)
def test_structured_lists(gpu, method):
    n = 200000
    ident = np.arange(1, n + 1, dtype=np.int32)
    ident[-1] = -1
    print(f"structured lists method={method}", flush=True)
    _check(gpu, ident, np.array([0], np.int32), method)
    rev = np.arange(-1, n - 1, dtype=np.int32)
    _check(gpu, rev, np.array([n - 1], np.int32), method)
    # Euler-tour-like order: all odd ids ascending, then all even ids descending
    order = np.concatenate([np.arange(1, n, 2), np.arange(n - 2, -1, -2)]).astype(np.int32)
    succ = np.full(n, -1, np.int32)
    succ[order[:-1]] = order[1:]
    _check(gpu, succ, order[:1], method)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n_lists,n_nodes", [(1, 10), (10, 10), (1000, 50000), (40000, 60000),
                                             (7, 200000), (5000, 2000000)])
def test_forests(gpu, method, n_lists, n_nodes):
    print(f"forest lists={n_lists} nodes={n_nodes} method={method}", flush=True)
    import psi
    succ, heads, _ = psi.gen_forest(n_lists * 31 + n_nodes, n_nodes, n_lists, 4096)
    _check(gpu, succ, heads, method)


@pytest.mark.parametrize("method", ["AUTO", "HJ", "WALK", "WYLLIE"])
def test_unreachable_nodes_read_minus_one(gpu, method):
    # list 0->1->2, a 3-cycle 3->4->5->3 and a list 6->7 whose head is not passed
    succ = np.array([1, 2, -1, 4, 5, 3, 7, -1], np.int32)
    heads = np.array([0], np.int32)
    got = gpu.list_rank_host(succ, heads, _m(gpu, method))
    exp = [0, 1, 2, -1, -1, -1, -1, -1]
    if method == "WYLLIE":       # Wyllie treats every node without predecessor as a head
        exp = [0, 1, 2, -1, -1, -1, 0, 1]
    assert got.tolist() == exp


@pytest.mark.parametrize("wy_max", [0, 16384, 131072, 1 << 22])
@pytest.mark.parametrize("n_lists,n_nodes", [(1, 16384), (1, 16385), (300, 16000), (3000, 200000),
                                             (1, 1 << 20)])
def test_auto_wyllie_with_given_heads(gpu, wy_max, n_lists, n_nodes):
    """AUTO ranks lists of at most PSI_LRT_WYLLIE_MAX nodes by Wyllie with the
    heads as given (in one workgroup's LDS up to 16,384 nodes, else by global
    rounds): the same ranks as the walks, also where some lists' heads are not
    passed (their nodes read -1) and with cycles no head reaches."""
    import psi
    succ, heads, _ = psi.gen_forest(n_lists * 7 + n_nodes, n_nodes, n_lists, 4096)
    succ = succ.copy()
    if n_lists > 1:
        heads = heads[::3].copy()                  # two lists in three lose their head
        dropped = np.setdiff1d(np.arange(n_lists), np.arange(0, n_lists, 3))
        h = int(psi.gen_forest(n_lists * 7 + n_nodes, n_nodes, n_lists, 4096)[1][dropped[0]])
        t = h                                      # close that list into a cycle
        while succ[t] != -1:
            t = succ[t]
        succ[t] = h
    print(f"AUTO/Wyllie lists={n_lists} nodes={n_nodes} wy_max={wy_max}", flush=True)
    gpu.list_rank_tune(gpu.LRT_WYLLIE_MAX, wy_max)
    try:
        _check(gpu, succ, heads, "AUTO")
        _check(gpu, succ, heads, "HJ")
    finally:
        gpu.list_rank_tune(gpu.LRT_WYLLIE_MAX, gpu.LRT_WYLLIE_MAX_DEFAULT)


@pytest.mark.parametrize("n_lists", [1024, 1025])
def test_top_level_by_sublist_walks_and_by_jumping(gpu, n_lists):
    """HJ on 60,000 nodes: the reduced list (about 8,600 slots) is ranked at
    once by the top level, by LDS sublist walks with up to 1,024 heads and by
    LDS pointer jumping beyond; about one head slot in 16 is itself a top-level
    splitter. A third of the lists lose their head and one of those is closed
    into a cycle, so unreached slots are part of the top level too."""
    import psi
    n = 60000
    succ0, heads, _ = psi.gen_forest(n_lists * 13 + 5, n, n_lists, 4096)
    print(f"top level lists={n_lists}", flush=True)
    _check(gpu, succ0, heads, "HJ")
    succ = succ0.copy()
    keep = np.ones(n_lists, bool)
    keep[1::3] = False
    t = h = int(heads[1])                          # a dropped head's list becomes a cycle
    while succ[t] != -1:
        t = succ[t]
    succ[t] = h
    _check(gpu, succ, heads[keep], "HJ")


@pytest.mark.parametrize("n", [10, 20000])
def test_auto_wyllie_reports_two_predecessors(gpu, n):
    """A node with two predecessors (not a forest of lists) is reported by
    the Wyllie paths as an error, not ranked silently."""
    gpu.list_rank_tune(gpu.LRT_WYLLIE_MAX, 1 << 20)
    try:
        succ = np.arange(1, n + 1, dtype=np.int32)
        succ[-1] = -1
        succ[0] = 5                                # nodes 0 and 4 both lead to 5
        with pytest.raises(gpu.PsiError):
            gpu.list_rank_host(succ, np.array([0], np.int32), gpu.LR_AUTO)
    finally:
        gpu.list_rank_tune(gpu.LRT_WYLLIE_MAX, gpu.LRT_WYLLIE_MAX_DEFAULT)


def test_long_list_head_walk_ranks_directly(gpu):
    # one list of 50,000 nodes ranked by head walks only: each node's rank is
    # its hop count, written during the walk (no overflow sublists, no reduced
    # level)
    succ, heads = perm_list(np.random.default_rng(9), 50000)
    _check(gpu, succ, heads, "WALK")
    bits, over = gpu.list_rank_status()
    assert bits == 0 and over == 0


def test_head_walk_reports_cycle_reached_from_head(gpu):
    # 0 -> 1 -> 2 -> 0: the head's walk never ends; it stops after n hops
    # with the "walk did not end" error instead of ranking silently
    succ = np.array([1, 2, 0, -1], np.int32)
    with pytest.raises(gpu.PsiError):
        gpu.list_rank_host(succ, np.array([0], np.int32), gpu.LR_WALK)


@pytest.mark.parametrize("log_d0,local_bits", [(3, 8), (6, 12), (8, 10)])
def test_tuning_knobs_keep_results(gpu, log_d0, local_bits):
    succ, heads = perm_list(np.random.default_rng(5), 1 << 20)
    expect = oracle.list_rank(succ, heads)
    gpu.list_rank_tune(gpu.LRT_LOG_D0, log_d0)
    gpu.list_rank_tune(gpu.LRT_LOCAL_BITS, local_bits)
    try:
        _check(gpu, succ, heads, "HJ", expect)
    finally:
        gpu.list_rank_tune(gpu.LRT_LOG_D0, 5)
        gpu.list_rank_tune(gpu.LRT_LOCAL_BITS, 12)


def test_full_size_list_16m(gpu):
    """Configuration 1(i): one list of 16,777,216 nodes (device API, unchecked)."""
    import psi
    n = 1 << 24
    succ, heads = psi.gen_list(1, n)
    expect = oracle.list_rank(succ, heads)
    for method in ("HJ", "WYLLIE"):
        print(f"16M list method={method}", flush=True)
        ds = psi.DeviceArray.from_host(succ)
        dh = psi.DeviceArray.from_host(heads)
        dr = psi.DeviceArray(4 * n)
        psi.list_rank_device(ds, n, dh, 1, dr, _m(psi, method))
        psi.synchronize()
        got = dr.to_host(np.int32, n)
        assert np.array_equal(got, expect), method
        # property: ranks are a permutation of 0..n-1 and follow succ
        nz = succ >= 0
        assert np.all(got[succ[nz]] == got[nz] + 1)


def test_full_size_forest(gpu):
    """Configuration 1(ii): 1,048,576 lists, 16,777,216 nodes."""
    import psi
    succ, heads, lens = psi.gen_forest(2, 1 << 24, 1 << 20, 4096)
    expect = oracle.list_rank(succ, heads)
    for method in ("AUTO", "HJ", "WYLLIE"):
        print(f"forest 16M method={method}", flush=True)
        _check(psi, succ, heads, method, expect)


def test_concurrent_calls_take_turns(gpu):
    """psi_list_rank_host from several host threads at once (ctypes releases
    the GIL): the calls share one workspace and take turns; every result is
    exact."""
    import threading
    rng = np.random.default_rng(77)
    cases = [perm_list(rng, n) for n in (50000, 200000, 1000, 300001, 4097, 120000)]
    expect = [oracle.list_rank(s, h) for s, h in cases]
    got = [None] * len(cases)
    errors = []

    def work(k):
        try:
            for _ in range(3):
                got[k] = gpu.list_rank_host(cases[k][0], cases[k][1], gpu.LR_AUTO)
        except Exception as e:          # reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(len(cases))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k in range(len(cases)):
        assert np.array_equal(got[k], expect[k]), k
