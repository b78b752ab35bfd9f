"""Euler-tour path sums (the GPUList valuation engine) on the GPU: exact
sums against a brute-force CPU walk on random link forests with cycles and
detached chains, and against the oracle's valuation (path lengths and
valuation digests) of arbitrary SI states, up to the 4M-vertex game."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = x
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def brute(link, w):
    n = len(link)
    s = np.zeros(n, np.uint64)
    r = np.zeros(n, np.int32)
    for v in range(n):
        acc, x, seen = 0, v, set()
        while x >= 0 and x not in seen:
            seen.add(x)
            acc = (acc + int(w[x])) & M64
            x = int(link[x])
        if x == -1:
            s[v], r[v] = acc, 1
    return s, r


def forest(seed, n, p_cycle, p_detach):
    rng = np.random.default_rng(seed)
    link = np.where(rng.random(n) < 0.15, -1, rng.integers(0, n, n)).astype(np.int32)
    # mostly a DAG toward lower ids (a forest), with some back links forming cycles
    perm = rng.permutation(n)
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    for v in range(n):
        if link[v] >= 0 and inv[link[v]] >= inv[v] and rng.random() > p_cycle:
            link[v] = -1 if inv[v] == 0 else perm[rng.integers(0, inv[v])]
    link[rng.random(n) < p_detach] = -2
    return link


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 7), (3, 100), (4, 1000), (5, 3000)])
@pytest.mark.parametrize("method", ["AUTO", "HJ", "WALK"])
def test_tour_sums_brute_force(gpu, seed, n, method):
    print(f"tour brute seed={seed} n={n} {method}", flush=True)
    link = forest(seed, n, 0.05, 0.05)
    w = splitmix64(np.arange(n) * 7 + seed)
    m = {"AUTO": gpu.LR_AUTO, "HJ": gpu.LR_HJ, "WALK": gpu.LR_WALK}[method]
    s, r = gpu.tour_path_sums(link, w, m)
    es, er = brute(link, w)
    assert np.array_equal(r, er)
    assert np.array_equal(s, es)


@pytest.mark.parametrize("seed", range(6))
def test_tour_equals_valuation(gpu, seed):
    """Depths and digests of an arbitrary strategy/solved state: the tour over
    the valuation's links equals the oracle's valuation for every finite
    vertex, and reaches exactly those."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(10, 5000))
    g = gpu.gen_random_game(700 + seed, n, int(rng.integers(1, 60)), 1, 4)
    strat = np.array([g.targets[rng.integers(g.offsets[v], g.offsets[v + 1])] for v in range(n)],
                     np.int32)
    strat[(g.owner == 0) & (rng.random(n) < 0.3)] = -1
    solved = (rng.random(n) < 0.05).astype(np.int32)
    lab_in = np.where(solved == 1, 1, 0).astype(np.int32)
    inf, dig, depth = oracle.valuate(g, strat, solved, lab_in)
    link = np.where(solved == 1, -2, strat).astype(np.int32)
    finite = (inf == 0) & (solved == 0)
    d, reach = gpu.tour_path_sums(link, np.ones(n, np.uint64))
    assert np.array_equal(reach.astype(bool), finite)
    assert np.array_equal(d[finite].astype(np.int64), depth[finite].astype(np.int64))
    h, _ = gpu.tour_path_sums(link, splitmix64(g.priority.astype(np.uint64)))
    assert np.array_equal(h[finite], dig[finite])


def test_tour_full_size_final_strategy(gpu):
    """The 4,194,304-vertex game's final strategy: the tour's path lengths and
    digests equal the pointer-jumping valuation's (trace mode)."""
    g = gpu.gen_random_game(2, 1 << 22, 1 << 20, 2, 8)
    s = gpu.Solver(g, trace=True, trace_cap=64)
    s.solve()
    lab, dep, dig, solved, st = s.labels(), s.depths(), s.digests(), s.solved(), s.strategy()
    s.close()
    link = np.where(solved == 1, -2, st).astype(np.int32)
    finite = (lab == 0) & (solved == 0)
    d, reach = gpu.tour_path_sums(link, np.ones(g.n, np.uint64))
    assert np.array_equal(reach.astype(bool), finite)
    assert np.array_equal(d[finite].astype(np.int64), dep[finite].astype(np.int64))
    h, _ = gpu.tour_path_sums(link, splitmix64(g.priority.astype(np.uint64)))
    assert np.array_equal(h[finite], dig[finite])
