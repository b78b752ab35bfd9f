"""The full-size workloads on the GPU against the oracle's fixtures
(tests/golden/large/, written by tests/golden/make_large.py with the CPUParity
restatement on the CPU):

* the 4,194,304-vertex random game (priority U[0,2^20), out-degree U[2,8],
  seed 2): iteration and valuation-round counts, every round's hashes of the
  strategy valued, solved flags, labels and valuation digests, and the final
  strategy, labels and solved flags;
* the 1,024-game mixed-size set (n log-uniform in [1000, 10^6], priority
  U[0,n), out-degree U[1,6]) in one batched handle: every game's counts,
  final strategy and labels;
* the first requests of the 4,096-request serving stream through the
  serving pool, arriving between steps: every request's counts, final
  strategy and labels.
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
LARGE = os.path.join(GOLDEN, "large")


def _load(name):
    with open(os.path.join(LARGE, name)) as fh:
        return json.load(fh)


def test_cfg2_4M_solve_equals_oracle(gpu):
    fx = _load("cfg2_4M.json")
    gen = fx["gen"]
    print("full-size 4M game, trace mode", flush=True)
    g = gpu.gen_random_game(gen["seed"], gen["n"], gen["pri_bound"], gen["deg_lo"], gen["deg_hi"])
    assert (g.n, g.m) == (fx["n"], fx["m"])
    s = gpu.Solver(g, trace=True, trace_cap=256)
    s.solve()
    it, br, capped = s.stats()
    assert not capped
    assert (it, br) == (fx["iter"], fx["br"])
    got = [[str(int(x)) for x in row] for row in s.trace_hashes()]
    assert got == fx["round_hashes"]
    assert str(gpu.hash_i32(s.strategy())) == fx["strat"]
    assert str(gpu.hash_i32(s.labels())) == fx["labels"]
    assert str(gpu.hash_i32(s.solved())) == fx["solved"]
    assert int((s.labels() == 1).sum()) == fx["w0"]
    s.close()
    # the plain (untraced) solve the benchmark times gives the same result
    s = gpu.Solver(g)
    s.solve()
    assert s.stats() == (fx["iter"], fx["br"], False)
    assert str(gpu.hash_i32(s.strategy())) == fx["strat"]
    assert str(gpu.hash_i32(s.labels())) == fx["labels"]
    s.close()


def test_cfg3_batch_every_game_equals_oracle(gpu):
    fx = _load("cfg3_batch.json")
    gen = fx["gen"]
    sizes = gpu.gen_sizes(gen["size_seed"], gen["games"], gen["lo"], gen["hi"])
    print(f"1,024-game batch, sum n = {int(sizes.sum())}", flush=True)

    def make(k):
        n = int(sizes[k])
        return gpu.gen_random_game(gen["size_seed"] * 100000 + k, n, n, gen["deg_lo"], gen["deg_hi"])

    with ThreadPoolExecutor(16) as ex:
        games = list(ex.map(make, range(gen["games"])))
    b = gpu.Solver(games)
    b.solve()
    res = b.results()
    b.close()
    exp = {r["k"]: r for r in fx["games"]}
    assert len(exp) == len(games)
    bad = []
    for k, (g, (stats, st, lb)) in enumerate(zip(games, res)):
        e = exp[k]
        ok = (g.n == e["n"] and stats == (e["iter"], e["br"], False) and
              str(gpu.hash_i32(st)) == e["strat"] and str(gpu.hash_i32(lb)) == e["labels"])
        if not ok:
            bad.append(k)
    assert not bad, f"games differing from the oracle: {bad[:20]}"


def test_cfg4_stream_prefix_through_pool(gpu):
    """The serving stream's first 64 requests (random games of the 4M
    generator up to 3.1M vertices and chain games; the four chains above
    400,000 vertices, hours of oracle time, are left out) arrive one by one
    between pool steps into a pool smaller than their total, queue for
    memory, run mixed with each other, and each result equals the oracle's
    solve of that game alone."""
    import numpy as np
    fx = _load("cfg4_stream.json")
    gen = fx["gen"]
    sizes = gpu.gen_sizes(gen["seed"], gen["requests"], gen["lo"], gen["hi"])
    chain = np.random.default_rng(gen["seed"]).random(gen["requests"]) < gen["chain_frac"]
    reqs = sorted(fx["requests"], key=lambda r: r["k"])
    print(f"serving stream prefix: {len(reqs)} requests, sum n = {sum(r['n'] for r in reqs)}",
          flush=True)

    def make(k):
        n = int(sizes[k])
        if chain[k]:
            return gpu.gen_chain_game(gen["seed"] * 100000 + k, n)
        return gpu.gen_random_game(gen["seed"] * 100000 + k, n, 1 << 20, 2, 8)

    with ThreadPoolExecutor(16) as ex:
        games = list(ex.map(make, [r["k"] for r in reqs]))
    # room for ~6M vertices at once: large requests wait for earlier ones
    bpv = gpu.pool_bytes_per_vertex(1 << 22)
    p = gpu.Pool(int((bpv + 4 * 5.0) * 6_500_000), edges_per_vertex=5.0, max_games=16,
                 max_n=1 << 22)
    tickets, done, max_running, max_queued = [], [], 0, 0
    for g in games:
        tickets.append(p.submit(g))
        done += p.step(2)
        info = p.info()
        max_running = max(max_running, info["running"])
        max_queued = max(max_queued, info["queued"])
    done += p.drain(2)
    assert sorted(done) == sorted(tickets)
    assert max_running > 1 and max_queued > 0          # mixed and queued, not one at a time
    bad = []
    for t, g, e in zip(tickets, games, reqs):
        stats, st, lb = p.result(t)
        if not (g.n == e["n"] and stats == (e["iter"], e["br"], False) and
                str(gpu.hash_i32(st)) == e["strat"] and str(gpu.hash_i32(lb)) == e["labels"]):
            bad.append(e["k"])
    p.close()
    assert not bad, f"requests differing from the oracle: {bad}"
