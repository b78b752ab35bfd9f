"""Serving pool (continuous batching) on the GPU: every request's result is
the same in the mix as alone -- equal to its single-game GPU solve and to the
oracle -- while requests arrive, queue for memory, and reuse the ranges and
control slots of finished ones."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _requests(gpu, seed, count, n_hi):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(2), np.log(n_hi))))
        if k % 5 == 4:
            g = gpu.gen_chain_game(seed * 1000 + k, max(n, 2))
        else:
            g = gpu.gen_random_game(seed * 1000 + k, n, max(1, int(rng.integers(1, 2 * n))),
                                    1, int(rng.integers(1, 7)))
        out.append((g, bool(rng.random() < 0.2)))
    return out


def _check(pool, tickets, reqs):
    for t, (g, reset) in zip(tickets, reqs):
        (it, br, capped), st, lb = pool.result(t)
        r = oracle.solve(g, reset=reset)
        assert not capped
        assert (it, br) == (r.iter, r.br), (t, g.n)
        assert np.array_equal(st, r.strat), (t, g.n)
        assert np.array_equal(lb, r.inf), (t, g.n)


def test_pool_single_equals_solver(gpu):
    g = gpu.gen_random_game(11, 5000, 300, 1, 5)
    s = gpu.Solver(g)
    s.solve()
    p = gpu.Pool(64 << 20, max_games=4, max_n=1 << 16)
    t = p.submit(g)
    assert p.poll(t) == gpu.RUNNING
    done = p.drain()
    assert done == [t]
    (it, br, capped), st, lb = p.result(t)
    assert (it, br, capped) == s.stats()
    assert np.array_equal(st, s.strategy()) and np.array_equal(lb, s.labels())
    p.close()


@pytest.fixture(params=[1, 0], ids=["program", "fixed"])
def pool_mode(gpu, request):
    """Pools step as the step program (admitted games join at a step
    boundary) or with the fixed launch sequence; fixed at pool creation."""
    gpu.si_tune(gpu.SIT_PROGRAM, request.param)
    yield request.param
    gpu.si_tune(gpu.SIT_PROGRAM, 1)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pool_mix_equals_alone(gpu, pool_mode, seed):
    """Requests arrive between steps into a pool too small to hold them all:
    they queue, are admitted as others finish, and each result equals the
    oracle's solve of that game alone."""
    print(f"pool mix seed={seed} program={pool_mode}", flush=True)
    reqs = _requests(gpu, seed, 40, 6000)
    bpv = gpu.pool_bytes_per_vertex(8192)
    p = gpu.Pool((bpv + 20) * 20000, edges_per_vertex=5.0, max_games=12, max_n=8192)
    tickets, done = [], []
    rng = np.random.default_rng(seed)
    i = 0
    while i < len(reqs):
        for _ in range(int(rng.integers(0, 5))):
            if i < len(reqs):
                g, reset = reqs[i]
                tickets.append(p.submit(g, reset=reset))
                i += 1
        done += p.step(int(rng.integers(1, 4)))
    done += p.drain()
    assert sorted(done) == sorted(tickets)
    _check(p, tickets, reqs)
    p.close()


def test_pool_mix_equals_batch_and_solo(gpu, pool_mode):
    """The same games through the pool, as one batch handle and one at a time
    on the GPU: identical results (mid-size games, beyond the oracle's quick
    range at scale)."""
    reqs = [(gpu.gen_random_game(70 + k, n, n, 2, 8), False)
            for k, n in enumerate([30000, 120000, 5000, 250000, 60000, 900, 180000])]
    reqs.append((gpu.gen_chain_game(5, 40000), False))
    p = gpu.Pool(1 << 30, max_games=8, max_n=1 << 18)
    tickets = [p.submit(g) for g, _ in reqs]
    p.drain(2)
    b = gpu.Solver([g for g, _ in reqs])
    b.solve()
    bulk = b.results()
    for k, (t, (g, _)) in enumerate(zip(tickets, reqs)):
        stats, st, lb = p.result(t)
        s = gpu.Solver(g)
        s.solve()
        assert stats == s.stats() == bulk[k][0], k
        assert np.array_equal(st, s.strategy()) and np.array_equal(st, bulk[k][1]), k
        assert np.array_equal(lb, s.labels()) and np.array_equal(lb, bulk[k][2]), k
        s.close()
    p.close()


def test_pool_many_tiny_reuse_slots(gpu, pool_mode):
    reqs = [(gpu.gen_random_game(500 + k, 2 + k % 50, 8, 1, 3), k % 7 == 0) for k in range(300)]
    p = gpu.Pool(8 << 20, max_games=8, max_n=64)
    tickets = [p.submit(g, reset=r) for g, r in reqs]
    done = p.drain(1)
    assert sorted(done) == sorted(tickets)
    _check(p, tickets, reqs)
    p.close()


def test_pool_round_cap(gpu, pool_mode):
    g = gpu.gen_random_game(5, 3000, 100, 1, 4)
    r = oracle.solve(g, max_rounds=3)
    p = gpu.Pool(32 << 20, max_games=2, max_n=4096)
    t = p.submit(g, max_rounds=3)
    p.drain()
    (it, br, capped), st, _ = p.result(t)
    assert capped and br == 3 and it == r.iter
    assert np.array_equal(st, r.strat)
    p.close()


def test_pool_refusals(gpu):
    p = gpu.Pool(16 << 20, max_games=2, max_n=1000)
    with pytest.raises(gpu.PsiError) as e:
        p.submit(gpu.gen_random_game(1, 2000, 10, 1, 3))
    assert e.value.rc == gpu.PSI_ERR_BUDGET
    with pytest.raises(gpu.PsiError, match="successor out of range"):
        p.submit(gpu.Game([0, 1], [1, 2], [0, 1, 2], [1, 5]))
    with pytest.raises(gpu.PsiError) as e:
        gpu.Pool(1 << 10, max_games=2, max_n=1000)
    assert e.value.rc == gpu.PSI_ERR_BUDGET
    p.close()
