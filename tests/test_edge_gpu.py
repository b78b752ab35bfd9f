"""Edge cases of the SI path on the GPU against the oracle: one-vertex games,
one-player games, duplicate successors (the switch scans CSR order, so a
repeated edge must not switch twice), vertices with many successors (past the
8-wide batches), priorities at the largest value the records hold (2^30 - 1),
all-zero priorities, and the same through a batch, the serving pool and
solve_bf."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _game(gpu, owner, pri, succ_lists):
    off, tgt = [0], []
    for s in succ_lists:
        tgt += list(s)
        off.append(len(tgt))
    return gpu.Game(np.array(owner, np.int32), np.array(pri, np.int32), np.array(off, np.int32),
                    np.array(tgt, np.int32))


def _same(gpu, g, reset=False):
    r = oracle.solve(g, rep=oracle.SPARSE, reset=reset, trace=True)
    s = gpu.Solver(g, trace=True, trace_cap=max(r.br + 1, 16))
    s.solve(reset=reset)
    it, br, capped = s.stats()
    assert (it, br) == (r.iter, r.br)
    assert np.array_equal(s.strategy(), r.strat)
    assert np.array_equal(s.labels(), r.inf)
    assert np.array_equal(s.trace_hashes(), r.hashes)
    s.close()
    return r


@pytest.mark.parametrize("owner,pri", [(0, 0), (0, 1), (1, 2), (1, 3), (0, (1 << 30) - 1),
                                       (1, (1 << 30) - 2)])
def test_one_vertex(gpu, owner, pri):
    g = _game(gpu, [owner], [pri], [[0]])
    r = _same(gpu, g)
    assert (r.inf[0] == 1) == (pri % 2 == 0)      # player 0 wins exactly the even loop


@pytest.mark.parametrize("player", [0, 1])
def test_one_player_games(gpu, player):
    g = gpu.gen_random_game(40 + player, 3000, 200, 1, 5)
    g.owner[:] = player
    _same(gpu, g)


def test_duplicate_successors(gpu):
    rng = np.random.default_rng(5)
    n = 500
    owner = rng.integers(0, 2, n)
    pri = rng.integers(0, 40, n)
    succ = []
    for v in range(n):
        a, b = rng.integers(0, n, 2)
        succ.append([a, b, a, a, b] if v % 3 else [a, a])
    _same(gpu, _game(gpu, owner, pri, succ))


def test_high_degree(gpu):
    rng = np.random.default_rng(6)
    n = 2000
    owner = rng.integers(0, 2, n)
    pri = rng.integers(0, 300, n)
    succ = [rng.choice(n, size=int(rng.integers(1, 120 if v % 50 == 0 else 4)), replace=False)
            for v in range(n)]
    _same(gpu, _game(gpu, owner, pri, succ))


@pytest.mark.parametrize("seed", range(3))
def test_largest_priorities(gpu, seed):
    """Priorities in [2^30 - 64, 2^30): the record's 30-bit field and the
    changed bit above it must not mix."""
    g = gpu.gen_random_game(70 + seed, 4000, 64, 1, 4)
    g.priority[:] = g.priority + ((1 << 30) - 64)
    _same(gpu, g, reset=seed == 1)


def test_all_zero_priorities(gpu):
    g = gpu.gen_random_game(80, 3000, 1, 1, 4)
    assert int(g.priority.max()) == 0
    _same(gpu, g)


def test_edge_games_in_batch_pool_and_bf(gpu):
    games = [_game(gpu, [0], [0], [[0]]), _game(gpu, [1], [5], [[0]]),
             gpu.gen_random_game(90, 300, 50, 1, 3)]
    games[2].priority[:] = games[2].priority + ((1 << 30) - 64)
    exp = [oracle.solve(g) for g in games]
    b = gpu.Solver(games)
    b.solve()
    for k, r in enumerate(exp):
        assert b.stats(k)[:2] == (r.iter, r.br)
        assert np.array_equal(b.strategy(k), r.strat)
    p = gpu.Pool(64 << 20, max_games=2, max_n=1024)
    ts = [p.submit(g) for g in games]
    p.drain()
    for t, r in zip(ts, exp):
        (it, br, _), st, lb = p.result(t)
        assert (it, br) == (r.iter, r.br) and np.array_equal(st, r.strat)
    p.close()
    g = _game(gpu, [0, 1, 0], [2, 4, 6], [[1, 2], [0, 2], [2]])
    rb = oracle.solve_bf(g)
    s = gpu.Solver(g)
    s.solve_bf()
    assert s.stats()[0] == rb.iter and np.array_equal(s.labels(), rb.inf)
