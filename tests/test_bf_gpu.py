"""solve_bf (the cpu_bf role) on the GPU against the oracle's CPU_BellmanFord
restatement and the reference's cpubf fixtures: labels, solved sets, final
valuations (digests), Even's strategy, iteration counts and winning regions
exactly; Odd's final strategy (succ) exactly where priorities are distinct,
so that every minimum successor is unique."""
import numpy as np
import pytest

import oracle
from conftest import fixture_game, load_fixtures

pytestmark = pytest.mark.gpu

BF_FIXTURES = [fx for fx in load_fixtures() if "cpubf" in fx["modes"]]


def _check(gpu, g, r, distinct):
    s = gpu.Solver(g, trace=True)
    s.solve_bf()
    it, br, capped = s.stats()
    assert not capped and s.bf_status() == 0
    assert it == r.iter
    st = s.strategy()
    even = g.owner == 0
    assert np.array_equal(st[even], r.strat[even])
    if distinct:
        assert np.array_equal(st, r.strat)
    assert np.array_equal(s.labels(), r.inf)
    assert np.array_equal(s.solved(), r.solved)
    assert np.array_equal(s.digests(), r.dig)
    s.close()


@pytest.mark.parametrize("fx", BF_FIXTURES, ids=[fx["name"] for fx in BF_FIXTURES])
def test_bf_fixtures(gpu, fx):
    print(f"bf fixture {fx['name']}", flush=True)
    g = fixture_game(fx)
    exp = fx["modes"]["cpubf"]
    if exp["returncode"] != 0:
        s = gpu.Solver(g)
        with pytest.raises(gpu.PsiError, match="odd cycle"):
            s.solve_bf()
        assert s.bf_status() == 1
        return
    fin = exp["final"]
    s = gpu.Solver(g)
    s.solve_bf()
    assert s.stats()[0] == fin["iter"]
    assert s.labels().tolist() == fin["inf"]
    assert s.winning(0).tolist() == fin["w0"]
    even = g.owner == 0
    assert np.array_equal(s.strategy()[even], np.array(fin["strat"])[even])
    if len(set(g.priority.tolist())) == g.n:
        assert s.strategy().tolist() == fin["strat"]


def _bf_games(gpu, seed, count, n_hi, distinct):
    rng = np.random.default_rng(seed)
    out = []
    k = 0
    while len(out) < count and k < 40 * count:
        k += 1
        n = int(rng.integers(4, n_hi))
        g = gpu.gen_random_game(seed * 1000 + k, n, int(rng.integers(2, 4 * n)), 1,
                                int(rng.integers(2, 6)))
        if distinct:
            g.priority[:] = rng.permutation(n).astype(np.int32)
        else:
            g.priority[:] = rng.integers(0, 12, n).astype(np.int32)
            g.priority[g.owner == 1] &= ~1
        r = oracle.solve_bf(g)
        if r.status == oracle.BF_OK and r.iter >= 2:
            out.append((g, r))
    return out


@pytest.mark.parametrize("seed,distinct", [(1, True), (2, True), (3, False), (4, False)])
def test_bf_random(gpu, seed, distinct):
    print(f"bf random seed={seed}", flush=True)
    games = _bf_games(gpu, seed, 12, 300, distinct)
    assert len(games) >= 6
    for g, r in games:
        _check(gpu, g, r, distinct)


def test_bf_odd_cycle_aborts(gpu):
    rng = np.random.default_rng(9)
    seen = 0
    for k in range(60):
        n = int(rng.integers(4, 40))
        g = gpu.gen_random_game(9100 + k, n, 2 * n, 1, 3)
        r = oracle.solve_bf(g)
        if r.status != oracle.BF_ODD_CYCLE:
            continue
        s = gpu.Solver(g)
        with pytest.raises(gpu.PsiError, match="odd cycle"):
            s.solve_bf()
        assert s.bf_status() == 1
        seen += 1
    assert seen >= 3


def test_bf_batch_equals_alone(gpu):
    games = _bf_games(gpu, 5, 10, 400, True)
    b = gpu.Solver([g for g, _ in games])
    b.solve_bf()
    for k, (g, r) in enumerate(games):
        assert b.stats(k)[0] == r.iter
        assert np.array_equal(b.strategy(k), r.strat)
        assert np.array_equal(b.labels(k), r.inf)
