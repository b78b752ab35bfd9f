"""SI solver on the GPU against the oracle and the reference's golden vectors:
bit-exact per-round strategies, solved sets, labels and valuation digests,
final strategies, winning regions and iteration / valuation-round counts."""
import numpy as np
import pytest

import oracle
from conftest import fixture_game, load_fixtures

pytestmark = pytest.mark.gpu

FIXTURES = load_fixtures()
CASES = [(fx["name"], mode) for fx in FIXTURES for mode in ("cpu", "cpu-reset")
         if mode in fx["modes"]]


def _fx(name):
    return next(f for f in FIXTURES if f["name"] == name)


@pytest.mark.parametrize("name,mode", CASES)
def test_golden(gpu, name, mode):
    print(f"golden {name} {mode}", flush=True)
    fx = _fx(name)
    g = fixture_game(fx)
    exp = fx["modes"][mode]
    fin = exp["final"]
    s = gpu.Solver(g, trace=True)
    s.solve(reset=(mode == "cpu-reset"))
    it, br, capped = s.stats()
    assert not capped
    assert (it, br) == (fin["iter"], fin["br"])
    assert s.strategy().tolist() == fin["strat"]
    assert s.labels().tolist() == fin["inf"]
    assert s.winning(0).tolist() == fin["w0"]
    assert s.winning(1).tolist() == fin["w1"]
    got = [[str(int(x)) for x in row] for row in s.trace_hashes()]
    assert got == exp["round_hashes"]
    s.close()


def _solve_both(gpu, g, reset=False):
    r = oracle.solve(g, rep=oracle.SPARSE, reset=reset, trace=True)
    s = gpu.Solver(g, trace=True, trace_cap=max(r.br + 1, 16))
    s.solve(reset=reset)
    return r, s


def _assert_same(r, s):
    it, br, capped = s.stats()
    assert (it, br) == (r.iter, r.br)
    assert np.array_equal(s.strategy(), r.strat)
    assert np.array_equal(s.labels(), r.inf)
    assert np.array_equal(s.solved(), r.solved)
    assert np.array_equal(s.trace_hashes(), r.hashes)


@pytest.mark.parametrize("seed", range(30))
def test_random_small(gpu, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    g = gpu.gen_random_game(9000 + seed, n, int(rng.integers(1, 3 * n)), 1, int(rng.integers(1, 7)))
    print(f"random n={n} seed={seed}", flush=True)
    r, s = _solve_both(gpu, g, reset=bool(seed % 3 == 0))
    _assert_same(r, s)


@pytest.mark.parametrize("n,pri,dlo,dhi", [(20000, 100, 1, 4), (65536, 1 << 14, 2, 8),
                                            (100000, 100000, 1, 6), (200000, 16, 1, 3)])
def test_random_mid(gpu, n, pri, dlo, dhi):
    print(f"random mid n={n}", flush=True)
    g = gpu.gen_random_game(n + pri, n, pri, dlo, dhi)
    r, s = _solve_both(gpu, g)
    _assert_same(r, s)


@pytest.mark.parametrize("n", [50, 1000, 20000])
def test_chain_games(gpu, n):
    print(f"chain n={n}", flush=True)
    g = gpu.gen_chain_game(77 + n, n)
    r, s = _solve_both(gpu, g)
    _assert_same(r, s)


@pytest.mark.parametrize("seed", range(12))
def test_valuation_of_given_state(gpu, seed):
    """One valuation of an arbitrary strategy with an arbitrary solved set:
    labels (cycle parity, inheritance from solved vertices), path lengths and
    digests equal the oracle's (pins the jump kernels alone)."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 3000))
    g = gpu.gen_random_game(400 + seed, n, int(rng.integers(1, 50)), 1, 4)
    strat = np.array([g.targets[rng.integers(g.offsets[v], g.offsets[v + 1])] for v in range(n)],
                     np.int32)
    strat[(g.owner == 0) & (rng.random(n) < 0.3)] = -1
    solved = (rng.random(n) < 0.1).astype(np.int32)
    lab_in = np.where(solved == 1, np.where(rng.random(n) < 0.5, 1, -1), 0).astype(np.int32)
    inf, dig, depth = oracle.valuate(g, strat, solved, lab_in)
    s = gpu.Solver(g, trace=True)
    s.set_state(strat, solved, lab_in)
    s.compute_valuation()
    assert np.array_equal(s.labels(), inf)
    assert np.array_equal(s.depths(), depth)
    assert np.array_equal(s.digests(), dig)


@pytest.mark.parametrize("seed", range(8))
def test_switch_of_given_state(gpu, seed):
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(5, 3000))
    g = gpu.gen_random_game(600 + seed, n, int(rng.integers(1, 20)), 1, 5)
    strat = np.array([g.targets[rng.integers(g.offsets[v], g.offsets[v + 1])] for v in range(n)],
                     np.int32)
    strat[(g.owner == 0) & (rng.random(n) < 0.5)] = -1
    solved = np.zeros(n, np.int32)
    for player in (0, 1):
        cnt, exp = oracle.switch(g, strat, solved, np.zeros(n, np.int32), player)
        s = gpu.Solver(g)
        s.set_state(strat, solved, None)
        s.compute_valuation()
        sw = s.switch_strategy(player)
        assert sw == (1 if cnt else 0)
        assert np.array_equal(s.strategy(), exp)


@pytest.mark.parametrize("seed", range(4))
def test_state_loaded_after_solve(gpu, seed):
    """A state loaded into a handle that has already solved (the game
    finished): the game's run restarts from that state, and the next
    valuation and switch equal the oracle's for it, whatever the earlier
    valuations stored."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(200, 3000))
    g = gpu.gen_random_game(700 + seed, n, int(rng.integers(1, 30)), 1, 4)
    s = gpu.Solver(g, trace=True)
    s.solve()
    assert s.stats()[1] > 0
    for rep in range(2):
        strat = np.array([g.targets[rng.integers(g.offsets[v], g.offsets[v + 1])]
                          for v in range(n)], np.int32)
        strat[(g.owner == 0) & (rng.random(n) < 0.3)] = -1
        solved = (rng.random(n) < 0.1).astype(np.int32)
        lab_in = rng.integers(-1, 2, n).astype(np.int32)
        lab_in[(solved == 1) & (lab_in == 0)] = 1
        inf, dig, depth = oracle.valuate(g, strat, solved, lab_in)
        s.set_state(strat, solved, lab_in)
        assert s.stats() == (0, 0, False)
        s.compute_valuation()
        assert np.array_equal(s.labels(), inf), rep
        assert np.array_equal(s.depths(), depth), rep
        assert np.array_equal(s.digests(), dig), rep
        for player in (0, 1):
            cnt, exp = oracle.switch(g, strat, solved, lab_in, player)
            s.set_state(strat, solved, lab_in)
            s.compute_valuation()
            assert s.switch_strategy(player) == (1 if cnt else 0)
            assert np.array_equal(s.strategy(), exp), (rep, player)
    s.close()


def test_state_loaded_into_one_game_of_a_batch(gpu):
    """Loading a state into one game of a solved batch restarts that game
    alone: its valuation equals the oracle's, the other games keep their
    results and counts."""
    rng = np.random.default_rng(77)
    games = [gpu.gen_random_game(810 + k, n, 40, 1, 4) for k, n in enumerate((900, 2500, 1300))]
    b = gpu.Solver(games, trace=True)
    b.solve()
    before = b.results()
    g = games[1]
    n = g.n
    strat = np.array([g.targets[rng.integers(g.offsets[v], g.offsets[v + 1])] for v in range(n)],
                     np.int32)
    solved = (rng.random(n) < 0.1).astype(np.int32)
    lab_in = np.where(solved == 1, 1, rng.integers(-1, 2, n)).astype(np.int32)
    inf, dig, depth = oracle.valuate(g, strat, solved, lab_in)
    b.set_state(strat, solved, lab_in, game=1)
    assert b.stats(1) == (0, 0, False)
    b.compute_valuation()
    assert np.array_equal(b.labels(1), inf)
    assert np.array_equal(b.depths(1), depth)
    assert np.array_equal(b.digests(1), dig)
    for k in (0, 2):
        assert b.stats(k) == before[k][0], k
        assert np.array_equal(b.strategy(k), before[k][1]), k
        assert np.array_equal(b.labels(k), before[k][2]), k
    b.close()


def test_stepwise_api_matches_solve(gpu):
    """Driving init_strat / compute_valuation / mark_solved / switch_strategy
    from the host (the reference's solve_si loop) gives the same result."""
    g = gpu.gen_random_game(31337, 5000, 300, 1, 5)
    r = oracle.solve(g)
    s = gpu.Solver(g)
    s.init_strat()
    br = it = 0
    while True:
        while True:
            s.compute_valuation()
            br += 1
            s.mark_solved(1)
            if s.switch_strategy(1) == 0:
                break
        s.mark_solved(0)
        it += 1
        if s.switch_strategy(0) == 0:
            break
    assert (it, br) == (r.iter, r.br)
    assert np.array_equal(s.strategy(), r.strat)
    assert np.array_equal(s.labels(), r.inf)


def test_batch_equals_alone(gpu):
    """Many games of uneven size in one handle: each result equals its solve
    alone (per-game control blocks)."""
    sizes = [3, 17, 400, 2500, 10, 9000, 1, 64]
    games = []
    for i, n in enumerate(sizes):
        n = max(n, 2)
        if i % 3 == 2:
            games.append(gpu.gen_chain_game(i, n))
        else:
            games.append(gpu.gen_random_game(50 + i, n, max(1, n // 3), 1, 4))
    b = gpu.Solver(games, trace=True)
    b.solve()
    bulk = b.results()
    for k, g in enumerate(games):
        r = oracle.solve(g, trace=True)
        it, br, capped = b.stats(k)
        assert (it, br) == (r.iter, r.br), k
        assert np.array_equal(b.strategy(k), r.strat), k
        assert np.array_equal(b.labels(k), r.inf), k
        assert np.array_equal(b.trace_hashes(k), r.hashes), k
        assert bulk[k][0] == (it, br, capped)
        assert np.array_equal(bulk[k][1], r.strat) and np.array_equal(bulk[k][2], r.inf)


def test_round_cap(gpu):
    g = gpu.gen_random_game(5, 3000, 100, 1, 4)
    r = oracle.solve(g, max_rounds=3)
    assert r.capped
    s = gpu.Solver(g)
    s.solve(max_rounds=3)
    it, br, capped = s.stats()
    assert capped and br == 3 and it == r.iter
    assert np.array_equal(s.strategy(), r.strat)


def test_invalid_games_rejected(gpu):
    G = gpu.Game([0, 1], [1, 2], [0, 1, 2], [1, 5])
    with pytest.raises(gpu.PsiError, match="successor out of range"):
        gpu.Solver(G)
    G = gpu.Game([0, 1], [-1, 2], [0, 1, 2], [1, 0])
    with pytest.raises(gpu.PsiError, match="priority"):
        gpu.Solver(G)


def _ladder(n, seed, podd):
    """Mostly-Even ladder v -> v+1, v+1+r with priorities rising along it and a
    small odd priority at the tail: sink-tree depths reach ~n (beyond the
    2^12 stored lifting levels, so deep compares take repeated top jumps)."""
    rng = np.random.default_rng(seed)
    import psi
    owner = np.zeros(n, np.int32)
    pri = (2 * (np.arange(n) // 7)).astype(np.int32)
    pri[rng.random(n) < podd] += 1
    pri[-1] = 1
    off, tgt = [0], []
    for v in range(n):
        if v == n - 1:
            tgt.append(v)
        else:
            tgt.append(v + 1)
            j = v + 1 + int(rng.integers(1, 3))
            if j < n:
                tgt.append(j)
        off.append(len(tgt))
    return psi.Game(owner, pri, np.array(off, np.int32), np.array(tgt, np.int32))


@pytest.mark.parametrize("podd", [0.0, 0.02, 0.1])
def test_deep_sink_trees(gpu, podd):
    g = _ladder(12000, 3, podd)
    print(f"deep ladder podd={podd}", flush=True)
    r, s = _solve_both(gpu, g)
    _assert_same(r, s)


def test_handles_solve_concurrently(gpu):
    """Distinct handles on distinct host threads (each handle has its own
    stream and step graphs): every result equals the oracle's."""
    import threading
    games = [gpu.gen_random_game(500 + k, 20000 + 3000 * k, 500, 1, 5) for k in range(4)]
    exp = [oracle.solve(g) for g in games]
    out = [None] * len(games)

    def run(k):
        s = gpu.Solver(games[k])
        for _ in range(3):
            s.solve()
        out[k] = (s.stats(), s.strategy(), s.labels())
        s.close()
    th = [threading.Thread(target=run, args=(k,)) for k in range(len(games))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (st, strat, lab), r in zip(out, exp):
        assert st[:2] == (r.iter, r.br)
        assert np.array_equal(strat, r.strat) and np.array_equal(lab, r.inf)
