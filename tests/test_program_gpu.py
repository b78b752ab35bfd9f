"""The solve loop as a step program (identical launches, each running the next
phase) against the same loop launched phase by phase, and both against the
oracle: identical iteration and valuation-round counts, strategies, labels,
solved sets and per-round trace hashes, whatever the number of launches per
graph; the launch-by-launch profile of the program solves the same."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture
def modes(gpu):
    gpu.si_tune(gpu.SIT_PUSH_MIN_N, 0)      # change marks on these small games too
    yield gpu
    gpu.si_tune(gpu.SIT_PUSH_MIN_N, gpu.SIT_PUSH_MIN_N_DEFAULT)
    gpu.si_tune(gpu.SIT_PROGRAM, 1)
    gpu.si_tune(gpu.SIT_TWO_LEVELS, 1)
    gpu.si_tune(gpu.SIT_PUSH, gpu.SIT_PUSH_DEFAULT)


def _run(gpu, games, program, steps_per_sync=4, reset=False):
    gpu.si_tune(gpu.SIT_PROGRAM, program)
    s = gpu.Solver(games, trace=True, trace_cap=512)
    s.solve(reset=reset, steps_per_sync=steps_per_sync)
    G = len(games) if isinstance(games, list) else 1
    out = [(s.stats(k), s.strategy(k), s.labels(k), s.solved(k), s.trace_hashes(k)) for k in range(G)]
    s.close()
    return out


def _same(a, b):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert x[0] == y[0], k
        for u, v in zip(x[1:], y[1:]):
            assert np.array_equal(u, v), k


def _games(gpu):
    return [gpu.gen_random_game(11, 30000, 20000, 2, 8),
            gpu.gen_random_game(12, 20000, 60, 1, 4),
            gpu.gen_chain_game(13, 5000)]


@pytest.mark.parametrize("idx", range(3))
def test_program_equals_phase_launches(modes, idx):
    gpu = modes
    g = _games(gpu)[idx]
    print(f"program vs phases, game {idx}", flush=True)
    a = _run(gpu, g, 1)
    b = _run(gpu, g, 0)
    _same(a, b)
    r = oracle.solve(g, trace=True)
    assert a[0][0][:2] == (r.iter, r.br)
    assert np.array_equal(a[0][1], r.strat) and np.array_equal(a[0][2], r.inf)
    assert np.array_equal(a[0][4], r.hashes[: len(a[0][4])])


def test_program_reset_schedule(modes):
    gpu = modes
    g = gpu.gen_random_game(21, 8000, 500, 1, 5)
    r = oracle.solve(g, reset=True)
    a = _run(gpu, g, 1, reset=True)
    _same(a, _run(gpu, g, 0, reset=True))
    assert a[0][0][:2] == (r.iter, r.br)
    assert np.array_equal(a[0][1], r.strat)


@pytest.mark.parametrize("steps", [1, 2, 3, 7])
def test_program_any_graph_length(modes, steps):
    """Graphs of 16 x steps launches stop anywhere in a step; the next graph
    continues it."""
    gpu = modes
    g = _games(gpu)[0]
    _same(_run(gpu, g, 1, steps_per_sync=steps), _run(gpu, g, 0))


def test_program_recaptures_on_one_handle(modes):
    """One handle solving again with other graph lengths (a new capture each
    time) and after the fixed-sequence form: the same results every time."""
    gpu = modes
    g = gpu.gen_random_game(31, 20000, 3000, 1, 5)
    r = oracle.solve(g)
    s = gpu.Solver(g)
    for program, steps in [(1, 4), (1, 1), (0, 4), (1, 3), (1, 4)]:
        gpu.si_tune(gpu.SIT_PROGRAM, program)
        s.solve(steps_per_sync=steps)
        assert s.stats()[:2] == (r.iter, r.br), (program, steps)
        assert np.array_equal(s.strategy(), r.strat) and np.array_equal(s.labels(), r.inf)
    s.close()


def test_program_batch(modes):
    gpu = modes
    sizes = [3, 17, 400, 2500, 10, 9000, 1, 64, 30000]
    games = []
    for i, n in enumerate(sizes):
        n = max(n, 2)
        games.append(gpu.gen_chain_game(i, n) if i % 3 == 2
                     else gpu.gen_random_game(70 + i, n, max(1, n // 3), 1, 4))
    _same(_run(gpu, games, 1), _run(gpu, games, 0))


def test_program_round_cap(modes):
    gpu = modes
    g = gpu.gen_random_game(5, 3000, 100, 1, 4)
    r = oracle.solve(g, max_rounds=3)
    gpu.si_tune(gpu.SIT_PROGRAM, 1)
    s = gpu.Solver(g)
    s.solve(max_rounds=3)
    it, br, capped = s.stats()
    assert capped and br == 3 and it == r.iter
    assert np.array_equal(s.strategy(), r.strat)


def test_program_profiled(modes):
    gpu = modes
    g = _games(gpu)[1]
    gpu.si_tune(gpu.SIT_PROGRAM, 1)
    s = gpu.Solver(g)
    s.solve()
    st, strat, lab = s.stats(), s.strategy(), s.labels()
    p = gpu.solve_program_profiled(s)
    assert s.stats() == st
    assert np.array_equal(s.strategy(), strat) and np.array_equal(s.labels(), lab)
    assert p["total"][1] == st[1]                 # one INIT per valuation round
    assert p["val_init"][1] == p["label"][1] == p["switch1"][1] == st[1]
    assert p["jump_idle"][1] == 0 and p["jump"][1] >= st[1]
    s.close()


def test_solve_then_bf_on_one_handle(modes):
    """solve_bf allocates its successor record after a plain solve captured
    its graphs: the captured launches are rebuilt with it."""
    gpu = modes
    rng = np.random.default_rng(3)
    found = 0
    for k in range(200):
        n = int(rng.integers(20, 300))
        g = gpu.gen_random_game(4400 + k, n, 4 * n, 1, 4)
        g.priority[:] = rng.permutation(n).astype(np.int32)
        r = oracle.solve_bf(g)
        if r.status != oracle.BF_OK or r.iter < 2:
            continue
        for program in (1, 0):
            gpu.si_tune(gpu.SIT_PROGRAM, program)
            s = gpu.Solver(g)
            s.solve()
            s.solve_bf()
            assert s.stats()[0] == r.iter
            assert np.array_equal(s.strategy(), r.strat)
            assert np.array_equal(s.labels(), r.inf)
            s.close()
        found += 1
        if found == 4:
            break
    assert found == 4


@pytest.mark.parametrize("two", [0, 1])
@pytest.mark.parametrize("idx", range(3))
def test_program_two_levels(modes, idx, two):
    """Jump launches advancing one or two levels: the same strategies, labels,
    solved sets, counts and per-round trace hashes as the rounds launched one
    by one."""
    gpu = modes
    g = _games(gpu)[idx]
    print(f"two levels {two}, game {idx}", flush=True)
    gpu.si_tune(gpu.SIT_TWO_LEVELS, two)
    _same(_run(gpu, g, 1), _run(gpu, g, 0))


def test_program_two_levels_batch_reset(modes):
    gpu = modes
    sizes = [3, 17, 400, 2500, 10, 9000, 1, 64, 30000]
    games = []
    for i, n in enumerate(sizes):
        n = max(n, 2)
        games.append(gpu.gen_chain_game(i, n) if i % 3 == 2
                     else gpu.gen_random_game(90 + i, n, max(1, n // 3), 1, 4))
    gpu.si_tune(gpu.SIT_TWO_LEVELS, 1)
    _same(_run(gpu, games, 1, reset=True), _run(gpu, games, 0, reset=True))
    _same(_run(gpu, games, 1, steps_per_sync=3), _run(gpu, games, 0))


def test_program_two_levels_long_cycles(modes):
    """Valuations whose cycle-bound rounds run past the lifting levels (into
    the rotating buffers, two levels per launch): a 4096-vertex single cycle
    with tails, and chains long enough for deep rounds."""
    gpu = modes
    games = [gpu.gen_chain_game(41, 70000), gpu.gen_random_game(42, 200000, 20, 1, 2)]
    for g in games:
        r = oracle.solve(g)
        gpu.si_tune(gpu.SIT_TWO_LEVELS, 1)
        a = _run(gpu, g, 1)
        assert a[0][0][:2] == (r.iter, r.br)
        assert np.array_equal(a[0][1], r.strat) and np.array_equal(a[0][2], r.inf)
        _same(a, _run(gpu, g, 0))


def test_program_grid_retuned_on_one_handle(modes):
    """The step-program grid retuned between solves of one handle (the
    captured launches are rebuilt with it): the same results every time."""
    gpu = modes
    g = gpu.gen_random_game(51, 60000, 4000, 1, 5)
    r = oracle.solve(g)
    gpu.si_tune(gpu.SIT_PROGRAM, 1)
    s = gpu.Solver(g)
    try:
        for grid in (2048, 64, 100, 2048):        # 118, 64, 100, 118 workgroups
            gpu.si_tune(gpu.SIT_PHASE_GRID, grid)
            s.solve()
            assert s.stats()[:2] == (r.iter, r.br), grid
            assert np.array_equal(s.strategy(), r.strat) and np.array_equal(s.labels(), r.inf), grid
    finally:
        gpu.si_tune(gpu.SIT_PHASE_GRID, 2048)
        s.close()


# Change marks (one game): push 1 marks after every valuation, 16 (default)
# after those with at most n/16 changes, 1 << 20 almost never, 0 never; the
# switch passes then visit marked entries only.  Results equal the oracle's
# and the program's without marks, round by round.
@pytest.mark.parametrize("idx", range(3))
@pytest.mark.parametrize("push", [1, 16, 1 << 20])
def test_program_change_marks(modes, idx, push):
    gpu = modes
    g = _games(gpu)[idx]
    print(f"change marks push={push}, game {idx}", flush=True)
    gpu.si_tune(gpu.SIT_PUSH, 0)
    a = _run(gpu, g, 1)
    gpu.si_tune(gpu.SIT_PUSH, push)
    b = _run(gpu, g, 1)
    _same(a, b)
    r = oracle.solve(g, trace=True)
    assert b[0][0][:2] == (r.iter, r.br)
    assert np.array_equal(b[0][1], r.strat) and np.array_equal(b[0][2], r.inf)
    assert np.array_equal(b[0][4], r.hashes[: len(b[0][4])])


@pytest.mark.parametrize("push", [1, 16])
def test_program_change_marks_reset_mode(modes, push):
    gpu = modes
    g = gpu.gen_random_game(21, 40000, 30000, 2, 6)
    gpu.si_tune(gpu.SIT_PUSH, push)
    b = _run(gpu, g, 1, reset=True)
    r = oracle.solve(g, reset=True, trace=True)
    assert b[0][0][:2] == (r.iter, r.br)
    assert np.array_equal(b[0][1], r.strat) and np.array_equal(b[0][2], r.inf)
    assert np.array_equal(b[0][4], r.hashes[: len(b[0][4])])


def test_program_change_marks_repeated_solves(modes):
    """Marks left by one solve are harmless to the next on the same handle,
    with the marks threshold retuned in between."""
    gpu = modes
    g = gpu.gen_random_game(22, 50000, 1 << 20, 2, 8)
    r = oracle.solve(g)
    s = gpu.Solver(g)
    for push in (1, 16, 1, 0, 4):
        gpu.si_tune(gpu.SIT_PUSH, push)
        s.solve()
        assert s.stats()[:2] == (r.iter, r.br), push
        assert np.array_equal(s.strategy(), r.strat), push
    s.close()


def test_program_marks_ties_and_duplicates(modes):
    """Few priorities (many equal keys: exact compares) and a game with
    repeated successors, with and without change marks."""
    gpu = modes
    games = [gpu.gen_random_game(31, 20000, 4, 2, 8), gpu.gen_random_game(32, 30000, 16, 1, 3)]
    dup = gpu.gen_random_game(33, 5000, 50, 2, 5)
    t = dup.targets.copy()
    t[1::3] = t[0::3][: len(t[1::3])]        # some successors listed twice
    games.append(gpu.Game(dup.owner, dup.priority, dup.offsets, t))
    for g in games:
        r = oracle.solve(g, trace=True)
        for push in (0, 1, 16):
            gpu.si_tune(gpu.SIT_PUSH, push)
            b = _run(gpu, g, 1)
            assert b[0][0][:2] == (r.iter, r.br), push
            assert np.array_equal(b[0][1], r.strat) and np.array_equal(b[0][2], r.inf), push
            assert np.array_equal(b[0][4], r.hashes[: len(b[0][4])]), push
