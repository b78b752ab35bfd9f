"""Games whose strategy graphs are mostly cycles with long tails, solved by the
step program and compared with the oracle round by round: these valuations
end by the cycle-label check (every vertex still jumping after the stop has
the label bit of its successor), not by a window as long as the count of
cycle-bound vertices.  Labels, solved sets, strategies, per-round hashes and
counts are bit-exact; the check is also shown to cut the jump launches."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def cycle_game(seed, n, n_cycles, max_cycle, chain_bias, pri_bound, p_even):
    """Odd vertices follow a random functional graph (cycles over the first
    vertices of a random order, every later vertex linked to an earlier one,
    to the one just before it with probability chain_bias: long tails); Even
    vertices (probability p_even) also get two random successors."""
    import psi
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    f = np.empty(n, np.int64)
    pos = 0
    for _ in range(n_cycles):
        L = min(int(rng.integers(1, max_cycle + 1)), n - pos)
        if L <= 0:
            break
        cyc = order[pos:pos + L]
        f[cyc] = np.roll(cyc, -1)
        pos += L
    prev = order[np.maximum(np.arange(n) - 1, 0)]
    pick = order[(rng.random(n) * np.maximum(np.arange(n), 1)).astype(np.int64)]
    later = np.where(rng.random(n) < chain_bias, prev, pick)
    f[order[pos:]] = later[pos:]
    owner = np.where(rng.random(n) < p_even, 0, 1).astype(np.int32)   # 0 Even, 1 Odd
    pri = rng.integers(0, pri_bound, n).astype(np.int32)
    off = [0]
    tgt = []
    extra = rng.integers(0, n, (n, 2))
    for v in range(n):
        tgt.append(int(f[v]))
        if owner[v] == 0:
            tgt.extend(int(x) for x in extra[v])
        off.append(len(tgt))
    return psi.Game(owner, pri, np.array(off, np.int32), np.array(tgt, np.int32))


CASES = [
    # seed, n, cycles, longest cycle, tail bias, priority bound, share of Even vertices
    (1, 3000, 10, 200, 0.95, 50, 0.05),
    (2, 20000, 40, 2000, 0.99, 1000, 0.02),
    (3, 60000, 5, 30000, 0.9, 7, 0.01),
    (4, 100000, 300, 64, 0.995, 100000, 0.03),
    (5, 150000, 2, 3, 0.999, 5, 0.0),
    (6, 50000, 100, 500, 0.97, 200, 0.3),
    (7, 200000, 1000, 100, 0.98, 1 << 16, 0.15),
    (8, 100000, 20, 5000, 0.999, 9, 0.5),
]


@pytest.mark.parametrize("seed,n,cycles,maxc,bias,pb,pe", CASES)
def test_cycle_games_equal_oracle(gpu, seed, n, cycles, maxc, bias, pb, pe):
    print(f"cycle game seed={seed} n={n}", flush=True)
    g = cycle_game(seed, n, cycles, maxc, bias, pb, pe)
    r = oracle.solve(g, rep=oracle.SPARSE, trace=True)
    s = gpu.Solver(g, trace=True, trace_cap=max(r.br + 1, 16))
    s.solve()
    it, br, capped = s.stats()
    assert (it, br) == (r.iter, r.br)
    assert np.array_equal(s.strategy(), r.strat)
    assert np.array_equal(s.labels(), r.inf)
    assert np.array_equal(s.solved(), r.solved)
    assert np.array_equal(s.trace_hashes(), r.hashes)
    s.close()


def test_cycle_games_batched_equal_oracle(gpu):
    """The check's count is over the whole batch: each game as alone."""
    print("cycle game batch", flush=True)
    games = [cycle_game(20 + k, 5000 + 7000 * k, 3 + k, 50 * (k + 1), 0.98, 30 + k, 0.03)
             for k in range(6)]
    games.append(gpu.gen_random_game(77, 30000, 1000, 1, 5))
    exp = [oracle.solve(g) for g in games]
    s = gpu.Solver(games)
    s.solve()
    for k, (g, r) in enumerate(zip(games, exp)):
        it, br, capped = s.stats(k)
        assert (it, br) == (r.iter, r.br), k
        assert np.array_equal(s.strategy(k), r.strat), k
        assert np.array_equal(s.labels(k), r.inf), k
    s.close()


def test_cycle_check_cuts_jump_launches(gpu):
    """A 2^18-vertex game whose single valuation is cycle-bound everywhere
    (tails of at most a few hundred vertices, short cycles): by the window
    rule the rounds would run until 2^r covers the 2^18 cycle-bound
    vertices (nine two-level launches); the label check stops them once the
    windows have reached the cycles."""
    n = 1 << 18
    g = cycle_game(9, n, 4000, 16, 0.5, 1 << 20, 0.0)
    r = oracle.solve(g)
    s = gpu.Solver(g)
    gpu.solve_program_profiled(s)
    tl = gpu.program_launches(s)
    jumps = [x for x in tl if x[0] == "jump"]
    assert 0 < len(jumps) <= 7, [x[1] for x in jumps]
    assert np.array_equal(s.labels(), r.inf)
    assert np.array_equal(s.strategy(), r.strat)
    s.close()
