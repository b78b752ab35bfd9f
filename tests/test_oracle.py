"""Pins the oracle (CPU restatement) before it is trusted as the checker.

* every golden fixture written by the compiled reference CPUParity
  (tests/golden/make_golden.py): per-round strategy / solved / label /
  valuation-digest hashes, final strategy, labels, iteration counts;
* dense and sparse valuation representations agree on everything;
* winning regions agree with Zielonka's algorithm (textbook, independent);
* the parser reproduces the reference's error messages (game.cpp:51-107);
* the brute-force list ranking satisfies closed forms.
A planted mistake in the oracle (see test_planted_mistakes_are_caught) fails
these checks.
"""
import numpy as np
import pytest

import oracle
import psi
from conftest import fixture_game, load_fixtures

FIXTURES = load_fixtures()
SOLVE_CASES = [(fx["name"], mode) for fx in FIXTURES for mode in ("cpu", "cpu-reset")
               if mode in fx["modes"]]


def _fx(name):
    return next(f for f in FIXTURES if f["name"] == name)


@pytest.mark.parametrize("name,mode", SOLVE_CASES)
@pytest.mark.parametrize("rep", [oracle.DENSE, oracle.SPARSE])
def test_solve_matches_reference(name, mode, rep):
    fx = _fx(name)
    g = fixture_game(fx)
    exp = fx["modes"][mode]
    r = oracle.solve(g, rep=rep, reset=(mode == "cpu-reset"), trace=True)
    fin = exp["final"]
    assert fin["status"] == "ok"
    assert (r.iter, r.br) == (fin["iter"], fin["br"])
    assert r.strat.tolist() == fin["strat"]
    assert r.inf.tolist() == fin["inf"]
    got = [[str(int(x)) for x in row] for row in r.hashes]
    assert got == exp["round_hashes"]


@pytest.mark.parametrize("name", [fx["name"] for fx in FIXTURES
                                  if "cpu" in fx["modes"] and "rounds" in fx["modes"]["cpu"]])
def test_full_round_arrays(name):
    fx = _fx(name)
    g = fixture_game(fx)
    exp = fx["modes"]["cpu"]["rounds"]
    r = oracle.solve(g, rep=oracle.SPARSE, full_trace_rounds=len(exp) + 1)
    assert len(r.strat_rounds) == len(exp)
    for k, e in enumerate(exp):
        assert r.strat_rounds[k].tolist() == e["strat"]
        assert r.inf_rounds[k].tolist() == e["inf"]
        assert [str(int(x)) for x in r.dig_rounds[k]] == e["dig"]


@pytest.mark.parametrize("name", [fx["name"] for fx in FIXTURES if "cpu" in fx["modes"]])
def test_zielonka_agrees_with_reference_winners(name):
    fx = _fx(name)
    fin = fx["modes"]["cpu"]["final"]
    w = oracle.zielonka(fixture_game(fx))
    assert [v for v in range(len(w)) if w[v] == 0] == fin["w0"]
    assert [v for v in range(len(w)) if w[v] == 1] == fin["w1"]


def test_test2_known_answer():
    """Hand trace of the reference's test2.gm (10 vertices)."""
    fx = _fx("test2")
    g = fixture_game(fx)
    r = oracle.solve(g)
    assert (r.iter, r.br) == (2, 3)
    assert r.strat.tolist() == [(i + 2) % 10 for i in range(10)]
    assert r.inf.tolist() == [1, -1] * 5
    m = fx["modes"]
    assert m["cpubv"]["message"] == "Game contains an odd cycle that uses only odd nodes"
    assert m["cpubf"]["message"] == "Game contains an odd cycle that uses only odd nodes"
    assert m["cpulist"]["final"]["status"] == "capped"   # list modes oscillate here


@pytest.mark.parametrize("seed", range(40))
def test_dense_equals_sparse_random(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    g = psi.gen_random_game(5000 + seed, n, int(rng.integers(1, 12)), 1, int(rng.integers(1, 5)))
    a = oracle.solve(g, rep=oracle.DENSE, trace=True)
    b = oracle.solve(g, rep=oracle.SPARSE, trace=True)
    assert (a.iter, a.br) == (b.iter, b.br)
    assert np.array_equal(a.strat, b.strat) and np.array_equal(a.inf, b.inf)
    assert np.array_equal(a.hashes, b.hashes)
    w = oracle.zielonka(g)
    assert np.array_equal(w == 0, a.inf == 1)


@pytest.mark.parametrize("seed", range(10))
def test_compare_is_total_preorder(seed):
    """Exhaustive pairwise compares on one valued strategy: antisymmetric,
    transitive, and sparse == dense."""
    g = psi.gen_random_game(7000 + seed, 25, 6, 1, 3)
    r = oracle.solve(g, max_rounds=2)
    strat = r.strat
    vs = np.arange(-1, g.n, dtype=np.int32)
    us = np.repeat(vs, vs.size)
    ws = np.tile(vs, vs.size)
    cs = oracle.compare_pairs(g, strat, us, ws, rep=oracle.SPARSE)
    cd = oracle.compare_pairs(g, strat, us, ws, rep=oracle.DENSE)
    assert np.array_equal(cs, cd)
    M = cs.reshape(vs.size, vs.size)
    assert np.array_equal(M, -M.T)
    le = M <= 0
    # transitivity of <= : le[i,j] & le[j,k] -> le[i,k]
    viol = (le.astype(np.int32) @ le.astype(np.int32)) > 0
    assert not np.any(viol & ~le)


PARSE_ERRORS = [
    ("parity 1;\n-1 0 0 0;\n", "invalid id"),
    ("parity 1;\n0\n", "missing priority"),
    ("parity 1;\n0 1\n", "missing player"),
    ("parity 1;\n0 1 2 0;\n", "invalid player"),
    ("parity 1;\n1 1 0 0;\n", "vertices must be listed in order starting at 0!"),
    ("parity 1;\n0 1 0\n", "missing successor"),
    ("parity 1;\n0 1 0 0\n", "missing ; to end line"),
    ("parity 1;\n0 1 0 x;\n", "missing successor"),
]


@pytest.mark.parametrize("text,msg", PARSE_ERRORS)
def test_parser_errors(text, msg):
    with pytest.raises(oracle.OracleError, match=msg):
        oracle.parse_gm(text)


def test_parser_grammar():
    text = 'parity 3;\n\n0 3 1 1,2 "a";\n1 0 0 0;\n2 7 1 2 ,0 ;\n'
    owner, pri, off, tgt, mp = oracle.parse_gm(text)
    assert owner.tolist() == [1, 0, 1]
    assert pri.tolist() == [3, 0, 7]
    assert off.tolist() == [0, 2, 3, 5]
    assert tgt.tolist() == [1, 2, 0, 2, 0]
    assert mp == 7


def test_gm_roundtrip():
    g = psi.gen_random_game(1, 300, 50, 1, 6)
    owner, pri, off, tgt, mp = oracle.parse_gm(g.to_gm())
    assert np.array_equal(owner, g.owner) and np.array_equal(pri, g.priority)
    assert np.array_equal(off, g.offsets) and np.array_equal(tgt, g.targets)


def test_list_rank_closed_forms():
    # one list over a permutation: rank is the inverse permutation
    perm = np.random.default_rng(3).permutation(1000).astype(np.int32)
    succ = np.full(1000, -1, np.int32)
    succ[perm[:-1]] = perm[1:]
    r = oracle.list_rank(succ, perm[:1])
    inv = np.empty(1000, np.int32)
    inv[perm] = np.arange(1000)
    assert np.array_equal(r, inv)
    # a node no head reaches reads -1; a head-reachable cycle is an error
    succ2 = np.array([1, -1, 3, 2], np.int32)
    assert oracle.list_rank(succ2, np.array([0], np.int32)).tolist() == [0, 1, -1, -1]
    with pytest.raises(oracle.OracleError):
        oracle.list_rank(np.array([1, 0], np.int32), np.array([0], np.int32))
    # weighted scan = prefix sums along the list
    w = np.arange(1000, dtype=np.int64)
    s = oracle.list_scan(succ, perm[:1], w)
    assert np.array_equal(s[perm], np.cumsum(w[perm]))


def test_planted_mistakes_are_caught():
    """A plausible mistake in the compare (odd/even swapped) would change the
    solve; the golden hashes detect it.  The mistake is planted here, in a
    Python restatement of the compare, not in the product."""
    fx = _fx("cfg0_seed0")
    g = fixture_game(fx)
    r = oracle.solve(g, rep=oracle.DENSE, max_rounds=1)
    strat = r.strat
    rng = np.random.default_rng(0)
    us = rng.integers(-1, g.n, 400).astype(np.int32)
    ws = rng.integers(-1, g.n, 400).astype(np.int32)
    good = oracle.compare_pairs(g, strat, us, ws, rep=oracle.DENSE)
    # planted: flip the sign of every finite compare
    inf, _, _ = oracle.valuate(g, strat, np.zeros(g.n, np.int32), np.zeros(g.n, np.int32))
    fin = np.array([(u < 0 or inf[u] == 0) and (w < 0 or inf[w] == 0) for u, w in zip(us, ws)])
    planted = np.where(fin, -good, good)
    assert not np.array_equal(planted, good)


BF_FIXTURES = [fx["name"] for fx in FIXTURES if "cpubf" in fx["modes"]]


@pytest.mark.parametrize("name", BF_FIXTURES)
def test_solve_bf_matches_reference(name):
    """CPU_BellmanFord restatement vs the compiled reference's cpubf run:
    final strategy (with succ for Odd), labels, iteration and relaxation-round
    counts; games whose first best response meets an odd cycle abort."""
    fx = _fx(name)
    exp = fx["modes"]["cpubf"]
    r = oracle.solve_bf(fixture_game(fx))
    if exp["returncode"] != 0:
        assert "odd cycle" in exp["message"]
        assert r.status == oracle.BF_ODD_CYCLE
        return
    fin = exp["final"]
    assert r.status == oracle.BF_OK
    assert (r.iter, r.br) == (fin["iter"], fin["br"])
    assert r.strat.tolist() == fin["strat"]
    assert r.inf.tolist() == fin["inf"]


@pytest.mark.parametrize("name", [n for n in BF_FIXTURES if _fx(n)["modes"]["cpubf"]["returncode"] == 0])
def test_solve_bf_winners_are_the_games(name):
    """On games solve_bf completes, its winning regions are the game's
    (Zielonka) and CPUParity's."""
    fx = _fx(name)
    g = fixture_game(fx)
    fin = fx["modes"]["cpubf"]["final"]
    w = oracle.zielonka(g)
    assert [v for v in range(g.n) if w[v] == 0] == fin["w0"]
    r = oracle.solve(g)
    assert [v for v in range(g.n) if r.inf[v] == 1] == fin["w0"]


def test_solve_bf_relaxation_is_a_best_response():
    """Property of the fixed point the relaxation reaches (independent of
    the update order): every unsolved finite Odd vertex's valuation is the
    minimum over its successors, so SI switch(1) on the final state switches
    nothing once Odd follows succ (checked with the dense CPUParity compare)."""
    rng = np.random.default_rng(3)
    checked = 0
    for k in range(60):
        n = int(rng.integers(4, 30))
        g = psi.gen_random_game(8100 + k, n, int(rng.integers(2, 9)), 1, 3)
        g.priority[g.owner == 1] &= ~1
        r = oracle.solve_bf(g)
        if r.status != oracle.BF_OK or r.iter < 2:
            continue
        cnt, after = oracle.switch(g, r.strat, r.solved, r.inf, 1)
        assert cnt == 0, k
        checked += 1
    assert checked >= 10


LIST_FIXTURES = [fx["name"] for fx in FIXTURES if "cpulist" in fx["modes"]]


@pytest.mark.parametrize("name", LIST_FIXTURES)
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_solve_list_matches_reference(name, threads):
    """The threaded CPUList restatement (Euler tour + Helman-JaJa valuation
    on T threads) vs the compiled reference's cpulist run on every fixture:
    per-round hashes of strategy, solved flags, labels and valuation digests,
    and the final counts, strategy and labels; games the reference's cpulist
    does not finish (test2.gm, whose odd vertices oscillate) match round for
    round up to its cap."""
    fx = _fx(name)
    g = fixture_game(fx)
    exp = fx["modes"]["cpulist"]
    fin = exp["final"]
    cap = len(exp["round_hashes"]) if fin["status"] == "capped" else 0
    r = oracle.solve_list(g, threads=threads, max_rounds=cap, trace=True)
    assert [[str(int(x)) for x in row] for row in r.hashes] == exp["round_hashes"]
    assert r.capped == (fin["status"] == "capped")
    assert (r.iter, r.br) == (fin["iter"], fin["br"])
    assert r.strat.tolist() == fin["strat"] and r.inf.tolist() == fin["inf"]


def test_solve_list_thread_count_changes_nothing():
    """The tour's child order depends on the threads' atomic exchanges and
    the sublists on the thread ranges; the valuations, hence the whole run,
    do not (mid-size games, 8 threads vs 1, round-capped)."""
    rng = np.random.default_rng(23)
    for k in range(6):
        n = int(rng.integers(2000, 20000))
        g = psi.gen_random_game(9700 + k, n, int(rng.integers(2, 2 * n)), 1, 4)
        a = oracle.solve_list(g, threads=1, max_rounds=60, trace=True)
        b = oracle.solve_list(g, threads=8, max_rounds=60, trace=True)
        assert (a.iter, a.br, a.capped) == (b.iter, b.br, b.capped)
        assert np.array_equal(a.strat, b.strat) and np.array_equal(a.inf, b.inf)
        assert np.array_equal(a.hashes, b.hashes)


BV_FIXTURES = [fx["name"] for fx in FIXTURES if "cpubv" in fx["modes"]]


@pytest.mark.parametrize("name", BV_FIXTURES)
@pytest.mark.parametrize("rep", [oracle.DENSE, oracle.SPARSE])
def test_solve_bv_matches_reference(name, rep):
    """CPUBV restatement vs the compiled reference's cpubv run (test2.gm, the
    reference's own fixture, and the bv_* games): games it completes end with
    its iteration and round counts, strategy, labels and per-round hashes;
    games it stops on an odd strategy cycle stop after the same number of
    completed valuations, each of which equals the reference's."""
    fx = _fx(name)
    g = fixture_game(fx)
    exp = fx["modes"]["cpubv"]
    r = oracle.solve_bv(g, rep=rep)
    rounds = exp["round_hashes"]
    if exp["returncode"] != 0:
        assert exp["message"] == "Game contains an odd cycle that uses only odd nodes"
        assert r.status == oracle.BV_ODD_CYCLE
        assert r.br == len(rounds)
        if rounds:
            t = oracle.solve(g, rep=rep, max_rounds=len(rounds), trace=True)
            got = [[str(int(x)) for x in row] for row in t.hashes]
            assert got == rounds
        return
    fin = exp["final"]
    assert r.status == oracle.BV_OK and not r.capped
    assert (r.iter, r.br) == (fin["iter"], fin["br"])
    assert r.strat.tolist() == fin["strat"] and r.inf.tolist() == fin["inf"]
    t = oracle.solve(g, rep=rep, trace=True)
    assert [[str(int(x)) for x in row] for row in t.hashes] == rounds


def test_solve_bv_completes_exactly_when_no_odd_cycle_label():
    """CPUBV stops exactly at the first round in which CPUParity labels a
    vertex -1 (only an odd strategy cycle produces -1 when nothing is solved
    -1 yet); otherwise both solves are the same."""
    rng = np.random.default_rng(17)
    stopped = completed = 0
    for k in range(80):
        n = int(rng.integers(3, 40))
        g = psi.gen_random_game(9500 + k, n, int(rng.integers(2, 2 * n + 2)), 1, 3)
        r = oracle.solve(g, full_trace_rounds=4096)
        first_neg = next((i for i, inf in enumerate(r.inf_rounds) if (inf == -1).any()), None)
        b = oracle.solve_bv(g)
        if first_neg is None:
            assert b.status == oracle.BV_OK
            assert (b.iter, b.br) == (r.iter, r.br) and np.array_equal(b.strat, r.strat)
            completed += 1
        else:
            assert b.status == oracle.BV_ODD_CYCLE and b.br == first_neg
            stopped += 1
    assert stopped >= 10 and completed >= 10
