"""solve_si with CPUList semantics (the cpulist mode, psi_solve_list) on the
GPU against the compiled reference's cpulist fixtures and the oracle's
threaded CPUList restatement: counts, strategies, labels and per-round hashes
bit-exact, including test2.gm, on which the reference's list modes never
finish (compared round for round up to the cap).  The Euler-tour valuation of
GPUList (tour of the strategy forest, Helman-JaJa ranking, broadcast:
psi_tour_path_sums) gives the same labels, path lengths and valuation digests
as the list-mode solve's final valuation."""
import numpy as np
import pytest

import oracle
from conftest import fixture_game, load_fixtures

pytestmark = pytest.mark.gpu

LIST_FIXTURES = [fx for fx in load_fixtures() if "cpulist" in fx["modes"]]


@pytest.mark.parametrize("fx", LIST_FIXTURES, ids=[fx["name"] for fx in LIST_FIXTURES])
def test_list_fixtures(gpu, fx):
    print(f"cpulist fixture {fx['name']}", flush=True)
    g = fixture_game(fx)
    exp = fx["modes"]["cpulist"]
    fin = exp["final"]
    capped = fin["status"] == "capped"
    cap = len(exp["round_hashes"]) if capped else 20000
    s = gpu.Solver(g, trace=True, trace_cap=max(cap + 1, 16) if capped else 4096)
    done = s.solve_list(max_rounds=cap)
    assert done == (not capped)
    it, br, cp = s.stats()
    assert cp == capped and (it, br) == (fin["iter"], fin["br"])
    assert s.strategy().tolist() == fin["strat"]
    assert s.labels().tolist() == fin["inf"]
    assert [[str(int(x)) for x in row] for row in s.trace_hashes()] == exp["round_hashes"]
    s.close()


def _games(gpu, seed, count, n_hi):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(2, n_hi))
        if k % 4 == 3:
            out.append(gpu.gen_chain_game(seed * 1000 + k, n))
        else:
            out.append(gpu.gen_random_game(seed * 1000 + k, n, int(rng.integers(2, 2 * n + 2)), 1,
                                           int(rng.integers(1, 5))))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_list_random_games_alone_and_batched(gpu, seed):
    games = _games(gpu, 60 + seed, 16, 4000)
    cap = 300
    exp = [oracle.solve_list(g, threads=4, max_rounds=cap, trace=True) for g in games]
    for g, r in zip(games, exp):
        s = gpu.Solver(g, trace=True, trace_cap=cap + 1)
        s.solve_list(max_rounds=cap)
        assert s.stats() == (r.iter, r.br, r.capped)
        assert np.array_equal(s.strategy(), r.strat) and np.array_equal(s.labels(), r.inf)
        assert np.array_equal(s.trace_hashes(), r.hashes)
        s.close()
    b = gpu.Solver(games)
    b.solve_list(max_rounds=cap)
    for (stats, st, lb), r in zip(b.results(), exp):
        assert stats == (r.iter, r.br, r.capped)
        assert np.array_equal(st, r.strat) and np.array_equal(lb, r.inf)
    b.close()


def test_list_mid_size_game_equals_threaded_restatement(gpu):
    g = gpu.gen_random_game(71, 200000, 1 << 20, 2, 8)
    r = oracle.solve_list(g, threads=8, max_rounds=200, trace=True)
    s = gpu.Solver(g, trace=True, trace_cap=256)
    s.solve_list(max_rounds=200)
    assert s.stats() == (r.iter, r.br, r.capped)
    assert np.array_equal(s.strategy(), r.strat) and np.array_equal(s.labels(), r.inf)
    assert np.array_equal(s.trace_hashes(), r.hashes)
    s.close()


@pytest.mark.parametrize("case", ["test2", "cfg0_seed0", "chain_01", "mid_03"])
def test_cli_cpulist(tmp_path, case):
    """gpuparsi --cpulist[=N]: the cpulist mode's line, winners and strategy;
    on test2.gm it stops at the cap in the reference's state after that many
    valuations."""
    import json
    import os
    import subprocess
    from conftest import ROOT
    fx = next(f for f in LIST_FIXTURES if f["name"] == case)
    fin = fx["modes"]["cpulist"]["final"]
    path = tmp_path / "g.gm"
    path.write_text(fixture_game(fx).to_gm())
    opt = "--cpulist"
    if fin["status"] == "capped":
        opt = f"--cpulist={len(fx['modes']['cpulist']['round_hashes'])}"
    cli = os.path.join(ROOT, "parallel-si_amd", "gpuparsi")
    p = subprocess.run([cli, opt, "--json", "--winning", "--strategy", "gpulist", str(path)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    res = json.loads(p.stdout)
    assert (res["iter"], res["br"]) == (fin["iter"], fin["br"])
    assert res["strategy"] == fin["strat"] and res["w0"] == fin["w0"] and res["w1"] == fin["w1"]
    assert ("round cap" in p.stderr) == (fin["status"] == "capped")


@pytest.mark.parametrize("kind,n", [("random", 3000), ("random", 300000), ("chain", 50000)])
def test_tour_valuation_equals_list_solve(gpu, kind, n):
    """The final strategy of a finished list-mode solve is the one its last
    valuation valued: the tour engine over it reaches exactly the finite
    vertices, with their path lengths and digests.  (Games whose Odd
    vertices have even priorities, and chain games: cpulist finishes on
    them; on the 4M generator's games it keeps switching.)"""
    if kind == "chain":
        g = gpu.gen_chain_game(805, n)
    else:
        g = gpu.gen_random_game(805, n, 64, 2, 8)
        g.priority[g.owner == 1] &= ~1
    s = gpu.Solver(g, trace=True, trace_cap=4096)
    assert s.solve_list(max_rounds=4000)
    st, lab, dep, dig = s.strategy(), s.labels(), s.depths(), s.digests()
    s.close()
    x = g.priority.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    w = x ^ (x >> np.uint64(31))
    sums, reach = gpu.tour_path_sums(st, w)
    ones, _ = gpu.tour_path_sums(st, np.ones(g.n, np.uint64))
    fin = lab == 0
    assert np.array_equal(reach == 1, fin) and np.all(lab[~fin] == 1)
    assert np.array_equal(ones[fin].astype(np.int64), dep[fin].astype(np.int64))
    assert np.array_equal(sums[fin], dig[fin])
