"""solve_si with CPUBV semantics (the cpubv mode, psi_solve_bv) on the GPU
against the compiled reference's cpubv fixtures and the oracle's CPUBV
restatement: completed games bit-exact (counts, strategies, labels, per-round
hashes); games with an odd strategy cycle stop at the same valuation, also
inside a batch where the other games run on to their own results."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, fixture_game, load_fixtures

pytestmark = pytest.mark.gpu

BV_FIXTURES = [fx for fx in load_fixtures() if "cpubv" in fx["modes"]]
CLI = os.path.join(ROOT, "parallel-si_amd", "gpuparsi")


@pytest.mark.parametrize("fx", BV_FIXTURES, ids=[fx["name"] for fx in BV_FIXTURES])
def test_bv_fixtures(gpu, fx):
    print(f"bv fixture {fx['name']}", flush=True)
    g = fixture_game(fx)
    exp = fx["modes"]["cpubv"]
    s = gpu.Solver(g, trace=True)
    ok = s.solve_bv()
    it, br, capped = s.stats()
    if exp["returncode"] != 0:
        assert not ok and s.bf_status() == 1
        assert br == len(exp["round_hashes"])
        got = [[str(int(x)) for x in row] for row in s.trace_hashes()][:br]
        assert got == exp["round_hashes"]
    else:
        fin = exp["final"]
        assert ok and s.bf_status() == 0 and not capped
        assert (it, br) == (fin["iter"], fin["br"])
        assert s.strategy().tolist() == fin["strat"]
        assert s.labels().tolist() == fin["inf"]
        assert [[str(int(x)) for x in row] for row in s.trace_hashes()] == exp["round_hashes"]
    s.close()


def _games(gpu, seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(3, 3000))
        g = gpu.gen_random_game(seed * 1000 + k, n, int(rng.integers(2, 2 * n + 2)), 1,
                                int(rng.integers(1, 5)))
        if k % 2:
            g.priority[:] &= ~1                     # even priorities: never stops
        out.append(g)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_bv_random_games_alone_and_batched(gpu, seed):
    games = _games(gpu, 40 + seed, 24)
    exp = [oracle.solve_bv(g) for g in games]
    assert any(r.status for r in exp) and not all(r.status for r in exp)
    for g, r in zip(games, exp):
        s = gpu.Solver(g)
        ok = s.solve_bv()
        assert ok == (r.status == 0) and s.bf_status() == r.status
        it, br, capped = s.stats()
        assert br == r.br
        if r.status == 0:
            assert it == r.iter
            assert np.array_equal(s.strategy(), r.strat) and np.array_equal(s.labels(), r.inf)
        s.close()
    b = gpu.Solver(games)
    ok = b.solve_bv()
    assert not ok
    for k, ((stats, st, lb), r) in enumerate(zip(b.results(), exp)):
        assert b.bf_status(k) == r.status, k
        assert stats[1] == r.br, k
        if r.status == 0:
            assert stats[0] == r.iter and np.array_equal(st, r.strat) and np.array_equal(lb, r.inf)
    b.close()


def test_cli_gpubv(tmp_path):
    fx = next(f for f in BV_FIXTURES if f["name"] == "test2")
    path = tmp_path / "test2.gm"
    path.write_text(fixture_game(fx).to_gm())
    p = subprocess.run([CLI, "gpubv", str(path)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert p.stdout.strip() == "Game contains an odd cycle that uses only odd nodes"
    ok = next(f for f in BV_FIXTURES if f["modes"]["cpubv"]["returncode"] == 0)
    path = tmp_path / "ok.gm"
    path.write_text(fixture_game(ok).to_gm())
    p = subprocess.run([CLI, "gpubv", str(path)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    sec, it, br = p.stdout.split()
    fin = ok["modes"]["cpubv"]["final"]
    assert (int(it), int(br)) == (fin["iter"], fin["br"])
