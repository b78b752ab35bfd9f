"""The game with no vertices (a PGSolver file with a header line only).

The compiled reference solves it in every mode (tests/golden/edge/empty.json,
written by tests/golden/make_golden.py --empty-only): solve_si reports iter 1,
br 1 (cpu, cpu-reset, cpubv, cpulist: one valuation and one outer step,
gpumain.cpp:18-53) and solve_bf iter 1, br 0 (main.cpp:57-77, whose first best
response counts no round), with empty strategies and winning sets.  A
one-game handle of it has no device state; batches and serving pools take
games of at least one vertex and refuse it."""
import json
import os
import subprocess

import pytest

import psi
from conftest import GOLDEN, ROOT

CLI = os.path.join(ROOT, "parallel-si_amd", "gpuparsi")


def _fixture():
    with open(os.path.join(GOLDEN, "edge", "empty.json")) as fh:
        return json.load(fh)


def _empty():
    return psi.Game([], [], [0], [])


def test_fixture_is_the_references_empty_solve():
    fx = _fixture()
    for mode, r in fx["modes"].items():
        assert r["returncode"] == 0 and r["final"]["status"] == "ok", mode
        assert r["final"]["strat"] == [] and r["final"]["w0"] == [] and r["final"]["w1"] == []
        assert (r["final"]["iter"], r["final"]["br"]) == ((1, 0) if mode == "cpubf" else (1, 1)), mode


def test_loader_reads_a_header_only_file():
    g = psi.game_from_gm(_fixture()["gm"])
    assert (g.n, g.m) == (0, 0)
    assert g.offsets.tolist() == [0]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["cpu", "cpu-reset", "cpubv", "cpulist", "cpubf"])
def test_empty_game_solves_as_the_reference(gpu, mode):
    final = _fixture()["modes"][mode]["final"]
    print(f"empty game, mode {mode}", flush=True)
    s = psi.Solver(_empty(), trace=True)
    if mode in ("cpu", "cpu-reset"):
        s.solve(reset=mode == "cpu-reset")
    elif mode == "cpubv":
        assert s.solve_bv()
    elif mode == "cpulist":
        assert s.solve_list(max_rounds=10)
    else:
        s.solve_bf()
        assert s.bf_status() == 0
    assert s.stats() == (final["iter"], final["br"], 0)
    assert s.strategy().tolist() == final["strat"]
    assert s.labels().tolist() == final["inf"]
    assert s.winning(0).tolist() == final["w0"] and s.winning(1).tolist() == final["w1"]
    th = s.trace_hashes()
    assert th.shape == (final["br"], 4) and not th.any()
    [(stats, strat, labels)] = s.results()
    assert len(strat) == 0 and len(labels) == 0
    assert stats == (final["iter"], final["br"], False)
    s.close()


@pytest.mark.gpu
def test_empty_game_stepwise_and_refused_in_batches(gpu):
    s = psi.Solver(_empty())
    s.init_strat()
    s.compute_valuation()
    s.mark_solved(1)
    assert s.switch_strategy(1) == 0
    assert s.switch_strategy(0) == 0
    s.sync()
    s.close()
    with pytest.raises(psi.PsiError):
        psi.Solver([_empty(), psi.gen_random_game(1, 100, 10, 1, 3)])
    pool = psi.Pool(1 << 26, edges_per_vertex=4.0, max_games=4, max_n=1000)
    with pytest.raises(psi.PsiError):
        pool.submit(_empty())
    pool.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["gpu", "gpulist"])
def test_cli_on_a_header_only_file(tmp_path, mode):
    path = tmp_path / "empty.gm"
    path.write_text(_fixture()["gm"])
    p = subprocess.run([CLI, mode, str(path)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    t, it, br = p.stdout.split()
    float(t)
    assert (int(it), int(br)) == (1, 1)
