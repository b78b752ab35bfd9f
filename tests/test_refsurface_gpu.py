"""A driver written only against the reference's C++ class headers
(parallel-si_amd/host/si_driver.cpp: game.h, gpu_game.h, gpu_parity.h,
gpu_list_ranking.h, si.h; the solve_si loop and test() of gpumain.cpp:18-71)
runs GPUParity and GPUList through StrategyImprovement& and prints the
iteration and valuation-round counts and winning sets the compiled
reference's CPUParity wrote into the golden fixtures; GPUGame's public device
arrays hold the game's CSR."""
import os
import subprocess

import pytest

from conftest import ROOT, fixture_game, load_fixtures

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "parallel-si_amd", "si_driver")
CASES = ["test2", "cfg0_seed0", "tiny_03", "tiny_17", "mid_02", "chain_02"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", ["gpu", "gpulist"])
def test_reference_surface_driver(tmp_path, case, mode):
    fx = next(f for f in load_fixtures() if f["name"] == case)
    g = fixture_game(fx)
    path = tmp_path / "g.gm"
    path.write_text(g.to_gm())
    print(f"si_driver {mode} {case}", flush=True)
    p = subprocess.run([DRIVER, mode, str(path)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = p.stdout.splitlines()
    sec, it, br = lines[0].split()
    fin = fx["modes"]["cpu"]["final"]
    assert float(sec) >= 0 and (int(it), int(br)) == (fin["iter"], fin["br"])
    assert lines[1] == "W0:" + "".join(f" {v}" for v in fin["w0"])
    assert lines[2] == "W1:" + "".join(f" {v}" for v in fin["w1"])
    assert lines[3] == "csr 1"


def test_reference_surface_driver_parse_error(tmp_path):
    path = tmp_path / "bad.gm"
    path.write_text("parity 1;\n0 1 0 1;\n1 2 5 0;\n")
    p = subprocess.run([DRIVER, "gpu", str(path)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 255
    assert p.stdout.startswith("parsing error: ")
