"""gpuparsi, the command line of the reference's gpumain.cpp, on the GPU:
'<seconds> <iter> <br>' plus winning sets and strategies equal the oracle's,
for plain, gzip and bzip2 input, device loop and host-driven solve_si."""
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
import psi
from conftest import ROOT, fixture_game, load_fixtures

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "parallel-si_amd", "gpuparsi")


def run(args):
    p = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("case", ["test2", "cfg0_seed0", "tiny_05", "chain_01"])
@pytest.mark.parametrize("mode", ["gpu", "gpulist"])
@pytest.mark.parametrize("loop", ["device", "host"])
def test_cli_matches_oracle(tmp_path, case, mode, loop):
    fx = next(f for f in load_fixtures() if f["name"] == case)
    g = fixture_game(fx)
    path = tmp_path / "g.gm"
    path.write_text(g.to_gm())
    args = ["--json", "--winning", "--strategy", mode, str(path)]
    if loop == "host":
        args.insert(0, "--host-loop")
    print(f"cli {case} {mode} {loop}", flush=True)
    rc, out, err = run(args)
    assert rc == 0, err
    res = json.loads(out)
    r = oracle.solve(g)
    assert (res["iter"], res["br"]) == (r.iter, r.br)
    assert res["w0"] == [v for v in range(g.n) if r.inf[v] == 1]
    assert res["strategy"] == r.strat.tolist()


def test_cli_plain_output_and_gzip(tmp_path):
    g = psi.gen_random_game(4, 2000, 60, 1, 4)
    path = tmp_path / "g.gm.gz"
    with gzip.open(path, "wt") as fh:
        fh.write(g.to_gm())
    rc, out, err = run(["gpu", str(path)])
    assert rc == 0, err
    sec, it, br = out.split()
    r = oracle.solve(g)
    assert float(sec) > 0 and (int(it), int(br)) == (r.iter, r.br)


def test_cli_bzip2(tmp_path):
    """.bz2 input (gpumain.cpp:118) gives the plain file's result; a damaged
    stream is a parsing error."""
    import bz2
    g = psi.gen_random_game(7, 3000, 80, 1, 4)
    path = tmp_path / "g.gm.bz2"
    path.write_bytes(bz2.compress(g.to_gm().encode()))
    rc, out, err = run(["--json", "--winning", "gpu", str(path)])
    assert rc == 0, err
    res = json.loads(out)
    r = oracle.solve(g)
    assert (res["iter"], res["br"]) == (r.iter, r.br)
    assert res["w0"] == [v for v in range(g.n) if r.inf[v] == 1]
    bad = tmp_path / "bad.gm.bz2"
    bad.write_bytes(path.read_bytes()[:40])
    rc, out, err = run(["gpu", str(bad)])
    assert rc == 255 and "parsing error" in err


def test_cli_errors(tmp_path):
    path = tmp_path / "bad.gm"
    path.write_text("parity 1;\n0 1 3 0;\n")
    rc, out, err = run(["gpu", str(path)])
    assert rc == 255 and "parsing error: invalid player" in err
    rc, out, err = run(["gpu"])
    assert rc == 1 and "usage" in out


@pytest.mark.parametrize("case", ["bf_00", "bf_03", "bf_06", "bf_abort_00"])
def test_cli_gpubf_matches_reference(tmp_path, case):
    """gpubf: solve_bf on the device against the reference's cpubf run."""
    fx = next(f for f in load_fixtures() if f["name"] == case)
    g = fixture_game(fx)
    path = tmp_path / "g.gm"
    path.write_text(g.to_gm())
    print(f"cli gpubf {case}", flush=True)
    rc, out, err = run(["--json", "--winning", "--strategy", "gpubf", str(path)])
    exp = fx["modes"]["cpubf"]
    if exp["returncode"] != 0:
        assert rc == 1 and "odd cycle that uses only odd nodes" in out
        return
    assert rc == 0, err
    res = json.loads(out)
    fin = exp["final"]
    assert res["iter"] == fin["iter"]
    assert res["w0"] == fin["w0"] and res["w1"] == fin["w1"]
    even = g.owner == 0
    assert np.array_equal(np.array(res["strategy"])[even], np.array(fin["strat"])[even])
