"""Writes the golden fixtures in tests/golden/ from the reference's own CPU
solvers, compiled from /root/reference by oracle/Makefile (target `ref`) into
oracle/_ref/ref_harness.  Run in the build container only:

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture is data: a game (CSR arrays) and what the reference computed for
it.  Per valuation round r the fixture holds hashes (oracle.hash_i32 /
hash_u64) of the strategy being valued, the solved flags, the labels
(infinite[]) and the per-vertex valuation digests
    dig[v] = sum_p count_v[p] * splitmix64(p)  (mod 2^64), 0 if v is not finite,
and, for small games, the arrays themselves.  Final strategy, labels, winning
sets, iteration and valuation-round counts are stored in full.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "parallel-si_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import psi  # noqa: E402  (host-only generators)
import oracle  # noqa: E402  (hash functions)

HARNESS = oracle.REF_HARNESS
TEST2 = "/root/reference/test2.gm"


def run_ref(mode, gm_path, max_rounds, threads=1):
    p = subprocess.run([HARNESS, mode, gm_path, str(max_rounds), str(threads)],
                       capture_output=True, text=True, timeout=600)
    rounds, final, aborted = [], None, None
    for line in p.stdout.splitlines():
        line = line.strip()
        if not line:
            continue
        if line.startswith("{"):
            obj = json.loads(line)
            if "round" in obj:
                rounds.append(obj)
            elif "final" in obj:
                final = obj
            elif "parse_error" in obj:
                aborted = "parse error: " + obj["parse_error"]
        else:
            aborted = line
    return {"returncode": p.returncode, "rounds": rounds, "final": final, "message": aborted}


def summarize(res, full):
    out = {"returncode": res["returncode"], "message": res["message"]}
    hashes = []
    for r in res["rounds"]:
        hashes.append([
            str(oracle.hash_i32(np.array(r["strat"], np.int32))),
            str(oracle.hash_i32(np.array(r["solved"], np.int32))),
            str(oracle.hash_i32(np.array(r["inf"], np.int32))),
            str(oracle.hash_u64(np.array([int(x) for x in r["dig"]], np.uint64))),
        ])
    out["round_hashes"] = hashes
    if full:
        out["rounds"] = [{"strat": r["strat"], "solved": r["solved"], "inf": r["inf"],
                          "dig": [str(x) for x in r["dig"]]} for r in res["rounds"]]
    f = res["final"]
    if f is not None:
        out["final"] = {k: f[k] for k in ("status", "iter", "br", "strat", "inf", "w0", "w1")}
    return out


def game_dict(g):
    return {"n": g.n, "owner": g.owner.tolist(), "priority": g.priority.tolist(),
            "offsets": g.offsets.tolist(), "targets": g.targets.tolist()}


def make(name, g, modes, full, max_rounds=2000, gm_text=None):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, name + ".gm")
        with open(path, "w") as fh:
            fh.write(gm_text if gm_text is not None else g.to_gm())
        fx = {"name": name, "game": game_dict(g), "modes": {}}
        for mode in modes:
            res = run_ref(mode, path, max_rounds)
            fx["modes"][mode] = summarize(res, full)
    with open(os.path.join(HERE, name + ".json"), "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    st = {m: (v.get("final") or {}).get("status", v["message"]) for m, v in fx["modes"].items()}
    print(name, g.n, st)


def make_bf():
    """cpubf fixtures: games the reference's solve_bf completes, and games where
    its first best response meets an odd strategy cycle and exits."""
    rng = np.random.default_rng(77)
    done_ok, done_abort, k = 0, 0, 0
    while done_ok < 8 or done_abort < 2:
        k += 1
        n = int(rng.integers(4, 40))
        if k % 2 == 0:
            g = psi.gen_random_game(7000 + k, n, int(rng.integers(2, 12)), 1, 3)
            g.priority[g.owner == 1] &= ~1      # Odd vertices even: no Odd-only odd cycles
        else:
            g = psi.gen_random_game(7000 + k, n, 4 * n, 1, 4)
            g.priority[:] = rng.permutation(n).astype(np.int32)   # distinct priorities
        with tempfile.NamedTemporaryFile("w", suffix=".gm", delete=False) as fh:
            fh.write(g.to_gm())
            path = fh.name
        res = run_ref("cpubf", path, 100000)
        os.unlink(path)
        ok = res["final"] is not None and res["final"]["status"] == "ok"
        if ok and done_ok < 8:
            name, done_ok = f"bf_{done_ok:02d}", done_ok + 1
        elif not ok and res["returncode"] != 0 and done_abort < 2:
            name, done_abort = f"bf_abort_{done_abort:02d}", done_abort + 1
        else:
            continue
        fx = {"name": name, "game": game_dict(g), "modes": {"cpubf": summarize(res, False)}}
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(fx, fh, separators=(",", ":"))
        print(name, g.n, res["final"]["status"] if res["final"] else res["message"])


def make_bv():
    """cpubv fixtures (CPUBV<ArrayVal> under solve_si): games the reference
    solves, and games where a valuation meets an odd strategy cycle and the
    reference exits (cpu_bv.cpp:73-84)."""
    rng = np.random.default_rng(91)
    done_ok, done_abort, k = 0, 0, 0
    while done_ok < 8 or done_abort < 4:
        k += 1
        n = int(rng.integers(4, 60))
        g = psi.gen_random_game(9100 + k, n, int(rng.integers(2, 3 * n)), 1, int(rng.integers(1, 5)))
        if k % 3 == 0:
            g.priority[:] &= ~1          # even priorities only: no cycle is odd
        with tempfile.NamedTemporaryFile("w", suffix=".gm", delete=False) as fh:
            fh.write(g.to_gm())
            path = fh.name
        res = run_ref("cpubv", path, 100000)
        os.unlink(path)
        ok = res["final"] is not None and res["final"]["status"] == "ok"
        if ok and done_ok < 8:
            name, done_ok = f"bv_{done_ok:02d}", done_ok + 1
        elif not ok and res["returncode"] != 0 and done_abort < 4:
            name, done_abort = f"bv_abort_{done_abort:02d}", done_abort + 1
        else:
            continue
        fx = {"name": name, "game": game_dict(g), "modes": {"cpubv": summarize(res, False)}}
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(fx, fh, separators=(",", ":"))
        print(name, g.n, res["final"]["status"] if res["final"] else res["message"])


def make_empty():
    """The game of a file with a header line only, in every mode; written to
    edge/ so that the fixture sweeps over this folder do not take it."""
    fx = {"name": "empty", "gm": "parity -1;\n", "modes": {}}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "empty.gm")
        with open(path, "w") as fh:
            fh.write(fx["gm"])
        for mode in ["cpu", "cpu-reset", "cpubv", "cpulist", "cpubf"]:
            res = run_ref(mode, path, 10)
            fx["modes"][mode] = {"returncode": res["returncode"], "final": res["final"],
                                 "rounds": len(res["rounds"])}
    os.makedirs(os.path.join(HERE, "edge"), exist_ok=True)
    with open(os.path.join(HERE, "edge", "empty.json"), "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    print("empty", {m: (v["final"]["iter"], v["final"]["br"]) for m, v in fx["modes"].items()})


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    if "--bf-only" in sys.argv:
        make_bf()
        return
    if "--bv-only" in sys.argv:
        make_bv()
        return
    if "--empty-only" in sys.argv:
        make_empty()
        return
    all_modes = ["cpu", "cpu-reset", "cpubv", "cpulist", "cpubf"]
    # 1. the reference's own input fixture
    text = open(TEST2).read()
    owner, pri, off, tgt, _ = oracle.parse_gm(text)
    make("test2", psi.Game(owner, pri, off, tgt), all_modes, full=True, max_rounds=8,
         gm_text=text)
    # 2. configuration 0: seed 0, n=1000, priority U[0,100), out-degree U[1,4]
    g0 = psi.gen_random_game(0, 1000, 100, 1, 4)
    make("cfg0_seed0", g0, ["cpu", "cpu-reset", "cpulist"], full=False, max_rounds=400)
    # 3. tiny random games, full per-round arrays
    for seed in range(24):
        n = 3 + seed % 12
        g = psi.gen_random_game(1000 + seed, n, 2 + seed % 7, 1, 1 + seed % 3)
        make(f"tiny_{seed:02d}", g, ["cpu", "cpu-reset", "cpulist"], full=True, max_rounds=200)
    # 4. mid-size random games (per-round hashes)
    for seed, (n, p, dlo, dhi) in enumerate([(200, 8, 1, 3), (300, 50, 2, 5), (500, 500, 1, 6),
                                             (400, 4, 1, 2), (256, 1000, 2, 8)]):
        g = psi.gen_random_game(2000 + seed, n, p, dlo, dhi)
        make(f"mid_{seed:02d}", g, ["cpu", "cpulist"], full=False, max_rounds=150)
    # 5. chain games
    for seed, n in enumerate([12, 40, 150]):
        g = psi.gen_chain_game(3000 + seed, n)
        make(f"chain_{seed:02d}", g, ["cpu", "cpulist"], full=n <= 40, max_rounds=2000)
    # 6. solve_bf (cpubf) fixtures
    make_bf()
    # 7. CPUBV (cpubv) fixtures
    make_bv()
    # 8. the game with no vertices
    make_empty()


if __name__ == "__main__":
    main()
