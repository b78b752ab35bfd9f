"""Writes the oracle fixtures of the full-size workloads into tests/golden/large/:

    cfg2_4M.json     the 4,194,304-vertex random game (priority U[0,2^20),
                     out-degree U[2,8], seed 2): iter, br, the per-round hashes
                     (strategy valued, solved flags, labels, valuation digests)
                     and hashes of the final strategy, labels and solved flags;
    cfg3_batch.json  the 1,024-game mixed-size set (sizes log-uniform in
                     [1,000, 1,000,000] from seed 3, game g = seed 300000 + g,
                     priority U[0,n), out-degree U[1,6]): per game n, m, iter,
                     br and hashes of the final strategy, labels, solved flags;
    cfg4_stream.json the serving stream (seed 4, 4,096 requests: sizes
                     log-uniform in [10^4, 4,194,304], 1 in 5 a chain game,
                     request k = game seed 400000 + k): the same per request
                     for every request of the first `--cfg4-prefix` whose
                     oracle solve fits the per-game limit (large chain games
                     take hours on the CPU and are listed as skipped).

The expected values come from the oracle (oracle/oracle.cpp, the CPUParity
restatement that tests/test_oracle.py pins against the compiled reference's
own fixtures), run here on the CPU, one game per process.  Hashes are
oracle.hash_i32 of the int32 arrays.  Usage (build container, ~10-20 CPU
minutes on 8 cores):

    python tests/golden/make_large.py cfg2 cfg3 cfg4
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor, as_completed

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "large")
sys.path.insert(0, os.path.join(ROOT, "parallel-si_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CFG2 = dict(seed=2, n=4194304, pri_bound=1 << 20, deg_lo=2, deg_hi=8)
CFG3 = dict(size_seed=3, games=1024, lo=1000, hi=1000000, deg_lo=1, deg_hi=6)
CFG4 = dict(seed=4, requests=4096, lo=1e4, hi=4194304, chain_frac=0.2)


def cfg3_game(psi, g, sizes):
    n = int(sizes[g])
    return psi.gen_random_game(CFG3["size_seed"] * 100000 + g, n, n, CFG3["deg_lo"], CFG3["deg_hi"])


def cfg4_plan():
    """(sizes, chain flags) of the serving stream, as bench.py draws them."""
    import psi
    sizes = psi.gen_sizes(CFG4["seed"], CFG4["requests"], CFG4["lo"], CFG4["hi"])
    rng = np.random.default_rng(CFG4["seed"])
    chain = rng.random(CFG4["requests"]) < CFG4["chain_frac"]
    return sizes, chain


def cfg4_game(psi, k, sizes, chain):
    n = int(sizes[k])
    if chain[k]:
        return psi.gen_chain_game(CFG4["seed"] * 100000 + k, n)
    return psi.gen_random_game(CFG4["seed"] * 100000 + k, n, 1 << 20, 2, 8)


def summary(oracle, g, r):
    return {"n": g.n, "m": g.m, "iter": r.iter, "br": r.br, "capped": bool(r.capped),
            "strat": str(oracle.hash_i32(r.strat)), "labels": str(oracle.hash_i32(r.inf)),
            "solved": str(oracle.hash_i32(r.solved)), "w0": int((r.inf == 1).sum()),
            "cpu_s": round(r.seconds, 3)}


def job(kind, k):
    import psi
    import oracle
    if kind == "cfg3":
        g = cfg3_game(psi, k, psi.gen_sizes(CFG3["size_seed"], CFG3["games"], CFG3["lo"],
                                             CFG3["hi"]))
    else:
        sizes, chain = cfg4_plan()
        g = cfg4_game(psi, k, sizes, chain)
    r = oracle.solve(g, rep=oracle.SPARSE)
    out = summary(oracle, g, r)
    out["k"] = k
    return out


def make_cfg2():
    import psi
    import oracle
    g = psi.gen_random_game(CFG2["seed"], CFG2["n"], CFG2["pri_bound"], CFG2["deg_lo"],
                            CFG2["deg_hi"])
    t = time.time()
    r = oracle.solve(g, rep=oracle.SPARSE, trace=True)
    fx = {"workload": "random game n=4194304, owner Bernoulli(1/2), priority U[0,2^20), "
                      "out-degree U[2,8], seed 2", "gen": CFG2}
    fx.update(summary(oracle, g, r))
    fx["round_hashes"] = [[str(int(x)) for x in row] for row in r.hashes]
    with open(os.path.join(OUT, "cfg2_4M.json"), "w") as fh:
        json.dump(fx, fh, indent=0)
    print(f"cfg2: iter={r.iter} br={r.br} in {time.time() - t:.1f}s", flush=True)


def run_pool(kind, ks, workers, limit_n_chain=None):
    res = {}
    with ProcessPoolExecutor(workers) as ex:
        futs = {ex.submit(job, kind, k): k for k in ks}
        for i, f in enumerate(as_completed(futs)):
            r = f.result()
            res[r["k"]] = r
            if i % 32 == 0 or r["cpu_s"] > 30:
                print(f"{kind}: {i + 1}/{len(ks)} done (k={r['k']} n={r['n']} br={r['br']} "
                      f"{r['cpu_s']}s)", flush=True)
    return [res[k] for k in sorted(res)]


def make_cfg3(workers):
    import psi
    sizes = psi.gen_sizes(CFG3["size_seed"], CFG3["games"], CFG3["lo"], CFG3["hi"])
    # largest games first so the pool drains evenly
    ks = [int(k) for k in np.argsort(-sizes)]
    games = run_pool("cfg3", ks, workers)
    fx = {"workload": "1,024 random games, n log-uniform in [1000, 1000000] (gen_sizes seed 3), "
                      "game g: seed 300000+g, priority U[0,n), out-degree U[1,6]",
          "gen": CFG3, "games": games}
    with open(os.path.join(OUT, "cfg3_batch.json"), "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    print(f"cfg3: {len(games)} games", flush=True)


def make_cfg4(workers, prefix, chain_max, random_max):
    sizes, chain = cfg4_plan()
    ks, skipped = [], []
    for k in range(prefix):
        lim = chain_max if chain[k] else random_max
        (ks if sizes[k] <= lim else skipped).append(k)
    ks.sort(key=lambda k: -int(sizes[k]) * (40 if chain[k] else 1))
    reqs = run_pool("cfg4", ks, workers)
    for r in reqs:
        r["chain"] = bool(chain[r["k"]])
    fx = {"workload": "serving stream seed 4: 4,096 requests, n log-uniform in [10^4, 4194304], "
                      "1 in 5 chain games (default_rng(4).random < 0.2), request k: game seed "
                      "400000+k (chain: gen_chain_game; random: priority U[0,2^20), out-degree "
                      "U[2,8])",
          "gen": CFG4, "prefix": prefix, "chain_max_n": chain_max, "random_max_n": random_max,
          "skipped": [{"k": k, "n": int(sizes[k]), "chain": bool(chain[k])} for k in skipped],
          "requests": reqs}
    with open(os.path.join(OUT, "cfg4_stream.json"), "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    print(f"cfg4: {len(reqs)} requests, {len(skipped)} skipped", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", nargs="+", choices=["cfg2", "cfg3", "cfg4"])
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--cfg4-prefix", type=int, default=64)
    ap.add_argument("--cfg4-chain-max", type=int, default=400000)
    ap.add_argument("--cfg4-random-max", type=int, default=4194304)
    a = ap.parse_args()
    os.makedirs(OUT, exist_ok=True)
    if "cfg2" in a.what:
        make_cfg2()
    if "cfg3" in a.what:
        make_cfg3(a.workers)
    if "cfg4" in a.what:
        make_cfg4(a.workers, a.cfg4_prefix, a.cfg4_chain_max, a.cfg4_random_max)


if __name__ == "__main__":
    main()
