"""The oracle fixtures of the full-size workloads (tests/golden/large/) are
consistent with their generators and with the oracle: the recorded sizes
match the seeded generators, round counts match the per-round hash rows, and
the oracle re-run on the smallest games of each set reproduces the recorded
results.  The hash the benchmark and GPU tests use on the host equals the
oracle's."""
import json
import os

import numpy as np
import pytest

import oracle
import psi
from conftest import GOLDEN

LARGE = os.path.join(GOLDEN, "large")


def _load(name):
    path = os.path.join(LARGE, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not written")
    with open(path) as fh:
        return json.load(fh)


def test_host_hash_equals_oracle_hash():
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 17, 4096):
        a = rng.integers(-3, 1 << 30, n).astype(np.int32)
        assert psi.hash_i32(a) == oracle.hash_i32(a)


def test_cfg2_fixture_consistent():
    fx = _load("cfg2_4M.json")
    gen = fx["gen"]
    g = psi.gen_random_game(gen["seed"], gen["n"], gen["pri_bound"], gen["deg_lo"], gen["deg_hi"])
    assert (g.n, g.m) == (fx["n"], fx["m"])
    assert len(fx["round_hashes"]) == fx["br"] and fx["iter"] <= fx["br"]
    assert not fx["capped"]
    # the strategy valued in round 0 is init_strat's (cpu_parity.cpp:276-286)
    init = np.where(g.owner == 0, -1, g.targets[g.offsets[:-1]]).astype(np.int32)
    assert fx["round_hashes"][0][0] == str(oracle.hash_i32(init))


def test_cfg3_fixture_reproduces_on_small_games():
    fx = _load("cfg3_batch.json")
    gen = fx["gen"]
    sizes = psi.gen_sizes(gen["size_seed"], gen["games"], gen["lo"], gen["hi"])
    assert [r["k"] for r in fx["games"]] == list(range(gen["games"]))
    assert all(r["n"] == int(sizes[r["k"]]) for r in fx["games"])
    for k in [int(k) for k in np.argsort(sizes)[:6]]:
        n = int(sizes[k])
        g = psi.gen_random_game(gen["size_seed"] * 100000 + k, n, n, gen["deg_lo"], gen["deg_hi"])
        r = oracle.solve(g)
        e = fx["games"][k]
        assert (g.m, r.iter, r.br) == (e["m"], e["iter"], e["br"])
        assert str(oracle.hash_i32(r.strat)) == e["strat"]
        assert str(oracle.hash_i32(r.inf)) == e["labels"]


def test_cfg4_fixture_reproduces_on_small_requests():
    fx = _load("cfg4_stream.json")
    gen = fx["gen"]
    sizes = psi.gen_sizes(gen["seed"], gen["requests"], gen["lo"], gen["hi"])
    chain = np.random.default_rng(gen["seed"]).random(gen["requests"]) < gen["chain_frac"]
    reqs = sorted(fx["requests"], key=lambda r: r["n"])
    done = {r["k"] for r in fx["requests"]} | {s["k"] for s in fx["skipped"]}
    assert done == set(range(fx["prefix"]))
    picked = [r for r in reqs if r["chain"]][:1] + [r for r in reqs if not r["chain"]][:2]
    for e in picked:
        k = e["k"]
        assert e["chain"] == bool(chain[k]) and e["n"] == int(sizes[k])
        if chain[k]:
            g = psi.gen_chain_game(gen["seed"] * 100000 + k, int(sizes[k]))
        else:
            g = psi.gen_random_game(gen["seed"] * 100000 + k, int(sizes[k]), 1 << 20, 2, 8)
        r = oracle.solve(g)
        assert (r.iter, r.br) == (e["iter"], e["br"])
        assert str(oracle.hash_i32(r.strat)) == e["strat"]
        assert str(oracle.hash_i32(r.inf)) == e["labels"]
