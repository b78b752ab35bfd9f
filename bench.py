#!/usr/bin/env python3
"""Benchmark of the MI355X strategy-improvement solver (one GPU).

Headline: SI vertex-iterations/s on the 4,194,304-vertex random game
(owner Bernoulli(1/2), priority U[0, 2^20), out-degree U[2, 8], seed 2):
value = n * br / t_solve summed over the K timed solves, where br counts
valuation rounds (gpumain.cpp:34) and t_solve is one full solve_si
(init_strat to convergence) with the game already resident in HBM.

One JSON line on stdout.  Besides the contract keys it carries
  roofline     - the dominant kernel of the solve, k_phase (every launch of
                 the step program): the solve's algorithmic bytes over the
                 HIP-event time of its launches; the pointer-jumping round
                 k_jump beside it (jump_round);
  cpu_baseline - the oracle's sparse CPUParity restatement (1 core) on the
                 first rounds of the same game (a bounded sample);
  listrank     - 16,777,216-node list and 1,048,576-list forest rankings;
  batch        - 1,024 mixed-size games in one handle vs one at a time;
  serve        - Poisson arrivals into the serving pool: p50/p99 latency and
                 throughput;
  serve_loaded - the same requests arriving 8x as fast (several games on the
                 GPU at once);
  bf           - solve_bf (the cpu_bf role) at 1M vertices, and a 16K-vertex
                 sample timed on the GPU and with the oracle on one core.
Progress goes to stderr.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-si_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


LARGE_FIXTURES = os.path.join(ROOT, "tests", "golden", "large")


def load_fixture(name):
    """Oracle fixture data (expected results written on the CPU by
    tests/golden/make_large.py); None if absent."""
    path = os.path.join(LARGE_FIXTURES, name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def si_parity(psi, fx, s, n, seed):
    """The timed solve's result against the oracle's fixture of this game."""
    if fx is None or fx["n"] != n or fx["gen"]["seed"] != seed:
        return None
    it, br, capped = s.stats()
    return bool((it, br, capped) == (fx["iter"], fx["br"], False) and
                str(psi.hash_i32(s.strategy())) == fx["strat"] and
                str(psi.hash_i32(s.labels())) == fx["labels"] and
                str(psi.hash_i32(s.solved())) == fx["solved"])


def si_bench(psi, n, steps, warmup, seed):
    t0 = time.time()
    g = psi.gen_random_game(seed, n, 1 << 20, 2, 8)
    log(f"[bench] generated n={g.n} m={g.m} in {time.time() - t0:.1f}s")
    s = psi.Solver(g)
    for _ in range(warmup):
        s.solve()
    s.sync()
    timer = psi.Timer(s)            # events on the handle's stream, where its kernels run
    brs = []
    t_wall = time.perf_counter()
    timer.start()
    for _ in range(steps):
        s.solve()
        brs.append(s.stats()[1])
    timer.stop()
    s.sync()
    wall = time.perf_counter() - t_wall
    dev_ms = timer.ms()
    it, br, capped = s.stats()
    log(f"[bench] {steps} solves: iter={it} br={br} wall={wall:.3f}s device={dev_ms:.1f}ms")
    parity = si_parity(psi, load_fixture("cfg2_4M.json"), s, g.n, seed)
    log(f"[bench] timed solve vs oracle fixture: {parity}")
    # Attribution: the loop launched kernel by kernel (k_jump etc., the
    # roofline's source), then the step program launch by launch (the phases
    # the timed solves ran).
    prof = psi.solve_profiled(s)
    log("[bench] per-kernel ms/launches: " + ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in prof.items()))
    prounds = psi.profile_rounds(s)
    pprof = psi.solve_program_profiled(s)
    log("[bench] step-program phases ms/launches: " +
        ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in pprof.items() if v[1]))
    # The phase of every launch in order: the timed graph solves run the same
    # launch sequence (then idle launches to the end of their last graph), so
    # this attributes per-dispatch profiler records to phases.
    launches = psi.program_launches(s)
    log("[bench] step-program launch phases: " + " ".join(p for p, _, _ in launches))
    # levels each valuation of the step program reached (its jump launches
    # advance two levels each; the cycle-label check stops cycle-bound
    # valuations early, which the kernel-by-kernel loop above does not)
    levels = []
    for ph, r, _ in launches:
        if ph == "init":
            levels.append(0)
        elif ph == "jump" and levels:
            levels[-1] = max(levels[-1], r + 1)
    return g, s, it, br, wall, dev_ms, parity, prof, pprof, (levels, prounds)


def listrank_bench(psi, reps=10, tune=True, sweep=True):
    out = {}
    n = 1 << 24
    if tune:
        import autotune
        t0 = time.perf_counter()
        cfg, sec, trials = autotune.tune_list_rank(n)
        out["autotuned"] = {"config": cfg, "trials": len(trials),
                            "tuning_s": round(time.perf_counter() - t0, 2)}
        log(f"[bench] list ranking autotuned: {cfg} ({len(trials)} trials)")
    succ, heads = psi.gen_list(1, n)
    cases = [("list_16M", succ, heads, psi.LR_HJ)]
    fs, fh, _ = psi.gen_forest(2, n, 1 << 20, 4096)
    cases.append(("forest_1M_lists", fs, fh, psi.LR_AUTO))
    for name, sc, hd, method in cases:
        ds = psi.DeviceArray.from_host(sc)
        dh = psi.DeviceArray.from_host(hd)
        dr = psi.DeviceArray(4 * n)
        for _ in range(3):
            psi.list_rank_device(ds, n, dh, hd.size, dr, method)
        psi.synchronize()
        t = psi.Timer()
        t.start()
        for _ in range(reps):
            psi.list_rank_device(ds, n, dh, hd.size, dr, method)
        t.stop()
        ms = t.ms() / reps
        gbs = 8.0 * n / (ms * 1e-3) / 1e9
        out[name] = {"nodes": n, "lists": int(hd.size), "us": round(ms * 1e3, 2),
                     "nodes_per_s": round(n / (ms * 1e-3), 1), "alg_GBps": round(gbs, 1),
                     "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"[bench] {name}: {ms * 1e3:.1f} us, {gbs:.0f} GB/s algorithmic")
        for d in (ds, dh, dr):
            d.free()
    if sweep:
        out["by_length"] = listrank_sweep(psi)
    return out


def listrank_sweep(psi, reps=20):
    """One list of n nodes by each method (device time per ranking, us): the
    measurements behind AUTO's choice by list length."""
    res = {}
    for n in (1024, 4096, 8192, 16384, 65536, 262144, 1 << 20):
        succ, heads = psi.gen_list(7, n)
        ds, dh = psi.DeviceArray.from_host(succ), psi.DeviceArray.from_host(heads)
        dr = psi.DeviceArray(4 * n)
        row = {}
        for name, method in (("HJ", psi.LR_HJ), ("WYLLIE", psi.LR_WYLLIE), ("AUTO", psi.LR_AUTO)):
            for _ in range(3):
                psi.list_rank_device(ds, n, dh, 1, dr, method)
            psi.synchronize()
            t = psi.Timer()
            t.start()
            for _ in range(reps):
                psi.list_rank_device(ds, n, dh, 1, dr, method)
            t.stop()
            row[name] = round(t.ms() / reps * 1e3, 2)
        res[str(n)] = row
        for d in (ds, dh, dr):
            d.free()
    log("[bench] list ranking by length (us): " + json.dumps(res))
    return res


def tour_bench(psi, g, reps=5):
    """Euler-tour path sums (the GPUList valuation engine: tour build, HJ
    ranking of the 2n+2-node tour, scan) over the final strategy forest of the
    4M game, weights = splitmix64(priority) (the valuation digests)."""
    s = psi.Solver(g)
    s.solve()
    st, solved = s.strategy(), s.solved()
    s.close()
    link = np.where(solved == 1, -2, st).astype(np.int32)
    x = g.priority.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    w = x ^ (x >> np.uint64(31))
    n = g.n
    d_link, d_w = psi.DeviceArray.from_host(link), psi.DeviceArray.from_host(w)
    d_sum, d_reach = psi.DeviceArray(8 * n), psi.DeviceArray(4 * n)
    run = lambda: psi.check(psi.lib.psi_tour_path_sums(d_link.ptr, n, d_w.ptr, d_sum.ptr,
                                                        d_reach.ptr, psi.LR_AUTO), "tour")
    run()
    psi.synchronize()
    timer = psi.Timer()
    timer.start()
    for _ in range(reps):
        run()
    timer.stop()
    psi.synchronize()
    ms = timer.ms() / reps
    reach = int(d_reach.to_host(np.int32, n).sum())
    # algorithmic bytes: links 4 + weights 8 + sums 8 per vertex, tour succ 8 + rank 8 per vertex
    # (2 nodes), scan in/out 32 per vertex
    alg = 68.0 * n
    return {"n": n, "tour_nodes": 2 * n + 2, "reached": reach, "us": round(ms * 1e3, 1),
            "alg_GBps": round(alg / (ms * 1e-3) / 1e9, 1),
            "roofline_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def gen_games_parallel(psi, specs, threads=16):
    """specs: list of (kind, seed, n, pri_bound, deg_lo, deg_hi); generated on a
    thread pool (the generator releases the GIL inside the C call)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(sp):
        kind, seed, n, pb, dlo, dhi = sp
        if kind == "chain":
            return psi.gen_chain_game(seed, n)
        return psi.gen_random_game(seed, n, pb, dlo, dhi)

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, specs))


def batch_bench(psi, n_games=1024, seed=3, lo=1000, hi=1000000, check_alone=True):
    """Configuration 3: n_games random games, sizes log-uniform in [lo, hi],
    priorities U[0, n), out-degree U[1, 6], per-game seed = seed * 100000 + g.
    games/s = games / wall time including CSR upload and result download,
    excluding generation.  One-at-a-time = the plain solver per game in turn."""
    sizes = psi.gen_sizes(seed, n_games, lo, hi)
    t0 = time.time()
    games = gen_games_parallel(psi, [("random", seed * 100000 + g, int(n), int(n), 1, 6)
                                     for g, n in enumerate(sizes)])
    log(f"[bench] batch: generated {n_games} games, sum n={int(sizes.sum())} in "
        f"{time.time() - t0:.1f}s")
    # launch shapes and steps per sync from the autotuner, on every 16th game
    # (outside the timed regions; both forms below run with the result)
    import autotune
    t0 = time.perf_counter()
    tuned, _, trials = autotune.tune_si(games[::16])
    tuning_s = time.perf_counter() - t0
    steps = tuned["steps_per_sync"]
    log(f"[bench] batch: autotuned {tuned} ({len(trials)} trials, {tuning_s:.1f}s)")
    # batched: one handle, every game with its own convergence
    t = time.perf_counter()
    b = psi.Solver(games)
    t1 = time.perf_counter()
    b.solve(steps_per_sync=steps)
    t2 = time.perf_counter()
    res_b = b.results()
    t3 = time.perf_counter()
    b.close()
    t_batch = time.perf_counter() - t
    if os.environ.get("PSI_BATCH_PROFILE"):
        b2 = psi.Solver(games)
        prof = psi.solve_profiled(b2)
        log("[bench] batch per-kernel ms/launches: " +
            ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in prof.items()))
        b2.close()
    log(f"[bench] batch: batched {t_batch:.2f}s (create+upload {t1 - t:.2f}, solve {t2 - t1:.2f}, "
        f"download {t3 - t2:.2f}, free {t_batch - (t3 - t):.2f})")
    # one at a time
    t = time.perf_counter()
    res_a = []
    for g in games:
        s = psi.Solver(g)
        s.solve(steps_per_sync=steps)
        res_a.append((s.stats(), s.strategy(), s.labels()))
        s.close()
    t_alone = time.perf_counter() - t
    log(f"[bench] batch: one-at-a-time {t_alone:.2f}s")
    autotune.apply_si(autotune.SI_DEFAULT)        # later legs run with the defaults
    same = all(x[0] == y[0] and np.array_equal(x[1], y[1]) and np.array_equal(x[2], y[2])
               for x, y in zip(res_b, res_a))
    # every game of the batch against the oracle's fixture of this set
    fx = load_fixture("cfg3_batch.json")
    vs_oracle, differing = None, None
    if fx is not None and fx["gen"]["games"] == n_games and fx["gen"]["size_seed"] == seed:
        exp = {r["k"]: r for r in fx["games"]}
        differing = [k for k, (st_, s_, l_) in enumerate(res_b)
                     if st_ != (exp[k]["iter"], exp[k]["br"], False) or
                     str(psi.hash_i32(s_)) != exp[k]["strat"] or
                     str(psi.hash_i32(l_)) != exp[k]["labels"]]
        vs_oracle = not differing
        log(f"[bench] batch vs oracle fixture: {len(res_b) - len(differing)}/{len(res_b)} equal")
    n_tot = int(sizes.sum())
    br = [x[0][1] for x in res_b]
    return {"games": n_games, "sum_n": n_tot, "sum_m": int(sum(g.m for g in games)),
            "batched_s": round(t_batch, 3), "one_at_a_time_s": round(t_alone, 3),
            "games_per_s_batched": round(n_games / t_batch, 1),
            "games_per_s_one_at_a_time": round(n_games / t_alone, 1),
            "speedup": round(t_alone / t_batch, 2), "br_max": int(max(br)),
            "vertex_iterations": int(sum(int(sz) * b_ for sz, b_ in zip(sizes, br))),
            "batched_equals_alone": bool(same),
            "batched_equals_oracle": vs_oracle,
            "games_differing_from_oracle": differing,
            "autotuned": {"config": tuned, "trials": len(trials), "tuning_s": round(tuning_s, 2),
                          "sample": "every 16th game of the batch"}}


STREAM_REQUESTS = 4096


def serve_bench(psi, requests=96, rate=2.0, seed=4, budget_gib=64.0, steps=4, lookahead=64,
                first=0, dump=None):
    """Serving: requests [first, first + requests) of a seeded 4,096-request
    Poisson stream of `rate` per second.  Request k: with probability 1/5 a
    chain game, else a random game of the 4M-game generator (priority
    U[0, 2^20), out-degree U[2, 8]); n log-uniform in [10^4, 4,194,304]; game
    seed = seed * 100000 + k.  The whole stream's sizes, kinds and arrival
    gaps are drawn first, so a slice is the same requests at the same
    spacing whatever its length (its clock starts at its first arrival).
    Games are generated ahead on host threads (outside the timing; a request
    whose game is not ready at its arrival time waits, and the wait counts in
    its latency).  Latency = harvest time - arrival time.  A sample of
    requests is solved again alone afterwards and compared bit for bit.
    dump: a path that receives every request's (k, chain, n, br, arrival,
    latency) as JSON, so that slices run apart can be pooled."""
    from concurrent.futures import ThreadPoolExecutor

    total = max(STREAM_REQUESTS, first + requests)
    sizes_all = psi.gen_sizes(seed, total, 1e4, 4194304)
    rng = np.random.default_rng(seed)
    chain_all = rng.random(total) < 0.2
    arrival_all = np.cumsum(rng.exponential(1.0 / rate, total))
    ks = np.arange(first, first + requests)
    sizes, chain = sizes_all[ks], chain_all[ks]
    arrival = arrival_all[ks] - arrival_all[first]

    def make(i):
        n, k = int(sizes[i]), first + i
        if chain[i]:
            return psi.gen_chain_game(seed * 100000 + k, n)
        return psi.gen_random_game(seed * 100000 + k, n, 1 << 20, 2, 8)

    ex = ThreadPoolExecutor(16)
    futs = [None] * requests
    made = 0

    def top_up(upto):
        nonlocal made
        while made < min(requests, upto):
            futs[made] = ex.submit(make, made)
            made += 1

    top_up(lookahead)
    pool = psi.Pool(int(budget_gib * (1 << 30)), edges_per_vertex=5.0, max_games=256,
                    max_n=4194304)
    # warm-up: graph capture and first launches on a tiny game
    pool.submit(psi.gen_random_game(1, 1000, 100, 1, 4))
    pool.drain(steps)
    for f in futs[:min(lookahead, requests)]:
        f.result()
    # requests with an oracle fixture (the stream's first requests, large
    # chain games excepted) are checked against it; a few others are solved
    # again alone and compared
    fx = load_fixture("cfg4_stream.json")
    expect = {}
    if fx is not None and fx["gen"]["seed"] == seed and fx["gen"]["requests"] == STREAM_REQUESTS:
        expect = {r["k"] - first: r for r in fx["requests"] if first <= r["k"] < first + requests}
    order = np.argsort(sizes)
    sample = set([int(k) for k in order if chain[k] and k not in expect][:2] +
                 [int(k) for k in order[::-1] if not chain[k] and k not in expect][:1])
    vs_oracle = {"checked": 0, "equal": 0}
    kept, n_of, br_of, finish, ticket_of = {}, {}, {}, {}, {}
    sub, late, max_running, max_queued, n_steps = 0, 0, 0, 0, 0
    t0 = time.perf_counter()
    t_log = 0.0
    while len(finish) < requests:
        now = time.perf_counter() - t0
        while sub < requests and arrival[sub] <= now:
            if not futs[sub].done():
                late += 1
                futs[sub].result()
            g = futs[sub].result()
            futs[sub] = None
            ticket_of[pool.submit(g)] = sub
            n_of[sub] = g.n
            if sub in sample:
                kept[sub] = g
            sub += 1
            top_up(sub + lookahead)
        done = pool.step(steps)
        n_steps += 1
        tn = time.perf_counter() - t0
        for t in done:
            k = ticket_of.pop(t)
            finish[k] = tn
            stats, st, lb = pool.result(t)
            br_of[k] = stats[1]
            if k in kept:
                kept[k] = (kept[k], stats, st, lb)
            if k in expect:
                e = expect[k]
                vs_oracle["checked"] += 1
                vs_oracle["equal"] += int(stats == (e["iter"], e["br"], False) and
                                          str(psi.hash_i32(st)) == e["strat"] and
                                          str(psi.hash_i32(lb)) == e["labels"])
        info = pool.info()
        if tn - t_log > 10.0:
            t_log = tn
            log(f"[bench] serve: t={tn:.0f}s submitted {sub} finished {len(finish)} "
                f"running {info['running']} queued {info['queued']}")
        max_running = max(max_running, info["running"])
        max_queued = max(max_queued, info["queued"])
        if not done and info["running"] == 0 and info["queued"] == 0 and sub < requests:
            time.sleep(max(0.0, arrival[sub] - (time.perf_counter() - t0)))
    makespan = max(finish.values())
    pool.close()
    ex.shutdown()
    lat = np.array([finish[k] - arrival[k] for k in range(requests)])
    same = True
    for k, (g, stats, st, lb) in kept.items():
        s = psi.Solver(g)
        s.solve()
        same &= stats == s.stats() and np.array_equal(st, s.strategy()) and \
            np.array_equal(lb, s.labels())
        s.close()
    vit = sum(n_of[k] * br_of[k] for k in range(requests))
    lat_c = lat[chain]
    lat_r = lat[~chain]
    if dump:
        with open(dump, "w") as fh:
            json.dump({"seed": seed, "rate_per_s": rate, "first": first, "requests": requests,
                       "makespan_s": makespan, "vertex_iterations": int(vit),
                       "served_equals_alone": bool(same), "checked_alone": len(kept),
                       "checked_vs_oracle": vs_oracle["checked"],
                       "equal_vs_oracle": vs_oracle["equal"],
                       "rows": [[int(first + i), bool(chain[i]), int(n_of[i]), int(br_of[i]),
                                 float(arrival[i]), float(lat[i])] for i in range(requests)]}, fh)
    return {"requests": requests, "requests_run": requests, "requests_in_stream": STREAM_REQUESTS,
            "reduced": requests < STREAM_REQUESTS,
            "prefix_note": (f"requests {first}..{first + requests - 1} of the seeded "
                            f"4,096-request stream"),
            "rate_per_s": rate, "seed": seed, "chain_requests": int(chain.sum()),
            "sum_n": int(sizes.sum()), "budget_gib": budget_gib, "steps_per_sync": steps,
            "makespan_s": round(makespan, 3),
            "throughput_req_per_s": round(requests / makespan, 3),
            "vertex_iterations_per_s": round(vit / makespan, 1),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "latency_mean_s": round(float(lat.mean()), 4),
            "latency_p50_random_s": round(float(np.percentile(lat_r, 50)), 4) if len(lat_r) else None,
            "latency_p99_random_s": round(float(np.percentile(lat_r, 99)), 4) if len(lat_r) else None,
            "latency_p50_chain_s": round(float(np.percentile(lat_c, 50)), 4) if len(lat_c) else None,
            "latency_p99_chain_s": round(float(np.percentile(lat_c, 99)), 4) if len(lat_c) else None,
            "br_max": int(max(br_of.values())), "max_running": max_running,
            "max_queued": max_queued, "late_generation": late, "pool_steps": n_steps,
            "served_equals_alone": bool(same), "checked_alone": len(kept),
            "served_equals_oracle": (vs_oracle["equal"] == vs_oracle["checked"]
                                     if vs_oracle["checked"] else None),
            "checked_vs_oracle": vs_oracle["checked"]}


def serve_merge(paths):
    """Pools the per-request dumps of stream slices served in separate runs
    (serve_bench(dump=...)): latency percentiles over every request, and the
    sums of the slices' makespans and vertex-iterations."""
    parts = []
    for p in paths:
        with open(p) as fh:
            parts.append(json.load(fh))
    rows = [r for d in parts for r in d["rows"]]
    ks = sorted(r[0] for r in rows)
    lat = np.array([r[5] for r in rows])
    ch = np.array([r[1] for r in rows])
    span = sum(d["makespan_s"] for d in parts)
    vit = sum(d["vertex_iterations"] for d in parts)

    def pct(x, q):
        return round(float(np.percentile(x, q)), 4) if len(x) else None

    return {"requests": len(rows), "distinct": len(set(ks)), "first": ks[0], "last": ks[-1],
            "slices": [[d["first"], d["requests"], round(d["makespan_s"], 1)] for d in parts],
            "rate_per_s": sorted(set(d["rate_per_s"] for d in parts)),
            "chain_requests": int(ch.sum()), "sum_n": int(sum(r[2] for r in rows)),
            "vertex_iterations": int(vit), "br_max": int(max(r[3] for r in rows)),
            "summed_makespan_s": round(span, 1),
            "throughput_req_per_s": round(len(rows) / span, 3),
            "vertex_iterations_per_s": round(vit / span, 1),
            "latency_p50_s": pct(lat, 50), "latency_p99_s": pct(lat, 99),
            "latency_mean_s": round(float(lat.mean()), 4),
            "latency_p50_random_s": pct(lat[~ch], 50), "latency_p99_random_s": pct(lat[~ch], 99),
            "latency_p50_chain_s": pct(lat[ch], 50), "latency_p99_chain_s": pct(lat[ch], 99),
            "served_equals_alone": all(d["served_equals_alone"] for d in parts),
            "checked_alone": sum(d["checked_alone"] for d in parts),
            "checked_vs_oracle": sum(d["checked_vs_oracle"] for d in parts),
            "equal_vs_oracle": sum(d["equal_vs_oracle"] for d in parts)}


def bf_game(psi, n, seed):
    """Random game of the 4M generator shape (out-degree U[2, 8]) with 64
    priorities, Odd vertices' priorities made even: no strategy cycle of Odd
    vertices alone is odd, so solve_bf never stops on one (the reference
    exits there)."""
    g = psi.gen_random_game(seed, n, 64, 2, 8)
    g.priority[g.owner == 1] &= ~1
    return g


def bf_bench(psi, n=1 << 20, seed=6, reps=3, cpu_n=1 << 14, cpu=True):
    """solve_bf (the cpu_bf role) on the device; the same solve of a smaller
    game of the same generator timed on the GPU and with the oracle's
    CPU_BellmanFord restatement (one core) for a like-for-like ratio."""
    g = bf_game(psi, n, seed)
    s = psi.Solver(g)
    s.solve_bf()
    s.sync()
    t = time.perf_counter()
    for _ in range(reps):
        s.solve_bf()
    s.sync()
    dt = (time.perf_counter() - t) / reps
    it, br, _ = s.stats()
    s.close()
    out = {"workload": f"solve_bf, n={n}, priority U[0,64) with Odd's made even, out-degree U[2,8], "
                       f"seed {seed}", "n": n, "m": g.m, "iter": it, "br_device_rounds": br,
           "ms": round(dt * 1e3, 3), "vertex_iterations_per_s": round(n * br / dt, 1)}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        gs = bf_game(psi, cpu_n, seed)
        s = psi.Solver(gs)
        s.solve_bf()
        s.sync()
        t = time.perf_counter()
        s.solve_bf()
        s.sync()
        gpu_s = time.perf_counter() - t
        t = time.perf_counter()
        r = oracle.solve_bf(gs)
        cpu_s = time.perf_counter() - t
        same = r.iter == s.stats()[0] and bool(np.array_equal(r.inf, s.labels()))
        s.close()
        out["sample"] = {"n": cpu_n, "gpu_ms": round(gpu_s * 1e3, 3), "cpu_ms": round(cpu_s * 1e3, 1),
                         "cpu_kind": "port (oracle CPU_BellmanFord, dense rows, one core)",
                         "iter": r.iter, "cpu_relaxation_sweeps": r.br, "gpu_equals_cpu": same}
    return out


def measured_traffic(n, seed, kernel):
    """HBM bytes per launch of `kernel` on this workload (k_phase: the
    average over one timed solve's launches; k_jump: a running round), from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated);
    None for other workloads or if the summary is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "si_solve_4M_hbm_bytes.md")
    if n != 1 << 22 or seed != 2 or not os.path.exists(path):
        return None
    m = re.search(kernel + r" HBM traffic per launch: (\d+) bytes", open(path).read())
    return int(m.group(1)) if m else None


def cpu_baseline(g, max_rounds):
    """CPUParity (the reference's cpu mode, single-threaded) restated with
    exact sparse valuations, on the first max_rounds valuation rounds of the
    same game, or the whole solve when max_rounds is 0."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    r = oracle.solve(g, rep=oracle.SPARSE, max_rounds=max_rounds)
    what = (f"first {r.br} valuation rounds of the same game" if r.capped else
            f"the whole solve of the same game (iter {r.iter}, br {r.br})")
    return {"value": round(g.n * r.br / r.seconds, 1), "unit": "vertex-iterations/s", "cores": 1,
            "kind": "port", "seconds": round(r.seconds, 2), "rounds": r.br,
            "sample": f"{what} (oracle sparse CPUParity restatement, single thread, "
                      f"{r.seconds:.1f}s, {cpu_model()})"}


def cpu_baseline_threaded(g, max_rounds, threads):
    """CPUList (the reference's cpulist mode: Euler tour + Helman-JaJa
    valuation, every phase on `threads` std::threads) restated with exact
    sparse valuations, on the first max_rounds rounds of the same game.  Its
    labels follow cpulist's rule (a vertex not reaching the sink is +inf), so
    its trajectory differs from CPUParity's after the first round."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    r = oracle.solve_list(g, threads=threads, max_rounds=max_rounds)
    return {"value": round(g.n * r.br / r.seconds, 1), "unit": "vertex-iterations/s",
            "cores": threads, "kind": "port", "seconds": round(r.seconds, 2), "rounds": r.br,
            "sample": f"first {r.br} valuation rounds of the same game (oracle CPUList restatement, "
                      f"{threads} threads, {r.seconds:.1f}s, {cpu_model()})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-rounds", type=int, default=6,
                    help="valuation rounds of the CPU baselines' bounded sample")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the whole CPUParity solve for cpu_baseline (minutes)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPUList baseline (default: min(16, cores))")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-listrank", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-games", type=int, default=1024)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-bf", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="list ranking: default splitter densities, no tuning trials (for profiling)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="list ranking: no sweep over list lengths (for profiling)")
    ap.add_argument("--requests", type=int, default=96, help="serving: requests")
    ap.add_argument("--rate", type=float, default=2.0, help="serving: arrivals per second")
    ap.add_argument("--first", type=int, default=0,
                    help="serving: index of the first request of the stream served")
    ap.add_argument("--serve-dump", default=None,
                    help="serving (--only serve): write per-request latencies to this JSON file")
    ap.add_argument("--load-rate", type=float, default=16.0,
                    help="serving under load: arrivals per second (0: skip)")
    ap.add_argument("--only", choices=["si", "listrank", "batch", "serve", "bf", "copy", "cpu"],
                    default=None,
                    help="run one workload only (for profiling)")
    ap.add_argument("--serve-merge", nargs="+", default=None,
                    help="pool the --serve-dump files of stream slices (no GPU work)")
    a = ap.parse_args()
    if a.serve_merge:
        print(json.dumps({"serve_merged": serve_merge(a.serve_merge)}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(independent games would shard by game across GPUs, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)
    import psi
    psi.init()
    if a.only == "listrank":
        print(json.dumps({"listrank": listrank_bench(psi, tune=not a.no_autotune,
                                                     sweep=not a.no_sweep)}), flush=True)
        return
    if a.only == "copy":
        # calibration for hardware byte counters: one device-to-device copy of 1 GiB
        nbytes = 1 << 30
        src, dst = psi.DeviceArray(nbytes), psi.DeviceArray(nbytes)
        psi.check(psi.lib.psi_memcpy_d2d(dst.ptr, src.ptr, nbytes), "d2d")
        psi.synchronize()
        print(json.dumps({"copy_bytes": nbytes}), flush=True)
        return
    if a.only == "batch":
        print(json.dumps({"batch": batch_bench(psi, a.batch_games)}), flush=True)
        return
    if a.only == "serve":
        print(json.dumps({"serve": serve_bench(psi, a.requests, a.rate, first=a.first,
                                               dump=a.serve_dump)}), flush=True)
        return
    if a.only == "bf":
        print(json.dumps({"bf": bf_bench(psi)}), flush=True)
        return
    if a.only == "si":
        a.no_listrank = a.no_cpu = a.no_batch = a.no_serve = a.no_bf = True
    if a.only == "cpu":            # the SI leg and the CPU baselines beside it
        a.no_listrank = a.no_batch = a.no_serve = a.no_bf = True
    g, s, it, br, wall, dev_ms, parity, prof, pprof, prounds = si_bench(psi, a.n, a.steps,
                                                                        a.warmup, a.seed)
    value = g.n * br * a.steps / wall
    # Roofline of the dominant kernel, k_phase: the timed solves are graphs
    # of k_phase launches only (the step program).  Algorithmic bytes per
    # valuation round (one valuation + one switch pass over all n vertices):
    # tour build 4n + 4(2n+2), tour rank 8(2n+2), key materialisation 12n,
    # switch 4n + 4(n+1) + 4m (owner, offsets, targets) + 4m + 4n (key
    # gathers) + 4n (strategy write): B_round = 56n + 8m + 28.  Per solve:
    # br * B_round over the k_phase launches of one solve, each launch timed
    # live by HIP events attached to its own dispatch on the handle's stream
    # (its kernel start and end, as rocprofv3's kernel trace reads them).
    phase_keys = [k for k in pprof if k not in ("total", "gathers", "ties", "scans", "scan_steps")]
    k_launches = sum(pprof[k][1] for k in phase_keys)
    k_ms = sum(pprof[k][0] for k in phase_keys)
    b_round = 56 * g.n + 8 * g.m + 28
    alg_solve = float(br) * b_round
    achieved = alg_solve / (k_ms * 1e-3) / 1e9 if k_ms else 0.0
    timed_gbs = alg_solve / (wall / a.steps) / 1e9
    # the pointer-jumping round alone (k_jump launched round by round, the
    # kernel-by-kernel solve): 16 B per vertex visited + 8 B per gather
    jump_ms, jump_launches = prof["jump"]
    gathers = prof["gathers"][1]
    jump_bytes = (16.0 * g.n * jump_launches + 8.0 * gathers) / max(jump_launches, 1)
    jump_gbs = jump_bytes / (jump_ms / jump_launches * 1e-3) / 1e9 if jump_launches else 0.0
    roof = {"bound": "hbm", "kernel": "k_phase", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(g.n, a.seed, "k_phase"),
            "alg_bytes_per_launch": round(alg_solve / max(k_launches, 1)),
            "b_round_bytes": b_round, "alg_bytes_per_solve": round(alg_solve),
            "launches_per_solve": k_launches,
            "avg_launch_us": round(k_ms / max(k_launches, 1) * 1e3, 2),
            "kernel_ms_per_solve": round(k_ms, 3),
            "timed_region": {"ms_per_solve": round(wall / a.steps * 1e3, 3),
                             "device_ms_per_solve": round(dev_ms / a.steps, 3),
                             "achieved": round(timed_gbs, 1),
                             "frac": round(timed_gbs / HBM_PEAK_GBS, 4)},
            "jump_round": {"kernel": "k_jump", "achieved": round(jump_gbs, 1),
                           "frac": round(jump_gbs / HBM_PEAK_GBS, 4),
                           "alg_bytes_per_launch": round(jump_bytes),
                           "avg_launch_us": round(jump_ms / max(jump_launches, 1) * 1e3, 2),
                           "launches": jump_launches,
                           "traffic": measured_traffic(g.n, a.seed, "k_jump")},
            "rounds_per_valuation": prounds[0],
            "rounds_per_valuation_kernel_by_kernel": prounds[1]}
    line = {
        "metric": "SI vertex-iterations/s on 4M-vertex game",
        "value": round(value, 1),
        "unit": "vertex-iterations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded random parity game, exact integer solve)",
        "config": {"workload": "plain SI solve, random game n=4194304, priority U[0,2^20), "
                               "out-degree U[2,8], seed 2",
                   "n": g.n, "m": g.m, "iter": it, "br": br, "parallelism": "single GPU"},
        "parity": parity,
        "parity_check": "timed solve's iter, br and hashes of the final strategy, labels and "
                        "solved flags equal the oracle's (tests/golden/large/cfg2_4M.json)",
        "roofline": roof,
        "kernels_ms": {k: {"ms": round(v[0], 3), "launches": v[1]} for k, v in prof.items()
                       if k != "gathers"},
        "program_phases_ms": {k: {"ms": round(v[0], 3), "launches": v[1]} for k, v in pprof.items()
                              if v[1]},
    }
    s.close()
    if not a.no_listrank:
        line["listrank"] = listrank_bench(psi, tune=not a.no_autotune, sweep=not a.no_sweep)
        line["listrank"]["tour_path_sums_4M"] = tour_bench(psi, g)
    if not a.no_batch:
        line["batch"] = batch_bench(psi, a.batch_games)
    if not a.no_bf:
        log("[bench] solve_bf ...")
        line["bf"] = bf_bench(psi)
    if not a.no_serve:
        log("[bench] serving ...")
        line["serve"] = serve_bench(psi, a.requests, a.rate)
        if a.load_rate > 0:
            log(f"[bench] serving under load ({a.load_rate:g} requests/s) ...")
            line["serve_loaded"] = serve_bench(psi, a.requests, a.load_rate)
    if not a.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(g, 0 if a.cpu_full else a.cpu_rounds)
        log(f"[bench] cpu baseline: {line['cpu_baseline']['sample']}")
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline_threaded"] = cpu_baseline_threaded(g, a.cpu_rounds, threads)
        log(f"[bench] threaded cpu baseline: {line['cpu_baseline_threaded']['sample']}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
