"""Autotuner for launch shapes and splitter densities on the GPU at hand.

Batched SI: times solves of a sample of the batch for each candidate
step-program grid, steps per synchronisation (launches per graph: 16 per
step) and whether jump launches advance two levels, one knob at a time
from the defaults (coordinate descent, keeping a
change only if it is at least 2% faster).  List ranking: the same over the level-0 and
reduced-level splitter densities and the level-0 hop cap, on a list of the
target length.  Results are plain dicts; `apply` sets them through
psi_si_tune / psi_list_rank_tune, `save`/`load` keep them as JSON.
"""
import json
import time

import numpy as np

import psi

SI_DEFAULT = {"phase_grid": 2048, "steps_per_sync": 4, "two_levels": 1}
SI_CANDIDATES = {"phase_grid": [1024, 1536, 2048], "steps_per_sync": [2, 4, 8],
                 "two_levels": [0, 1]}
LR_DEFAULT = {"log_d0": 5, "local_bits": 12, "log_dk": 2}
LR_CANDIDATES = {"log_d0": [3, 4, 5, 6], "local_bits": [10, 12, 14], "log_dk": [2, 3, 4]}
_LR_PARAM = {"log_d0": psi.LRT_LOG_D0, "local_bits": psi.LRT_LOCAL_BITS, "log_dk": psi.LRT_LOG_DK}


def apply_si(cfg):
    psi.si_tune(psi.SIT_PHASE_GRID, int(cfg["phase_grid"]))
    psi.si_tune(psi.SIT_TWO_LEVELS, int(cfg.get("two_levels", SI_DEFAULT["two_levels"])))


def apply_lr(cfg):
    for k, p in _LR_PARAM.items():
        psi.list_rank_tune(p, int(cfg[k]))


def _descend(default, candidates, measure):
    best = dict(default)
    base = measure(best)
    trials = [(dict(best), base)]
    for knob, values in candidates.items():
        for v in values:
            if v == best[knob]:
                continue
            cfg = dict(best, **{knob: v})
            t = measure(cfg)
            trials.append((cfg, t))
            if t < 0.98 * base:
                best, base = cfg, t
    return best, base, trials


def tune_si(games, reps=2):
    """games: psi.Game list (a sample of the batch).  Returns (config, seconds
    per batched solve of the sample, trials)."""
    def measure(cfg):
        apply_si(cfg)
        s = psi.Solver(games)                 # a new handle captures graphs with these shapes
        s.solve(steps_per_sync=cfg["steps_per_sync"])
        s.sync()
        t = time.perf_counter()
        for _ in range(reps):
            s.solve(steps_per_sync=cfg["steps_per_sync"])
        s.sync()
        dt = (time.perf_counter() - t) / reps
        s.close()
        return dt
    best, base, trials = _descend(SI_DEFAULT, SI_CANDIDATES, measure)
    apply_si(best)
    return best, base, trials


def tune_list_rank(n, seed=1, reps=3):
    """Splitter densities and hop cap for one list of n nodes (HJ)."""
    succ, heads = psi.gen_list(seed, n)
    d_succ = psi.DeviceArray.from_host(succ)
    d_head = psi.DeviceArray.from_host(heads)
    d_rank = psi.DeviceArray(4 * n)
    timer = psi.Timer()

    def measure(cfg):
        apply_lr(cfg)
        run = lambda: psi.list_rank_device(d_succ, n, d_head, 1, d_rank, psi.LR_HJ)
        run()
        psi.synchronize()
        timer.start()
        for _ in range(reps):
            run()
        timer.stop()
        psi.synchronize()
        return timer.ms() / reps * 1e-3
    best, base, trials = _descend(LR_DEFAULT, LR_CANDIDATES, measure)
    apply_lr(best)
    return best, base, trials


def save(path, si=None, lr=None):
    with open(path, "w") as fh:
        json.dump({"si": si, "list_rank": lr}, fh)


def load(path):
    with open(path) as fh:
        d = json.load(fh)
    if d.get("si"):
        apply_si(d["si"])
    if d.get("list_rank"):
        apply_lr(d["list_rank"])
    return d
