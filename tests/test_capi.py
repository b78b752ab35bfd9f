"""The C-ABI library loads and exports every function include/psi.h declares
(no device calls: this runs on machines without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "psi.h")
LIB = os.path.join(ROOT, "parallel-si_amd", "libpsi.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psi_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    assert len(declared()) > 10


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libpsi.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_library_loads_with_ctypes():
    import psi
    for s in declared():
        assert hasattr(psi.lib, s)


def test_generators_are_deterministic():
    import numpy as np
    import psi
    a = psi.gen_random_game(7, 500, 100, 1, 4)
    b = psi.gen_random_game(7, 500, 100, 1, 4)
    assert np.array_equal(a.targets, b.targets) and np.array_equal(a.priority, b.priority)
    c = psi.gen_random_game(8, 500, 100, 1, 4)
    assert not np.array_equal(a.priority, c.priority)
    # out-degree in range, successors distinct, owners 0/1
    deg = np.diff(a.offsets)
    assert deg.min() >= 1 and deg.max() <= 4
    for v in range(a.n):
        s = a.targets[a.offsets[v]:a.offsets[v + 1]]
        assert len(set(s.tolist())) == len(s)
    assert set(a.owner.tolist()) <= {0, 1}


def test_list_and_forest_generators():
    import numpy as np
    import oracle
    import psi
    succ, head = psi.gen_list(1, 5000)
    r = oracle.list_rank(succ, head)
    assert sorted(r.tolist()) == list(range(5000))
    succ, heads, lens = psi.gen_forest(2, 100000, 4096, 4096)
    assert lens.sum() == 100000 and lens.min() >= 1
    r = oracle.list_rank(succ, heads)
    assert (r >= 0).all()
    assert np.bincount(r).max() == 4096


def test_size_lists_are_log_uniform_bounds():
    import psi
    s = psi.gen_sizes(11, 1024, 1000, 1000000)
    assert s.min() >= 1000 and s.max() <= 1000000


def test_autotune_descent_keeps_only_clear_wins():
    """Coordinate descent of the autotuner: a knob value is kept only if it
    is at least 2% faster than the best so far."""
    import autotune
    cost = {("a", 1): 1.0, ("a", 2): 0.99, ("a", 3): 0.5, ("b", 1): 1.0, ("b", 2): 1.2}

    def measure(cfg):
        return cost[("a", cfg["a"])] * cost[("b", cfg["b"])]
    best, t, trials = autotune._descend({"a": 1, "b": 1}, {"a": [1, 2, 3], "b": [1, 2]}, measure)
    assert best == {"a": 3, "b": 1} and t == 0.5 and len(trials) == 4


def test_si_tune_rejects_bad_values():
    """Launch-shape settings are host state (no device call)."""
    import psi
    with pytest.raises(psi.PsiError):
        psi.si_tune(psi.SIT_PROGRAM, 2)
    with pytest.raises(psi.PsiError):
        psi.si_tune(psi.SIT_PHASE_GRID, 3)
    with pytest.raises(psi.PsiError):
        psi.si_tune(99, 2048)
    with pytest.raises(psi.PsiError):
        psi.si_tune(psi.SIT_TWO_LEVELS, 2)
    with pytest.raises(psi.PsiError):
        psi.si_tune(psi.SIT_TWO_LEVELS, -1)
    psi.si_tune(psi.SIT_PROGRAM, 0)
    psi.si_tune(psi.SIT_PROGRAM, 1)
    psi.si_tune(psi.SIT_PHASE_GRID, 2048)
    psi.si_tune(psi.SIT_TWO_LEVELS, 0)
    psi.si_tune(psi.SIT_TWO_LEVELS, 1)
