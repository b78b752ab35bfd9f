"""psi_list_rank_noalloc (fixed, statically allocated workspace) against
psi_list_rank (grown workspace) through the C ABI: bit-identical ranks on two
calls in a row, the second with a smaller input (and unreachable nodes, so
stale contents of the shared array would show), a third back on a large
input; inputs whose workspace exceeds the fixed array are refused with a
clear error and leave the array usable."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _rank(psi, fn, succ, heads, method):
    n = succ.shape[0]
    ds, dh = psi.DeviceArray.from_host(succ), psi.DeviceArray.from_host(heads)
    dr = psi.DeviceArray(4 * n)
    rc = getattr(psi.lib, fn)(ds.ptr, n, dh.ptr, heads.shape[0], dr.ptr, method)
    if rc != 0:
        msg = psi.lib.psi_last_error().decode()
        for d in (ds, dh, dr):
            d.free()
        return rc, msg
    psi.synchronize()
    bits, _ = psi.list_rank_status()
    out = dr.to_host(np.int32, n)
    for d in (ds, dh, dr):
        d.free()
    assert bits == 0
    return 0, out


def test_noalloc_equals_allocating_entry(gpu):
    psi = gpu
    psi.lib.psi_list_rank_noalloc.restype = C.c_int
    psi.lib.psi_list_rank_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                              C.c_int]
    checked = psi.LR_CHECKED
    # 1) the 16,777,216-node list (HJ)
    print("noalloc: 16M list HJ", flush=True)
    succ, heads = psi.gen_list(11, 1 << 24)
    rc, a = _rank(psi, "psi_list_rank_noalloc", succ, heads, psi.LR_HJ)
    assert rc == 0, a
    rc, b = _rank(psi, "psi_list_rank", succ, heads, psi.LR_HJ)
    assert rc == 0 and np.array_equal(a, b)
    # 2) a smaller forest in checked mode, with nodes no head reaches: every
    # position the first call wrote must be rewritten or cleared
    print("noalloc: 1M-node forest, checked, unreachable nodes", flush=True)
    fs, fh, _ = psi.gen_forest(12, 1 << 20, 1 << 14, 256)
    fh_part = fh[::2].copy()                       # half of the lists have no head given
    for method in (psi.LR_AUTO, psi.LR_HJ, psi.LR_WALK, psi.LR_WYLLIE):
        m = method | checked if method != psi.LR_WYLLIE else method
        heads_m = fh_part if method != psi.LR_WYLLIE else fh
        rc, a = _rank(psi, "psi_list_rank_noalloc", fs, heads_m, m)
        assert rc == 0, a
        rc, b = _rank(psi, "psi_list_rank", fs, heads_m, m)
        assert rc == 0 and np.array_equal(a, b), method
        assert np.array_equal(a, oracle.list_rank(fs, heads_m)), method
    # 3) refused: Wyllie on 16M nodes needs 384 MiB of workspace
    rc, msg = _rank(psi, "psi_list_rank_noalloc", succ, heads, psi.LR_WYLLIE)
    assert rc == psi.PSI_ERR_INVALID
    assert "psi_list_rank_noalloc" in msg and "bytes" in msg
    # 4) the array still works afterwards: the 1,048,576-list forest (AUTO)
    print("noalloc: 1M-list forest", flush=True)
    fs, fh, _ = psi.gen_forest(2, 1 << 24, 1 << 20, 4096)
    rc, a = _rank(psi, "psi_list_rank_noalloc", fs, fh, psi.LR_AUTO)
    assert rc == 0, a
    rc, b = _rank(psi, "psi_list_rank", fs, fh, psi.LR_AUTO)
    assert rc == 0 and np.array_equal(a, b)
