"""kd_quickstat_index_device_noalloc (the selection with its workspace in a
fixed device array inside the library, for callers that cannot allocate)
against kd_quickstat_index_device (the same selection in a caller-owned
workspace), through the C ABI: identical index arrays, bit for bit, on two
calls in a row, the second with a smaller input; and the size limit."""
import ctypes as C

import numpy as np
import pytest

from multilanguagekdtree_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_LEN = 4_000_000  # KD_NOALLOC_SELECT_MAX_LEN


def select_with_workspace(idx, vals, goal):
    L = _lib.lib()
    n = idx.numel()
    wsb = int(L.kd_quickstat_workspace_bytes(n))
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    out = idx.clone()
    torch.cuda.synchronize()
    _lib.check(L.kd_quickstat_index_device(n, C.c_void_p(out.data_ptr()), goal,
                                           C.c_void_p(vals.data_ptr()), vals.numel(),
                                           C.c_void_p(ws.data_ptr()), wsb))
    return out


def select_noalloc(idx, vals, goal):
    L = _lib.lib()
    out = idx.clone()
    torch.cuda.synchronize()
    _lib.check(L.kd_quickstat_index_device_noalloc(idx.numel(), C.c_void_p(out.data_ptr()),
                                                   goal, C.c_void_p(vals.data_ptr()),
                                                   vals.numel()))
    return out


def case(n, nvals, seed, ties):
    rng = np.random.default_rng(seed)
    vals = rng.integers(0, 1000, nvals).astype(np.float64) if ties else rng.random(nvals)
    idx = rng.choice(nvals, n, replace=False).astype(np.int64)
    goal = int(rng.integers(0, n))
    return (torch.as_tensor(idx, device="cuda"), torch.as_tensor(vals, device="cuda"), goal)


def test_noalloc_equals_workspace_selection_twice():
    # first call near the limit, the second smaller (the array keeps the
    # first call's contents), the third with ties (the index passes run)
    for n, nvals, seed, ties in ((MAX_LEN, MAX_LEN + 500, 1, False),
                                 (100_000, 150_000, 2, False),
                                 (250_000, 300_000, 3, True)):
        idx, vals, goal = case(n, nvals, seed, ties)
        ref = select_with_workspace(idx, vals, goal)
        got = select_noalloc(idx, vals, goal)
        assert torch.equal(got, ref), (n, ties)
        # the reference postcondition (quickstat.rs:199-215) on the result
        g = got.cpu().numpy()
        v = vals.cpu().numpy()
        key = (v[g[goal]], g[goal])
        assert all((v[a], a) < key for a in g[:goal][:1000])
        assert all((v[a], a) > key for a in g[goal + 1:][:1000])


def test_noalloc_refuses_beyond_the_limit():
    # refused on the host before any device work
    L = _lib.lib()
    idx = torch.zeros(1, dtype=torch.int64, device="cuda")
    vals = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = L.kd_quickstat_index_device_noalloc(MAX_LEN + 1, C.c_void_p(idx.data_ptr()), 0,
                                             C.c_void_p(vals.data_ptr()), 1)
    assert rc == _lib.KD_ERR_CAPACITY
    assert b"KD_NOALLOC_SELECT_MAX_LEN" in L.kd_last_error()
