"""Shared comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np


U = 2.0 ** -53  # unit roundoff of FP64


def node_counts(t):
    """Bodies under every node (leaf: num_parts; internal: left + right)."""
    cnt = np.where(t.is_leaf == 1, t.num_parts, 0).astype(np.int64)
    # children have larger indices than their parent in both reference
    # layouts, so one pass from the back sees every child before its parent
    internal = np.nonzero(t.is_leaf == 0)[0][::-1]
    left, right = t.left, t.right
    for i in internal:
        cnt[i] = cnt[left[i]] + cnt[right[i]]
    return cnt


def compare_trees(gpu, ref, pscale=1.0):
    """Bit-exact comparison of two node vectors in the same layout.

    Exact: vector length, node kinds, leaf sizes, leaf members (canonical
    order) and the particle order, split dimension, split value, size, child
    indices.  Node mass m and centre of mass cm are sums whose rounding
    depends on summation order, which the reference leaves to its random
    pivots; two orders of the same c positive terms differ by at most
    2*(c-1)*u*sum, so that is the bound used (pscale = largest |coordinate|
    bounds |cm|)."""
    nn = len(gpu.is_leaf)
    assert len(ref.is_leaf) == nn, (len(ref.is_leaf), nn)
    np.testing.assert_array_equal(gpu.is_leaf, ref.is_leaf)
    np.testing.assert_array_equal(gpu.num_parts, ref.num_parts)
    np.testing.assert_array_equal(gpu.perm, ref.perm)
    # equal particle order, leaf starts and leaf sizes: equal leaf members
    leaves = gpu.is_leaf == 1
    np.testing.assert_array_equal(gpu.first[leaves & (gpu.num_parts > 0)],
                                  ref.first[leaves & (ref.num_parts > 0)])
    internal = gpu.is_leaf == 0
    for name in ("split_dim", "left", "right"):
        np.testing.assert_array_equal(getattr(gpu, name)[internal],
                                      getattr(ref, name)[internal], err_msg=name)
    for name in ("split_val", "size"):
        a = getattr(gpu, name)[internal]
        b = getattr(ref, name)[internal]
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), name
    c = node_counts(ref)[internal].astype(np.float64)
    gm, rm = gpu.m[internal], ref.m[internal]
    assert np.all(np.abs(gm - rm) <= 2 * c * U * np.abs(rm)), "m"
    gc, rc = gpu.cm[internal], ref.cm[internal]
    # cm = sum(m p)/m: sum error <= 2 c u m pscale, plus the two roundings of
    # the division and of m itself
    tol = (2 * c + 4)[:, None] * U * pscale
    assert np.all(np.abs(gc - rc) <= tol), ("cm", (np.abs(gc - rc) / tol).max())


def accel_errors(a, ref, absref):
    """Plain relative error |da|/|a| and conditioned error |da|/sum_j|a_ij|
    (norms over the three components)."""
    da = np.linalg.norm(a - ref, axis=1)
    plain = da / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    cond = da / np.maximum(np.linalg.norm(absref, axis=1), 1e-300)
    return plain, cond


def pscale(b):
    return float(max(np.abs(b["x"]).max(), np.abs(b["y"]).max(), np.abs(b["z"]).max()))
