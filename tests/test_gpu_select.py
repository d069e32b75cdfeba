"""Standalone selection (kd_quickstat_index) against the reference's tests
of quickstat_index (Parallel/RustVersion/src/quickstat.rs:191-253) and the
oracle: the element at goal is the rank-goal element under (value, index),
both sides satisfy the reference's postcondition and keep their input order."""
import numpy as np
import pytest

from multilanguagekdtree_amd import _lib
from multilanguagekdtree_amd import kdtree as K


def check_post(ix, goal, vals, before_ix):
    vals = np.asarray(vals, np.float64)
    pv, pi = vals[ix[goal]], ix[goal]
    lo, hi = ix[:goal], ix[goal + 1:]
    assert np.all((vals[lo] < pv) | ((vals[lo] == pv) & (lo < pi)))
    assert np.all((vals[hi] > pv) | ((vals[hi] == pv) & (hi > pi)))
    np.testing.assert_array_equal(np.sort(ix), np.sort(before_ix))
    # stable: each side keeps the input order
    pos = np.empty(before_ix.max() + 1, np.int64)
    pos[before_ix] = np.arange(len(before_ix))
    for side in (lo, hi):
        assert np.all(np.diff(pos[side]) > 0)


def test_host_rejects_bad_arguments():
    with pytest.raises(_lib.KdError):
        K.quickstat_index([0, 1, 5], 1, np.zeros(3))  # index outside vals
    with pytest.raises(_lib.KdError):
        K.quickstat_index([0, 1, 2], 3, np.zeros(3))  # goal outside the slice


@pytest.mark.gpu
def test_reference_small_test():
    # quickstat.rs:191-197
    vals = [2.3, 9.8, 3.1, 1.6, 6.7, 7.8, 8.6]
    ix = K.quickstat_index(np.arange(7), 3, vals)
    assert ix[3] == 4


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_random_against_oracle(orc, seed):
    rng = np.random.default_rng(seed)
    n = 100000
    vals = rng.random(n)
    for _ in range(5):
        goal = int(rng.integers(0, n))
        before = rng.permutation(n).astype(np.int64)
        ix = K.quickstat_index(before, goal, vals)
        ref = orc.quickstat_index(before, goal, vals)
        assert ix[goal] == ref[goal]
        check_post(ix, goal, vals, before)


@pytest.mark.gpu
def test_slices_and_ties():
    # slices of a larger value array (quickstat.rs:221-253), with many ties
    rng = np.random.default_rng(7)
    vals = np.floor(rng.random(50000) * 20.0)
    vals[::7] = -0.0
    vals[::11] = 0.0
    for _ in range(10):
        start = int(rng.integers(0, 10000))
        end = start + int(rng.integers(1, 30000))
        goal = int(rng.integers(0, end - start))
        before = np.arange(start, end, dtype=np.int64)[::-1].copy()
        ix = K.quickstat_index(before, goal, np.where(vals == 0.0, 0.0, vals))
        check_post(ix, goal, np.where(vals == 0.0, 0.0, vals), before)


@pytest.mark.gpu
def test_all_equal_and_single():
    vals = np.full(5000, 3.5)
    before = np.random.default_rng(1).permutation(5000).astype(np.int64)
    ix = K.quickstat_index(before, 1234, vals)
    assert ix[1234] == 1234  # ties go by index
    check_post(ix, 1234, vals, before)
    assert K.quickstat_index([0], 0, [1.0]).tolist() == [0]
