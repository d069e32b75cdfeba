"""The build's deeper device-wide levels as one launch of one workgroup per
segment (segments of at most 65,536 bodies, from level 8; builds from the
previous tree order only): the same splits
and stable partitions as the kernel-per-step levels
(Parallel/RustVersion/src/array_kd_tree.rs:63-130), so trees must equal the
oracle's and the kernel-per-step build's bit for bit.  Cases that reach those
levels: a lattice where equal coordinates straddle nearly every median (the
workgroup orders each run by original index), a uniform cube split in every
dimension, and a thin ring whose third dimension is lazy."""
import numpy as np
import pytest

from helpers import compare_trees, pscale
from multilanguagekdtree_amd import kdtree as K

pytestmark = pytest.mark.gpu


def bodies(p, m):
    n = len(p)
    return dict(x=p[:, 0].copy(), y=p[:, 1].copy(), z=p[:, 2].copy(), vx=np.zeros(n),
                vy=np.zeros(n), vz=np.zeros(n), m=m)


def lattice(n, seed):
    rng = np.random.default_rng(seed)
    return bodies(rng.integers(0, 64, (n, 3)).astype(np.float64), np.full(n, 1.0 / n))


def cube(n, seed):
    rng = np.random.default_rng(seed)
    return bodies(rng.uniform(-1.0, 1.0, (n, 3)), np.full(n, 1.0 / n))


def trees(b, mid, builds=2):
    """The tree of the last of `builds` builds of the same positions: the
    first builds from the upload's order (always level by level), the later
    ones from tree order, where the workgroup levels run."""
    n = len(b["x"])
    with K.Simulation(n) as s:
        s.set_tuning("mid_levels", mid)
        assert s.get_tuning("mid_levels") == mid
        s.upload(b)
        for _ in range(builds):
            s.build()
        assert s.get_tuning("mid_builds") == (builds - 1 if mid else 0)
        return s.export_tree(K.KD_LAYOUT_COMPACT)


@pytest.mark.parametrize("name,make,n", [("lattice", lattice, 2_500_000),
                                         ("cube", cube, 3_000_000)])
def test_mid_levels_match_oracle(orc, name, make, n):
    b = make(n, seed=97)
    g = trees(b, 1)
    r = orc.build_tree(b)
    compare_trees(g, r, pscale(b))
    compare_trees(trees(b, 0), r, pscale(b))


def test_mid_levels_lazy_ring_equal_kernel_per_step():
    # three builds in one context: the later ones skip z (the lazy dimension)
    # in the workgroup levels too
    b = K.ring(5_000_000, seed=98)
    a = trees(b, 1, builds=3)
    c = trees(b, 0, builds=3)
    for k in ("is_leaf", "split_dim", "num_parts", "first", "split_val", "size", "left",
              "right", "perm", "m", "cm"):
        np.testing.assert_array_equal(getattr(a, k), getattr(c, k), err_msg=k)


def test_mid_levels_step_equal():
    b = K.ring(3_000_000, seed=99)
    out = []
    for mid in (1, 0):
        with K.Simulation(3_000_000, precision=K.KD_PREC_MIXED) as s:
            s.set_tuning("mid_levels", mid)
            s.upload(b)
            s.step(3)
            assert s.get_tuning("mid_builds") == 2 * mid
            out.append(s.download())
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg=k)
