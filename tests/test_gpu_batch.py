"""Batches of independent systems (kd_batch): every system's tree is the tree
a single-system context builds (and so the oracle's); accelerations and
trajectories equal the single-system path bit for bit, in every precision;
the graph-replayed step equals the eagerly launched one on every replay."""
import numpy as np
import pytest

from helpers import accel_errors, compare_trees, pscale
from multilanguagekdtree_amd import _lib
from multilanguagekdtree_amd import kdtree as K

pytestmark = pytest.mark.gpu

# uneven systems: one-leaf, a few leaves, subtree-only (<= 2048, the subtree
# builder's bound), and ones that need device-wide levels; not multiples of 64
SIZES = [5, 70, 1000, 2048, 2049, 4096, 4097, 9000, 30001, 2000, 65, 120000, 1, 777]


def single(b, prec, fn):
    with K.Simulation(len(b["x"]), precision=prec) as s:
        if prec == K.KD_PREC_FP64:
            s.set_tuning("fp64_batched_walk", 1)  # the walk the batch runs
        s.upload(b)
        return fn(s)


@pytest.fixture(scope="module")
def bodies():
    return K.ensemble(SIZES, seed0=100)


def test_trees_match_oracle(orc, bodies):
    with K.Batch(SIZES, precision=K.KD_PREC_MIXED) as bt:
        bt.upload(bodies)
        bt.accel()
        for i in range(len(SIZES)):
            b = bt.system(bodies, i)
            for layout in (K.KD_LAYOUT_COMPACT, K.KD_LAYOUT_PADDED):
                compare_trees(bt.export_tree(i, layout), orc.build_tree(b, 7, layout), pscale(b))


@pytest.mark.parametrize("prec", [K.KD_PREC_MIXED, K.KD_PREC_FP64, K.KD_PREC_MIXED_D64])
def test_accel_equals_single_system(bodies, prec):
    with K.Batch(SIZES, precision=prec) as bt:
        bt.upload(bodies)
        a = bt.accel()
    for i in range(len(SIZES)):
        b = bt.system(bodies, i)
        ref = single(b, prec, lambda s: s.accel())
        lo, hi = bt.offsets[i], bt.offsets[i + 1]
        np.testing.assert_array_equal(a[lo:hi], ref, err_msg=f"system {i}")


def test_fp64_accel_matches_oracle(orc, bodies):
    with K.Batch(SIZES, precision=K.KD_PREC_FP64) as bt:
        bt.upload(bodies)
        a = bt.accel()
    for i in (0, 2, 5, 9):
        b = bt.system(bodies, i)
        ref, absref, _ = orc.accel(b, want_abs=True)
        _, cond = accel_errors(a[bt.offsets[i]:bt.offsets[i + 1]], ref, absref)
        assert cond.max() <= 1e-12, (i, cond.max())


@pytest.mark.parametrize("leaf,theta", [(2, 0.3), (8, 0.5), (32, 0.3)])
def test_other_leaf_sizes_and_angles_match_oracle(orc, bodies, leaf, theta):
    # the forest's trees (both layouts) and FP64 accelerations at the leaf
    # sizes and opening angle the one-system tests also cover
    with K.Batch(SIZES, theta=theta, leaf_size=leaf, precision=K.KD_PREC_FP64) as bt:
        bt.upload(bodies)
        a = bt.accel()
        for i in range(len(SIZES)):
            b = bt.system(bodies, i)
            for layout in (K.KD_LAYOUT_COMPACT, K.KD_LAYOUT_PADDED):
                compare_trees(bt.export_tree(i, layout), orc.build_tree(b, leaf, layout),
                              pscale(b))
    for i in (0, 2, 3, 5, 9, 11):
        b = bt.system(bodies, i)
        ref, absref, _ = orc.accel(b, theta=theta, leaf=leaf, want_abs=True)
        _, cond = accel_errors(a[bt.offsets[i]:bt.offsets[i + 1]], ref, absref)
        assert cond.max() <= 1e-12, (i, cond.max())


@pytest.mark.parametrize("shape", [(256, 0), (64, 8), (512, 32), (128, 16)])
def test_launch_shapes_agree(bodies, shape):
    with K.Batch(SIZES) as bt:
        bt.upload(bodies)
        ref = bt.accel()
        bt.set_tuning("walk_threads", shape[0])
        bt.set_tuning("walk_waves_per_cu", shape[1])
        np.testing.assert_array_equal(bt.accel(), ref)


def test_graph_replays_equal_eager_steps(bodies):
    # the graph is replayed with host work and synchronisation between
    # replays; every replay must equal the eager step
    with K.Batch(SIZES) as g, K.Batch(SIZES) as e:
        g.upload(bodies)
        e.upload(bodies)
        e.set_tuning("graph", 0)
        for k in range(5):
            g.step(1)
            e.step(1)
            sg, se = g.download(), e.download()
            for f in ("x", "y", "z", "vx", "vy", "vz"):
                np.testing.assert_array_equal(sg[f], se[f], err_msg=f"replay {k} {f}")


def test_steps_equal_single_system(bodies):
    with K.Batch(SIZES) as bt:
        bt.upload(bodies)
        bt.step(3)
        got = bt.download()
    for i in (0, 3, 4, 6, 9, 10):
        b = bt.system(bodies, i)

        def run(s):
            s.step(3)
            return s.download()
        ref = single(b, K.KD_PREC_MIXED, run)
        for f in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(bt.system(got, i)[f], ref[f], err_msg=f"{i} {f}")


@pytest.mark.parametrize("graph", [1, 0])
def test_fused_kick_drift_equals_separate_pass(bodies, graph):
    # the walk applying the kick-drift (and the next build taking the range
    # its wavefronts reduce) against the separate pass, in both step forms
    out = []
    for fused in (1, 0):
        with K.Batch(SIZES) as bt:
            bt.set_tuning("graph", graph)
            bt.set_tuning("fused_kick_drift", fused)
            bt.upload(bodies)
            bt.step(2)
            bt.step(2)
            out.append(bt.download())
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        np.testing.assert_array_equal(out[0][f], out[1][f], err_msg=f)


def cubes(sizes, seed):
    rng = np.random.default_rng(seed)
    n = int(np.sum(sizes))
    p = rng.uniform(-1.0, 1.0, (n, 3))
    return dict(x=p[:, 0].copy(), y=p[:, 1].copy(), z=p[:, 2].copy(), vx=np.zeros(n),
                vy=np.zeros(n), vz=np.zeros(n), m=np.full(n, 1.0 / n))


@pytest.mark.parametrize("graph", [1, 0])
def test_lazy_dimension_equals_full_builds(bodies, graph):
    # thin ring systems: after the first (full) step every build skips z,
    # and the check after each step holds; then cubes in the same batch: the
    # first step after the upload skips z again, its check fails, and the
    # step is taken again with a full build
    out = []
    for lazy in (1, 0):
        with K.Batch(SIZES) as bt:
            bt.set_tuning("graph", graph)
            bt.set_tuning("lazy_dimension", lazy)
            bt.upload(bodies)
            bt.step(3)
            if lazy:
                assert bt.get_tuning("lazy_builds") == 2 and bt.get_tuning("lazy_redos") == 0
                assert bt.get_tuning("lazy_next") == 2
            first = bt.download()
            bt.upload(cubes(SIZES, 7))
            bt.step(2)
            if lazy:
                assert bt.get_tuning("lazy_redos") == 1
                assert bt.get_tuning("lazy_next") == 3
            out.append((first, bt.download()))
    for a, b in zip(out[0], out[1]):
        for f in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(a[f], b[f], err_msg=f)


def test_autotune_keeps_results(bodies):
    with K.Batch(SIZES) as bt:
        bt.upload(bodies)
        ref = bt.accel()
        best = bt.autotune(reps=2)
        assert best["walk_threads"] in (64, 128, 256, 512)
        assert 0 <= best["walk_waves_per_cu"] <= 64 and best["walk_ms"] > 0
        np.testing.assert_array_equal(bt.accel(), ref)


@pytest.mark.parametrize("threads,waves", [(64, 8), (256, 32), (512, 16), (128, 1)])
def test_persistent_walk_matches_one_launch_per_task(bodies, threads, waves):
    # the persistent walk takes tasks from one counter per XCD (chunks dealt
    # round robin); every task must run exactly once, so the accelerations
    # equal those of one wavefront per task bit for bit
    with K.Batch(SIZES) as bt:
        bt.upload(bodies)
        bt.set_tuning("walk_threads", 64)
        bt.set_tuning("walk_waves_per_cu", 0)
        ref = bt.accel()
        bt.set_tuning("walk_threads", threads)
        bt.set_tuning("walk_waves_per_cu", waves)
        np.testing.assert_array_equal(bt.accel(), ref)


def test_counters_sum_over_systems(bodies):
    with K.Batch(SIZES, count_interactions=True) as bt:
        bt.upload(bodies)
        bt.accel()
        c = bt.counters()
    tot = np.zeros(3, np.int64)
    for i in range(len(SIZES)):
        b = bt.system(bodies, i)
        with K.Simulation(len(b["x"]), precision=K.KD_PREC_MIXED, count_interactions=True) as s:
            s.upload(b)
            s.accel()
            tot += s.counters()[:3]
    np.testing.assert_array_equal(c[:3], tot)


def test_batch_errors(bodies):
    with pytest.raises(_lib.KdError):
        K.Batch([10, 0, 5])
    with K.Batch([10, 20]) as bt:
        with pytest.raises(_lib.KdError):
            bt.step(1)  # nothing uploaded
        with pytest.raises(_lib.KdError):
            bt.set_tuning("walk_threads", 100)
        bad = K.ensemble([10, 20])
        bad["x"][13] = np.nan
        with pytest.raises(_lib.KdError) as e:
            bt.upload(bad)
        assert e.value.code == _lib.KD_ERR_NONFINITE


def test_graph_replays_equal_eager_steps_large():
    # large enough that the sort library takes its multi-pass path (whose
    # memset nodes the capture turns into fill kernels)
    sizes = np.random.default_rng(5).integers(2000, 200000, size=48)
    bodies = K.ensemble(sizes, seed0=500)
    with K.Batch(sizes) as g, K.Batch(sizes) as e:
        g.upload(bodies)
        e.upload(bodies)
        e.set_tuning("graph", 0)
        for k in range(3):
            g.step(1)
            e.step(1)
            sg, se = g.download(), e.download()
            for f in ("x", "vx", "z"):
                np.testing.assert_array_equal(sg[f], se[f], err_msg=f"replay {k} {f}")
