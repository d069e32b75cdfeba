"""The FP64 direct sum on the GPU (kd_direct_sum, the truth every
mixed-precision force error is measured against) against the oracle's
direct sum (orc_direct_comp): calc_pp_accel summed over every other body (Sequential/RustVersion/src/array_particle.rs:67-76).

The GPU sums the terms in 256 strided partial sums combined by a tree; the
reference sums here are the oracle's compensated sums of the same terms
(exact to about one rounding), so the difference is the GPU's own summation
error: the conditioned error |da| / sum_j |a_ij| is held to 1e-12, the FP64
path's stated bound.  (A left-to-right sum over 10 M terms is itself off
by up to ~1e-12 conditioned: that order is not the yardstick.)"""
import numpy as np
import pytest

from helpers import accel_errors
from multilanguagekdtree_amd import kdtree as K

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu_direct(b, idx):
    with K.Simulation(len(b["x"])) as s:
        s.upload(b)
        return s.direct_sum(idx)


def check(orc, b, idx):
    a, ab = gpu_direct(b, idx)
    ref, absref = orc.direct(b, idx, compensated=True)
    plain, cond = accel_errors(a, ref, absref)
    assert cond.max() <= TOL, cond.max()
    # the sums of |a_ij| have no cancellation: plain relative agreement
    rel_abs = np.abs(ab - absref) / absref
    assert rel_abs.max() <= TOL, rel_abs.max()
    return plain


def test_direct_sum_ring_10m(orc):
    # the error-sample truth of the benchmark workload: 64 seeded bodies of
    # the 10 M ring, the central star among them
    n = 10_000_000
    b = K.ring(n, seed=0)
    idx = np.sort(np.random.default_rng(31).choice(n, 64, replace=False))
    idx[0] = 0
    plain = check(orc, b, idx)
    # ring bodies: |a| is the star's pull, no cancellation
    assert plain[1:].max() <= TOL, plain[1:].max()


def test_direct_sum_equal_mass_cube(orc):
    # no dominant mass: every force is a sum of 20k comparable terms with
    # cancellation
    n = 20_000
    rng = np.random.default_rng(32)
    p = rng.uniform(-1.0, 1.0, (n, 3))
    b = dict(x=p[:, 0].copy(), y=p[:, 1].copy(), z=p[:, 2].copy(), vx=np.zeros(n),
             vy=np.zeros(n), vz=np.zeros(n), m=np.full(n, 1.0 / n))
    idx = np.sort(rng.choice(n, 256, replace=False))
    check(orc, b, idx)


def test_direct_sum_two_bodies_closed_form():
    # two_bodies (array_particle.rs:10-17): a_1 = (-1, 0, 0) and
    # a_0 = (1e-20, 0, 0) exactly
    b = K.two_bodies()
    a, _ = gpu_direct(b, np.array([0, 1]))
    np.testing.assert_array_equal(a[1], [-1.0, 0.0, 0.0])
    np.testing.assert_array_equal(a[0], [1e-20, 0.0, 0.0])


def test_direct_sum_equals_theta0_walk(orc):
    # theta = 0 opens every node, so the oracle walk is a direct sum in leaf
    # order: a third summation order of the same terms
    n = 30_000
    b = K.ring(n, seed=33)
    idx = np.sort(np.random.default_rng(34).choice(n, 128, replace=False))
    a, _ = gpu_direct(b, idx)
    ref, absref, _ = orc.accel(b, theta=0.0, want_abs=True, idx=idx)
    _, cond = accel_errors(a, ref, absref)
    assert cond.max() <= TOL, cond.max()
