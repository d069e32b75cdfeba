"""Pins the CPU oracle (CPU only, no GPU).

Checks, each chosen so that a plausible mistake fails one of them:
  * golden vectors of the reference's own C build (tests/golden/config0.npz):
    tree (membership, split dim/value, size, child indices), node moments,
    accelerations and a 10-step trajectory;
  * the reference's only known-answer test (quickstat.rs:191-197) and the
    quickselect postcondition (quickstat.rs:199-219);
  * closed forms: node-vector sizing of both layouts, two_bodies, theta = 0
    equals the O(N^2) direct sum;
  * brute force: membership by full sort of every segment on small inputs;
  * the spatial invariant of recur_test_tree_struct (array_kd_tree.rs:834-877),
    here on all three dimensions;
  * the live reference C build on fresh inputs (when oracle/_ref exists).
"""
from pathlib import Path

import numpy as np
import pytest

from helpers import accel_errors, compare_trees, pscale

GOLD = Path(__file__).resolve().parent / "golden" / "config0.npz"


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def bodies_of(g):
    return {k: g[k] for k in ("x", "y", "z", "vx", "vy", "vz", "m")}


def rand_bodies(n, seed, dims=3):
    rng = np.random.default_rng(seed)
    b = {k: rng.standard_normal(n) for k in ("x", "y", "z")}
    if dims == 2:
        b["z"] = np.zeros(n)
    b.update(vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), m=rng.uniform(0.5, 1.5, n))
    return b


# ---- golden vectors of the reference C build ------------------------------

def test_tree_matches_reference_golden(orc, gold):
    t = orc.build_tree(bodies_of(gold), leaf=7)
    nn = len(gold["tree_num_parts"])
    assert t.used == nn
    np.testing.assert_array_equal(t.num_parts[:nn], gold["tree_num_parts"])
    internal = gold["tree_num_parts"] == 0
    for i in np.nonzero(~internal)[0]:
        np.testing.assert_array_equal(t.leaf_parts(i),
                                      gold["tree_parts"][i][:gold["tree_num_parts"][i]])
    for k in ("split_dim", "left", "right"):
        np.testing.assert_array_equal(getattr(t, k)[:nn][internal], gold[f"tree_{k}"][internal])
    for k in ("split_val", "size"):
        a, r = getattr(t, k)[:nn][internal], gold[f"tree_{k}"][internal]
        assert np.array_equal(a.view(np.int64), r.view(np.int64)), k
    rm = gold["tree_m"][internal]
    assert np.all(np.abs(t.m[:nn][internal] - rm) <= 2 * 1000 * 2.0 ** -53 * rm)
    assert np.abs(t.cm[:nn][internal] - gold["tree_cm"][internal]).max() <= 1e-13


def test_accel_matches_reference_golden(orc, gold):
    b = bodies_of(gold)
    a, ab, _ = orc.accel(b, want_abs=True)
    plain, cond = accel_errors(a, gold["acc_step0"].T, ab)
    assert cond.max() <= 1e-13
    assert plain[1:].max() <= 1e-13


def test_trajectory_matches_reference_golden(orc, gold):
    b = bodies_of(gold)
    state, hist = orc.simple_sim(b, dt=1e-3, steps=10, keep_acc=True)
    for s in (4, 9):
        _, cond = accel_errors(hist[s].T, gold[f"acc_step{s}"].T, np.abs(gold[f"acc_step{s}"].T) + 1e-300)
        # positions differ by O(1e-16) between the two runs; compare relatively
        da = np.linalg.norm(hist[s].T - gold[f"acc_step{s}"].T, axis=1)
        na = np.linalg.norm(gold[f"acc_step{s}"].T, axis=1)
        assert np.all(da[1:] <= 1e-12 * na[1:])
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        ref = gold[f"final_{k}"]
        assert np.abs(state[k] - ref).max() <= 1e-13 * np.abs(ref).max(), k


# ---- quickselect ------------------------------------------------------------

@pytest.mark.parametrize("seed", range(8))
def test_quickstat_known_answer(orc, gold, seed):
    ix = orc.quickstat_index(np.arange(7), int(gold["qs_goal"]), gold["qs_keys"], seed=seed)
    assert ix[int(gold["qs_goal"])] == int(gold["qs_expect"])


@pytest.mark.parametrize("seed", range(5))
def test_quickstat_postcondition(orc, seed):
    rng = np.random.default_rng(seed)
    n = 5000
    keys = rng.random(n)
    for _ in range(20):
        goal = int(rng.integers(0, n))
        ix = orc.quickstat_index(np.arange(n), goal, keys, seed=int(rng.integers(1 << 30)))
        assert sorted(ix) == list(range(n))
        kg = keys[ix[goal]]
        assert np.all(keys[ix[:goal]] < kg) and np.all(keys[ix[goal:]] >= kg)


def test_quickstat_ties_by_index(orc):
    keys = np.array([1.0, 0.0, 1.0, 1.0, 0.0, 1.0])
    ix = orc.quickstat_index(np.arange(6), 3, keys, seed=3)
    # order (key, index): 1, 4, 0, 2, 3, 5 -> rank 3 is index 2
    assert ix[3] == 2 and set(ix[:3]) == {1, 4, 0}


# ---- closed forms -----------------------------------------------------------

def test_node_vector_sizes(orc):
    # Sequential: 2*(n/(L/2)+1) (array_kd_tree.rs:37-39); the reference's own
    # test asserts 6 for 12 bodies, which the formula does not give (10)
    assert orc.nodes_needed(12, 7, 0) == 10
    assert orc.nodes_needed(2, 7, 0) == 2
    # Parallel: 1 if n <= 8 else 2*2^ceil(log2(n/4))-1 (array_kd_tree.rs:45-53)
    assert orc.nodes_needed(2, 8, 1) == 1
    assert orc.nodes_needed(9, 8, 1) == 3
    assert orc.nodes_needed(1_000_000, 8, 1) == 2 * 262144 - 1
    assert orc.nodes_needed(10_000_000, 8, 1) == 2 * 4194304 - 1
    # compact nodes used: 2*leaves-1 with leaves fixed by the floor(n/2) recursion
    assert orc.compact_count(1000, 7) == 463
    assert orc.compact_count(1_000_000, 7) == 427135
    assert orc.compact_count(1_000_000, 8) == 262143


def test_two_bodies_closed_form(orc):
    b = dict(x=np.array([0.0, 1.0]), y=np.zeros(2), z=np.zeros(2), vx=np.zeros(2),
             vy=np.array([0.0, 1.0]), vz=np.zeros(2), m=np.array([1.0, 1e-20]))
    a = orc.accel(b)
    assert a[1].tolist() == [-1.0, 0.0, 0.0]
    assert a[0].tolist() == [1e-20, 0.0, 0.0]


def test_theta_zero_is_direct_sum(orc, gold):
    b = bodies_of(gold)
    a = orc.accel(b, theta=0.0)
    ref, ab = orc.direct(b, np.arange(1000))
    _, cond = accel_errors(a, ref, ab)
    assert cond.max() <= 1e-14


def test_theta_controls_the_error(orc):
    # a dropped or inverted opening test would make theta irrelevant
    b = rand_bodies(2000, 9)
    ref, _ = orc.direct(b, np.arange(2000))
    med = []
    for theta in (0.1, 0.3, 0.6):
        plain, _ = accel_errors(orc.accel(b, theta=theta), ref, np.abs(ref))
        med.append(np.median(plain))
    assert med[0] < med[1] < med[2] < 0.05 and med[1] > 1e-6, med


# ---- brute force and invariants ---------------------------------------------

def brute_membership(b, leaf):
    """Leaves (as sorted index tuples) in DFS order by full sort per segment."""
    c = np.stack([b["x"], b["y"], b["z"]], 1)
    out = []

    def rec(ix):
        if len(ix) <= leaf:
            out.append(tuple(sorted(ix.tolist())))
            return
        ext = c[ix].max(0) - c[ix].min(0)
        sd = 0
        for d in (1, 2):
            if ext[d] > ext[sd]:
                sd = d
        order = np.lexsort((ix, c[ix, sd]))  # key, then original index
        s = ix[order]
        mid = len(ix) // 2
        rec(s[:mid])
        rec(s[mid:])

    rec(np.arange(len(b["x"])))
    return out


@pytest.mark.parametrize("n,leaf,seed", [(300, 7, 0), (1000, 8, 1), (777, 3, 2), (64, 2, 3)])
def test_membership_brute_force(orc, n, leaf, seed):
    b = rand_bodies(n, seed)
    t = orc.build_tree(b, leaf=leaf)
    leaves = [tuple(t.leaf_parts(i).tolist()) for i in range(t.used) if t.is_leaf[i]]
    assert leaves == brute_membership(b, leaf)


def test_membership_brute_force_with_ties(orc):
    rng = np.random.default_rng(4)
    n = 600
    b = rand_bodies(n, 4)
    for k in ("x", "y", "z"):
        b[k] = np.round(b[k] * 2) / 2  # many equal coordinates
    t = orc.build_tree(b, leaf=7)
    leaves = [tuple(t.leaf_parts(i).tolist()) for i in range(t.used) if t.is_leaf[i]]
    assert leaves == brute_membership(b, 7)


def check_spatial_invariant(t, b):
    c = np.stack([b["x"], b["y"], b["z"]], 1)
    stack = [(0, np.full(3, -np.inf), np.full(3, np.inf))]
    while stack:
        i, lo, hi = stack.pop()
        if t.is_leaf[i]:
            p = c[t.leaf_parts(i)]
            assert np.all(p >= lo) and np.all(p <= hi)
            continue
        d, v = t.split_dim[i], t.split_val[i]
        hl, lr = hi.copy(), lo.copy()
        hl[d] = v
        lr[d] = v
        stack.append((t.left[i], lo, hl))
        stack.append((t.right[i], lr, hi))


@pytest.mark.parametrize("layout", [0, 1])
def test_spatial_invariant(orc, gold, layout):
    b = bodies_of(gold)
    check_spatial_invariant(orc.build_tree(b, leaf=7 if layout == 0 else 8, layout=layout), b)


def test_pivot_seed_independence(orc, gold):
    b = bodies_of(gold)
    compare_trees(orc.build_tree(b, seed=1), orc.build_tree(b, seed=12345), pscale(b))


# ---- the live reference build ------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_vs_reference_c(orc, seed):
    if orc.ref_lib("seq") is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")
    b = rand_bodies(3000, 100 + seed, dims=2 if seed == 2 else 3)
    t = orc.build_tree(b, leaf=7)
    r = orc.ref_build(b, seed=seed)
    assert r["used"] == t.used
    for i in range(t.used):
        if t.is_leaf[i]:
            assert sorted(r["parts"][i][:r["num_parts"][i]]) == t.leaf_parts(i).tolist()
        else:
            assert (r["split_dim"][i], r["split_val"][i], r["size"][i], r["left"][i],
                    r["right"][i]) == (t.split_dim[i], t.split_val[i], t.size[i], t.left[i],
                                       t.right[i])
    a, ab, _ = orc.accel(b, want_abs=True)
    _, cond = accel_errors(a, orc.ref_accel(b, seed=seed), ab)
    assert cond.max() <= 1e-13


def test_compensated_direct_sum_is_exactly_rounded(orc):
    # orc_direct_comp against math.fsum of the same FP64 terms (numpy forms
    # each term as calc_pp_accel does, array_particle.rs:67-76: no fused
    # operations, correctly rounded sqrt and divide), which is the correctly
    # rounded sum; and orc_direct's left-to-right sums within the classic
    # (n-1) u sum|a_ij| bound of the exact ones
    import math
    n = 3000
    rng = np.random.default_rng(61)
    r, phi = rng.uniform(1.0, 2.0, n), rng.uniform(0.0, 2 * np.pi, n)
    b = dict(x=r * np.cos(phi), y=r * np.sin(phi), z=rng.normal(0.0, 1e-3, n),
             m=np.full(n, 1e-14))
    for k in "xyz":
        b[k][0] = 0.0
    b["m"][0] = 1.0
    idx = np.array([0, 1, 17, 2999])
    comp, cabs = orc.direct(b, idx, compensated=True)
    plain, pabs = orc.direct(b, idx)
    x, y, z, m = (b[k] for k in "xyzm")
    for r, i in enumerate(idx):
        j = np.arange(n) != i
        dx, dy, dz = x[i] - x[j], y[i] - y[j], z[i] - z[j]
        dist = np.sqrt(dx * dx + dy * dy + dz * dz)
        magi = -m[j] / (dist * dist * dist)
        for c, d in enumerate((dx, dy, dz)):
            t = magi * d
            exact = math.fsum(t)
            assert abs(comp[r, c] - exact) <= 2 * np.spacing(abs(exact)), (i, c)
            assert abs(cabs[r, c] - math.fsum(np.abs(t))) <= 2 * np.spacing(cabs[r, c])
            assert abs(plain[r, c] - exact) <= (n - 1) * 2.0 ** -53 * pabs[r, c]
