"""CPU oracle for the kD-tree Barnes-Hut step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker.  It wraps:

* liboracle.so  -- the C restatement in kdtree_oracle.c (see its header for
  the reference lines each function follows);
* _ref/libkdref_seq.so, _ref/libkdref_par.so -- the reference's own C
  implementation compiled from its sources (Makefile in this directory),
  driven through ref_shim.c.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)

LAYOUT_COMPACT = 0
LAYOUT_PADDED = 1


class OrcNode(C.Structure):
    _fields_ = [("is_leaf", C.c_int32), ("split_dim", C.c_int32),
                ("num_parts", C.c_int64), ("first", C.c_int64),
                ("split_val", C.c_double), ("m", C.c_double),
                ("cm", C.c_double * 3), ("size", C.c_double),
                ("left", C.c_int64), ("right", C.c_int64)]


class OrcTree(C.Structure):
    _fields_ = [("n", C.c_int64), ("n_nodes", C.c_int64), ("used", C.c_int64),
                ("leaf_size", C.c_int32), ("layout", C.c_int32),
                ("nodes", C.POINTER(OrcNode)), ("perm", _i64)]


class OrcCounters(C.Structure):
    _fields_ = [("node_tests", C.c_int64), ("m2p", C.c_int64), ("p2p", C.c_int64)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    args = ["make", "-s", "-C", str(HERE)]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)


_ORC = None
_REF = {}


def _d_(a):
    return a.ctypes.data_as(_d)


def _i_(a):
    return a.ctypes.data_as(_i64)


def lib():
    global _ORC
    if _ORC is None:
        p = HERE / "liboracle.so"
        if not p.exists():
            build()
        L = C.CDLL(str(p))
        L.orc_build.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_int, C.c_int,
                                C.c_uint64, C.POINTER(OrcTree)]
        L.orc_build.restype = C.c_int
        L.orc_tree_free.argtypes = [C.POINTER(OrcTree)]
        L.orc_accel.argtypes = [C.POINTER(OrcTree), C.c_int64, _d, _d, _d, _d, C.c_double,
                                _d, _d, _d, _d, _d, _d, C.POINTER(OrcCounters)]
        L.orc_accel.restype = C.c_int
        L.orc_accel_subset.argtypes = [C.POINTER(OrcTree), C.c_int64, _d, _d, _d, _d,
                                       C.c_double, C.c_int64, _i64, _d, _d, _d, _d, _d, _d,
                                       C.POINTER(OrcCounters)]
        L.orc_accel_subset.restype = C.c_int
        L.orc_simple_sim.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _d, C.c_double,
                                     C.c_int64, C.c_double, C.c_int, C.c_int, C.c_uint64, _d]
        L.orc_simple_sim.restype = C.c_int
        L.orc_direct.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_int64, _i64, _d, _d, _d,
                                 _d, _d, _d]
        L.orc_direct_comp.argtypes = L.orc_direct.argtypes
        L.orc_quickstat_index.argtypes = [_i64, C.c_int64, C.c_int64, _d, C.c_uint64]
        L.orc_nodes_needed.argtypes = [C.c_int64, C.c_int, C.c_int]
        L.orc_nodes_needed.restype = C.c_int64
        L.orc_compact_count.argtypes = [C.c_int64, C.c_int]
        L.orc_compact_count.restype = C.c_int64
        L.orc_kick_drift.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _d, _d, _d, C.c_double]
        L.orc_adj_forward.argtypes = [C.c_int64] + [_d] * 7 + [C.c_double, C.c_int64, C.c_double,
                                                              C.c_int, C.POINTER(C.c_void_p),
                                                              C.c_void_p]
        L.orc_adj_forward.restype = C.c_int
        L.orc_adj_backward.argtypes = [C.c_void_p] + [_d] * 14
        L.orc_adj_backward.restype = C.c_int
        L.orc_tape_free.argtypes = [C.c_void_p]
        L.orc_adj_debug_drop.argtypes = [C.c_int]
        _ORC = L
    return _ORC


def ref_lib(kind: str = "seq"):
    """The reference C build (kind 'seq' or 'par'); None if not built."""
    if kind not in _REF:
        p = HERE / "_ref" / f"libkdref_{kind}.so"
        if not p.exists():
            return None
        L = C.CDLL(str(p))
        L.ref_build.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_uint, C.c_int64, _i64, _i64,
                                _i64, _d, _d, _d, _d, _d, _d, _i64, _i64, _i64]
        L.ref_build.restype = C.c_int64
        L.ref_accel.argtypes = [C.c_int64, _d, _d, _d, _d, C.c_uint, _d, _d, _d]
        L.ref_sim.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _d, C.c_double, C.c_int64,
                              C.c_uint, _d]
        L.ref_cpu_sample.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, _d, C.c_int64, _i64,
                                     C.c_int, C.c_double, _d, _d]
        L.ref_max_parts.restype = C.c_int64
        L.ref_simple_sim_timed.argtypes = [C.c_int64] + [_d] * 7 + [C.c_double, C.c_int64,
                                                                   C.c_int, _d]
        _REF[kind] = L
    return _REF[kind]


def _f(b, k):
    return np.ascontiguousarray(b[k], dtype=np.float64)


class OracleTree:
    """Python copy of an orc_tree (node vector + particle order)."""

    def __init__(self, t: OrcTree):
        nn = t.n_nodes
        self.n_nodes = nn
        self.used = t.used
        nodes = np.ctypeslib.as_array(t.nodes, shape=(nn,))
        self.is_leaf = nodes["is_leaf"].astype(np.int32)
        self.split_dim = nodes["split_dim"].astype(np.int32)
        self.num_parts = nodes["num_parts"].astype(np.int64)
        self.first = nodes["first"].astype(np.int64)
        self.split_val = nodes["split_val"].copy()
        self.m = nodes["m"].copy()
        self.cm = np.array(nodes["cm"]).reshape(nn, 3).copy()
        self.size = nodes["size"].copy()
        self.left = nodes["left"].astype(np.int64)
        self.right = nodes["right"].astype(np.int64)
        self.perm = np.ctypeslib.as_array(t.perm, shape=(t.n,)).copy()

    def leaf_parts(self, node):
        return self.perm[self.first[node]:self.first[node] + self.num_parts[node]]


def build_tree(b: dict, leaf: int = 7, layout: int = LAYOUT_COMPACT, seed: int = 1):
    L = lib()
    x, y, z, m = (_f(b, k) for k in "xyzm")
    t = OrcTree()
    rc = L.orc_build(len(x), _d_(x), _d_(y), _d_(z), _d_(m), leaf, layout, seed, C.byref(t))
    if rc:
        L.orc_tree_free(C.byref(t))
        raise RuntimeError(f"orc_build failed: {rc}")
    out = OracleTree(t)
    L.orc_tree_free(C.byref(t))
    return out


def accel(b: dict, theta: float = 0.3, leaf: int = 7, layout: int = LAYOUT_COMPACT,
          seed: int = 1, want_abs: bool = False, idx=None):
    """Accelerations (n, 3) of one build (or of the bodies idx only); with
    want_abs also sum |a_ij| per component and the (tests, M2P, P2P) counters."""
    L = lib()
    x, y, z, m = (_f(b, k) for k in "xyzm")
    n = len(x)
    t = OrcTree()
    rc = L.orc_build(n, _d_(x), _d_(y), _d_(z), _d_(m), leaf, layout, seed, C.byref(t))
    if rc:
        L.orc_tree_free(C.byref(t))
        raise RuntimeError(f"orc_build failed: {rc}")
    cnt = OrcCounters()
    if idx is None:
        a = [np.empty(n) for _ in range(3)]
        ab = [np.empty(n) for _ in range(3)]
        L.orc_accel(C.byref(t), n, _d_(x), _d_(y), _d_(z), _d_(m), theta,
                    *[_d_(v) for v in a], *[_d_(v) for v in ab], C.byref(cnt))
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        a = [np.empty(len(idx)) for _ in range(3)]
        ab = [np.empty(len(idx)) for _ in range(3)]
        L.orc_accel_subset(C.byref(t), n, _d_(x), _d_(y), _d_(z), _d_(m), theta, len(idx),
                           _i_(idx), *[_d_(v) for v in a], *[_d_(v) for v in ab],
                           C.byref(cnt))
    L.orc_tree_free(C.byref(t))
    acc = np.stack(a, axis=1)
    if want_abs:
        return acc, np.stack(ab, axis=1), (cnt.node_tests, cnt.m2p, cnt.p2p)
    return acc


def simple_sim(b: dict, dt=1e-3, steps=1, theta=0.3, leaf=7, layout=LAYOUT_COMPACT, seed=1,
               keep_acc=False):
    L = lib()
    arrs = {k: _f(b, k).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    n = len(arrs["x"])
    hist = np.empty((steps, 3, n)) if keep_acc else None
    rc = L.orc_simple_sim(n, *[_d_(arrs[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")],
                          dt, steps, theta, leaf, layout, seed,
                          _d_(hist) if keep_acc else None)
    if rc:
        raise RuntimeError(f"orc_simple_sim failed: {rc}")
    return (arrs, hist) if keep_acc else arrs


class Tape:
    """A recorded frozen-topology run (orc_tape)."""

    def __init__(self, ptr, n, m):
        self.ptr, self.n, self.m = ptr, n, m

    def __del__(self):
        if self.ptr:
            lib().orc_tape_free(self.ptr)
            self.ptr = None


def adj_forward(b: dict, dt=1e-3, steps=1, theta=0.3, leaf=7, replay: Tape | None = None):
    """Frozen-topology simple_sim.  Without replay: records and returns
    (final state, tape); with replay: (final state, None) using that tape's
    trees and interaction lists."""
    L = lib()
    arrs = {k: _f(b, k).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    n = len(arrs["x"])
    tp = C.c_void_p()
    rc = L.orc_adj_forward(n, *[_d_(arrs[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")],
                           dt, steps, theta, leaf, None if replay else C.byref(tp),
                           replay.ptr if replay else None)
    if rc:
        raise RuntimeError(f"orc_adj_forward failed: {rc}")
    return arrs, (None if replay else Tape(tp, n, arrs["m"]))


def adj_backward(tape: Tape, lp, lv):
    """Gradients (d/dp0 (n,3), d/dv0 (n,3), d/dm (n,)) from dL/dp_N and
    dL/dv_N (both (n,3))."""
    L = lib()
    n = tape.n
    lp = np.ascontiguousarray(np.asarray(lp, np.float64).T)
    lv = np.ascontiguousarray(np.asarray(lv, np.float64).T)
    out = [np.zeros(n) for _ in range(7)]
    rc = L.orc_adj_backward(tape.ptr, _d_(np.ascontiguousarray(tape.m)),
                            *[_d_(lp[c]) for c in range(3)], *[_d_(lv[c]) for c in range(3)],
                            *[_d_(o) for o in out])
    if rc:
        raise RuntimeError(f"orc_adj_backward failed: {rc}")
    return np.stack(out[0:3], 1), np.stack(out[3:6], 1), out[6]


def direct(b: dict, idx, compensated: bool = False):
    """FP64 direct sums (calc_pp_accel terms) for the bodies idx and the sums
    of |a_ij|; compensated: Neumaier-summed, exact to about one rounding."""
    L = lib()
    x, y, z, m = (_f(b, k) for k in "xyzm")
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    k = len(idx)
    a = [np.empty(k) for _ in range(6)]
    fn = L.orc_direct_comp if compensated else L.orc_direct
    fn(len(x), _d_(x), _d_(y), _d_(z), _d_(m), k, _i_(idx), *[_d_(v) for v in a])
    return np.stack(a[:3], axis=1), np.stack(a[3:], axis=1)


def quickstat_index(indices, goal, keys, seed=1):
    ix = np.ascontiguousarray(indices, dtype=np.int64).copy()
    keys = np.ascontiguousarray(keys, dtype=np.float64)
    lib().orc_quickstat_index(_i_(ix), len(ix), goal, _d_(keys), seed)
    return ix


def nodes_needed(n, leaf, layout):
    return int(lib().orc_nodes_needed(n, leaf, layout))


def compact_count(n, leaf):
    return int(lib().orc_compact_count(n, leaf))


def ref_build(b: dict, seed: int = 1, kind: str = "seq"):
    """The reference C build_tree on the same positions: dict of node arrays."""
    R = ref_lib(kind)
    if R is None:
        return None
    x, y, z, m = (_f(b, k) for k in "xyzm")
    n = len(x)
    mp = int(R.ref_max_parts())
    cap = 4 * n + 16
    out = dict(num_parts=np.zeros(cap, np.int64), parts=np.zeros(cap * mp, np.int64),
               split_dim=np.zeros(cap, np.int64), split_val=np.zeros(cap), m=np.zeros(cap),
               cmx=np.zeros(cap), cmy=np.zeros(cap), cmz=np.zeros(cap), size=np.zeros(cap),
               left=np.zeros(cap, np.int64), right=np.zeros(cap, np.int64),
               indices=np.zeros(n, np.int64))
    used = R.ref_build(n, _d_(x), _d_(y), _d_(z), _d_(m), seed, cap, _i_(out["num_parts"]),
                       _i_(out["parts"]), _i_(out["split_dim"]), _d_(out["split_val"]),
                       _d_(out["m"]), _d_(out["cmx"]), _d_(out["cmy"]), _d_(out["cmz"]),
                       _d_(out["size"]), _i_(out["left"]), _i_(out["right"]),
                       _i_(out["indices"]))
    if used < 0:
        raise RuntimeError("ref_build capacity")
    for k in list(out):
        if k == "parts":
            out[k] = out[k][:used * mp].reshape(used, mp)
        elif k != "indices":
            out[k] = out[k][:used]
    out["used"] = used
    return out


def ref_accel(b: dict, seed: int = 1, kind: str = "seq"):
    R = ref_lib(kind)
    if R is None:
        return None
    x, y, z, m = (_f(b, k) for k in "xyzm")
    n = len(x)
    a = [np.empty(n) for _ in range(3)]
    R.ref_accel(n, _d_(x), _d_(y), _d_(z), _d_(m), seed, *[_d_(v) for v in a])
    return np.stack(a, axis=1)


def ref_simple_sim_timed(b: dict, dt=1e-3, steps=1, threads=0):
    """The reference's Parallel/CVersion simple_sim (OpenMP) run whole on a
    copy of b; returns (final state, wall seconds of the call), or None when
    the reference build is absent."""
    R = ref_lib("par")
    if R is None:
        return None
    arrs = {k: _f(b, k).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    t = np.zeros(1)
    R.ref_simple_sim_timed(len(arrs["x"]), *[_d_(arrs[k]) for k in ("x", "y", "z", "vx", "vy",
                                                                    "vz", "m")],
                           dt, steps, threads, _d_(t))
    return arrs, float(t[0])


def ref_sim(b: dict, dt=1e-3, steps=1, seed=1, kind="seq"):
    R = ref_lib(kind)
    if R is None:
        return None
    arrs = {k: _f(b, k).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
    n = len(arrs["x"])
    hist = np.empty((steps, 3, n))
    R.ref_sim(n, *[_d_(arrs[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")], dt, steps,
              seed, _d_(hist))
    return arrs, hist
