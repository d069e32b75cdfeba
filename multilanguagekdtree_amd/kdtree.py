"""Python host interface over the C ABI (tests, benchmarks, torch integration).

Mirrors the reference crate's API (Parallel/RustVersion/src/lib.rs:1-3):
``simple_sim`` (array_kd_tree.rs:623), ``build_tree`` (:63 / Sequential :49),
``calc_accel`` (:619), ``allocate_node_vec`` (:55) and the initial-condition
generators of array_particle.rs, backed by the gfx950 kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (KD_LAYOUT_COMPACT, KD_LAYOUT_PADDED, KD_PREC_FP64,
                   KD_PREC_MIXED, KD_PREC_MIXED_D64, KdTreeView, check, dptr,
                   i32ptr, iptr)

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m")


@dataclass
class Tree:
    """A node vector in one of the reference layouts, as flat arrays."""
    is_leaf: np.ndarray
    split_dim: np.ndarray
    num_parts: np.ndarray
    first: np.ndarray
    split_val: np.ndarray
    m: np.ndarray
    cm: np.ndarray          # (n_nodes, 3)
    size: np.ndarray
    left: np.ndarray
    right: np.ndarray
    perm: np.ndarray        # particle order (original indices)

    def leaf_parts(self, node: int) -> np.ndarray:
        f = self.first[node]
        return self.perm[f:f + self.num_parts[node]]

    def view(self) -> tuple[KdTreeView, list]:
        """A kd_tree_view over these arrays (and the arrays it points into,
        which must stay alive while the view is used)."""
        keep = dict(
            is_leaf=np.ascontiguousarray(self.is_leaf, np.int32),
            split_dim=np.ascontiguousarray(self.split_dim, np.int32),
            num_parts=np.ascontiguousarray(self.num_parts, np.int64),
            first=np.ascontiguousarray(self.first, np.int64),
            split_val=np.ascontiguousarray(self.split_val, np.float64),
            m=np.ascontiguousarray(self.m, np.float64),
            cmx=np.ascontiguousarray(self.cm[:, 0], np.float64),
            cmy=np.ascontiguousarray(self.cm[:, 1], np.float64),
            cmz=np.ascontiguousarray(self.cm[:, 2], np.float64),
            size=np.ascontiguousarray(self.size, np.float64),
            left=np.ascontiguousarray(self.left, np.int64),
            right=np.ascontiguousarray(self.right, np.int64),
            perm=np.ascontiguousarray(self.perm, np.int64))
        v = KdTreeView()
        v.n_nodes = len(keep["is_leaf"])
        v.is_leaf = i32ptr(keep["is_leaf"])
        v.split_dim = i32ptr(keep["split_dim"])
        for k in ("num_parts", "first", "left", "right", "perm"):
            setattr(v, k, iptr(keep[k]))
        for k in ("split_val", "m", "cmx", "cmy", "cmz", "size"):
            setattr(v, k, dptr(keep[k]))
        return v, list(keep.values())


class Simulation:
    """One device context (kd_ctx) holding up to ``capacity`` bodies."""

    def __init__(self, capacity: int, theta: float = 0.3, dt: float = 1e-3,
                 leaf_size: int = 7, precision: int = KD_PREC_FP64,
                 count_interactions: bool = False, profile: bool = False):
        L = _lib.lib()
        self._L = L
        p = _lib.default_params(theta=theta, dt=dt, leaf_size=leaf_size,
                                precision=precision,
                                count_interactions=int(count_interactions),
                                profile=int(profile))
        ctx = C.c_void_p()
        check(L.kd_create(capacity, C.byref(p), C.byref(ctx)))
        self._ctx = ctx
        self.capacity = capacity

    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._L.kd_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self):
        p = _lib.KdParams()
        check(self._L.kd_get_params(self._ctx, C.byref(p)))
        return p

    def set_params(self, **kw):
        p = self.params
        for k, v in kw.items():
            setattr(p, k, v)
        check(self._L.kd_set_params(self._ctx, C.byref(p)))

    @property
    def n(self) -> int:
        return int(self._L.kd_size(self._ctx))

    def upload(self, bodies: dict):
        arrs = [np.ascontiguousarray(bodies[k], dtype=np.float64) for k in FIELDS]
        n = len(arrs[0])
        check(self._L.kd_upload(self._ctx, n, *[dptr(a) for a in arrs]))

    def upload_device(self, tensors: dict):
        """Copy from torch CUDA tensors (float64, contiguous) already in HBM."""
        ts = [tensors[k] for k in FIELDS]
        n = ts[0].numel()
        check(self._L.kd_upload_device(self._ctx, n, *[C.c_void_p(t.data_ptr()) for t in ts]))

    def download(self) -> dict:
        n = self.n
        out = {k: np.empty(n) for k in FIELDS[:6]}
        check(self._L.kd_download(self._ctx, *[dptr(out[k]) for k in FIELDS[:6]]))
        return out

    def build(self):
        check(self._L.kd_build(self._ctx))

    def accel(self) -> np.ndarray:
        n = self.n
        a = [np.empty(n) for _ in range(3)]
        check(self._L.kd_accel(self._ctx, *[dptr(x) for x in a]))
        return np.stack(a, axis=1)

    def step(self, steps: int = 1):
        check(self._L.kd_step(self._ctx, steps))

    def sync(self):
        check(self._L.kd_sync(self._ctx))

    def counters(self) -> np.ndarray:
        """[node tests, M2P, P2P] summed over bodies, then [node visits, leaf
        visits] summed over wavefronts."""
        out = np.zeros(5, dtype=np.int64)
        check(self._L.kd_counters(self._ctx, iptr(out)))
        return out

    def set_tuning(self, key: str, value: int):
        check(self._L.kd_set_tuning(self._ctx, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        out = np.zeros(1, np.int64)
        check(self._L.kd_get_tuning(self._ctx, key.encode(), iptr(out)))
        return int(out[0])

    def timings(self) -> dict:
        """Summed HIP-event times (ms) of kd_step phases since the last reset."""
        out = np.zeros(4)
        check(self._L.kd_timings(self._ctx, dptr(out)))
        return dict(build_ms=out[0], traverse_ms=out[1], kick_drift_ms=out[2],
                    steps=int(out[3]))

    def reset_timings(self):
        check(self._L.kd_reset_timings(self._ctx))

    def direct_sum(self, idx) -> tuple[np.ndarray, np.ndarray]:
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        k = len(idx)
        a = [np.empty(k) for _ in range(6)]
        check(self._L.kd_direct_sum(self._ctx, k, iptr(idx), *[dptr(x) for x in a]))
        return np.stack(a[:3], axis=1), np.stack(a[3:], axis=1)

    def export_tree(self, layout: int = KD_LAYOUT_COMPACT) -> Tree:
        return _export(lambda v: self._L.kd_export_tree(self._ctx, layout, v), self.n)


def _export(call, n: int) -> Tree:
    """Run an export entry point twice: once for the length, once to fill."""
    v = KdTreeView()
    v.n_nodes = 0
    rc = call(C.byref(v))
    if rc != _lib.KD_ERR_CAPACITY:
        check(rc)
    nn = v.n_nodes
    a = dict(
        is_leaf=np.zeros(nn, np.int32), split_dim=np.zeros(nn, np.int32),
        num_parts=np.zeros(nn, np.int64), first=np.zeros(nn, np.int64),
        split_val=np.zeros(nn), m=np.zeros(nn), cmx=np.zeros(nn),
        cmy=np.zeros(nn), cmz=np.zeros(nn), size=np.zeros(nn),
        left=np.zeros(nn, np.int64), right=np.zeros(nn, np.int64),
        perm=np.zeros(n, np.int64))
    v.n_nodes = nn
    v.is_leaf = i32ptr(a["is_leaf"])
    v.split_dim = i32ptr(a["split_dim"])
    for k in ("num_parts", "first", "left", "right", "perm"):
        setattr(v, k, iptr(a[k]))
    for k in ("split_val", "m", "cmx", "cmy", "cmz", "size"):
        setattr(v, k, dptr(a[k]))
    check(call(C.byref(v)))
    return Tree(a["is_leaf"], a["split_dim"], a["num_parts"], a["first"],
                a["split_val"], a["m"],
                np.stack([a["cmx"], a["cmy"], a["cmz"]], axis=1), a["size"],
                a["left"], a["right"], a["perm"])


class Batch:
    """Many independent systems in one context (kd_batch): built and stepped
    together.  Bodies are concatenated in system order."""

    def __init__(self, sizes, theta: float = 0.3, dt: float = 1e-3, leaf_size: int = 7,
                 precision: int = KD_PREC_MIXED, count_interactions: bool = False):
        L = _lib.lib()
        self._L = L
        self.sizes = np.ascontiguousarray(sizes, np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        p = _lib.default_params(theta=theta, dt=dt, leaf_size=leaf_size, precision=precision,
                                count_interactions=int(count_interactions))
        b = C.c_void_p()
        check(L.kd_batch_create(len(self.sizes), iptr(self.sizes), C.byref(p), C.byref(b)))
        self._b = b
        self.n = int(self.offsets[-1])

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            self._L.kd_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, bodies: dict):
        cols = [np.ascontiguousarray(bodies[f], np.float64) for f in FIELDS]
        assert all(len(c) == self.n for c in cols)
        check(self._L.kd_batch_upload(self._b, *[dptr(c) for c in cols]))

    def download(self) -> dict:
        out = {f: np.zeros(self.n) for f in FIELDS[:6]}
        check(self._L.kd_batch_download(self._b, *[dptr(out[f]) for f in FIELDS[:6]]))
        return out

    def accel(self) -> np.ndarray:
        a = [np.zeros(self.n) for _ in range(3)]
        check(self._L.kd_batch_accel(self._b, *[dptr(x) for x in a]))
        return np.stack(a, axis=1)

    def step(self, steps: int = 1):
        check(self._L.kd_batch_step(self._b, steps))

    def export_tree(self, sys: int, layout: int = KD_LAYOUT_COMPACT) -> Tree:
        return _export(lambda v: self._L.kd_batch_export_tree(self._b, sys, layout, v),
                       int(self.sizes[sys]))

    def set_tuning(self, key: str, value: int):
        check(self._L.kd_batch_set_tuning(self._b, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        out = np.zeros(1, np.int64)
        check(self._L.kd_batch_get_tuning(self._b, key.encode(), iptr(out)))
        return int(out[0])

    def autotune(self, reps: int = 3) -> dict:
        out = np.zeros(3)
        check(self._L.kd_batch_autotune(self._b, reps, dptr(out)))
        return {"walk_threads": int(out[0]), "walk_waves_per_cu": int(out[1]),
                "walk_ms": float(out[2])}

    def counters(self) -> np.ndarray:
        out = np.zeros(5, np.int64)
        check(self._L.kd_batch_counters(self._b, iptr(out)))
        return out

    def system(self, arrays: dict, i: int) -> dict:
        """The slice of system i from concatenated arrays."""
        lo, hi = int(self.offsets[i]), int(self.offsets[i + 1])
        return {k: v[lo:hi] for k, v in arrays.items()}


def ensemble(sizes, seed0: int = 0) -> dict:
    """Concatenated ring systems, system i from seed seed0 + i."""
    parts = [ring(int(n), seed=seed0 + i) for i, n in enumerate(sizes)]
    return {f: np.concatenate([p[f] for p in parts]) for f in FIELDS}


def allocate_node_vec(n: int, leaf_size: int = 7, layout: int = KD_LAYOUT_COMPACT) -> int:
    """Node-vector length the reference allocates (array_kd_tree.rs:37-46 /
    Parallel :45-60)."""
    return int(_lib.lib().kd_nodes_needed(n, leaf_size, layout))


def simple_sim(bodies: dict, dt: float = 1e-3, steps: int = 1, theta: float = 0.3,
               leaf_size: int = 7, precision: int = KD_PREC_FP64) -> dict:
    """simple_sim (Parallel/RustVersion/src/array_kd_tree.rs:623-664): returns the
    bodies after ``steps`` kick-drift steps (positions and velocities updated)."""
    n = len(bodies["x"])
    with Simulation(n, theta=theta, dt=dt, leaf_size=leaf_size, precision=precision) as s:
        s.upload(bodies)
        s.step(steps)
        out = s.download()
    out["m"] = np.asarray(bodies["m"], dtype=np.float64).copy()
    return out


def build_tree(bodies: dict, leaf_size: int = 7, layout: int = KD_LAYOUT_COMPACT) -> Tree:
    """build_tree over all bodies (Sequential/RustVersion/src/array_kd_tree.rs:49)."""
    n = len(bodies["x"])
    with Simulation(n, leaf_size=leaf_size) as s:
        s.upload(bodies)
        s.build()
        return s.export_tree(layout)


def calc_accel(bodies: dict, theta: float = 0.3, leaf_size: int = 7,
               precision: int = KD_PREC_FP64) -> np.ndarray:
    """calc_accel(p) for every body p (array_kd_tree.rs:619), shape (n, 3)."""
    n = len(bodies["x"])
    with Simulation(n, theta=theta, leaf_size=leaf_size, precision=precision) as s:
        s.upload(bodies)
        return s.accel()


def calc_accel_tree(bodies: dict, tree: Tree, theta: float = 0.3,
                    idx: np.ndarray | None = None) -> np.ndarray:
    """calc_accel(p, particles, nodes) over a given node vector
    (array_kd_tree.rs:619 / Sequential :174) for the bodies ``idx`` (default:
    all), FP64; the tree is traversed as given, not rebuilt.  Shape (k, 3)."""
    n = len(bodies["x"])
    cols = [np.ascontiguousarray(bodies[f], np.float64) for f in ("x", "y", "z", "m")]
    v, keep = tree.view()
    if idx is None:
        k, ip = n, None
    else:
        idx = np.ascontiguousarray(idx, np.int64)
        k, ip = len(idx), iptr(idx)
    out = [np.zeros(k) for _ in range(3)]
    check(_lib.lib().kd_accel_tree(n, *[dptr(c) for c in cols], C.byref(v), theta, k, ip,
                                   *[dptr(o) for o in out]))
    del keep
    return np.stack(out, axis=1)


def quickstat_index(indices, goal: int, vals) -> np.ndarray:
    """quickstat_index(indices, goal, |i, j| vals[i] < vals[j])
    (Parallel/RustVersion/src/quickstat.rs:9-34) on the GPU; returns the
    reordered indices (ties in vals ordered by index)."""
    ix = np.ascontiguousarray(indices, np.int64).copy()
    v = np.ascontiguousarray(vals, np.float64)
    check(_lib.lib().kd_quickstat_index(len(ix), iptr(ix), goal, dptr(v), len(v)))
    return ix


def print_tree(step: int, tree: Tree, bodies: dict, directory: str = ".") -> str:
    """print_tree (Parallel/RustVersion/src/array_kd_tree.rs:666-692): writes
    tree{step}.txt in `directory` and returns its path."""
    import os
    path = os.path.join(directory, f"tree{step}.txt")
    cols = [np.ascontiguousarray(bodies[f], np.float64) for f in ("x", "y", "z")]
    v, keep = tree.view()
    check(_lib.lib().kd_print_tree(C.byref(v), len(cols[0]), *[dptr(c) for c in cols],
                                   path.encode()))
    del keep
    return path


def circular_orbits(n: int, seed: int = 0) -> dict:
    """circular_orbits(n) of array_particle.rs:19-44 (n+1 bodies), seeded."""
    return _lib.ic("circular_orbits", n, seed)


def two_bodies() -> dict:
    """two_bodies() of array_particle.rs:10-17."""
    return _lib.ic("two_bodies", 2)


def ring(n: int, seed: int = 0) -> dict:
    """The benchmark ring distribution: star of mass 1 plus n-1 bodies."""
    return _lib.ic("ring", n, seed)


__all__ = ["Simulation", "Batch", "ensemble", "Tree", "allocate_node_vec", "simple_sim", "build_tree",
           "calc_accel", "calc_accel_tree", "print_tree", "quickstat_index", "circular_orbits", "two_bodies", "ring",
           "KD_PREC_FP64", "KD_PREC_MIXED", "KD_PREC_MIXED_D64",
           "KD_LAYOUT_COMPACT", "KD_LAYOUT_PADDED"]
