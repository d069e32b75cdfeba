"""Differentiable simulation: torch.autograd over the C ABI's reverse sweep.

``simulate(x0, v0, m, sim)`` runs ``sim.steps`` steps of simple_sim
(Parallel/RustVersion/src/array_kd_tree.rs:623-664) with the tree topology
frozen within each step and returns the final positions and velocities; its
backward is the hand-written adjoint kernel (kd_adj_backward), with a
checkpoint of the state every ``checkpoint_every`` steps and each segment's
trees recomputed from its checkpoint.  Tensors are float64 on the GPU, shape
(n, 3) for positions and velocities and (n,) for masses.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import KD_PREC_MIXED, check


class Adjoint:
    """One kd_adj context: n bodies, a fixed number of steps."""

    def __init__(self, n: int, steps: int, checkpoint_every: int = 20, theta: float = 0.3,
                 dt: float = 1e-3, leaf_size: int = 7, precision: int = KD_PREC_MIXED):
        self._L = _lib.lib()
        p = _lib.default_params(theta=theta, dt=dt, leaf_size=leaf_size, precision=precision)
        h = C.c_void_p()
        check(self._L.kd_adj_create(n, steps, checkpoint_every, C.byref(p), C.byref(h)))
        self._h = h
        self.n, self.steps = n, steps
        # one context holds the checkpoints of its latest forward only; every
        # forward (successful or not) starts a new generation, and a backward
        # recorded against an older one is refused
        self.generation = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.kd_adj_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _cols(t: torch.Tensor):
        assert t.is_cuda and t.dtype == torch.float64
        return [t[:, c].contiguous() for c in range(3)]

    def forward(self, x0: torch.Tensor, v0: torch.Tensor, m: torch.Tensor):
        n = self.n
        assert x0.shape == (n, 3) and v0.shape == (n, 3) and m.shape == (n,)
        ins = self._cols(x0) + self._cols(v0) + [m.contiguous()]
        outs = [torch.empty(n, dtype=torch.float64, device=x0.device) for _ in range(6)]
        torch.cuda.current_stream().synchronize()
        self.generation += 1
        check(self._L.kd_adj_forward(self._h, *[t.data_ptr() for t in ins],
                                     *[t.data_ptr() for t in outs]))
        return torch.stack(outs[:3], 1), torch.stack(outs[3:], 1)

    def backward(self, gp: torch.Tensor | None, gv: torch.Tensor | None):
        n = self.n
        dev = (gp if gp is not None else gv).device
        gins = (self._cols(gp) if gp is not None else [None] * 3) + \
            (self._cols(gv) if gv is not None else [None] * 3)
        outs = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(7)]
        torch.cuda.current_stream().synchronize()
        check(self._L.kd_adj_backward(self._h, *[t.data_ptr() if t is not None else None
                                                 for t in gins],
                                      *[t.data_ptr() for t in outs]))
        return torch.stack(outs[0:3], 1), torch.stack(outs[3:6], 1), outs[6]


class _Simulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, sim):
        ctx.sim = sim
        out = sim.forward(x0, v0, m)
        ctx.generation = sim.generation
        return out

    @staticmethod
    def backward(ctx, gp, gv):
        if ctx.sim.generation != ctx.generation:
            raise RuntimeError(
                "this Adjoint context ran another forward after the simulate() call being "
                "differentiated, so its checkpoints belong to the later run; use one Adjoint "
                "per differentiated simulate() call")
        gx, gvv, gm = ctx.sim.backward(gp, gv)
        return gx, gvv, gm, None


def simulate(x0: torch.Tensor, v0: torch.Tensor, m: torch.Tensor, sim: Adjoint):
    """Final (positions, velocities); differentiable in x0, v0 and m."""
    return _Simulate.apply(x0, v0, m, sim)


__all__ = ["Adjoint", "simulate"]
