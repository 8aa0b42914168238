"""Gradients of receiver traces with respect to per-node rho, lambda, mu.

The discrete adjoint of the fused RK step (wql_adjoint_* in libwql3d_hip.so)
behind a ``torch.autograd.Function``.  The reference has no adjoint; the
forward map is the reference's ``time_step_RK`` (time_step.f90:68-150) with
receivers sampled at stage 1 of every step (time_step.f90:137), i.e. the
library's ``wql_time_step``.

Memory: checkpoint-and-recompute.  The forward run keeps the state (F, eta) at
every C-th step start (C = ceil(sqrt(N)) by default) in library-held slots.
The reverse sweep walks the segments backwards: it recomputes the step starts
of one segment from its checkpoint, then for each step (last first) recomputes
the five stage inputs and reverses the five stages; the adjoint of the step's
receiver sample (taken from the step's input state) is added after the five
stages of that step have been reversed.  Slots:
ceil(N/C) checkpoints + C step starts + 5 stage inputs; forward work 3N steps
in total.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import check, lib
from .block import RK_A, RK_B, RK_C


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class AdjointRunner:
    """Forward run with checkpoints and the reverse sweep on one block.

    ``blk`` must come from ``problems.make_device_block(..., state_mirror=True)``
    (device material and state mirrors) and have no PML.  ``receivers`` are
    1-based node indices (nrec, 3)."""

    def __init__(self, blk, dt, nsteps, receivers, t0=0.0, every=None):
        self.blk = blk
        self.dt = float(dt)
        self.N = int(nsteps)
        self.t0 = float(t0)
        self.C = int(every) if every else max(1, int(math.ceil(math.sqrt(self.N))))
        self.nck = (self.N + self.C - 1) // self.C
        self.s_seg = self.nck
        self.s_stage = self.nck + self.C
        self.receivers = np.asarray(receivers)
        self.nrec = self.receivers.shape[0]
        self.ts = None
        check(lib().wql_adjoint_create(blk._h, self.s_stage + 5))

    def set_material(self, rho, lam, mu):
        """Per-node material (torch float64 tensors (nz, ny, nx) on the GPU)."""
        M = self.blk.M
        M[0].copy_(lam)
        M[1].copy_(mu)
        M[2].copy_(rho)
        self.blk.upload(L.WQL_MATERIAL)

    # -- forward -------------------------------------------------------------
    def forward(self):
        """Run N steps from rest; returns traces (N, nrec, 9) on the GPU."""
        import torch
        blk = self.blk
        blk.init_state()
        blk.set_receivers(self.receivers, self.N)
        t = self.t0
        self.ts = []
        for n in range(self.N):
            if n % self.C == 0:
                check(lib().wql_state_save(blk._h, n // self.C))
            self.ts.append(t)
            t = blk.time_step(self.dt, 1, t)
        return torch.from_numpy(blk.seismograms()).to(blk.M.device)

    # -- reverse sweep -------------------------------------------------------
    def _stage_adjoint(self, n):
        """Reverse the five stages of step n (its start state is current).
        A reverse stage reads F of its input state only, so the stage inputs
        are kept without eta."""
        h, dt, t = self.blk._h, self.dt, self.ts[n]
        for s in range(5):
            check(lib().wql_state_save_part(h, self.s_stage + s, L.WQL_F))
            check(lib().wql_stage_fused(h, RK_A[s], RK_B[s] * dt, t + RK_C[s] * dt))
        for s in reversed(range(5)):
            check(lib().wql_state_load_part(h, self.s_stage + s, L.WQL_F))
            check(lib().wql_adjoint_stage(h, RK_A[s], RK_B[s] * dt))

    def backward(self, g_traces):
        """Adjoint of the traces -> (g_rho, g_lambda, g_mu), each (nz, ny, nx)."""
        import torch
        blk = self.blk
        h = blk._h
        dev = blk.M.device
        g = g_traces.detach().to(device=dev, dtype=torch.float64).contiguous()
        check(lib().wql_adjoint_reset(h))
        for seg in reversed(range(self.nck)):
            n0, n1 = seg * self.C, min(self.N, (seg + 1) * self.C)
            check(lib().wql_state_load(h, seg))
            for n in range(n0, n1):
                check(lib().wql_state_save(h, self.s_seg + n - n0))
                blk.time_step(self.dt, 1, self.ts[n])
            for n in reversed(range(n0, n1)):
                check(lib().wql_state_load(h, self.s_seg + n - n0))
                self._stage_adjoint(n)
                check(lib().wql_adjoint_add_receivers(h, _vp(g[n])))
        nz, ny, nx = blk.ns, blk.nr, blk.nq
        out = [torch.zeros((nz, ny, nx), dtype=torch.float64, device=dev) for _ in range(3)]
        check(lib().wql_adjoint_gradient(h, *[_vp(o) for o in out]))
        return tuple(out)

    # -- the bare operators, for the dot-product test ---------------------------
    def adjoint_steps(self, Fbar, etabar, nsteps):
        """Apply the transpose of nsteps source-free RK steps to (Fbar, etabar)
        (dense device tensors) in place; the material gradient accumulated on
        the way is not meaningful here."""
        h = self.blk._h
        check(lib().wql_adjoint_reset(h))
        check(lib().wql_adjoint_set_state(h, _vp(Fbar), _vp(etabar)))
        for _ in range(nsteps):
            for s in reversed(range(5)):
                check(lib().wql_adjoint_stage(h, RK_A[s], RK_B[s] * self.dt))
        check(lib().wql_adjoint_get_state(h, _vp(Fbar), _vp(etabar)))


def _function():
    import torch

    class ReceiverTraces(torch.autograd.Function):
        """traces = f(rho, lambda, mu): the receiver samples of N steps."""

        @staticmethod
        def forward(ctx, rho, lam, mu, runner):
            runner.set_material(rho, lam, mu)
            ctx.runner = runner
            return runner.forward()

        @staticmethod
        def backward(ctx, g):
            g_rho, g_lam, g_mu = ctx.runner.backward(g)
            return g_rho, g_lam, g_mu, None

    return ReceiverTraces


_RT = None


def receiver_traces(rho, lam, mu, runner):
    """Differentiable receiver traces (N, nrec, 9) of the block driven by
    ``runner`` with material (rho, lambda, mu)."""
    global _RT
    if _RT is None:
        _RT = _function()
    return _RT.apply(rho, lam, mu, runner)
