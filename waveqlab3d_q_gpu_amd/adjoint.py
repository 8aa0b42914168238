"""Gradients of receiver traces with respect to per-node rho, lambda, mu.

The discrete adjoint of the fused RK step (wql_adjoint_* in libwql3d_hip.so)
behind a ``torch.autograd.Function``.  The reference has no adjoint; the
forward map is the reference's ``time_step_RK`` (time_step.f90:68-150) with
receivers sampled at stage 1 of every step (time_step.f90:137), i.e. the
library's ``wql_time_step``.

Memory: checkpoint-and-recompute.  The forward run keeps the state (F, eta) at
every C-th step start (C = ceil(sqrt(N)) by default) in library-held slots.
The reverse sweep walks the segments backwards: it recomputes one segment from
its checkpoint stage by stage, keeping F of every stage input (a reverse stage
reads F of its input state only), then for each step (last first) reverses
the five stages from those inputs; the adjoint of the step's receiver sample
(taken from the step's input state) is added after the five stages of that
step have been reversed.  Slots: ceil(N/C) checkpoints (F and eta) and 5C
stage inputs (F only); forward work 2N steps in total.
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
        self.s_stage = self.nck          # first F-only slot: stage inputs
        self.receivers = np.asarray(receivers)
        self.nrec = self.receivers.shape[0]
        self.ts = None
        check(lib().wql_adjoint_create_f(blk._h, self.nck, 5 * self.C))

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
        for seg in range(self.nck):
            n0, n1 = seg * self.C, min(self.N, (seg + 1) * self.C)
            check(lib().wql_state_save(blk._h, seg))
            # one call per segment (its steps' last-stage rates are not stored
            # but the segment's last); the step start times are those the call
            # forms, t0 + C(5) dt per step
            tn = t
            for _ in range(n0, n1):
                self.ts.append(tn)
                tn = tn + RK_C[5] * self.dt
            t = blk.time_step(self.dt, n1 - n0, t)
            if t != tn:
                raise RuntimeError("step times of the segment differ from the library's")
        return torch.from_numpy(blk.seismograms()).to(blk.M.device)

    # -- reverse sweep -------------------------------------------------------
    def _segment_stage_inputs(self, n0, n1):
        """Recompute steps n0 .. n1-1 from the current state (the segment's
        checkpoint), keeping F of each stage input in the F-only slots (the
        same fused stages, coefficients and stage times as wql_time_step)."""
        h, dt = self.blk._h, self.dt
        for n in range(n0, n1):
            t = self.ts[n]
            for s in range(5):
                check(lib().wql_state_save_part(h, self.s_stage + 5 * (n - n0) + s, L.WQL_F))
                check(lib().wql_stage_fused(h, RK_A[s], RK_B[s] * dt, t + RK_C[s] * dt))

    def _stage_adjoint(self, n, n0):
        """Reverse the five stages of step n from their kept inputs."""
        h, dt = self.blk._h, self.dt
        for s in reversed(range(5)):
            check(lib().wql_state_load_part(h, self.s_stage + 5 * (n - n0) + s, L.WQL_F))
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
            self._segment_stage_inputs(n0, n1)
            for n in reversed(range(n0, n1)):
                self._stage_adjoint(n, n0)
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
