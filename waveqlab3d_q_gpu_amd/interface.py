"""Two-block domain: two :class:`Block` objects coupled across a q-interface.

Host-side mirror of the reference's two-block path: iface_type and init_iface
(datatypes.f90, iface.f90:13-55), enforce_bound_iface_conditions
(domain.f90:615-644), time_step_RK over a domain (time_step.f90:68-150) and the
fault output (write_fault_output / write_hat_output, domain.f90:782-847).  The
compute is ``wql_iface_*`` of ``libwql3d_hip.so``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib
from .block import Block, _ptr

# fault-output planes of one recorded step, in the order of the reference's
# files (fault_output.f90:30-46): name -> number of components
FRAME_KEYS = (("Uface", 9), ("Vface", 9), ("S", 4), ("Svel", 3), ("Uhat", 6), ("Vhat", 6),
              ("trup", 1), ("state", 1))
FRAME_PLANES = sum(n for _, n in FRAME_KEYS)


class Interface:
    """Couples the x-high face (bc 0) of ``left`` to the x-low face (bc 0) of
    ``right``.  ``y``, ``z``: coordinates of the interface nodes, arrays of
    shape (nz, ny) (host numpy or device torch float64).  ``coupling``:
    ``WQL_IFACE_LOCKED``, ``WQL_IFACE_SLIP_WEAKENING_TPV5``,
    ``WQL_IFACE_SLIP_WEAKENING_TPV28`` (``s0``: its initial stress
    (xx, yy, zz, xy, xz, yz), default the source's), or
    ``WQL_IFACE_SLIP_WEAKENING`` with ``tables`` = dict of host arrays T0L,
    T0R (3, nz, ny) (prestress in each side's (n, m, l) face frame) and fs,
    fd, Dc (nz, ny), or ``WQL_IFACE_RATE_STATE_TPV101`` / ``_TPV102`` /
    ``_RW`` (problem 'rate-weakening') with the initial state ``W0`` (nz, ny)
    (default 0, as init_iface sets it), or ``WQL_IFACE_LINEAR_FRICTION``
    with ``alpha`` (tau = alpha v; None = the reference's 1e14) and
    optionally ``tables`` = dict with T0L, T0R, the prestress the fault
    output adds.

    The blocks must outlive the interface; call :meth:`close` first."""

    def __init__(self, left: Block, right: Block, y, z, coupling=L.WQL_IFACE_SLIP_WEAKENING_TPV5,
                 s0=None, tables=None, W0=None, alpha=None):
        self.left, self.right = left, right
        self.ny, self.nz = left.nr, left.ns
        d = L.IfaceDesc()
        d.coupling = int(coupling)
        d.alpha = -1.0 if alpha is None else float(alpha)
        on_dev = not isinstance(y, np.ndarray)
        if not on_dev:
            y = np.ascontiguousarray(y, dtype=np.float64)
            z = np.ascontiguousarray(z, dtype=np.float64)
        if tuple(y.shape) != (self.nz, self.ny) or tuple(z.shape) != (self.nz, self.ny):
            raise ValueError(f"face coordinates must be ({self.nz}, {self.ny})")
        keep = [y, z]              # keep the arrays alive during create
        d.arrays_on_device = 1 if on_dev else 0
        d.y, d.z = _ptr(y), _ptr(z)
        if s0 is not None:
            s0a = np.ascontiguousarray(s0, dtype=np.float64).reshape(6)
            keep.append(s0a)
            d.s0 = s0a.ctypes.data_as(C.POINTER(C.c_double))
        if tables is not None:
            if on_dev:
                raise ValueError("tables are host arrays; pass host y, z with them")
            for name, shape in (("T0L", (3,)), ("T0R", (3,)), ("fs", ()), ("fd", ()), ("Dc", ())):
                if name not in tables and coupling == L.WQL_IFACE_LINEAR_FRICTION:
                    continue
                a = np.ascontiguousarray(tables[name], dtype=np.float64).reshape(
                    shape + (self.nz, self.ny))
                keep.append(a)
                setattr(d, name, _ptr(a))
        if W0 is not None:
            # rate-and-state: initial state psi (nz, ny), same array kind as y, z
            if on_dev:
                if tuple(W0.shape) != (self.nz, self.ny):
                    raise ValueError(f"W0 must be ({self.nz}, {self.ny})")
                w = W0
            else:
                w = np.ascontiguousarray(W0, dtype=np.float64).reshape(self.nz, self.ny)
            keep.append(w)
            d.W0 = _ptr(w)
        self._keep = keep
        h = C.c_void_p()
        check(lib().wql_iface_create(left._h, right._h, C.byref(d), C.byref(h)))
        self._h = h
        self._nslots = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wql_iface_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stage(self, A, Bdt, t, stage1=False):
        """One RK stage of both blocks and the interface."""
        check(lib().wql_iface_stage(self._h, A, Bdt, t, 1 if stage1 else 0))

    def time_step(self, dt, nsteps, t=0.0):
        """time_step_RK for ``nsteps`` steps; returns the advanced time."""
        tt = C.c_double(t)
        check(lib().wql_iface_time_step(self._h, dt, nsteps, C.byref(tt)))
        return tt.value

    def set_frames(self, nslots):
        """Record the fault output of the next ``nslots`` steps."""
        check(lib().wql_iface_set_frames(self._h, nslots))
        self._nslots = nslots

    def frames(self, nslots=None):
        """Recorded fault output: dict name -> array (nslots, ncomp, nz, ny)."""
        nslots = self._nslots if nslots is None else nslots
        raw = np.zeros((nslots, FRAME_PLANES, self.nz, self.ny))
        if nslots:
            check(lib().wql_iface_get_frames(self._h, nslots,
                                             raw.ctypes.data_as(C.POINTER(C.c_double))))
        out, k = {}, 0
        for name, n in FRAME_KEYS:
            out[name] = raw[:, k:k + n].copy()
            k += n
        return out

    def state(self):
        """(S (4, nz, ny), DS (4, nz, ny), rupture time (nz, ny))."""
        S = np.zeros((4, self.nz, self.ny))
        DS = np.zeros((4, self.nz, self.ny))
        tr = np.zeros((self.nz, self.ny))
        P = C.POINTER(C.c_double)
        check(lib().wql_iface_get_state(self._h, S.ctypes.data_as(P), DS.ctypes.data_as(P),
                                        tr.ctypes.data_as(P)))
        return S, DS, tr

    def friction_state(self):
        """Rate-and-state state variable psi (nz, ny) (iface_type%W)."""
        W = np.zeros((self.nz, self.ny))
        check(lib().wql_iface_get_friction_state(self._h, W.ctypes.data_as(C.POINTER(C.c_double))))
        return W

    def sync(self):
        self.left.sync()
        self.right.sync()
