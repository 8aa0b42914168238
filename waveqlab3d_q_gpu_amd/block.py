"""Host-side mirror of one block and of the reference's per-stage seams.

Mirrors the reference interfaces of the time-step hot path:

* ``device_data``: enter_domain_device_data / update_domain_host_data /
  exit_domain_device_data (device_data_cpu.f90:8-26) -> :class:`Block`
  create / :meth:`Block.download` / :meth:`Block.close`;
* ``rk_vector_kernels``: scale_rate_array / update_state_array
  (rk_vector_cpu.f90:7-19) -> :meth:`Block.scale_rates` /
  :meth:`Block.update_fields` over every rate/state array of the block, and
  :func:`scale_rate_array` / :func:`update_state_array` on single arrays;
* ``set_rates_elastic`` (elastic.f90:13-29) -> :meth:`Block.set_rates`;
* ``try_traditional_cartesian_rhs`` (traditional_cartesian_rhs_openacc.f90:10-123)
  -> :meth:`Block.try_interior_rhs`;
* ``set_moment_tensor`` (moment_tensor.f90:875-935) -> :meth:`Block.set_moment_tensor`;
* ``impose_boundary_condition`` (RHS_Interior.f90:1584-1768) ->
  :meth:`Block.impose_boundary_condition`.

Arrays follow the reference's Fortran layout: a numpy array of shape
``(ncomp, nz, ny, nx)`` in C order is the Fortran array ``A(x, y, z, c)``
with x fastest.  Host arrays are numpy float64; with ``on_device=True`` the
arrays are float64 torch tensors on the GPU (passed by device pointer).
All compute runs in libwql3d_hip.so; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import check, lib

TAU0 = 1.0 / 0.315949074074074  # block.f90:242

# Kennedy-Carpenter (5,4) low-storage coefficients (time_step.f90:53-62)
RK_A = (0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0)
RK_B = (1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
        2277821191437.0 / 14882151754819.0)
RK_C = (0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
        2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0, 1.0)


def init_RK():
    """init_RK (time_step.f90:20-65): order-4 KC(5,4) -> (A, B, C)."""
    return RK_A, RK_B, RK_C


@dataclass
class SourceSpec:
    """One moment-tensor point source (moment_tensor.f90:8-242)."""
    kind: int            # L.WQL_SRC_*
    t_init: float
    duration: float      # Gaussian width w / Ricker peak frequency
    near: tuple          # nearest node, 1-based
    near_phys: tuple
    location: tuple
    mxx: float = 0.0
    myy: float = 0.0
    mzz: float = 0.0
    mxy: float = 0.0
    mxz: float = 0.0
    myz: float = 0.0
    order: int = 4       # 2 or 4 (singular_source); 0: mollified (Gaussian, sigma = 2h)
    alpha: int = 0


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError("host arrays must be C-contiguous float64")
        return a.ctypes.data_as(C.POINTER(C.c_double))
    # torch tensor on the device
    if str(a.dtype) != "torch.float64" or not a.is_contiguous() or not a.is_cuda:
        raise ValueError("device arrays must be contiguous float64 CUDA tensors")
    return C.cast(C.c_void_p(a.data_ptr()), C.POINTER(C.c_double))


def _box(lo, hi):
    b = L.Box()
    b.lo[:] = list(lo) + [1] * (3 - len(lo))
    b.hi[:] = list(hi) + [1] * (3 - len(hi))
    return b


def pml_slab_box(face, n, npml):
    """Node range (1-based, inclusive) of PMLB(face)%Q (pml.f90:11-168)."""
    lo = [1, 1, 1]
    hi = list(n)
    ax = face // 2
    if face % 2 == 0:
        hi[ax] = npml
    else:
        lo[ax] = n[ax] - npml + 1
    return lo, hi


@dataclass
class Block:
    """One block mirrored on the GPU."""
    nq: int
    nr: int
    ns: int
    mesh: int
    metricx: object
    metricy: object
    metricz: object
    J: object
    M: object
    precision: int = L.WQL_FP64
    nmech: int = 0
    bc: tuple = (0, 0, 0, 0, 0, 0)
    pml: tuple = (0, 0, 0, 0, 0, 0)
    npml: tuple = (0, 0, 0, 0, 0, 0)
    tau: tuple = ()
    weight: tuple = ()
    F: object = None
    DF: object = None
    eta: object = None          # (6, nmech, nz, ny, nx)
    Deta: object = None
    Qs_inv: object = None
    Qp_inv: object = None
    pml_Q: list = field(default_factory=lambda: [None] * 6)
    pml_DQ: list = field(default_factory=lambda: [None] * 6)
    n_l: list = field(default_factory=lambda: [None] * 6)
    n_m: list = field(default_factory=lambda: [None] * 6)
    n_n: list = field(default_factory=lambda: [None] * 6)
    tau0: float = TAU0
    sat_form: int = L.WQL_SAT_BC
    on_device: bool = False
    _h: object = None

    def __post_init__(self):
        d = L.BlockDesc()
        n = (self.nq, self.nr, self.ns)
        d.nq, d.nr, d.ns = n
        d.mesh, d.precision, d.nmech = self.mesh, self.precision, self.nmech
        d.bc[:] = [int(v) for v in self.bc]
        d.pml[:] = [int(v) for v in self.pml]
        d.npml[:] = [int(v) for v in self.npml]
        d.arrays_on_device = 1 if self.on_device else 0
        d.hq, d.hr, d.hs = (1.0 / (self.nq - 1), 1.0 / (self.nr - 1), 1.0 / (self.ns - 1))
        d.tau0 = self.tau0
        d.sat_form = int(self.sat_form)
        d.tau[:] = list(self.tau) + [0.0] * (8 - len(self.tau))
        d.weight[:] = list(self.weight) + [0.0] * (8 - len(self.weight))
        d.body = _box((1, 1, 1), n)
        for name in ("metricx", "metricy", "metricz", "J", "M", "F", "DF", "Qs_inv", "Qp_inv"):
            setattr(d, name, _ptr(getattr(self, name)))
        if self.nmech:
            for c in range(6):
                d.eta[c] = _ptr(self.eta[c]) if self.eta is not None else None
                d.Deta[c] = _ptr(self.Deta[c]) if self.Deta is not None else None
        for f in range(6):
            if self.pml[f]:
                lo, hi = pml_slab_box(f, n, self.npml[f])
                d.pml_box[f] = _box(lo, hi)
                d.pml_Q[f] = _ptr(self.pml_Q[f])
                d.pml_DQ[f] = _ptr(self.pml_DQ[f])
            ax = f // 2
            fn = [(self.nr, self.ns), (self.nr, self.ns), (self.nq, self.ns), (self.nq, self.ns),
                  (self.nq, self.nr), (self.nq, self.nr)][f]
            d.face_box[f] = _box((1, 1), fn)
            d.n_l[f] = _ptr(self.n_l[f])
            d.n_m[f] = _ptr(self.n_m[f])
            d.n_n[f] = _ptr(self.n_n[f])
            del ax
        self._desc = d
        h = C.c_void_p()
        check(lib().wql_block_create(C.byref(d), C.byref(h)))
        self._h = h
        self._nslots = 0
        self._nrec = 0

    # -- device_data seam ----------------------------------------------------
    def upload(self, mask=L.WQL_ALL):
        check(lib().wql_block_upload(self._h, mask))

    def download(self, mask=L.WQL_F | L.WQL_DF | L.WQL_MEMORY | L.WQL_PML):
        check(lib().wql_block_download(self._h, mask))

    def init_state(self):
        check(lib().wql_block_init_state(self._h))

    def device_bytes(self):
        return int(lib().wql_block_device_bytes(self._h))

    def close(self):
        if self._h is not None and self._h.value:
            lib().wql_block_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- sources and receivers ----------------------------------------------
    def set_sources(self, sources):
        arr = (L.Source * max(1, len(sources)))()
        for i, s in enumerate(sources):
            o = arr[i]
            o.type, o.order, o.alpha = s.kind, s.order, s.alpha
            o.near[:] = [int(v) for v in s.near]
            o.duration, o.t_init = s.duration, s.t_init
            o.mXX, o.mYY, o.mZZ, o.mXY, o.mXZ, o.mYZ = s.mxx, s.myy, s.mzz, s.mxy, s.mxz, s.myz
            o.location[:] = list(s.location)
            o.near_phys[:] = list(s.near_phys)
        check(lib().wql_block_set_sources(self._h, len(sources), arr))

    def set_coordinates(self, X):
        """Node coordinates (3, nz, ny, nx) float64, a numpy array or a CUDA
        tensor, for mollified sources (order 0)."""
        if hasattr(X, "data_ptr"):
            if X.dtype != __import__("torch").float64 or not X.is_contiguous():
                raise ValueError("coordinates: contiguous float64 tensor expected")
            check(lib().wql_block_set_coordinates(self._h, C.c_void_p(X.data_ptr()),
                                                  1 if X.is_cuda else 0))
        else:
            X = np.ascontiguousarray(X, dtype=np.float64)
            check(lib().wql_block_set_coordinates(self._h, X.ctypes.data_as(C.c_void_p), 0))

    def set_receivers(self, ijk, nslots):
        ijk = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
        check(lib().wql_block_set_receivers(
            self._h, ijk.shape[0], ijk.ctypes.data_as(C.POINTER(C.c_int32)), nslots))
        self._nrec, self._nslots = ijk.shape[0], nslots

    def seismograms(self, nslots=None):
        nslots = self._nslots if nslots is None else nslots
        out = np.zeros((nslots, self._nrec, 9))
        if nslots and self._nrec:
            check(lib().wql_block_get_seismograms(self._h, nslots,
                                                  out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- stage seams ---------------------------------------------------------
    def scale_rates(self, A):
        check(lib().wql_rk_scale(self._h, A))

    def update_fields(self, inc):
        check(lib().wql_rk_update(self._h, inc))

    def set_rates(self):
        check(lib().wql_rhs(self._h))

    def try_interior_rhs(self):
        handled = C.c_int32(0)
        check(lib().wql_rhs_interior(self._h, C.byref(handled)))
        return bool(handled.value)

    def set_moment_tensor(self, t):
        check(lib().wql_set_moment_tensor(self._h, t))

    def impose_boundary_condition(self):
        check(lib().wql_sat(self._h))

    def stage_fused(self, A, Bdt, t):
        check(lib().wql_stage_fused(self._h, A, Bdt, t))

    def time_step(self, dt, nsteps, t=0.0):
        tt = C.c_double(t)
        check(lib().wql_time_step(self._h, dt, nsteps, C.byref(tt)))
        return tt.value

    def sync(self):
        check(lib().wql_block_sync(self._h))

    def stream(self):
        return lib().wql_block_stream(self._h)

    def set_timing(self, enable: bool = True):
        """Record per-stage kernel durations (HIP events on the block stream)."""
        check(lib().wql_block_set_timing(self._h, 1 if enable else 0))

    def max_state(self):
        """close_domain's shutdown check (domain.f90:520-545) on the device:
        (max|F|, max|eta|, number of non-finite values)."""
        out = (C.c_double * 3)()
        check(lib().wql_block_max_state(self._h, out))
        return out[0], out[1], int(out[2])

    def max_state_noalloc(self, out_dev):
        """The same check without a device allocation: the three results go
        to ``out_dev`` (a float64 CUDA tensor of 3), ordered on the block's
        stream (call :meth:`sync` before reading it)."""
        if str(out_dev.dtype) != "torch.float64" or out_dev.numel() < 3 or not out_dev.is_cuda:
            raise ValueError("out_dev must be a float64 CUDA tensor of at least 3 values")
        check(lib().wql_block_max_state_noalloc(self._h, C.c_void_p(out_dev.data_ptr())))

    def kernel_times(self):
        """([interior_ms, shell_ms, memory_ms], stages) summed since the last call."""
        ms, n = (C.c_double * 3)(), C.c_int32()
        check(lib().wql_block_kernel_times(self._h, ms, C.byref(n)))
        return list(ms), n.value


def time_step_RK(block: Block, dt: float, nsteps: int, t: float = 0.0, fused: bool = True):
    """time_step_RK (time_step.f90:68-150) for ``nsteps`` steps.

    ``fused=False`` runs the reference stage order with one kernel per seam:
    scale_rates -> set_rates -> set_moment_tensor -> impose_boundary_condition
    -> update_fields.  ``fused=True`` runs one fused kernel per stage.
    Receivers set on the block are sampled at stage 1 (fused path).
    Returns the advanced time."""
    if fused:
        return block.time_step(dt, nsteps, t)
    A, B, Cc = init_RK()
    for _ in range(nsteps):
        t0 = t
        for s in range(5):
            block.scale_rates(A[s])
            block.set_rates()
            block.set_moment_tensor(t)
            block.impose_boundary_condition()
            t = t0 + Cc[s + 1] * dt
            block.update_fields(B[s] * dt)
    return t


def scale_rate_array(rate_dev, coefficient):
    """scale_rate_array (rk_vector_cpu.f90:7-11) on a float64 CUDA tensor."""
    check(lib().wql_scale_rate_array(C.c_void_p(rate_dev.data_ptr()), rate_dev.numel(),
                                     coefficient))


def update_state_array(state_dev, rate_dev, increment):
    """update_state_array (rk_vector_cpu.f90:13-19) on float64 CUDA tensors."""
    if tuple(state_dev.shape) != tuple(rate_dev.shape):
        raise ValueError("RK state/rate shape mismatch")
    check(lib().wql_update_state_array(C.c_void_p(state_dev.data_ptr()),
                                       C.c_void_p(rate_dev.data_ptr()), state_dev.numel(),
                                       increment))


def grid_metrics(X_dev):
    """derivatives() (grid.f90:398-471) on the device: X (3, nz, ny, nx) float64
    CUDA tensor -> (metricx, metricy, metricz, J) CUDA tensors."""
    import torch
    _, nz, ny, nx = X_dev.shape
    mx = torch.empty_like(X_dev); my = torch.empty_like(X_dev); mz = torch.empty_like(X_dev)
    J = torch.empty((nz, ny, nx), dtype=X_dev.dtype, device=X_dev.device)
    torch.cuda.synchronize()
    check(lib().wql_grid_metrics(nx, ny, nz, C.c_void_p(X_dev.data_ptr()),
                                 C.c_void_p(mx.data_ptr()), C.c_void_p(my.data_ptr()),
                                 C.c_void_p(mz.data_ptr()), C.c_void_p(J.data_ptr())))
    return mx, my, mz, J
