"""ctypes wrapper of the CPU restatement in ``wql_oracle.c``.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP product
path is compared against; it is imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` and by
nothing in the product package.

Arrays use the layout documented in ``wql_oracle.h``: numpy arrays of shape
``(ncomp, nz, ny, nx)`` (x fastest), float64, C-contiguous.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libwql_oracle.so")

SRC_GAUSSIAN, SRC_COSBELL, SRC_RICKER = 0, 1, 2
TAU0 = 1.0 / 0.315949074074074  # block.f90:242


def build(force: bool = False) -> str:
    """Compile the restatement with ``make`` (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "libwql_oracle.so"],
                       check=True)
    return _LIB_PATH


class _Source(C.Structure):
    _fields_ = [("type", C.c_int), ("order", C.c_int), ("alpha", C.c_int),
                ("near", C.c_int * 3),
                ("duration", C.c_double), ("t_init", C.c_double),
                ("mXX", C.c_double), ("mYY", C.c_double), ("mZZ", C.c_double),
                ("mXY", C.c_double), ("mXZ", C.c_double), ("mYZ", C.c_double),
                ("loc", C.c_double * 3), ("near_phys", C.c_double * 3)]


_P = C.POINTER(C.c_double)


class _Block(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("curvilinear", C.c_int), ("nmech", C.c_int), ("nsrc", C.c_int),
                ("bc", C.c_int * 6), ("pml", C.c_int * 6), ("npml", C.c_int * 6),
                ("sat_form", C.c_int), ("hq", C.c_double), ("hr", C.c_double), ("hs", C.c_double),
                ("tau0", C.c_double),
                ("tau", C.c_double * 8), ("weight", C.c_double * 8),
                ("F", _P), ("DF", _P),
                ("metx", _P), ("mety", _P), ("metz", _P), ("J", _P),
                ("M", _P), ("eta", _P), ("Deta", _P),
                ("Qs_inv", _P), ("Qp_inv", _P),
                ("pQ", _P * 6), ("pDQ", _P * 6),
                ("nl", _P * 6), ("nm", _P * 6), ("nn", _P * 6),
                ("src", C.POINTER(_Source)), ("X", _P)]


class _Iface(C.Structure):
    _fields_ = [("ny", C.c_int), ("nz", C.c_int), ("coupling", C.c_int),
                ("s0", C.c_double * 6), ("yf", _P), ("zf", _P), ("S", _P), ("DS", _P), ("W", _P), ("DW", _P),
                ("Uhat", _P), ("Vhat", _P), ("Svel", _P), ("trup", _P), ("alpha", C.c_double)]


IFACE_LOCKED, IFACE_SLIP_WEAKENING_TPV5, IFACE_SLIP_WEAKENING_TPV28 = 0, 1, 2
IFACE_RATE_STATE_TPV101, IFACE_RATE_STATE_TPV102 = 4, 5   # the library's coupling codes
IFACE_LINEAR_FRICTION, IFACE_RATE_STATE_RW = 6, 7

_lib = None


def build_planted(plant: int, outdir: str) -> str:
    """Compile the restatement with the planted mistake ``ORC_PLANT == plant``
    (wql_oracle.c: 1 lambda for lambda + 2 mu in the Qp term, 2 the PML
    feedback sign flipped, 3 the PML auxiliary source sign flipped, 4 2 mu for
    mu in the shear memory source) into ``outdir``: the tests that must catch
    each mistake run against it."""
    out = os.path.join(outdir, f"libwql_oracle_plant{plant}.so")
    src = [os.path.join(_HERE, f) for f in ("wql_oracle.c", "wql_oracle_iface.c")]
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    f"-DORC_PLANT={int(plant)}", "-shared", "-o", out, *src, "-lm"], check=True)
    return out


@contextlib.contextmanager
def using_library(path: str):
    """Route every oracle call to the library at ``path`` (a planted build)
    inside the block."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _load(_LIB_PATH)
    return _lib


def _load(path):
    L = C.CDLL(path)
    L.orc_rk_coefficients.argtypes = [_P, _P, _P]
    L.orc_sbp_closure.argtypes = [_P]
    L.orc_apply_d1.argtypes = [C.c_int, C.c_double, _P, _P]
    L.orc_metrics.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]
    L.orc_normals.argtypes = [C.POINTER(_Block)]
    L.orc_q_coefficients.argtypes = [C.c_int, C.c_double, C.c_double, _P, _P]
    L.orc_q_modulus_correction.argtypes = [C.POINTER(_Block), C.c_double]
    L.orc_q_modulus_correction.restype = C.c_int
    for name in ("orc_set_rates", "orc_impose_boundary_condition"):
        getattr(L, name).argtypes = [C.POINTER(_Block)]
    L.orc_scale_rates.argtypes = [C.POINTER(_Block), C.c_double]
    L.orc_update_fields.argtypes = [C.POINTER(_Block), C.c_double]
    L.orc_set_moment_tensor.argtypes = [C.POINTER(_Block), C.c_double]
    L.orc_time_step.argtypes = [C.POINTER(_Block), C.c_double, C.c_int, _P,
                                C.c_int, C.POINTER(C.c_long), _P]
    L.orc_source_time.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.orc_source_time.restype = C.c_double
    L.orc_num_threads.restype = C.c_int
    L.orc_set_num_threads.argtypes = [C.c_int]
    L.orc_iface_enforce.argtypes = [C.POINTER(_Iface), C.POINTER(_Block), C.POINTER(_Block),
                                    C.c_double, C.c_int, C.c_int]
    L.orc_iface_scale.argtypes = [C.POINTER(_Iface), C.c_double]
    L.orc_iface_update.argtypes = [C.POINTER(_Iface), C.c_double]
    L.orc_rs_regula_falsi.argtypes = [C.c_double] * 7
    L.orc_rs_regula_falsi.restype = C.c_double
    L.orc_rs_state_rate.argtypes = [C.c_double] * 6
    L.orc_rs_state_rate.restype = C.c_double
    L.orc_rs_a.argtypes = [C.c_double, C.c_double]
    L.orc_rs_a.restype = C.c_double
    for name in ("orc_rw_a", "orc_rw_vw"):
        getattr(L, name).argtypes = [C.c_double, C.c_double]
        getattr(L, name).restype = C.c_double
    L.orc_rw_state_rate.argtypes = [C.c_double] * 4
    L.orc_rw_state_rate.restype = C.c_double
    L.orc_rw_prestress.argtypes = [_P, _P, _P] + [C.c_double] * 3 + [_P]
    L.orc_welded_hats.argtypes = [_P, _P] + [C.c_double] * 4 + [_P, _P]
    L.orc_linear_friction_hats.argtypes = [_P, _P] + [C.c_double] * 5 + [_P, _P]
    return L


def rs_regula_falsi(V, Phi, eta, sigma_n, psi, V0, a):
    """Regula_Falsi of the restatement: the slip rate from the guess V."""
    return lib().orc_rs_regula_falsi(V, Phi, eta, sigma_n, psi, V0, a)


def rs_state_rate(W, Vel, L0=0.02, f0=0.6, b=0.012, V0=1.0e-6):
    return lib().orc_rs_state_rate(W, Vel, L0, f0, b, V0)


def rs_a(y, z):
    return lib().orc_rs_a(y, z)


def rw_state_rate(W, Vel, a, Vw):
    """State rate of problem 'rate-weakening' (flash-weakened steady state)."""
    return lib().orc_rw_state_rate(W, Vel, a, Vw)


def rw_a(y, z):
    return lib().orc_rw_a(y, z)


def rw_vw(y, z):
    return lib().orc_rw_vw(y, z)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def rk_coefficients():
    A = np.zeros(5); B = np.zeros(5); Cc = np.zeros(6)
    lib().orc_rk_coefficients(_ptr(A), _ptr(B), _ptr(Cc))
    return A, B, Cc


def sbp_closure():
    t = np.zeros(54)
    lib().orc_sbp_closure(_ptr(t))
    return t.reshape(6, 9)


def apply_d1(u, h):
    """SBP-6 D1 of the restatement applied to a 1-D array."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.zeros_like(u)
    lib().orc_apply_d1(u.size, h, _ptr(u), _ptr(du))
    return du


def metrics(X):
    """derivatives() restatement. X: (3, nz, ny, nx). Returns metx, mety, metz, J."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    _, nz, ny, nx = X.shape
    metx = np.zeros((3, nz, ny, nx)); mety = np.zeros_like(metx); metz = np.zeros_like(metx)
    J = np.zeros((nz, ny, nx))
    lib().orc_metrics(nx, ny, nz, _ptr(X), _ptr(metx), _ptr(mety), _ptr(metz), _ptr(J))
    return metx, mety, metz, J


def q_coefficients(nmech, fmin=0.05, fmax=20.0):
    tau = np.zeros(8); w = np.zeros(8)
    lib().orc_q_coefficients(nmech, fmin, fmax, _ptr(tau), _ptr(w))
    return tau[:nmech].copy(), w[:nmech].copy()


def source_time(kind, t, t0, w):
    return lib().orc_source_time(kind, t, t0, w)


def num_threads():
    """OpenMP threads of the node loops (OMP_NUM_THREADS, or set_num_threads)."""
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


class OracleBlock:
    """One block of the restatement.  Fill the public numpy attributes, then
    call the stage routines or :meth:`time_step`."""

    def __init__(self, nx, ny, nz, *, curvilinear, hq=None, hr=None, hs=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.curvilinear = int(curvilinear)
        self.hq = 1.0 / (nx - 1) if hq is None else hq
        self.hr = 1.0 / (ny - 1) if hr is None else hr
        self.hs = 1.0 / (nz - 1) if hs is None else hs
        self.tau0 = TAU0
        shp = (nz, ny, nx)
        self.F = np.zeros((9,) + shp)
        self.DF = np.full((9,) + shp, 1.0e40)          # fields.f90:39
        self.metx = np.zeros((3,) + shp); self.mety = np.zeros((3,) + shp)
        self.metz = np.zeros((3,) + shp); self.J = np.zeros(shp)
        self.M = np.zeros((3,) + shp)
        self.bc = [-1] * 6
        self.pml = [0] * 6
        self.npml = [0] * 6
        # 0: the SAT penalty of the current source; 1: the alpha/theta form
        # behind the archived golden files (wql_oracle.h, orc_block.sat_form)
        self.sat_form = 0
        self.nmech = 0
        self.tau = np.zeros(8); self.weight = np.zeros(8)
        self.eta = None; self.Deta = None; self.Qs_inv = None; self.Qp_inv = None
        self.pQ = [None] * 6; self.pDQ = [None] * 6
        fshape = [(3, nz, ny), (3, nz, ny), (3, nz, nx), (3, nz, nx), (3, ny, nx), (3, ny, nx)]
        self.nl = [np.zeros(s) for s in fshape]
        self.nm = [np.zeros(s) for s in fshape]
        self.nn = [np.zeros(s) for s in fshape]
        self.sources = []
        self.X = None            # (3, nz, ny, nx) coordinates, for mollified sources
        self._struct = None

    # -- setup helpers --------------------------------------------------
    def set_anelastic(self, nmech, Qs_inv, Qp_inv, fref=1.0, fmin=0.05, fmax=20.0,
                      correct_modulus=True):
        tau, w = q_coefficients(nmech, fmin, fmax)
        self.nmech = nmech
        self.tau[:] = 0; self.weight[:] = 0
        self.tau[:nmech] = tau; self.weight[:nmech] = w
        shp = (self.nz, self.ny, self.nx)
        self.Qs_inv = np.ascontiguousarray(np.broadcast_to(Qs_inv, shp), dtype=np.float64)
        self.Qp_inv = np.ascontiguousarray(np.broadcast_to(Qp_inv, shp), dtype=np.float64)
        self.eta = np.zeros((6, nmech) + shp)
        self.Deta = np.zeros((6, nmech) + shp)   # material.f90:389-404
        if correct_modulus:
            rc = lib().orc_q_modulus_correction(C.byref(self._sync()), fref)
            if rc != 0:
                raise ValueError("invalid Q modulus correction: weight/Q sum >= 1")

    def set_pml(self, face, npml):
        shp = (9, self.nz, self.ny, self.nx)
        self.pml[face] = 1
        self.npml[face] = npml
        self.pQ[face] = np.zeros(shp)
        self.pDQ[face] = np.full(shp, 1.0e40)     # pml.f90:46

    def add_source(self, *, kind, t_init, duration, mxx=0., myy=0., mzz=0., mxy=0., mxz=0.,
                   myz=0., loc, near, near_phys, order=4, alpha=0):
        s = _Source()
        s.type = kind; s.order = order; s.alpha = alpha
        s.near[:] = [int(v) for v in near]
        s.duration = duration; s.t_init = t_init
        s.mXX, s.mYY, s.mZZ, s.mXY, s.mXZ, s.mYZ = mxx, myy, mzz, mxy, mxz, myz
        s.loc[:] = list(loc); s.near_phys[:] = list(near_phys)
        self.sources.append(s)

    def _sync(self):
        b = _Block()
        b.nx, b.ny, b.nz = self.nx, self.ny, self.nz
        b.curvilinear = self.curvilinear
        b.nmech = self.nmech
        b.nsrc = len(self.sources)
        b.bc[:] = list(self.bc); b.pml[:] = list(self.pml); b.npml[:] = list(self.npml)
        if self.sat_form and any(self.pml):
            raise ValueError("the alpha/theta SAT form is defined without PML faces")
        b.sat_form = int(self.sat_form)
        b.hq, b.hr, b.hs, b.tau0 = self.hq, self.hr, self.hs, self.tau0
        b.tau[:] = list(self.tau); b.weight[:] = list(self.weight)
        for name in ("F", "DF", "metx", "mety", "metz", "J", "M", "eta", "Deta",
                     "Qs_inv", "Qp_inv"):
            setattr(b, name, _ptr(getattr(self, name)))
        for f in range(6):
            b.pQ[f] = _ptr(self.pQ[f]); b.pDQ[f] = _ptr(self.pDQ[f])
            b.nl[f] = _ptr(self.nl[f]); b.nm[f] = _ptr(self.nm[f]); b.nn[f] = _ptr(self.nn[f])
        if self.sources:
            self._src_arr = (_Source * len(self.sources))(*self.sources)
            b.src = C.cast(self._src_arr, C.POINTER(_Source))
        if any(src.order == 0 for src in self.sources) and self.X is None:
            raise ValueError("a mollified source (order 0) needs the coordinates X")
        if self.X is not None:
            self.X = np.ascontiguousarray(self.X, dtype=np.float64)
        b.X = _ptr(self.X)
        self._struct = b
        return b

    # -- stage routines ---------------------------------------------------
    def normals(self):
        lib().orc_normals(C.byref(self._sync()))

    def scale_rates(self, A):
        lib().orc_scale_rates(C.byref(self._sync()), A)

    def set_rates(self):
        lib().orc_set_rates(C.byref(self._sync()))

    def set_moment_tensor(self, t):
        lib().orc_set_moment_tensor(C.byref(self._sync()), t)

    def impose_boundary_condition(self):
        lib().orc_impose_boundary_condition(C.byref(self._sync()))

    def update_fields(self, inc):
        lib().orc_update_fields(C.byref(self._sync()), inc)

    def time_step(self, dt, nsteps, t=0.0, receivers=None):
        """Advance nsteps; receivers: flat 0-based node indices.
        Returns (t, seismograms[nsteps, nrec, 9])."""
        tt = np.array([t], dtype=np.float64)
        if receivers is None:
            rec = np.zeros(1, dtype=np.int64); nrec = 0
        else:
            rec = np.ascontiguousarray(receivers, dtype=np.int64); nrec = rec.size
        seis = np.zeros((max(nsteps, 1), max(nrec, 1), 9))
        lib().orc_time_step(C.byref(self._sync()), dt, nsteps, _ptr(tt), nrec,
                            rec.ctypes.data_as(C.POINTER(C.c_long)), _ptr(seis))
        return float(tt[0]), seis[:, :nrec] if nrec else None


def rw_prestress(n, m, l, y, z, t):
    """Prestress (T0_n, T0_m, T0_l) of problem 'rate-weakening' on a face with
    unit frame n, m, l at (y, z) km, time t."""
    vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in (n, m, l)]
    T0 = np.zeros(3)
    lib().orc_rw_prestress(*[_ptr(v) for v in vecs], y, z, t, _ptr(T0))
    return T0


def welded_hats(U, V, Z1s, Z2s, Z1p, Z2p):
    """Welded_Interface hats (P, Q) of rotated grid values U, V (6 each)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    P, Q = np.zeros(6), np.zeros(6)
    lib().orc_welded_hats(_ptr(U), _ptr(V), Z1s, Z2s, Z1p, Z2p, _ptr(P), _ptr(Q))
    return P, Q


def linear_friction_hats(U, V, Z1s, Z2s, Z1p, Z2p, alpha):
    """Linear_Friction hats (P, Q) with tau = alpha v."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    P, Q = np.zeros(6), np.zeros(6)
    lib().orc_linear_friction_hats(_ptr(U), _ptr(V), Z1s, Z2s, Z1p, Z2p, alpha, _ptr(P), _ptr(Q))
    return P, Q


class OracleIface:
    """The q-interface between the x-high face of block L and the x-low face
    of block R (wql_oracle_iface.c).  yf, zf: coordinates of the face nodes
    (nz, ny)."""

    # TPV28 initial stress of the source (initial_stress_condition.f90:78-84)
    TPV28_STRESS = (-60.0, 0.0, 60.0, 0.0, 29.38, 0.0)

    def __init__(self, ny, nz, yf, zf, coupling=IFACE_SLIP_WEAKENING_TPV5, s0=None, W0=None,
                 alpha=1.0e14):
        self.ny, self.nz, self.coupling = ny, nz, coupling
        self.alpha = float(alpha)   # linear friction (the reference's value, :365)
        self.s0 = tuple(self.TPV28_STRESS if s0 is None else s0)
        shp = (nz, ny)
        self.yf = np.ascontiguousarray(yf, dtype=np.float64).reshape(shp)
        self.zf = np.ascontiguousarray(zf, dtype=np.float64).reshape(shp)
        self.S = np.zeros((4,) + shp)
        self.DS = np.full((4,) + shp, 1.0e40)       # iface.f90:38-39
        # state variable: 0 (iface.f90:38) or the caller's initial state
        self.W = np.zeros(shp) if W0 is None else np.array(W0, dtype=np.float64).reshape(shp)
        self.DW = np.full(shp, 1.0e40)
        self.Uhat = np.zeros((6,) + shp)
        self.Vhat = np.zeros((6,) + shp)
        self.Svel = np.zeros((3,) + shp)
        self.trup = np.full(shp, 1.0e9)

    def _sync(self):
        s = _Iface()
        s.ny, s.nz, s.coupling = self.ny, self.nz, self.coupling
        s.alpha = self.alpha
        s.s0[:] = list(self.s0)
        for name in ("yf", "zf", "S", "DS", "W", "DW", "Uhat", "Vhat", "Svel", "trup"):
            setattr(s, name, _ptr(getattr(self, name)))
        self._struct = s
        return s

    def enforce(self, L, R, t, stage, side):
        """side 2: block L's evaluation (SAT into L); side 1: block R's."""
        lib().orc_iface_enforce(C.byref(self._sync()), C.byref(L._sync()), C.byref(R._sync()),
                                t, stage, side)

    def scale_rates(self, A):
        lib().orc_iface_scale(C.byref(self._sync()), A)

    def update_fields(self, inc):
        lib().orc_iface_update(C.byref(self._sync()), inc)


def two_block_time_step(L, R, I, dt, nsteps, t=0.0, slices=False):
    """time_step_RK (time_step.f90:68-150) of a two-block domain: per stage
    scale_rates, set_rates with the moment tensor (domain.f90:951-979),
    boundary conditions (block L, then R) and the
    interface (enforce_bound_iface_conditions, domain.f90:615-644), the
    stage-1 output, update_fields (blocks, then interface).  Returns (t,
    frames) with the fault output of every step (write_fault_output /
    write_hat_output, domain.f90:782-847): Uface, Vface (block faces, 9),
    S (4), Svel (3), Uhat, Vhat (6), trup, state (the friction state W), each
    (nsteps, ncomp, nz, ny).  With slices, also block L's volume slices of
    the same stage (write_slice, slice_output.f90:80-94): Hslice = F on the
    plane r = 1, (nsteps, 9, nz, nx), and Vslice = F on s = 1,
    (nsteps, 9, ny, nx)."""
    A, B, Cc = rk_coefficients()
    keys = {"Uface": 9, "Vface": 9, "S": 4, "Svel": 3, "Uhat": 6, "Vhat": 6, "trup": 1,
            "state": 1}
    frames = {k: np.zeros((nsteps, n, I.nz, I.ny)) for k, n in keys.items()}
    if slices:
        nz, ny, nx = L.F.shape[1:]
        frames["Hslice"] = np.zeros((nsteps, 9, nz, nx))
        frames["Vslice"] = np.zeros((nsteps, 9, ny, nx))
    for n in range(nsteps):
        t0 = t
        for s in range(5):
            for b in (L, R):
                b.scale_rates(A[s])
            I.scale_rates(A[s])
            for b in (L, R):
                b.set_rates()
                if b.sources:
                    b.set_moment_tensor(t)
            L.impose_boundary_condition()
            I.enforce(L, R, t, s + 1, 2)
            R.impose_boundary_condition()
            I.enforce(L, R, t, s + 1, 1)
            if s == 0:
                frames["Uface"][n] = L.F[:, :, :, -1]
                frames["Vface"][n] = R.F[:, :, :, 0]
                frames["S"][n] = I.S
                frames["Svel"][n] = I.Svel
                frames["Uhat"][n] = I.Uhat
                frames["Vhat"][n] = I.Vhat
                frames["trup"][n, 0] = I.trup
                frames["state"][n, 0] = I.W
                if slices:
                    frames["Hslice"][n] = L.F[:, :, 0, :]
                    frames["Vslice"][n] = L.F[:, 0, :, :]
            t = t0 + Cc[s + 1] * dt
            for b in (L, R):
                b.update_fields(B[s] * dt)
            I.update_fields(B[s] * dt)
    return t, frames
