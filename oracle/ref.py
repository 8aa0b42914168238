"""ctypes wrapper of ``oracle/_ref/libwql_ref.so``: the reference's own
MPI-free Fortran modules (metrics.f90, anelastic_q{4,8}_model.f90,
rk_vector_cpu.f90, initial_stress_condition.f90, common.f90) compiled from
the reference sources with the cpu-release real-8 promotion, behind the
C-callable harness ``ref_harness.f90``.

TEST INFRASTRUCTURE ONLY: a second checker that pins parts of the CPU
restatement (``wql_oracle.c``) bit for bit against the reference's code.
Only ``tests/`` import it.  The library is built by ``make -C oracle ref``
where the reference sources exist; elsewhere :func:`available` is False.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "libwql_ref.so")
REF_SRC = "/root/reference/src"

_P = C.POINTER(C.c_double)
_lib = None


def build(force: bool = False) -> bool:
    """Build the library if the reference sources are present; True if it exists."""
    if (force or not os.path.exists(_LIB_PATH)) and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []) + ["ref"],
                       check=True)
    return os.path.exists(_LIB_PATH)


def available() -> bool:
    return build()


def lib():
    global _lib
    if _lib is None:
        if not build():
            raise RuntimeError("oracle/_ref/libwql_ref.so is not built (no reference sources)")
        L = C.CDLL(_LIB_PATH)
        L.ref_d6.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P]
        L.ref_metrics.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]
        L.ref_q_coefficients.argtypes = [C.c_int, _P, _P]
        L.ref_q_coefficients.restype = C.c_int
        L.ref_q_max_relative_error.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_q_max_relative_error.restype = C.c_double
        L.ref_q_dt_limit.argtypes = [C.c_int]
        L.ref_q_dt_limit.restype = C.c_double
        L.ref_rk_vector.argtypes = [C.c_int] * 6 + [_P, _P, C.c_double, C.c_double]
        L.ref_initial_stress.argtypes = [C.c_int, C.c_char_p, _P, _P]
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def d6(u, h):
    """D6x / D6y / D6z of the 3 components of u (3, nz, ny, nx) along each
    axis: returns (axis, comp, nz, ny, nx)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    _, nz, ny, nx = u.shape
    du = np.zeros((3, 3, nz, ny, nx))
    hh = np.ascontiguousarray(h, dtype=np.float64)
    lib().ref_d6(nx, ny, nz, _ptr(hh), _ptr(u), _ptr(du))
    return du


def metrics(X):
    """derivatives() of grid.f90:398-471 on X (3, nz, ny, nx)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    _, nz, ny, nx = X.shape
    metx = np.zeros((3, nz, ny, nx)); mety = np.zeros_like(metx); metz = np.zeros_like(metx)
    J = np.zeros((nz, ny, nx))
    lib().ref_metrics(nx, ny, nz, _ptr(X), _ptr(metx), _ptr(mety), _ptr(metz), _ptr(J))
    return metx, mety, metz, J


def q_coefficients(nmech):
    tau = np.zeros(8); w = np.zeros(8)
    if lib().ref_q_coefficients(nmech, _ptr(tau), _ptr(w)) != 0:
        raise ValueError("nmech must be 4 or 8")
    return tau[:nmech].copy(), w[:nmech].copy()


def q_max_relative_error(nmech, target_Q, fmin, fmax):
    return lib().ref_q_max_relative_error(nmech, target_Q, fmin, fmax)


def q_dt_limit(nmech):
    return lib().ref_q_dt_limit(nmech)


def rk_vector(state, rate, A, inc, lo=(-2, -2, -2)):
    """scale_rate_array(rate, A) then update_state_array(state, rate, inc) on
    arrays (9, n3, n2, n1) placed at the Fortran lower bounds lo; returns
    (state, rate)."""
    s = np.array(state, dtype=np.float64, order="C")
    r = np.array(rate, dtype=np.float64, order="C")
    _, n3, n2, n1 = s.shape
    lib().ref_rk_vector(n1, n2, n3, lo[0], lo[1], lo[2], _ptr(s), _ptr(r), A, inc)
    return s, r


def initial_stress(problem, x):
    s0 = np.zeros(6)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    name = problem.encode()
    lib().ref_initial_stress(len(name), name, _ptr(xx), _ptr(s0))
    return s0
