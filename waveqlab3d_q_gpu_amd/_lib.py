"""ctypes binding of libwql3d_hip.so (C ABI declared in include/wql3d.h).

The library is required: importing a GPU entry point without the built
library raises immediately.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwql3d_hip.so")

# enums of wql3d.h
WQL_OK, WQL_EINVAL, WQL_EHIP, WQL_ENOMEM, WQL_ESTATE, WQL_ENODEV = range(6)
WQL_CARTESIAN, WQL_CURVILINEAR = 0, 1
WQL_FP64, WQL_FP32_MIXED, WQL_FP32_KAHAN = 0, 1, 2
WQL_SRC_GAUSSIAN, WQL_SRC_COSBELL, WQL_SRC_RICKER = 0, 1, 2
WQL_SAT_BC, WQL_SAT_ALPHA_THETA = 0, 1
WQL_IFACE_LOCKED, WQL_IFACE_SLIP_WEAKENING_TPV5, WQL_IFACE_SLIP_WEAKENING_TPV28 = 0, 1, 2
WQL_IFACE_SLIP_WEAKENING = 3
WQL_IFACE_RATE_STATE_TPV101, WQL_IFACE_RATE_STATE_TPV102 = 4, 5
WQL_IFACE_LINEAR_FRICTION, WQL_IFACE_RATE_STATE_RW = 6, 7
WQL_F, WQL_DF, WQL_GEOMETRY, WQL_MATERIAL = 1, 2, 4, 8
WQL_MEMORY, WQL_QINV, WQL_PML, WQL_FACES, WQL_ALL = 16, 32, 64, 128, 255

_D = C.POINTER(C.c_double)


class Box(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("hi", C.c_int32 * 3)]


class BlockDesc(C.Structure):
    _fields_ = [
        ("nq", C.c_int32), ("nr", C.c_int32), ("ns", C.c_int32),
        ("mesh", C.c_int32), ("precision", C.c_int32), ("nmech", C.c_int32),
        ("bc", C.c_int32 * 6), ("pml", C.c_int32 * 6), ("npml", C.c_int32 * 6),
        ("arrays_on_device", C.c_int32),
        ("hq", C.c_double), ("hr", C.c_double), ("hs", C.c_double), ("tau0", C.c_double),
        ("tau", C.c_double * 8), ("weight", C.c_double * 8),
        ("body", Box),
        ("metricx", _D), ("metricy", _D), ("metricz", _D), ("J", _D), ("M", _D),
        ("F", _D), ("DF", _D),
        ("eta", _D * 6), ("Deta", _D * 6),
        ("Qs_inv", _D), ("Qp_inv", _D),
        ("pml_box", Box * 6), ("pml_Q", _D * 6), ("pml_DQ", _D * 6),
        ("face_box", Box * 6), ("n_l", _D * 6), ("n_m", _D * 6), ("n_n", _D * 6),
        ("sat_form", C.c_int32),
    ]


class IfaceDesc(C.Structure):
    _fields_ = [("coupling", C.c_int32), ("arrays_on_device", C.c_int32),
                ("y", _D), ("z", _D), ("s0", _D), ("T0L", _D), ("T0R", _D),
                ("fs", _D), ("fd", _D), ("Dc", _D), ("W0", _D), ("alpha", C.c_double)]


class Source(C.Structure):
    _fields_ = [("type", C.c_int32), ("order", C.c_int32), ("alpha", C.c_int32),
                ("near", C.c_int32 * 3),
                ("duration", C.c_double), ("t_init", C.c_double),
                ("mXX", C.c_double), ("mYY", C.c_double), ("mZZ", C.c_double),
                ("mXY", C.c_double), ("mXZ", C.c_double), ("mYZ", C.c_double),
                ("location", C.c_double * 3), ("near_phys", C.c_double * 3)]


# name -> (restype, argtypes); every entry point declared in include/wql3d.h
_P = C.c_void_p
SIGNATURES = {
    "wql_runtime_init": (C.c_int, [C.c_int32, C.POINTER(C.c_int32)]),
    "wql_runtime_finalize": (C.c_int, []),
    "wql_capacity_check": (C.c_int, [C.c_int64]),
    "wql_block_create": (C.c_int, [C.POINTER(BlockDesc), C.POINTER(_P)]),
    "wql_block_device_bytes": (C.c_int64, [_P]),
    "wql_block_upload": (C.c_int, [_P, C.c_uint32]),
    "wql_block_download": (C.c_int, [_P, C.c_uint32]),
    "wql_block_init_state": (C.c_int, [_P]),
    "wql_block_destroy": (None, [_P]),
    "wql_block_set_sources": (C.c_int, [_P, C.c_int32, C.POINTER(Source)]),
    "wql_block_set_coordinates": (C.c_int, [_P, _P, C.c_int32]),
    "wql_block_set_receivers": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.c_int32]),
    "wql_block_get_seismograms": (C.c_int, [_P, C.c_int32, _D]),
    "wql_rk_scale": (C.c_int, [_P, C.c_double]),
    "wql_rk_update": (C.c_int, [_P, C.c_double]),
    "wql_scale_rate_array": (C.c_int, [_P, C.c_int64, C.c_double]),
    "wql_update_state_array": (C.c_int, [_P, _P, C.c_int64, C.c_double]),
    "wql_array_enter": (C.c_int, [_P, C.c_int64]),
    "wql_array_exit": (C.c_int, [_P]),
    "wql_array_present": (C.c_int, [_P, C.c_int64]),
    "wql_array_update_device": (C.c_int, [_P, C.c_int64]),
    "wql_array_update_self": (C.c_int, [_P, C.c_int64]),
    "wql_scale_rate_present": (C.c_int, [_P, C.c_int64, C.c_double]),
    "wql_update_state_present": (C.c_int, [_P, _P, C.c_int64, C.c_double]),
    "wql_block_find": (C.c_int, [_P, C.POINTER(_P)]),
    "wql_rhs": (C.c_int, [_P]),
    "wql_rhs_interior": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "wql_set_moment_tensor": (C.c_int, [_P, C.c_double]),
    "wql_sat": (C.c_int, [_P]),
    "wql_stage_fused": (C.c_int, [_P, C.c_double, C.c_double, C.c_double]),
    "wql_time_step": (C.c_int, [_P, C.c_double, C.c_int32, _D]),
    "wql_grid_metrics": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, _P, _P]),
    "wql_last_error": (C.c_char_p, []),
    "wql_block_max_state": (C.c_int, [_P, _D]),
    "wql_block_max_state_noalloc": (C.c_int, [_P, _P]),
    "wql_block_sync": (C.c_int, [_P]),
    "wql_block_stream": (_P, [_P]),
    "wql_block_set_timing": (C.c_int, [_P, C.c_int32]),
    "wql_block_kernel_times": (C.c_int, [_P, _D, C.POINTER(C.c_int32)]),
    "wql_adjoint_create": (C.c_int, [_P, C.c_int32]),
    "wql_adjoint_create_f": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "wql_adjoint_reset": (C.c_int, [_P]),
    "wql_state_save": (C.c_int, [_P, C.c_int32]),
    "wql_state_load": (C.c_int, [_P, C.c_int32]),
    "wql_state_save_part": (C.c_int, [_P, C.c_int32, C.c_uint32]),
    "wql_state_load_part": (C.c_int, [_P, C.c_int32, C.c_uint32]),
    "wql_adjoint_stage": (C.c_int, [_P, C.c_double, C.c_double]),
    "wql_adjoint_add_receivers": (C.c_int, [_P, _P]),
    "wql_adjoint_get_state": (C.c_int, [_P, _P, _P]),
    "wql_adjoint_set_state": (C.c_int, [_P, _P, _P]),
    "wql_adjoint_gradient": (C.c_int, [_P, _P, _P, _P]),
    "wql_iface_create": (C.c_int, [_P, _P, C.POINTER(IfaceDesc), C.POINTER(_P)]),
    "wql_iface_destroy": (None, [_P]),
    "wql_iface_stage": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_int32]),
    "wql_iface_time_step": (C.c_int, [_P, C.c_double, C.c_int32, _D]),
    "wql_iface_set_frames": (C.c_int, [_P, C.c_int32]),
    "wql_iface_get_frames": (C.c_int, [_P, C.c_int32, _D]),
    "wql_iface_get_state": (C.c_int, [_P, _D, _D, _D]),
    "wql_iface_get_friction_state": (C.c_int, [_P, _D]),
}

_lib = None


class WqlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"wql error {code}: {msg}")
        self.code = code


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: when PyTorch is importable its bundled
        # libamdhip64.so.7 is loaded first, and the library's own dependency
        # on libamdhip64.so.7 then binds to that same runtime (two runtimes
        # in one process would each claim the device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m waveqlab3d_q_gpu_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != WQL_OK:
        msg = lib().wql_last_error()
        raise WqlError(rc, msg.decode() if msg else "")
    return rc
