"""MI355X-native per-block time step of WaveQLab3D-Q (HIP for gfx950).

The compute lives in ``libwql3d_hip.so`` (C ABI: ``include/wql3d.h``);
this package is its host-side mirror (ctypes binding, block descriptor,
stage driver and synthetic set-ups).
"""
from . import _lib
from ._lib import WqlError
from .block import (Block, SourceSpec, init_RK, time_step_RK, scale_rate_array,
                    update_state_array, grid_metrics, RK_A, RK_B, RK_C, TAU0)
from .interface import Interface

__all__ = ["Block", "SourceSpec", "init_RK", "time_step_RK", "scale_rate_array",
           "update_state_array", "grid_metrics", "WqlError", "RK_A", "RK_B", "RK_C", "TAU0",
           "Interface"]
