"""The arena search of wql_block_create: a large anelastic block allocates
candidate arenas one after another, times the memory-variable kernel on each
and keeps the fastest (the kernel's speed depends on where the arena lands in
HBM).  A block created through the search must step exactly as one created
on the first arena: same inputs, bit-identical fields and memory variables
after two fused steps (FP32-mixed, 257^3 Q8 + PML 20, a 12.7 GB arena)."""
import ctypes
import os

import numpy as np
import pytest

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB

pytestmark = pytest.mark.gpu


def test_arena_search_steps_like_the_first_arena(monkeypatch):
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    import torch
    prob = PB.curvilinear_q(n=257, h=50.0, nsteps=2, nmech=8, npml=20, nrec=8)
    out = []
    for tries in ("1", "6"):
        monkeypatch.setenv("WQL3D_ARENA_TRIES", tries)
        blk = PB.make_device_block(prob, precision=L.WQL_FP32_MIXED, state_mirror=True)
        assert L.lib().wql_block_device_bytes(blk._h) > 4 << 30
        t = blk.time_step(prob.dt, 2, prob.sources[0].t_init - 0.1)
        blk.download(L.WQL_F | L.WQL_MEMORY)
        out.append((blk.F.cpu().numpy().copy(), blk.eta.cpu().numpy().copy(), t))
        print(f"tries {tries}: max|F| {np.abs(out[-1][0]).max():.4e}", flush=True)
        blk.close()
        del blk
        torch.cuda.empty_cache()
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.abs(out[0][0]).max() > 0
