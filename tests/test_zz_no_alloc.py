"""The allocation-free shutdown check (wql_block_max_state_noalloc) against
wql_block_max_state and against the downloaded state, on two calls in a row
(the second on a smaller block, after the first left its partial results in
the library's scratch array)."""
import ctypes

import numpy as np
import pytest
import torch

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB

pytestmark = pytest.mark.gpu


def _init():
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))


def _both(blk):
    """(existing entry point, allocation-free entry point) results."""
    ref = blk.max_state()
    out = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    blk.max_state_noalloc(out)
    blk.sync()
    got = out.cpu().numpy()
    return ref, (float(got[0]), float(got[1]), int(got[2]))


def _host_max(a):
    a = np.asarray(a, dtype=np.float64)
    fin = np.isfinite(a)
    return (float(np.abs(a[fin]).max()) if fin.any() else 0.0), int((~fin).sum())


def test_max_state_noalloc_matches_existing_entry_point():
    _init()
    # call 1: curvilinear Q8 + PML block, 30^3, FP32-mixed, stepped through a Ricker source
    prob = PB.curvilinear_q(n=30, h=50.0 * 256 / 29 / 4, nsteps=12, nmech=8, npml=8, seed=5,
                            nrec=4, f0=4.0)
    print(f"no-alloc call 1: {prob.name} FP32-mixed", flush=True)
    blk = PB.make_device_block(prob, precision=L.WQL_FP32_MIXED, state_mirror=True)
    blk.time_step(prob.dt, prob.nsteps, prob.sources[0].t_init - 0.05)
    ref, got = _both(blk)
    assert got == ref, (got, ref)
    blk.download(L.WQL_F | L.WQL_MEMORY)
    F = blk.F.cpu().numpy()
    eta = blk.eta.cpu().numpy().astype(np.float32)   # FP32-mixed memory variables are float
    fmax, fbad = _host_max(F)
    emax, ebad = _host_max(eta)
    assert fmax > 0 and emax > 0
    assert got == (fmax, emax, fbad + ebad), (got, fmax, emax)
    blk.close()

    # call 2: a smaller elastic block (16^3, FP32-Kahan) with a NaN and an Inf planted in F
    prob2 = PB.cartesian_bump(n=16, h=100.0, nsteps=1)
    print(f"no-alloc call 2: {prob2.name} FP32-Kahan with planted non-finite values", flush=True)
    blk2 = PB.make_device_block(prob2, precision=L.WQL_FP32_KAHAN, state_mirror=True)
    F0 = np.array(prob2.F0)
    F0[3, 2, 5, 7] = np.nan
    F0[8, 10, 1, 0] = -np.inf
    blk2.F.copy_(torch.from_numpy(F0))
    blk2.upload(L.WQL_F)
    ref2, got2 = _both(blk2)
    assert got2 == ref2, (got2, ref2)
    fmax2, fbad2 = _host_max(F0.astype(np.float32))
    assert got2 == (fmax2, 0.0, 2) and fbad2 == 2, (got2, fmax2)
    blk2.close()
