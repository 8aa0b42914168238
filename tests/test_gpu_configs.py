"""FP64 parity of the HIP path with the oracle at the workloads' own sizes
(the oracle on the host's threads finishes each in well under a minute):

* config 1: 65^3 Cartesian elastic bump + noise, free surface on r-min,
  characteristic elsewhere, 200 steps, all fields and 8 surface receivers;
* config 2: 257^3 curvilinear block with topography, layered + von Karman
  medium, Q8, PML 20 on four sides and the bottom, Ricker source, from a
  random state (fields and memory variables), 2 fused steps: fields, memory
  variables and every PML slab;
* config 5 physics at reduced size: two curvilinear blocks (topography,
  planar fault) with Q8, PML and TPV5 slip-weakening friction, fields of both
  blocks and the whole fault output (slip, slip rate, hat variables, rupture
  time) of every step.

Tolerances: relative L2 <= 1e-12 per field (the bar of the reference's FP64
cpu-release oracle), seismograms sample by sample to 1e-12 of the trace
peak, rupture times exactly equal (stage times, or 1e9 where the fault has
not ruptured)."""
import ctypes

import numpy as np
import pytest

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB
from waveqlab3d_q_gpu_amd.block import pml_slab_box, time_step_RK
from waveqlab3d_q_gpu_amd.interface import Interface

from helpers import (field_errors, gpu_block, oracle_block, oracle_two_block, receiver_flat,
                     rel_l2)

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _runtime():
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))


def test_config1_65_cubed_200_steps(oracle):
    prob = PB.cartesian_bump(n=65, nsteps=200, nrec=8)
    print(f"config 1: {prob.name}, {prob.nsteps} steps", flush=True)
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk = gpu_block(prob, met)
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps)
    blk.download(L.WQL_F | L.WQL_DF)
    errs = field_errors(blk.F, orc.F)
    print("config 1 field rel L2:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= TOL64
    assert rel_l2(blk.DF, orc.DF) <= TOL64
    seis_g = blk.seismograms()
    worst = 0.0
    for r in range(prob.receivers.shape[0]):
        for c in range(9):
            ref = seis_o[:, r, c]
            peak = max(np.abs(ref).max(), 1e-30)
            worst = max(worst, np.abs(seis_g[:, r, c] - ref).max() / peak)
    print(f"config 1 seismograms: max |diff| / peak = {worst:.2e}", flush=True)
    assert worst <= TOL64
    blk.close()


def test_config2_257_cubed_two_steps(oracle):
    """Both sides start from the same random state, so every node's stencil,
    closure row, slab and face term contributes from the first stage (a
    wrong neighbour or offset that only appears at large n changes O(1) of
    the result)."""
    prob = PB.curvilinear_q(n=257, h=50.0, nsteps=2, nmech=8, npml=20)
    print(f"config 2: {prob.name}, 2 steps from a random state", flush=True)
    rng = np.random.default_rng(257)
    shp = prob.X.shape[1:]
    prob.F0 = rng.uniform(-1.0, 1.0, (9,) + shp)
    prob.F0[3:] *= 1.0e6
    eta0 = rng.uniform(-1.0e3, 1.0e3, (6, prob.nmech) + shp)
    # the zero-weight mechanisms (2 and 7 of Q8) stay as they start
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    orc.eta[...] = eta0
    t0 = prob.sources[0].t_init - 0.1
    orc.time_step(prob.dt, prob.nsteps, t0)
    blk = gpu_block(prob, met)
    blk.eta[...] = eta0
    blk.upload(L.WQL_MEMORY)
    time_step_RK(blk, prob.dt, prob.nsteps, t0)
    blk.download(L.WQL_F | L.WQL_MEMORY | L.WQL_PML)
    errs = field_errors(blk.F, orc.F)
    print("config 2 field rel L2:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= TOL64
    e_eta = rel_l2(blk.eta, orc.eta)
    print(f"config 2 eta rel L2: {e_eta:.2e}", flush=True)
    assert e_eta <= TOL64
    for f in range(6):
        if prob.pml[f]:
            lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
            sl = (slice(None), slice(lo[2] - 1, hi[2]), slice(lo[1] - 1, hi[1]),
                  slice(lo[0] - 1, hi[0]))
            e = rel_l2(blk.pml_Q[f], orc.pQ[f][sl])
            print(f"config 2 PML Q face {f} rel L2: {e:.2e}", flush=True)
            assert e <= TOL64, f
    blk.close()


def test_config5_physics_reduced_curvilinear_fault(oracle):
    tp = PB.fault_two_block(n=(32, 48, 32), h=0.3, nmech=8, npml=8, nsteps=60, nrec=8)
    print(f"config 5 reduced: {tp.name}, {tp.nsteps} steps", flush=True)
    assert tp.left.mesh == L.WQL_CURVILINEAR
    Lo, Ro, Io = oracle_two_block(oracle, tp)
    _, fr_o = oracle.two_block_time_step(Lo, Ro, Io, tp.dt, tp.nsteps, 0.0)
    bl = gpu_block(tp.left, oracle.metrics(tp.left.X))
    br = gpu_block(tp.right, oracle.metrics(tp.right.X))
    itf = Interface(bl, br, tp.yf, tp.zf, coupling=tp.coupling, s0=tp.s0)
    itf.set_frames(tp.nsteps)
    itf.time_step(tp.dt, tp.nsteps, 0.0)
    fr = itf.frames()
    for key in ("S", "Svel", "Uface", "Vface", "Uhat", "Vhat"):
        e = rel_l2(fr[key], fr_o[key])
        print(f"config 5 reduced {key} rel L2: {e:.2e}", flush=True)
        assert e <= TOL64, (key, e)
    assert np.array_equal(fr["trup"], fr_o["trup"])
    ruptured = int((fr_o["trup"][-1] < 1e8).sum())
    print(f"config 5 reduced: {ruptured} ruptured fault nodes, max slip "
          f"{np.abs(fr_o['S'][-1, :3]).max():.3e}", flush=True)
    assert ruptured > 0
    bl.download(L.WQL_F | L.WQL_MEMORY)
    br.download(L.WQL_F | L.WQL_MEMORY)
    el, er = field_errors(bl.F, Lo.F), field_errors(br.F, Ro.F)
    print("config 5 reduced fields L", max(el), "R", max(er), flush=True)
    assert max(el) <= TOL64 and max(er) <= TOL64
    assert rel_l2(bl.eta, Lo.eta) <= TOL64 and rel_l2(br.eta, Ro.eta) <= TOL64
    itf.close()
    bl.close()
    br.close()
