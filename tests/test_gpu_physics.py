"""The physics pins of the anelastic and PML terms on the HIP path (the CPU
versions against the oracle are in test_oracle_pml_q_physics.py): realized Q
of the fused stage's Q4/Q8 memory variables at 12 frequencies in 0.08-15 Hz
(FP64 and FP32-mixed), and plane-wave pulses into the PML slabs of the x, y
and z faces (FP64) with the same bounds; the slabs of every axis equal the
oracle's to 1e-12."""
import ctypes

import numpy as np
import pytest

from tests import pml_plane as PP
from tests import realized_q as RQ
from tests.helpers import gpu_block, oracle_block, rel_l2
from waveqlab3d_q_gpu_amd import _lib as L

pytestmark = pytest.mark.gpu

PRECISIONS = [L.WQL_FP64, L.WQL_FP32_MIXED]


@pytest.fixture(scope="module", autouse=True)
def runtime():
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))


def gpu_stepper(O, prob, precision):
    """step() of realized_q.drive on the fused stage: the velocity is set on
    the host copy of F (whose stresses are the downloaded state), uploaded,
    one RK step runs on the GPU, F comes back; node (1, 1, 1) is sampled."""
    blk = gpu_block(prob, O.metrics(prob.X), precision=precision)
    X = prob.X

    def step(a, dt, t):
        blk.F[0] = a * X[0]
        blk.F[1] = a * X[0]
        blk.upload(L.WQL_F)
        blk.time_step(dt, 1, t)
        blk.download(L.WQL_F)
        return blk.F[3, 0, 0, 0], blk.F[4, 0, 0, 0], blk.F[6, 0, 0, 0]
    return step, blk


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("nmech", [4, 8])
def test_realized_q_gpu(oracle, nmech, precision):
    # FP64 at all 12 frequencies, FP32-mixed at every other one
    tol = 1.0e-3 if precision == L.WQL_FP64 else 2.0e-3
    worst = 0.0
    for f in (RQ.FREQS if precision == L.WQL_FP64 else RQ.FREQS[::2]):
        prob = RQ.uniform_problem(nmech)
        step, blk = gpu_stepper(oracle, prob, precision)
        ts, es, ss = RQ.drive(step, f, prob.tau)
        blk.close()
        m = RQ.measured_moduli(ts, es, ss, f, prob.tau)
        an = RQ.analytic_moduli(f, prob)
        qp, qs = abs(m[0].real / m[0].imag), abs(m[2].real / m[2].imag)
        qpa, qsa = abs(an[0].real / an[0].imag), abs(an[2].real / an[2].imag)
        ep, es_ = abs(qp / qpa - 1.0), abs(qs / qsa - 1.0)
        worst = max(worst, ep, es_)
        assert ep <= tol and es_ <= tol, (f, qp, qpa, qs, qsa)
        if nmech == 8:
            assert abs(qs / 50.0 - 1.0) <= 0.05, (f, qs)
    print(f"realized Q, Q{nmech}, precision {precision}: worst |Q/Q(f) - 1| = {worst:.2e}")


def gpu_pml_run(O, axis, comp, precision, curvilinear=False, nmech=0):
    prob = PP.plane_problem(axis, curvilinear, True)
    if nmech:
        prob.nmech = nmech
        prob.tau, prob.weight = O.q_coefficients(nmech)
        shp = prob.X.shape[1:]
        prob.Qs_inv = np.full(shp, 1.0 / 50.0)
        prob.Qp_inv = np.full(shp, 1.0 / 100.0)
    prob.F0 = PP.initial_pulse(prob, axis, comp)
    met = O.metrics(prob.X)
    blk = gpu_block(prob, met, precision=precision)
    state = {"t": 0.0}

    def step():
        state["t"] = blk.time_step(prob.dt, 1, state["t"])
        blk.download(L.WQL_F)
        return blk.F
    out = PP.run(step, prob, axis, comp, met[3])
    return prob, met, blk, out


# FP64 only: the four faces across the axis carry no boundary condition (bc
# 0), which keeps an exactly planar wave planar but lets any variation across
# grow; the closure rows' sums of equal values are zero only to round-off, so
# in FP32 that round-off (1e-7) grows into the run within 60 steps (in FP64,
# 1e-16, it stays negligible).  The FP32-mixed PML path is checked against the
# FP64 oracle on bounded set-ups (test_gpu_parity.py: test_fp32_error_budget).
@pytest.mark.parametrize("precision", [L.WQL_FP64])
@pytest.mark.parametrize("axis,comp,curvilinear,nmech", [
    (2, 2, False, 0), (2, 0, False, 0), (0, 0, True, 0), (1, 2, False, 0), (2, 2, False, 8)])
def test_pml_absorbs_plane_waves_gpu(oracle, axis, comp, curvilinear, nmech, precision):
    prob, met, blk, (E0, tot, refl) = gpu_pml_run(oracle, axis, comp, precision, curvilinear,
                                                  nmech)
    blk.close()
    bound = 5.0e-3 if comp == axis else 5.0e-4
    assert tot.max() <= E0 * (1.0 + 1e-6), tot.max() / E0
    assert np.diff(np.r_[E0, tot]).max() <= 1e-4 * E0
    assert tot[-1] <= 1e-2 * E0, tot[-1] / E0
    assert refl <= bound, refl


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_pml_slabs_fp64_match_oracle(oracle, axis):
    """The plane-wave block's PML slabs and free-surface ends against the
    oracle, with characteristic SAT on the four faces across the axis (with
    no boundary condition there the round-off of two different operation
    orders grows without bound, so the comparison needs a well-posed block):
    fields and PML auxiliaries after the pulse has entered the slabs."""
    prob = PP.plane_problem(axis, curvilinear=(axis == 0), pml=True)
    prob.bc = tuple(2 if f // 2 == axis else 1 for f in range(6))
    prob.F0 = PP.initial_pulse(prob, axis, axis)
    met = oracle.metrics(prob.X)
    nsteps = 120
    orc = oracle_block(oracle, prob, met)
    orc.time_step(prob.dt, nsteps, 0.0)
    blk = gpu_block(prob, met)
    blk.time_step(prob.dt, nsteps, 0.0)
    blk.download(L.WQL_F | L.WQL_PML)
    err = rel_l2(blk.F, orc.F)
    print(f"PML slabs axis {axis}: field rel L2 {err:.2e}", flush=True)
    assert err <= 1e-12
    from waveqlab3d_q_gpu_amd.block import pml_slab_box
    for f in (2 * axis, 2 * axis + 1):
        lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
        sl = (slice(None), slice(lo[2] - 1, hi[2]), slice(lo[1] - 1, hi[1]),
              slice(lo[0] - 1, hi[0]))
        assert np.abs(orc.pQ[f][sl]).max() > 0
        assert rel_l2(blk.pml_Q[f], orc.pQ[f][sl]) <= 1e-12, f
    blk.close()
