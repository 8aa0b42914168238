"""Properties of the fused time step at the production sizes, where the oracle
is too slow to run: the 385^3 curvilinear Q8 + PML20 block and the two
201x401x201 blocks of the faulted domain.

* Linearity.  Without sources the semi-discrete system (SBP RHS, SAT with
  homogeneous boundary data, Q8 memory variables, PML auxiliaries) is linear
  in the state, so one RK step S satisfies S(a u) = a S(u) and
  S(u1 + u2) = S(u1) + S(u2) up to FP64 rounding; a dropped, doubled or
  stale term of the fused stage breaks both by O(1).
* Determinism.  The same state stepped twice gives bit-identical fields
  (no atomics, no order-dependent sums): FP64 and FP32-mixed at 385^3, and
  the two-block fault domain, whose interface kernel and two block streams
  are ordered by events.
"""
import ctypes

import numpy as np
import pytest

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB

pytestmark = pytest.mark.gpu

N = 385


@pytest.fixture(scope="module", autouse=True)
def _runtime():
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    yield


@pytest.fixture(scope="module")
def production():
    prob = PB.curvilinear_q(n=N, h=50.0, nsteps=1, nmech=8, npml=20, seed=2024, nrec=8, vk=False)
    print(f"problem {N}^3 curvilinear Q8 PML20 generated", flush=True)
    return prob


def _load_state(blk, terms, torch):
    """Write sum_k coeff_k * (random state of seed_k) into the block's device
    mirrors (F: velocities in [-1, 1], stresses in [-1e6, 1e6]; eta in
    [-1e3, 1e3]; Deta 0), one random array alive at a time."""
    blk.F.zero_()
    blk.eta.zero_()
    blk.Deta.zero_()
    for seed, coeff in terms:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        r = torch.rand(blk.F.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        r[3:] *= 1e6
        blk.F.add_(r, alpha=coeff)
        del r
        r = torch.rand(blk.eta.shape, generator=g, device="cuda", dtype=torch.float64) * 2e3 - 1e3
        blk.eta.add_(r, alpha=coeff)
        del r


def _step(blk, prob, terms, torch, keep_eta=False):
    """One fused RK step from the state of ``terms`` with the init_* rates and
    PML state; returns F and eta (all of it, or planes (c, i) = (0, 0) and
    (5, nmech-1))."""
    blk.init_state()
    _load_state(blk, terms, torch)
    blk.upload(L.WQL_F | L.WQL_MEMORY)
    blk.time_step(prob.dt, 1, 0.0)
    blk.download(L.WQL_F | L.WQL_MEMORY)
    torch.cuda.synchronize()
    eta = blk.eta.clone() if keep_eta else torch.stack([blk.eta[0, 0], blk.eta[5, -1]]).clone()
    return blk.F.clone(), eta


def _rel(a, b, torch):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def test_linearity_and_superposition_385(production):
    torch = pytest.importorskip("torch")
    prob = production
    blk = PB.make_device_block(prob, precision=L.WQL_FP64, state_mirror=True)
    blk.set_sources([])
    s1 = _step(blk, prob, [(1, 1.0)], torch)
    sa = _step(blk, prob, [(1, 2.5)], torch)
    lin = [_rel(sa[k], 2.5 * s1[k], torch) for k in range(2)]
    del sa
    s2 = _step(blk, prob, [(2, 1.0)], torch)
    s12 = _step(blk, prob, [(1, 1.0), (2, 1.0)], torch)
    sup = [_rel(s12[k], s1[k] + s2[k], torch) for k in range(2)]
    del s12, s2
    _load_state(blk, [(1, 1.0)], torch)
    change = _rel(s1[0], blk.F, torch)
    print(f"385^3 FP64: |S(u)-u|/|u| = {change:.3e}; linearity (F, eta) = "
          f"{lin[0]:.2e}, {lin[1]:.2e}; superposition = {sup[0]:.2e}, {sup[1]:.2e}", flush=True)
    blk.close()
    del blk, s1
    torch.cuda.empty_cache()
    assert change > 1e-6          # the step did something
    assert max(lin) <= 1e-12
    assert max(sup) <= 1e-12


@pytest.mark.parametrize("precision", [L.WQL_FP64, L.WQL_FP32_MIXED])
def test_repeat_is_bit_identical_385(production, precision):
    torch = pytest.importorskip("torch")
    prob = production
    blk = PB.make_device_block(prob, precision=precision, state_mirror=True)
    a = _step(blk, prob, [(3, 1.0)], torch, keep_eta=True)
    b = _step(blk, prob, [(3, 1.0)], torch, keep_eta=True)
    same = [bool(torch.equal(a[k], b[k])) for k in range(2)]
    print(f"385^3 precision={precision}: repeat identical (F, eta) = {same}", flush=True)
    blk.close()
    del a, b, blk
    torch.cuda.empty_cache()
    assert all(same)


def test_two_block_fault_repeat_is_bit_identical():
    torch = pytest.importorskip("torch")
    from waveqlab3d_q_gpu_amd.interface import Interface
    tp = PB.fault_two_block(nsteps=3)
    print(f"two-block fault {tp.left.n} x 2 generated", flush=True)
    bl = PB.make_device_block(tp.left, precision=L.WQL_FP32_MIXED, state_mirror=True)
    br = PB.make_device_block(tp.right, precision=L.WQL_FP32_MIXED, state_mirror=True)

    def run():
        bl.init_state()
        br.init_state()
        itf = Interface(bl, br, tp.yf, tp.zf, coupling=tp.coupling)
        itf.time_step(tp.dt, 3, 0.0)
        itf.sync()
        S, DS, tr = itf.state()
        itf.close()
        bl.download(L.WQL_F)
        br.download(L.WQL_F)
        torch.cuda.synchronize()
        return bl.F.clone(), br.F.clone(), S, DS, tr

    a = run()
    b = run()
    bl.close()
    br.close()
    slip = float(np.abs(a[2][:3]).max())
    same_f = bool(torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]))
    same_i = [bool(np.array_equal(a[k], b[k])) for k in range(2, 5)]
    print(f"two-block fault: max slip after 3 steps {slip:.3e} m; repeat identical "
          f"(fields, S/DS/trup) = {same_f}, {same_i}", flush=True)
    del a, b, bl, br
    torch.cuda.empty_cache()
    assert slip > 0
    assert same_f and all(same_i)
