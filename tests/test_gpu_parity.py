"""Parity of the HIP path (libwql3d_hip.so through its C ABI) with the oracle.

Tolerances: FP64 paths must match the oracle to relative L2 <= 1e-12 per
field after N steps (the bar the reference's own cpu-release oracle sets);
receiver seismograms sample by sample to 1e-12 relative to the trace peak.
"""
import math

import numpy as np
import pytest

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB
from waveqlab3d_q_gpu_amd.block import pml_slab_box, time_step_RK

from helpers import (field_errors, gpu_block, oracle_block, receiver_flat, rel_l2,
                     small_curvilinear)

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _runtime():
    import ctypes
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    yield


def _download_all(blk):
    blk.download(L.WQL_F | L.WQL_DF | L.WQL_MEMORY | L.WQL_PML)


@pytest.mark.parametrize("fused", [True, False])
def test_reference_kat_41(oracle, fused):
    prob = PB.kat_cartesian_41()
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    orc.time_step(prob.dt, prob.nsteps)
    blk = gpu_block(prob, met)
    time_step_RK(blk, prob.dt, prob.nsteps, fused=fused)
    blk.download(L.WQL_F)
    got = np.abs(blk.F).max()
    ref = 1.6623128367649574e+02
    print(f"KAT max|F| gpu={got!r} ref={ref!r}", flush=True)
    assert abs(got - ref) / ref < 1e-13
    errs = field_errors(blk.F, orc.F)
    assert max(errs) <= TOL64, errs


def test_cartesian_bump_fields_and_seismograms(oracle):
    prob = PB.cartesian_bump(n=33, nsteps=40, nrec=8)
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk = gpu_block(prob, met)
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps)
    _download_all(blk)
    errs = field_errors(blk.F, orc.F)
    print("field rel L2:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= TOL64
    assert rel_l2(blk.DF, orc.DF) <= TOL64
    seis_g = blk.seismograms()
    for r in range(prob.receivers.shape[0]):
        for c in range(9):
            ref = seis_o[:, r, c]
            peak = np.abs(ref).max()
            assert np.abs(seis_g[:, r, c] - ref).max() <= TOL64 * max(peak, 1e-30), (r, c)


@pytest.mark.parametrize("nmech,fused", [(8, True), (8, False), (4, True), (0, True)])
def test_curvilinear_q_pml(oracle, nmech, fused):
    prob = small_curvilinear(nmech=nmech)
    met = oracle.metrics(prob.X)
    assert met[3].min() > 0
    orc = oracle_block(oracle, prob, met)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk = gpu_block(prob, met)
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps, fused=fused)
    _download_all(blk)
    errs = field_errors(blk.F, orc.F)
    print(f"nmech={nmech} fused={fused} field rel L2:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= TOL64
    if nmech:
        assert rel_l2(blk.eta, orc.eta) <= TOL64
        assert rel_l2(blk.Deta, orc.Deta) <= TOL64
    for f in range(6):
        if prob.pml[f]:
            lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
            sl = (slice(None), slice(lo[2] - 1, hi[2]), slice(lo[1] - 1, hi[1]),
                  slice(lo[0] - 1, hi[0]))
            assert rel_l2(blk.pml_Q[f], orc.pQ[f][sl]) <= TOL64, f
    if fused:
        seis_g = blk.seismograms()
        assert rel_l2(seis_g, seis_o) <= TOL64


@pytest.mark.parametrize("precision", [L.WQL_FP64, L.WQL_FP32_MIXED])
@pytest.mark.parametrize("fused", [True, False])
def test_mollified_source_matches_oracle(oracle, fused, precision):
    """The mollified moment-tensor source (set_moment_tensor_smooth, the
    reference's mollify_source: a Gaussian of standard deviation 2h per axis
    from the node coordinates) on the curvilinear Q8 + PML block, fused
    stage (source_at) and reference-order seams (k_source): fields, memory
    variables and seismograms against the oracle's whole-block loop (FP64
    1e-12; FP32-mixed to its budget, 1e-5)."""
    if precision != L.WQL_FP64 and not fused:
        pytest.skip("the one-kernel-per-seam path is FP64")
    prob = small_curvilinear(nmech=8)
    for src in prob.sources:
        src.order = 0
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk = gpu_block(prob, met, precision=precision)
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps, fused=fused)
    _download_all(blk)
    errs = field_errors(blk.F, orc.F)
    tol = TOL64 if precision == L.WQL_FP64 else 1e-5
    print(f"mollified fused={fused} precision={precision} field rel L2:",
          ["%.2e" % e for e in errs], flush=True)
    assert np.abs(orc.F[3:]).max() > 0
    assert max(errs) <= tol
    assert rel_l2(blk.eta, orc.eta) <= tol
    if fused:
        assert rel_l2(blk.seismograms(), seis_o) <= tol


def test_mollified_source_needs_coordinates(oracle):
    """wql_block_set_sources refuses a mollified source before the node
    coordinates are set (WQL_ESTATE) and takes it afterwards."""
    prob = small_curvilinear(nmech=0, npml=0)
    for src in prob.sources:
        src.order = 0
    met = oracle.metrics(prob.X)
    blk = PB.make_block(prob, met, upload=False)
    with pytest.raises(L.WqlError):
        blk.set_sources(prob.sources)
    blk.set_coordinates(prob.X)
    blk.set_sources(prob.sources)
    blk.close()


@pytest.mark.parametrize("fused", [True, False])
def test_clamped_faces_match_oracle(oracle, fused):
    """Clamped faces (bc 3, BoundaryConditions.f90:151-161) on a curvilinear
    Q4 block: x-high, y-high and z-low clamped, free surface on y-low,
    characteristic elsewhere; fields, memory variables and seismograms."""
    prob = small_curvilinear(nmech=4, npml=0)
    prob.bc = (1, 3, 2, 3, 3, 1)
    # a random start, so every face carries velocity and stress from step 1
    prob.F0 = np.random.default_rng(33).uniform(-1.0, 1.0, (9,) + prob.X.shape[1:])
    prob.F0[3:] *= 1.0e7
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk = gpu_block(prob, met)
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps, fused=fused)
    _download_all(blk)
    errs = field_errors(blk.F, orc.F)
    print(f"clamped fused={fused} field rel L2:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= TOL64
    assert rel_l2(blk.eta, orc.eta) <= TOL64
    if fused:
        assert rel_l2(blk.seismograms(), seis_o) <= TOL64


# Launch shapes of the fused stage (WQL3D_NODE_TUNING = "nontemporal,xcd,band,
# split,interior,zchunk"): workgroup order over the XCDs round-robin (0),
# contiguous (1) or in groups of 3 or 16 (whose last workgroups keep their
# index); strict-interior kernel off, or tiles 64x8 / 32x16 / 64x4 with z
# chunks that cut the interior box unevenly; shell memory variables in their
# own kernel (split 1) or in the node kernel (split 0).
# WQL3D_STAGE_CHUNKS: z chunks of the two-stream pipeline of anelastic stages
# (1 = one stream; 13 chunks of 3 z levels = the halo width; 5 uneven chunks).
# WQL3D_MEM_NPT / WQL3D_MEM_NT / WQL3D_MEM_XCD: nodes per thread (1, 2, 4),
# non-temporal streams and workgroup order of the memory-variable kernel.
@pytest.mark.parametrize("tuning,chunks,npt,nt,mx", [
    ("0,1,0,1,0,64", "4", "2", "1", "16"), ("0,1,0,1,1,7", "4", "2", "1", "16"),
    ("0,1,0,1,2,5", "5", "2", "1", "16"), ("0,1,0,1,3,64", "4", "2", "1", "16"),
    ("0,1,0,0,1,16", "4", "2", "1", "16"), ("0,0,0,1,1,9", "4", "2", "1", "16"),
    ("0,1,0,1,1,32", "1", "2", "1", "16"), ("0,1,0,1,1,32", "13", "2", "1", "16"),
    ("0,1,0,1,0,32", "13", "2", "1", "16"), ("0,0,0,1,1,32", "1", "1", "0", "16"),
    ("0,0,0,1,1,32", "1", "4", "1", "16"), ("0,0,0,1,1,32", "1", "2", "0", "0"),
    ("0,16,0,1,1,7", "4", "2", "1", "16"), ("0,3,0,1,2,5", "1", "2", "1", "3")])
def test_fused_stage_launch_shapes(oracle, monkeypatch, tuning, chunks, npt, nt, mx):
    monkeypatch.setenv("WQL3D_NODE_TUNING", tuning)
    monkeypatch.setenv("WQL3D_STAGE_CHUNKS", chunks)
    monkeypatch.setenv("WQL3D_MEM_NPT", npt)
    monkeypatch.setenv("WQL3D_MEM_NT", nt)
    monkeypatch.setenv("WQL3D_MEM_XCD", mx)
    prob = small_curvilinear(n=40, nmech=8, nsteps=3, seed=11)
    met = oracle.metrics(prob.X)
    rng = np.random.default_rng(7)
    F0 = rng.uniform(-1, 1, (9,) + prob.X.shape[1:])
    F0[3:] *= 1e6
    prob.F0 = F0
    orc = oracle_block(oracle, prob, met)
    orc.eta[...] = rng.uniform(-1, 1, orc.eta.shape)
    for f in range(6):
        if orc.pQ[f] is not None:
            orc.pQ[f][...] = rng.uniform(-1, 1, orc.pQ[f].shape)
    blk = gpu_block(prob, met)
    blk.eta[...] = orc.eta
    for f in range(6):
        if prob.pml[f]:
            lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
            blk.pml_Q[f][...] = orc.pQ[f][:, lo[2] - 1:hi[2], lo[1] - 1:hi[1], lo[0] - 1:hi[0]]
    blk.upload(L.WQL_MEMORY | L.WQL_PML)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps, fused=True)
    _download_all(blk)
    errs = field_errors(blk.F, orc.F)
    print(f"tuning={tuning} chunks={chunks} npt={npt} nt={nt} field rel L2:",
          ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= TOL64
    assert rel_l2(blk.DF, orc.DF) <= TOL64
    assert rel_l2(blk.eta, orc.eta) <= TOL64
    assert rel_l2(blk.Deta, orc.Deta) <= TOL64
    for f in range(6):
        if prob.pml[f]:
            lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
            sl = (slice(None), slice(lo[2] - 1, hi[2]), slice(lo[1] - 1, hi[1]),
                  slice(lo[0] - 1, hi[0]))
            assert rel_l2(blk.pml_Q[f], orc.pQ[f][sl]) <= TOL64, f
    assert rel_l2(blk.seismograms(), seis_o) <= TOL64


def test_stage_level_rates(oracle):
    """One stage of the reference order: DF after scale+rhs+source+sat."""
    prob = small_curvilinear(nmech=8, nsteps=1)
    met = oracle.metrics(prob.X)
    rng = np.random.default_rng(1)
    F0 = rng.uniform(-1, 1, (9,) + prob.X.shape[1:])
    F0[3:] *= 1e6
    prob.F0 = F0
    orc = oracle_block(oracle, prob, met)
    orc.eta[...] = rng.uniform(-1, 1, orc.eta.shape)
    for f in range(6):
        if orc.pQ[f] is not None:
            orc.pQ[f][...] = rng.uniform(-1, 1, orc.pQ[f].shape)
    blk = gpu_block(prob, met)
    blk.eta[...] = orc.eta
    for f in range(6):
        if prob.pml[f]:
            lo, hi = pml_slab_box(f, prob.n, prob.npml[f])
            blk.pml_Q[f][...] = orc.pQ[f][:, lo[2] - 1:hi[2], lo[1] - 1:hi[1], lo[0] - 1:hi[0]]
    blk.upload(L.WQL_MEMORY | L.WQL_PML)
    t = 0.13
    orc.scale_rates(0.0); orc.set_rates(); orc.set_moment_tensor(t); orc.impose_boundary_condition()
    blk.scale_rates(0.0); blk.set_rates(); blk.set_moment_tensor(t); blk.impose_boundary_condition()
    _download_all(blk)
    assert rel_l2(blk.DF, orc.DF) <= TOL64
    assert rel_l2(blk.Deta, orc.Deta) <= TOL64


def test_interior_rhs_seam(oracle):
    """try_traditional_cartesian_rhs seam: strict interior only."""
    prob = PB.cartesian_bump(n=28, nsteps=1)
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    orc.DF[...] = 0.0
    orc.set_rates()
    blk = gpu_block(prob, met)
    blk.scale_rates(0.0)
    assert blk.try_interior_rhs()
    blk.download(L.WQL_DF)
    core = (slice(None), slice(6, -6), slice(6, -6), slice(6, -6))
    assert rel_l2(blk.DF[core], orc.DF[core]) <= TOL64
    shell = np.ones(blk.DF.shape[1:], bool)
    shell[6:-6, 6:-6, 6:-6] = False
    assert np.all(blk.DF[:, shell] == 0.0)


def test_rk_vector_kernels_bit_exact():
    """tests/rk_vector_kernel_test.f90:6-35: asymmetric bounds
    (-2:4, 3:7, -1:1, 4); the GPU result must equal the CPU expression."""
    import torch
    i = np.arange(-2, 5)[:, None, None, None]
    j = np.arange(3, 8)[None, :, None, None]
    k = np.arange(-1, 2)[None, None, :, None]
    c = np.arange(1, 5)[None, None, None, :]
    state = ((11 * i - 7 * j + 5 * k + 3 * c) / 13.0).astype(np.float64)
    rate = ((-2 * i + 9 * j - 4 * k + 7 * c) / 17.0).astype(np.float64)
    state, rate = np.broadcast_arrays(state, rate)
    state = np.asfortranarray(state); rate = np.asfortranarray(rate)
    exp_rate = -0.375 * rate
    exp_state = state + 0.125 * exp_rate
    from waveqlab3d_q_gpu_amd import scale_rate_array, update_state_array
    r = torch.from_numpy(rate.ravel(order="F").copy()).cuda()
    s = torch.from_numpy(state.ravel(order="F").copy()).cuda()
    scale_rate_array(r, -0.375)
    update_state_array(s, r, 0.125)
    assert np.array_equal(r.cpu().numpy(), exp_rate.ravel(order="F"))
    assert np.array_equal(s.cpu().numpy(), exp_state.ravel(order="F"))


def test_grid_metrics_kernel(oracle):
    import torch
    from waveqlab3d_q_gpu_amd import grid_metrics
    prob = small_curvilinear(n=26)
    ref = oracle.metrics(prob.X)
    got = grid_metrics(torch.from_numpy(prob.X).cuda())
    for g, r in zip(got, ref):
        assert rel_l2(g.cpu().numpy(), r) < 1e-14


@pytest.mark.parametrize("precision", [L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN])
def test_fp32_error_budget(oracle, precision):
    """FP32 stencil/metric/memory arithmetic with the RK state in FP64
    (FP32-mixed) or in FP32 with a Kahan-compensated update (FP32-Kahan):
    stated error against the FP64 oracle, per field and per receiver."""
    name = "fp32-mixed" if precision == L.WQL_FP32_MIXED else "fp32-kahan"
    prob = small_curvilinear(nmech=8, nsteps=30)
    met = oracle.metrics(prob.X)
    orc = oracle_block(oracle, prob, met)
    _, seis_o = orc.time_step(prob.dt, prob.nsteps, receivers=receiver_flat(prob))
    blk = gpu_block(prob, met, precision=precision)
    blk.set_receivers(prob.receivers, prob.nsteps)
    time_step_RK(blk, prob.dt, prob.nsteps)
    _download_all(blk)
    errs = field_errors(blk.F, orc.F)
    seis_g = blk.seismograms()
    serr = [rel_l2(seis_g[:, r], seis_o[:, r]) for r in range(seis_o.shape[1])]
    print(f"{name} field rel L2:", ["%.2e" % e for e in errs], flush=True)
    print(f"{name} seismogram rel L2:", ["%.2e" % e for e in serr], flush=True)
    assert max(errs) <= 1e-4
    assert max(serr) <= 1e-5


def _state_arrays(blk):
    arrs = [blk.F, blk.DF]
    if blk.nmech:
        arrs += [blk.eta, blk.Deta]
    arrs += [q for q in blk.pml_Q if q is not None] + [q for q in blk.pml_DQ if q is not None]
    return arrs


@pytest.mark.parametrize("precision", [L.WQL_FP64, L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN])
def test_state_changes_between_fused_stages(oracle, precision):
    """Fused steps after the RK state changed through the API (F uploaded,
    the reference-order RK update, init_state) equal, bit for bit, fused steps
    on a fresh block given the same state: the FP32-mixed stencils' float
    copy of F follows every change of F."""
    prob = small_curvilinear(n=30, nmech=8, nsteps=2, seed=3)
    met = oracle.metrics(prob.X)
    blk = gpu_block(prob, met, precision=precision)
    state = L.WQL_F | L.WQL_DF | L.WQL_MEMORY | L.WQL_PML

    def continues_like_fresh(t, label):
        blk.download(state)
        fresh = gpu_block(prob, met, precision=precision)
        for dst, src in zip(_state_arrays(fresh), _state_arrays(blk)):
            dst[...] = src
        fresh.upload(state)
        t_next = blk.time_step(prob.dt, 2, t)
        fresh.time_step(prob.dt, 2, t)
        blk.download(state)
        fresh.download(state)
        for k, (a, b) in enumerate(zip(_state_arrays(blk), _state_arrays(fresh))):
            assert np.array_equal(a, b), (label, k, np.abs(a - b).max())
        print(f"precision={precision} {label}: identical", flush=True)
        return t_next

    t = blk.time_step(prob.dt, 3)
    blk.download(state)
    blk.F[...] *= 1.25
    blk.upload(L.WQL_F)
    t = continues_like_fresh(t, "after F upload")
    if precision != L.WQL_FP32_KAHAN:
        blk.update_fields(0.37 * prob.dt)
        t = continues_like_fresh(t, "after RK update")
    blk.init_state()
    continues_like_fresh(0.0, "after init_state")


@pytest.mark.parametrize("precision", [L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN, L.WQL_FP64])
def test_launch_shapes_consistent_all_precisions(oracle, monkeypatch, precision):
    """Every precision (the FP32-mixed float copy of F, the Kahan
    compensation, FP64) under launch shapes that cut a stage differently.
    The two-stream z-chunk pipeline runs the same kernels on the same nodes,
    so it must reproduce the one-stream stage bit for bit (its event order
    covers F, the float copy and the memory variables).  Other tiles, and the
    shell kernel over the whole block (JJU form instead of the flux form),
    evaluate the same rates with another operation order: equal to rounding
    (FP64 1e-13, FP32 1e-5 relative)."""
    prob = small_curvilinear(n=40, nmech=8, nsteps=3, seed=11)
    met = oracle.metrics(prob.X)
    rng = np.random.default_rng(7)
    F0 = rng.uniform(-1, 1, (9,) + prob.X.shape[1:])
    F0[3:] *= 1e6

    def run(tuning, chunks):
        monkeypatch.setenv("WQL3D_NODE_TUNING", tuning)
        monkeypatch.setenv("WQL3D_STAGE_CHUNKS", chunks)
        prob.F0 = F0.copy()          # the block's host F aliases prob.F0
        blk = gpu_block(prob, met, precision=precision)
        blk.set_receivers(prob.receivers, prob.nsteps)
        time_step_RK(blk, prob.dt, prob.nsteps, fused=True)
        _download_all(blk)
        out = (blk.F.copy(), blk.DF.copy(), blk.eta.copy(), blk.seismograms())
        blk.close()
        return out

    tol = 1e-13 if precision == L.WQL_FP64 else 1e-5
    ref = run("0,0,0,1,1,32", "1")
    failures = []
    for tuning, chunks, exact in (("0,0,0,1,1,32", "5", True), ("0,0,0,1,1,32", "13", True),
                                  ("0,1,0,1,2,5", "1", False), ("0,0,0,1,0,32", "4", False)):
        got = run(tuning, chunks)
        diff = [float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
                for a, b in zip(got, ref)]
        print(f"precision={precision} tuning={tuning} chunks={chunks}: max rel diff "
              f"{['%.1e' % d for d in diff]}", flush=True)
        if (max(diff) > 0.0) if exact else (max(diff) > tol):
            failures.append((tuning, chunks, diff))
    assert not failures, failures


@pytest.mark.parametrize("precision", [L.WQL_FP64, L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN])
def test_first_stage_does_not_read_old_rates(oracle, precision):
    """The first RK stage has A = 0, so its result cannot depend on the old
    rate registers (DF, Deta, PML DQ), and the stage kernels skip reading
    them there.  Starting from NaN rates instead of the initial 1e40 gives
    bit-identical fields, memory variables and PML auxiliaries after three
    steps, all finite (curvilinear Q8 + PML on every face kind)."""
    prob = small_curvilinear(n=30, nmech=8, npml=8, nsteps=3)
    met = oracle.metrics(prob.X)
    runs = []
    for fill in (None, np.nan):
        blk = gpu_block(prob, met, precision=precision)
        if fill is not None:
            blk.DF[...] = fill
            blk.Deta[...] = fill
            for f in range(6):
                if blk.pml_DQ[f] is not None:
                    blk.pml_DQ[f][...] = fill
        blk.upload(L.WQL_DF | L.WQL_MEMORY | L.WQL_PML)
        t = prob.sources[0].t_init - 2 * prob.dt
        time_step_RK(blk, prob.dt, 3, t)
        _download_all(blk)
        runs.append([blk.F.copy(), blk.DF.copy(), blk.eta.copy(), blk.Deta.copy()]
                    + [q.copy() for q in blk.pml_Q if q is not None]
                    + [q.copy() for q in blk.pml_DQ if q is not None])
        blk.close()
    assert np.abs(runs[0][0]).max() > 0
    for a, b in zip(*runs):
        assert np.isfinite(a).all() and np.isfinite(b).all()
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [L.WQL_FP64, L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN])
def test_dead_rate_stores_change_nothing(oracle, precision):
    """Inside one wql_time_step call the fifth stage of every step but the
    last does not store its rates (DF, Deta, PML DQ): the next step's first
    stage (A = 0) reads none of them.  One call of 4 steps must leave the
    whole state (F, DF, eta, Deta, PML Q and DQ) and the seismograms
    bit-identical to 4 calls of one step, which store every rate."""
    prob = small_curvilinear(n=30, nmech=8, npml=8, nsteps=4)
    met = oracle.metrics(prob.X)
    runs = []
    for per_call in (4, 1):
        blk = gpu_block(prob, met, precision=precision)
        blk.set_receivers(prob.receivers, 4)
        t = prob.sources[0].t_init - 2 * prob.dt
        for _ in range(4 // per_call):
            t = blk.time_step(prob.dt, per_call, t)
        _download_all(blk)
        runs.append([blk.F.copy(), blk.DF.copy(), blk.eta.copy(), blk.Deta.copy(),
                     blk.seismograms()]
                    + [q.copy() for q in blk.pml_Q if q is not None]
                    + [q.copy() for q in blk.pml_DQ if q is not None])
        blk.close()
    assert np.abs(runs[0][0]).max() > 0
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
