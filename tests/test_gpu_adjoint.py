"""Discrete adjoint of the fused RK step (libwql3d_hip.so wql_adjoint_*).

There is no reference counterpart; the adjoint is pinned against the
oracle-pinned forward path by
  * the dot-product identity <Phi x, y> = <x, Phi^T y> for Phi = N source-free
    RK steps (FP64, relative 1e-10), and
  * FP64 central finite differences of a receiver misfit along seeded random
    material directions (relative 1e-6).
"""
import ctypes

import numpy as np
import pytest

from waveqlab3d_q_gpu_amd import _lib as L
from waveqlab3d_q_gpu_amd import problems as PB

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _runtime():
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    yield


def _problem(kind, n=24, nsteps=3, f0=20.0):
    if kind == "cart_elastic":
        prob = PB.curvilinear_q(n=n, h=100.0, nsteps=nsteps, nmech=0, npml=0, seed=31, nrec=4,
                                f0=f0, hills=False)
        prob.mesh = L.WQL_CARTESIAN
        return prob
    nmech = {"curv_elastic": 0, "curv_q4": 4, "curv_q8": 8}[kind]
    return PB.curvilinear_q(n=n, h=100.0, nsteps=nsteps, nmech=nmech, npml=0, seed=31, nrec=4,
                            f0=f0)


@pytest.mark.parametrize("kind", ["cart_elastic", "curv_elastic", "curv_q4", "curv_q8"])
def test_adjoint_dot_product(kind):
    import torch
    from waveqlab3d_q_gpu_amd.adjoint import AdjointRunner
    prob = _problem(kind)
    K = 3
    blk = PB.make_device_block(prob, precision=L.WQL_FP64, state_mirror=True)
    blk.set_sources([])
    run = AdjointRunner(blk, prob.dt, K, prob.receivers)
    rng = np.random.default_rng(17)
    shp = blk.F.shape
    dev = blk.F.device
    xF = torch.from_numpy(rng.uniform(-1, 1, shp)).to(dev)
    yF = torch.from_numpy(rng.uniform(-1, 1, shp)).to(dev)
    # stresses scaled so that every component contributes to the products
    xF[3:] *= 1e6
    yF[3:] *= 1e-6
    blk.F.copy_(xF)
    mask = L.WQL_F
    if prob.nmech:
        xE = torch.from_numpy(rng.uniform(-1, 1, blk.eta.shape) * 1e3).to(dev)
        yE = torch.from_numpy(rng.uniform(-1, 1, blk.eta.shape) * 1e-3).to(dev)
        blk.eta.copy_(xE)
        blk.Deta.zero_()
        mask |= L.WQL_MEMORY
    blk.upload(mask)
    blk.time_step(prob.dt, K, 0.0)
    blk.download(mask)
    lhs = float((blk.F * yF).sum())
    if prob.nmech:
        lhs += float((blk.eta * yE).sum())
    Fb = yF.clone().contiguous()
    Eb = yE.clone().contiguous() if prob.nmech else None
    run.adjoint_steps(Fb, Eb, K)
    rhs = float((xF * Fb).sum())
    if prob.nmech:
        rhs += float((xE * Eb).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"{kind}: <Phi x, y> = {lhs!r}  <x, Phi^T y> = {rhs!r}  rel {rel:.2e}", flush=True)
    assert rel <= 1e-10


def test_adjoint_gradient_matches_central_differences():
    import torch
    from waveqlab3d_q_gpu_amd.adjoint import AdjointRunner, receiver_traces
    N = 40
    prob = _problem("curv_q4", n=28, nsteps=N, f0=20.0)
    prob.sources[0].t_init = 0.06
    near = np.asarray(prob.sources[0].near)
    rec = np.asarray([near + o for o in ((5, 0, 0), (-5, 0, 0), (0, -5, 0), (0, 5, 0),
                                          (0, 0, 5), (0, 0, -5), (4, -4, 4), (-4, 4, -4))])
    blk = PB.make_device_block(prob, precision=L.WQL_FP64, state_mirror=True)
    run = AdjointRunner(blk, prob.dt, N, rec, every=7)
    dev = blk.M.device
    M0 = blk.M.clone()
    lam0, mu0, rho0 = M0[0].clone(), M0[1].clone(), M0[2].clone()
    rng = np.random.default_rng(5)
    pert = [torch.from_numpy(1.0 + 0.02 * rng.uniform(-1, 1, lam0.shape)).to(dev) for _ in range(3)]
    with torch.no_grad():
        obs = receiver_traces(rho0 * pert[0], lam0 * pert[1], mu0 * pert[2], run).clone()

    def misfit(rho, lam, mu):
        tr = receiver_traces(rho, lam, mu, run)
        return 0.5 * ((tr[..., :3] - obs[..., :3]) ** 2).sum()

    rho = rho0.clone().requires_grad_(True)
    lam = lam0.clone().requires_grad_(True)
    mu = mu0.clone().requires_grad_(True)
    J = misfit(rho, lam, mu)
    J.backward()
    assert float(J.detach()) > 0
    grads = (rho.grad, lam.grad, mu.grad)
    for k in range(4):
        d = [torch.from_numpy(rng.uniform(-1, 1, lam0.shape)).to(dev) * m
             for m in (rho0, lam0, mu0)]
        adj = sum(float((g * dk).sum()) for g, dk in zip(grads, d))
        eps = 1e-5
        with torch.no_grad():
            jp = float(misfit(rho0 + eps * d[0], lam0 + eps * d[1], mu0 + eps * d[2]))
            jm = float(misfit(rho0 - eps * d[0], lam0 - eps * d[1], mu0 - eps * d[2]))
        fd = (jp - jm) / (2 * eps)
        rel = abs(fd - adj) / abs(fd)
        print(f"direction {k}: fd {fd!r} adjoint {adj!r} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6


@pytest.mark.parametrize("precision", [L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN])
def test_adjoint_fp32_matches_fp64(precision):
    """The FP32-mixed block (float stencil and memory arithmetic, FP64 RK
    state) and the FP32-Kahan block (FP32 RK state with its compensation,
    both checkpointed) through the same adjoint runner: checkpoint reloads and
    the forward recomputation between them (which rebuild the stencils' float
    copy of F) give the FP64 misfit and gradient to FP32 accuracy."""
    import torch
    from waveqlab3d_q_gpu_amd.adjoint import AdjointRunner, receiver_traces
    N = 40
    prob = _problem("curv_q4", n=28, nsteps=N, f0=20.0)
    prob.sources[0].t_init = 0.06
    near = np.asarray(prob.sources[0].near)
    rec = np.asarray([near + o for o in ((5, 0, 0), (-5, 0, 0), (0, -5, 0), (0, 5, 0))])
    out = {}
    for prec in (L.WQL_FP64, precision):
        blk = PB.make_device_block(prob, precision=prec, state_mirror=True)
        run = AdjointRunner(blk, prob.dt, N, rec, every=7)
        dev = blk.M.device
        M0 = blk.M.clone()
        lam0, mu0, rho0 = M0[0].clone(), M0[1].clone(), M0[2].clone()
        rng = np.random.default_rng(5)
        pert = [torch.from_numpy(1.0 + 0.02 * rng.uniform(-1, 1, lam0.shape)).to(dev)
                for _ in range(3)]
        with torch.no_grad():
            obs = receiver_traces(rho0 * pert[0], lam0 * pert[1], mu0 * pert[2], run).clone()
        rho = rho0.clone().requires_grad_(True)
        lam = lam0.clone().requires_grad_(True)
        mu = mu0.clone().requires_grad_(True)
        tr = receiver_traces(rho, lam, mu, run)
        J = 0.5 * ((tr[..., :3] - obs[..., :3]) ** 2).sum()
        J.backward()
        d = [torch.from_numpy(rng.uniform(-1, 1, lam0.shape)).to(dev) * m
             for m in (rho0, lam0, mu0)]
        dJ = sum(float((g * dk).sum()) for g, dk in zip((rho.grad, lam.grad, mu.grad), d))
        out[prec] = (float(J.detach()), dJ)
        blk.close()
    (J64, g64), (J32, g32) = out[L.WQL_FP64], out[precision]
    print(f"misfit FP64 {J64!r} FP32 (precision {precision}) {J32!r}; "
          f"directional gradient {g64!r} {g32!r}", flush=True)
    assert J64 > 0 and abs(J32 - J64) / J64 <= 1e-4
    assert abs(g32 - g64) / abs(g64) <= 1e-3


@pytest.mark.parametrize("precision", [L.WQL_FP64, L.WQL_FP32_MIXED, L.WQL_FP32_KAHAN])
def test_checkpoint_reload_steps_bit_identically(precision):
    """wql_state_save / wql_state_load hold everything the next steps read:
    two steps from a reloaded checkpoint repeat the two steps taken after the
    save bit for bit (FP64, FP32-mixed with its float copy of F, FP32-Kahan
    with the compensation of its FP32 state)."""
    import torch
    prob = _problem("curv_q8", n=24, nsteps=5)
    blk = PB.make_device_block(prob, precision=precision, state_mirror=True)
    L.check(L.lib().wql_adjoint_create(blk._h, 1))
    blk.time_step(prob.dt, 3, 0.0)
    L.check(L.lib().wql_state_save(blk._h, 0))
    t0 = 3 * prob.dt
    blk.time_step(prob.dt, 2, t0)
    blk.download(L.WQL_F | L.WQL_MEMORY)
    torch.cuda.synchronize()
    a = (blk.F.clone(), blk.eta.clone())
    L.check(L.lib().wql_state_load(blk._h, 0))
    blk.time_step(prob.dt, 2, t0)
    blk.download(L.WQL_F | L.WQL_MEMORY)
    torch.cuda.synchronize()
    same = [bool(torch.equal(a[0], blk.F)), bool(torch.equal(a[1], blk.eta))]
    moved = float(torch.abs(a[0]).max())
    print(f"precision {precision}: reload repeat identical (F, eta) = {same}, max|F| {moved:.3e}",
          flush=True)
    blk.close()
    assert moved > 0 and all(same)


def test_f_only_slots():
    """wql_adjoint_create_f: the F-only slots after the full ones take and
    give back F (with eta from a full slot, a reloaded stage input repeats
    the stage bit for bit) and refuse the memory variables; slots past the
    last are refused."""
    import torch
    prob = _problem("curv_q8", n=24, nsteps=3)
    blk = PB.make_device_block(prob, precision=L.WQL_FP64, state_mirror=True)
    lib, h = L.lib(), blk._h
    L.check(lib.wql_adjoint_create_f(h, 1, 2))
    blk.time_step(prob.dt, 2, 0.0)
    L.check(lib.wql_state_save_part(h, 0, L.WQL_MEMORY))
    L.check(lib.wql_state_save_part(h, 2, L.WQL_F))
    assert lib.wql_state_save_part(h, 2, L.WQL_MEMORY) == L.WQL_EINVAL
    assert lib.wql_state_load(h, 1) == L.WQL_EINVAL
    assert lib.wql_state_save_part(h, 3, L.WQL_F) == L.WQL_EINVAL
    # a first stage (A = 0 reads no old rate) from the saved state, twice
    t = 2 * prob.dt
    L.check(lib.wql_stage_fused(h, 0.0, 0.1 * prob.dt, t))
    blk.download(L.WQL_F)
    torch.cuda.synchronize()
    first = blk.F.clone()
    L.check(lib.wql_state_load_part(h, 0, L.WQL_MEMORY))
    L.check(lib.wql_state_load_part(h, 2, L.WQL_F))
    L.check(lib.wql_stage_fused(h, 0.0, 0.1 * prob.dt, t))
    blk.download(L.WQL_F)
    torch.cuda.synchronize()
    same = bool(torch.equal(first, blk.F))
    blk.close()
    assert same and float(torch.abs(first).max()) > 0


def test_adjoint_refuses_pml_slabs():
    """A block with PML slabs has no adjoint: wql_adjoint_create and
    wql_adjoint_create_f return WQL_EINVAL naming the face, allocate nothing,
    and the block keeps stepping."""
    prob = PB.curvilinear_q(n=24, h=100.0, nsteps=2, nmech=4, npml=5, seed=31, nrec=4)
    blk = PB.make_device_block(prob, precision=L.WQL_FP64, state_mirror=True)
    lib, h = L.lib(), blk._h
    try:
        assert lib.wql_adjoint_create(h, 2) == L.WQL_EINVAL
        assert b"PML" in lib.wql_last_error()
        assert lib.wql_adjoint_create_f(h, 1, 5) == L.WQL_EINVAL
        assert lib.wql_state_save(h, 0) == L.WQL_ESTATE
        blk.time_step(prob.dt, 2, 0.0)
        blk.download(L.WQL_F)
        import torch
        torch.cuda.synchronize()
        assert bool(torch.isfinite(blk.F).all())
    finally:
        blk.close()


def test_config4_adjoint_gradient_129():
    """The 129^3 Q4, 400-step workload with checkpoints every 20 steps: the
    loss is the L2 velocity misfit at 64 surface receivers against synthetic
    observations of a seeded 2 % perturbed model; the FP64 adjoint gradient
    with respect to per-node rho, lambda, mu agrees with FP64 central
    differences along 16 seeded random directions (relative 1e-6; measured
    <= 3.3e-8), the dot-product identity holds over 10 source-free steps of
    the same block (1e-10; measured 2.3e-15), and the FP32-mixed gradient is
    within 1e-4 of the FP64 one (measured 3e-6)."""
    import time

    import torch

    from waveqlab3d_q_gpu_amd.adjoint import AdjointRunner, receiver_traces
    n, h, steps, every, eps, seed = 129, 100.0, 400, 20, 1e-5, 4
    prob = PB.differentiation_q4(n, h, steps, 1.5, seed)
    print(f"config 4: {n}^3 Q4, {steps} steps, checkpoints every {every}", flush=True)

    def runner_for(precision):
        blk = PB.make_device_block(prob, precision=precision, state_mirror=True)
        return blk, AdjointRunner(blk, prob.dt, steps, prob.receivers, every=every)

    blk, run = runner_for(L.WQL_FP64)
    dev = blk.M.device
    lam0, mu0, rho0 = (blk.M[k].clone() for k in range(3))
    shape = lam0.shape
    pert = [torch.from_numpy(1.0 + 0.4 * PB.von_karman(shape, h, seed + 100 + k)).to(dev)
            for k in range(3)]   # 5 % std x 0.4 = 2 %
    with torch.no_grad():
        obs = receiver_traces(rho0 * pert[0], lam0 * pert[1], mu0 * pert[2], run).clone()

    def misfit(rho, lam, mu, r=run):
        tr = receiver_traces(rho, lam, mu, r)
        return 0.5 * ((tr[..., :3] - obs[..., :3]) ** 2).sum()

    rho = rho0.clone().requires_grad_(True)
    lam = lam0.clone().requires_grad_(True)
    mu = mu0.clone().requires_grad_(True)
    t0 = time.perf_counter()
    J = misfit(rho, lam, mu)
    J.backward()
    torch.cuda.synchronize()
    print(f"config 4: FP64 loss {float(J.detach()):.10e}, gradient "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    g64 = [rho.grad.detach().clone(), lam.grad.detach().clone(), mu.grad.detach().clone()]

    # dot-product identity over 10 source-free steps of the same block
    K = 10
    rng = np.random.default_rng(seed)
    blk.set_sources([])
    xF = torch.from_numpy(rng.uniform(-1, 1, blk.F.shape)).to(dev)
    yF = torch.from_numpy(rng.uniform(-1, 1, blk.F.shape)).to(dev)
    xF[3:] *= 1e6
    yF[3:] *= 1e-6
    xE = torch.from_numpy(rng.uniform(-1, 1, blk.eta.shape) * 1e3).to(dev)
    yE = torch.from_numpy(rng.uniform(-1, 1, blk.eta.shape) * 1e-3).to(dev)
    blk.M[0].copy_(lam0); blk.M[1].copy_(mu0); blk.M[2].copy_(rho0)
    blk.upload(L.WQL_MATERIAL)
    blk.F.copy_(xF); blk.eta.copy_(xE); blk.Deta.zero_()
    torch.cuda.synchronize()
    blk.upload(L.WQL_F | L.WQL_MEMORY)
    blk.time_step(prob.dt, K, 0.0)
    blk.download(L.WQL_F | L.WQL_MEMORY)
    lhs = float((blk.F * yF).sum() + (blk.eta * yE).sum())
    Fb, Eb = yF.clone(), yE.clone()
    run.adjoint_steps(Fb, Eb, K)
    rhs = float((xF * Fb).sum() + (xE * Eb).sum())
    dp = abs(lhs - rhs) / abs(lhs)
    print(f"config 4: dot product {lhs!r} vs {rhs!r}, rel {dp:.2e}", flush=True)
    assert dp <= 1e-10
    blk.set_sources(prob.sources)

    # central differences along 16 seeded random directions
    rels = []
    for k in range(16):
        d = [torch.from_numpy(PB.von_karman(shape, h, seed + 1000 + 3 * k + j) / 0.05)
             .to(dev) * m for j, m in enumerate((rho0, lam0, mu0))]
        adj = sum(float((g * dk).sum()) for g, dk in zip(g64, d))
        with torch.no_grad():
            jp = float(misfit(rho0 + eps * d[0], lam0 + eps * d[1], mu0 + eps * d[2]))
            jm = float(misfit(rho0 - eps * d[0], lam0 - eps * d[1], mu0 - eps * d[2]))
        fd = (jp - jm) / (2 * eps)
        rels.append(abs(fd - adj) / abs(fd))
    print("config 4: FD rel errors", ["%.2e" % r for r in rels], flush=True)
    assert max(rels) <= 1e-6
    blk.close()
    del run

    blk32, run32 = runner_for(L.WQL_FP32_MIXED)
    rho = rho0.clone().requires_grad_(True)
    lam = lam0.clone().requires_grad_(True)
    mu = mu0.clone().requires_grad_(True)
    J32 = misfit(rho, lam, mu, run32)
    J32.backward()
    g32 = [rho.grad, lam.grad, mu.grad]
    errs = [float(torch.linalg.norm(a - b)) / float(torch.linalg.norm(b)) for a, b in zip(g32, g64)]
    print("config 4: FP32-mixed gradient rel L2 vs FP64", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= 1e-4
    blk32.close()
