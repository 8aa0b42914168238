"""Gradient check of the discrete adjoint on the differentiation workload.

``python -m waveqlab3d_q_gpu_amd.adjoint_check`` (defaults: 129^3 curvilinear
block, h = 100 m, seeded topography, 5 % von Karman perturbations of rho and of
the wave speeds, Q4, characteristic SAT on the sides and bottom, free surface
on top, Ricker source f0 = 1.5 Hz, 400 steps, 64 surface receivers,
checkpoints every 20 steps).  Loss: L2 misfit of the velocity traces against
synthetic observations of a seeded 2 % perturbed model.  Prints one JSON line:
  * dot-product test <Phi x, y> vs <x, Phi^T y> for Phi = 10 source-free steps
    (FP64);
  * the FP64 adjoint gradient against FP64 central differences along 16 seeded
    random directions (relative error per direction);
  * the FP32-mixed adjoint gradient against the FP64 one (relative L2 per
    parameter);
  * wall times of the forward run, the gradient and the directional checks.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import time

import numpy as np


def build_problem(n, h, nsteps, f0, seed):
    from . import problems as PB
    prob = PB.curvilinear_q(n=n, h=h, nsteps=nsteps, nmech=4, npml=0, seed=seed, nrec=64, f0=f0)
    # heterogeneous density: an independent 5 % von Karman field
    rho = prob.M[2] * (1.0 + PB.von_karman(prob.M.shape[1:], h, seed + 23))
    lam, mu = prob.M[0], prob.M[1]
    ratio = rho / prob.M[2]
    prob.M = np.stack([lam * ratio, mu * ratio, rho])   # keep the wave speeds
    return prob


def main():
    import torch

    from . import _lib as L
    from . import problems as PB
    from .adjoint import AdjointRunner, receiver_traces

    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=129)
    ap.add_argument("--h", type=float, default=100.0)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--f0", type=float, default=1.5)
    ap.add_argument("--dirs", type=int, default=16)
    ap.add_argument("--eps", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--no-fp32", action="store_true")
    args = ap.parse_args()

    dev_id = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev_id)))
    prob = build_problem(args.n, args.h, args.steps, args.f0, args.seed)
    out = {"workload": f"{args.n}^3 curvilinear, h={args.h} m, Q4, Ricker f0={args.f0} Hz, "
                       f"{args.steps} steps, 64 surface receivers, checkpoints every "
                       f"{args.every} steps", "dt": prob.dt}

    def runner_for(precision):
        blk = PB.make_device_block(prob, precision=precision, state_mirror=True)
        return blk, AdjointRunner(blk, prob.dt, args.steps, prob.receivers, every=args.every)

    blk, run = runner_for(L.WQL_FP64)
    dev = blk.M.device
    lam0, mu0, rho0 = (blk.M[k].clone() for k in range(3))
    shape = lam0.shape

    # observations: a seeded, smooth 2 % perturbation of every parameter
    pert = [torch.from_numpy(1.0 + 0.4 * PB.von_karman(shape, args.h, args.seed + 100 + k)).to(dev)
            for k in range(3)]   # von_karman has 5 % std -> 2 % after the 0.4 factor
    with torch.no_grad():
        t0 = time.perf_counter()
        obs = receiver_traces(rho0 * pert[0], lam0 * pert[1], mu0 * pert[2], run).clone()
        torch.cuda.synchronize()
        out["forward_s"] = round(time.perf_counter() - t0, 3)

    def misfit(rho, lam, mu):
        tr = receiver_traces(rho, lam, mu, run)
        return 0.5 * ((tr[..., :3] - obs[..., :3]) ** 2).sum()

    rho = rho0.clone().requires_grad_(True)
    lam = lam0.clone().requires_grad_(True)
    mu = mu0.clone().requires_grad_(True)
    t0 = time.perf_counter()
    J = misfit(rho, lam, mu)
    J.backward()
    torch.cuda.synchronize()
    out["loss"] = float(J.detach())
    out["gradient_s"] = round(time.perf_counter() - t0, 3)
    g64 = [rho.grad.detach().clone(), lam.grad.detach().clone(), mu.grad.detach().clone()]

    # dot-product test of 10 source-free steps on the same block
    K = 10
    rng = np.random.default_rng(args.seed)
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
    blk.upload(L.WQL_F | L.WQL_MEMORY)
    blk.time_step(prob.dt, K, 0.0)
    blk.download(L.WQL_F | L.WQL_MEMORY)
    lhs = float((blk.F * yF).sum() + (blk.eta * yE).sum())
    Fb, Eb = yF.clone(), yE.clone()
    run.adjoint_steps(Fb, Eb, K)
    rhs = float((xF * Fb).sum() + (xE * Eb).sum())
    out["dot_product"] = {"steps": K, "lhs": lhs, "rhs": rhs, "rel": abs(lhs - rhs) / abs(lhs)}
    blk.set_sources(prob.sources)

    # central differences along seeded random directions
    t0 = time.perf_counter()
    rels = []
    for k in range(args.dirs):
        d = [torch.from_numpy(PB.von_karman(shape, args.h, args.seed + 1000 + 3 * k + j) / 0.05)
             .to(dev) * m for j, m in enumerate((rho0, lam0, mu0))]
        adj = sum(float((g * dk).sum()) for g, dk in zip(g64, d))
        with torch.no_grad():
            jp = float(misfit(rho0 + args.eps * d[0], lam0 + args.eps * d[1], mu0 + args.eps * d[2]))
            jm = float(misfit(rho0 - args.eps * d[0], lam0 - args.eps * d[1], mu0 - args.eps * d[2]))
        fd = (jp - jm) / (2 * args.eps)
        rels.append(abs(fd - adj) / abs(fd))
    out["fd_rel_err"] = [float("%.3e" % r) for r in rels]
    out["fd_rel_err_max"] = max(rels)
    out["fd_s"] = round(time.perf_counter() - t0, 3)
    blk.close()
    del run

    if not args.no_fp32:
        blk32, run32 = runner_for(L.WQL_FP32_MIXED)
        with torch.no_grad():
            obs32 = obs
        rho = rho0.clone().requires_grad_(True)
        lam = lam0.clone().requires_grad_(True)
        mu = mu0.clone().requires_grad_(True)
        t0 = time.perf_counter()
        tr = receiver_traces(rho, lam, mu, run32)
        J32 = 0.5 * ((tr[..., :3] - obs32[..., :3]) ** 2).sum()
        J32.backward()
        torch.cuda.synchronize()
        g32 = [rho.grad, lam.grad, mu.grad]
        out["fp32_mixed"] = {
            "loss": float(J32.detach()), "gradient_s": round(time.perf_counter() - t0, 3),
            "grad_rel_l2_vs_fp64": [float("%.3e" % (float(torch.linalg.norm(a - b)) /
                                                    float(torch.linalg.norm(b))))
                                    for a, b in zip(g32, g64)]}
        blk32.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
