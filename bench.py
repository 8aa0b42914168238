"""Throughput of the fused per-block RK step on one MI355X.

Workload (default): 385^3 curvilinear block with topography, layered +
von Karman medium, anelastic Q8 (8 mechanisms), 20-node PML on four sides
and the bottom, free surface on top, Ricker moment-tensor source.  One
"step" = one Kennedy-Carpenter (5,4) RK step = 5 fused stage kernels.

Prints ONE JSON line: metric/value = FP32-mixed Mpt-RK-stage/s (grid points
x 5 stages x steps / wall time), with the FP64 path on the same block and
GPU timed beside it, the roofline of the fused stage kernel (HBM bound;
algorithmic bytes per node-stage / measured launch time), and the CPU
restatement timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpt-RK-stage/s & % HBM peak: FP32-mixed vs FP64 GPU at stated seismogram L2"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_launch(prob, precision):
    """Compulsory HBM bytes of one fused stage over the block: F read+write
    (ping-pong), DF read+write in FP64; geometry (9 metrics + J), material
    (3), memory variables eta/Deta read+write (4*6*nmech), Q^-1 (2) and PML
    auxiliaries Q/DQ read+write (36 per slab node) in the storage type."""
    nq, nr, ns = prob.n
    P = nq * nr * ns
    es = 8 if precision == 0 else 4
    ngeo = 10 if prob.mesh == 1 else 3
    per_node = 36 * 8 + (ngeo + 3) * es
    if prob.nmech:
        per_node += (24 * prob.nmech + 2) * es
    slab = 0
    for f in range(6):
        if prob.pml[f]:
            nn = [nq, nr, ns]
            nn[f // 2] = prob.npml[f]
            slab += nn[0] * nn[1] * nn[2]
    return P * per_node + slab * 36 * es, per_node


def timed_run(blk, prob, steps, warmup, t0):
    import torch
    st = torch.cuda.ExternalStream(blk.stream())
    t = blk.time_step(prob.dt, warmup, t0)
    blk.sync()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(st)
    t = blk.time_step(prob.dt, steps, t)
    e1.record(st)
    e1.synchronize()
    blk.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_time(e1)
    return ms, wall, t


def cpu_baseline(n_sample, nmech, npml, seed):
    """Oracle (scalar C, 1 thread) on a sub-block of the same set-up, 1 step."""
    from oracle import oracle as O
    from waveqlab3d_q_gpu_amd import problems as PB
    prob = PB.curvilinear_q(n=n_sample, h=50.0, nsteps=1, nmech=nmech, npml=npml, seed=seed,
                            nrec=4)
    met = O.metrics(prob.X)
    nq, nr, ns = prob.n
    b = O.OracleBlock(nq, nr, ns, curvilinear=1)
    b.metx, b.mety, b.metz, b.J = met
    b.M[...] = prob.M
    b.bc = list(prob.bc)
    for f in range(6):
        if prob.pml[f]:
            b.set_pml(f, prob.npml[f])
    if nmech:
        b.set_anelastic(nmech, prob.Qs_inv, prob.Qp_inv, correct_modulus=False)
    b.normals()
    t0 = time.perf_counter()
    b.time_step(prob.dt, 1, 0.0)
    el = time.perf_counter() - t0
    P = nq * nr * ns
    return {"value": P * 5 / el / 1e6, "unit": "Mpt-RK-stage/s", "cores": 1, "kind": "port",
            "sample": f"{n_sample}^3 sub-block of the same curvilinear Q{nmech} + PML{npml} "
                      f"set-up, 1 step (5 stages), oracle/wql_oracle.c (gcc -O2, 1 thread)",
            "seconds": el}


def read_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            return json.load(fh)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=385)
    ap.add_argument("--nmech", type=int, default=8)
    ap.add_argument("--npml", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=97)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device copy of this many MiB 5 times (scale factor "
                         "for FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X; multi-GPU runs are out of scope "
              "(one block per GPU, no halo exchange). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib * 2 ** 20 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        # the library's own streaming kernels on n/2 doubles: k_scale reads and
        # writes 4n bytes, k_update reads 8n and writes 4n bytes
        from waveqlab3d_q_gpu_amd import scale_rate_array, update_state_array
        x = torch.ones(n // 2, dtype=torch.float64, device="cuda")
        y = torch.ones(n // 2, dtype=torch.float64, device="cuda")
        for _ in range(5):
            scale_rate_array(x, 0.5)
            update_state_array(y, x, 0.25)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n,
                          "k_scale_bytes_read": 4 * n, "k_update_bytes_read": 8 * n,
                          "launches": 5}), flush=True)
        return

    import ctypes

    import torch

    from waveqlab3d_q_gpu_amd import _lib as L
    from waveqlab3d_q_gpu_amd import problems as PB

    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    print(f"[bench] building {args.n}^3 curvilinear Q{args.nmech} PML{args.npml} set-up",
          file=sys.stderr, flush=True)
    prob = PB.curvilinear_q(n=args.n, h=50.0, nsteps=args.steps, nmech=args.nmech,
                            npml=args.npml, seed=2024, nrec=64)
    nq, nr, ns = prob.n
    P = nq * nr * ns
    # start so that the Ricker peak falls inside the timed window
    t0 = prob.sources[0].t_init - 0.5 * (args.steps + args.warmup) * prob.dt

    results = {}
    finals = {}
    for prec, name in ((L.WQL_FP32_MIXED, "fp32"), (L.WQL_FP64, "fp64")):
        if name == "fp64" and args.no_fp64:
            continue
        print(f"[bench] {name}: creating block", file=sys.stderr, flush=True)
        blk = PB.make_device_block(prob, precision=prec)
        blk.set_receivers(prob.receivers, args.steps + args.warmup)
        print(f"[bench] {name}: device bytes {blk.device_bytes() / 2**30:.1f} GiB; timing "
              f"{args.warmup}+{args.steps} steps", file=sys.stderr, flush=True)
        ms, wall, _ = timed_run(blk, prob, args.steps, args.warmup, t0)
        nbytes, per_node = algorithmic_bytes_per_launch(prob, prec)
        launch_ms = ms / (5 * args.steps)
        results[name] = {
            "value": P * 5 * args.steps / (ms / 1e3) / 1e6,
            "ms_per_step": ms / args.steps,
            "wall_ms_per_step": wall * 1e3 / args.steps,
            "launch_ms": launch_ms,
            "bytes_per_launch": nbytes,
            "bytes_per_node_stage": nbytes / P,
            "achieved_GBs": nbytes / (launch_ms / 1e3) / 1e9,
        }
        print(f"[bench] {name}: {results[name]['value']:.1f} Mpt-RK-stage/s, "
              f"{results[name]['achieved_GBs']:.0f} GB/s", file=sys.stderr, flush=True)
        blk.F = np.zeros((9, ns, nr, nq))
        blk._desc.F = blk.F.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        blk._desc.arrays_on_device = 0
        finals[name] = (blk.seismograms(), blk)
        if name == "fp32":
            # keep only what the comparison needs
            pass

    # FP32-mixed vs FP64 on the same GPU, same state and time window
    acc = {}
    if "fp64" in finals:
        s32, b32 = finals["fp32"]
        s64, b64 = finals["fp64"]
        seis_err = []
        for r in range(s64.shape[1]):
            den = np.linalg.norm(s64[:, r].ravel())
            if den > 0:
                seis_err.append(float(np.linalg.norm((s32[:, r] - s64[:, r]).ravel()) / den))
        acc["seismogram_rel_l2_max_fp32_vs_fp64"] = max(seis_err) if seis_err else None
        acc["receivers_with_signal"] = len(seis_err)

    r32 = results["fp32"]
    traffic = read_traffic("traffic_stage_fused_fp32_q8_385.json")
    roof = {"bound": "hbm", "achieved": round(r32["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(r32["achieved_GBs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("bytes_per_launch") if traffic else None}
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_sample, args.nmech, args.npml, 2024)
    out = {
        "metric": METRIC, "value": round(r32["value"], 2), "unit": "Mpt-RK-stage/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 stencil/memory + f64 RK accumulation",
        "data": "synthetic (seeded topography, von Karman medium, Ricker source)",
        "config": {"workload": f"{nq}^3 curvilinear block, anelastic Q{args.nmech}, PML "
                               f"{args.npml} on 4 sides + bottom, free surface, h=50 m, "
                               "fused RHS+Q+PML+SAT+source+RK stage",
                   "nodes": P, "precision": "fp32-mixed"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "fp64" in results:
        r64 = results["fp64"]
        out["fp64"] = {"value": round(r64["value"], 2), "ms_per_step": round(r64["ms_per_step"], 4),
                       "achieved_GBs": round(r64["achieved_GBs"], 1),
                       "frac": round(r64["achieved_GBs"] / HBM_PEAK_GBS, 4),
                       "bytes_per_node_stage": r64["bytes_per_node_stage"]}
        out["speedup_fp32_mixed_vs_fp64"] = round(r32["value"] / r64["value"], 3)
    out["fp32_bytes_per_node_stage"] = r32["bytes_per_node_stage"]
    out["accuracy"] = acc
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
