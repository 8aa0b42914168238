"""Throughput of the per-block RK step on one MI355X.

Workload (default): 385^3 curvilinear block with topography, layered +
von Karman medium, anelastic Q8 (8 mechanisms), 20-node PML on four sides
and the bottom, free surface on top, Ricker moment-tensor source.  One
"step" = one Kennedy-Carpenter (5,4) RK step = 5 stages; one stage = the
strict-interior stencil kernel, the node kernel over the rest of the block
(closures, PML, SAT) and the streaming memory-variable + RK kernel.

Prints ONE JSON line: metric/value = FP32-mixed Mpt-RK-stage/s (grid points
x 5 stages x steps / time on the block stream), the FP64 path on the same
block and GPU beside it, the per-kernel durations (HIP events on the block
stream), the roofline of the dominant kernel (HBM bound; algorithmic bytes
per launch / measured launch time; traffic from rocprofv3 counters when the
committed profile is present), the FP32-mixed vs FP64 error per field and
per receiver seismogram, and the CPU restatement timed on a bounded sample
of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpt-RK-stage/s & % HBM peak: FP32-mixed vs FP64 GPU at stated seismogram L2"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = "traffic_memory_update_fp32_q8_385.json"


def active_mechanisms(prob):
    """Mechanisms whose state can be non-zero: a zero-weight mechanism that
    starts at zero stays zero and is never read or written."""
    if not prob.nmech:
        return 0
    return int(np.count_nonzero(np.asarray(prob.weight)))


def interior_box(prob):
    """Strict interior (0-based inclusive) per axis: max(npml, 6) from PML
    faces, 6 from the others."""
    lo = [max(prob.npml[2 * a], 6) if prob.pml[2 * a] else 6 for a in range(3)]
    hi = [prob.n[a] - 1 - (max(prob.npml[2 * a + 1], 6) if prob.pml[2 * a + 1] else 6)
          for a in range(3)]
    return lo, hi


def slab_nodes(prob):
    nq, nr, ns = prob.n
    tot = 0
    for f in range(6):
        if prob.pml[f]:
            nn = [nq, nr, ns]
            nn[f // 2] = prob.npml[f]
            tot += nn[0] * nn[1] * nn[2]
    return tot


# share of the stages of the KC(5,4) low-storage scheme with A = 0 (the
# first of five), which read no old rate
FIRST_STAGE_SHARE = 1 / 5


def byte_model(prob, precision, steps=None):
    """Algorithmic HBM bytes of one stage.

    End to end (what the step must move, intermediates excluded): F read and
    write (ping-pong) and DF read and write in FP64 (36 x 8 B); geometry
    (curvilinear 9 metrics + J, Cartesian 3) and material (3) in the storage
    type; eta and Deta read and write for the active mechanisms and Q^-1 (2);
    PML Q and DQ read and write (36 values per slab node).

    Per kernel (what each kernel of the stage must move, its W hand-off
    included): interior stencil = F read (9), geometry, material, the
    velocity components' F write and DF read and write (3), W write (6
    strain rates); shell = the same per node with W write 12 (six stress
    rates as well) plus the PML streams; memory-variable kernel = W read (6,
    and 6 more at shell nodes), lambda, mu, Q^-1 (4), eta/Deta read and
    write, and the six stress components' F read and write and DF read and
    write.  FP32-mixed: the stencils read F rounded to float
    (9 x 4 B) plus the velocity components' FP64 F (3 x 8 B), and every F
    write also writes the float copy (4 B per component).

    The first of the five stages has A = 0 and reads no old rate (DF, Deta,
    PML DQ), so the rates the fifth stage stores are dead whenever another
    step of the same wql_time_step call follows, and they are not stored:
    every figure is the average over the stages of a call of ``steps`` steps
    (None: a long run), i.e. the full-stage count less a fifth of the
    old-rate reads and (steps - 1) / steps of a fifth of the rate stores,
    which move the same planes (fields "rate_reads", "first_stage_share",
    "dead_store_share")."""
    nq, nr, ns = prob.n
    P = nq * nr * ns
    es = 8 if precision == 0 else 4
    # RK state bytes per component: F read + write and DF read + write; the
    # FP32-Kahan state adds the compensation plane (read + write)
    kahan = precision == 2
    ss = 4 if kahan else 8
    state = 4 * ss + (8 if kahan else 0)
    ngeo = 10 if prob.mesh == 1 else 3
    na = active_mechanisms(prob)
    mem = (24 * na + 2) * es if prob.nmech else 0
    slab = slab_nodes(prob)
    end_to_end = P * (9 * state + (ngeo + 3) * es + mem) + slab * 36 * es
    lo, hi = interior_box(prob)
    n_int = int(np.prod([max(0, hi[a] - lo[a] + 1) for a in range(3)]))
    # elastic stages have no hand-off: the node kernels write F and DF
    # velocity update in the stencil kernels: F write, DF read + write (and
    # the Kahan compensation), the F read being the stencil's own centre value
    vel = 3 * (state - ss)
    opcopy = precision == 1
    if opcopy:
        # float operands; the updated components' FP64 F read and float copy written
        nupd = 3 if prob.nmech else 9
        operands = 9 * 4 + nupd * (8 + 4)
    else:
        operands = 9 * ss
    # W hand-off: 12 planes (6 stress rates, 6 strain rates) from the shell;
    # the strict interior hands over the strain rates only (its stress rates
    # are Hooke's law of them, formed again by the memory-variable kernel)
    shell_node = operands + (ngeo + 3) * es + (12 * es + vel if prob.nmech else 9 * (state - ss))
    interior_node = operands + (ngeo + 3) * es + (6 * es + vel if prob.nmech else 9 * (state - ss))
    kernels = {
        "interior": n_int * interior_node,
        "shell": (P - n_int) * shell_node + slab * 36 * es,
        "memory_update": ((P * (6 * es + 4 * es + 24 * na * es + 6 * state + (6 * 4 if opcopy else 0))
                           + (P - n_int) * 6 * es) if prob.nmech else 0),
    }
    # old-rate reads that the first stage (A = 0) skips: DF (ss per
    # component), Deta (6 per active mechanism) and PML DQ (9 per slab node)
    vel_rd = 3 * ss if prob.nmech else 9 * ss
    rate_reads = {
        "interior": n_int * vel_rd,
        "shell": (P - n_int) * vel_rd + slab * 9 * es,
        "memory_update": P * (6 * ss + 6 * na * es) if prob.nmech else 0,
    }
    rate_reads_e2e = P * (9 * ss + (6 * na * es if prob.nmech else 0)) + slab * 9 * es
    share = FIRST_STAGE_SHARE
    dead = FIRST_STAGE_SHARE * (1.0 if steps is None else (steps - 1) / steps)
    # the rate stores move the planes the old-rate reads move
    kernels = {k: int(round(v - (share + dead) * rate_reads[k])) for k, v in kernels.items()}
    end_to_end = int(round(end_to_end - (share + dead) * rate_reads_e2e))
    # the same count with every mechanism's eta and Deta streamed at every
    # stage (the per-point figure quoted for the workload, 1116 B FP32-mixed
    # Q8 before PML), whatever the mechanisms' weights
    all_mech = (P * (9 * state + (ngeo + 3) * es + ((24 * prob.nmech + 2) * es if prob.nmech else 0))
                + slab * 36 * es)
    return {"P": P, "interior_nodes": n_int, "active_mechanisms": na,
            "end_to_end": end_to_end, "end_to_end_all_mechanisms": all_mech, "kernels": kernels,
            "rate_reads": dict(rate_reads, end_to_end=rate_reads_e2e),
            "first_stage_share": share, "dead_store_share": dead}


def near_source_receivers(prob, offsets=((4, 0, 0), (-4, 0, 0), (0, 4, 0), (0, -4, 0),
                                         (0, 0, 4), (0, 0, -4), (3, 3, 3), (-3, -3, -3))):
    near = np.asarray(prob.sources[0].near)
    return np.asarray([near + np.asarray(o) for o in offsets], dtype=np.int64)


def timed_run(blk, prob, steps, warmup, t0):
    import torch
    st = torch.cuda.ExternalStream(blk.stream())
    t = blk.time_step(prob.dt, warmup, t0)
    blk.sync()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(st)
    t = blk.time_step(prob.dt, steps, t)
    e1.record(st)
    e1.synchronize()
    blk.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    return e0.elapsed_time(e1), wall, t


def cpu_baseline(n_sample, n_single, nmech, npml, seed):
    """Oracle (C, OpenMP over its node loops) on sub-blocks of the same set-up,
    1 step each: all host threads (OMP_NUM_THREADS) on n_sample^3, and one
    thread on n_single^3."""
    from oracle import oracle as O
    from waveqlab3d_q_gpu_amd import problems as PB

    def run(n, threads):
        prob = PB.curvilinear_q(n=n, h=50.0, nsteps=1, nmech=nmech, npml=npml, seed=seed,
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
        O.set_num_threads(threads)
        t0 = time.perf_counter()
        b.time_step(prob.dt, 1, 0.0)
        el = time.perf_counter() - t0
        return nq * nr * ns * 5 / el / 1e6, el

    threads = O.num_threads()
    value, el = run(n_sample, threads)
    v1, el1 = run(n_single, 1)
    O.set_num_threads(threads)
    return {"value": round(value, 4), "unit": "Mpt-RK-stage/s", "cores": threads,
            "kind": "port",
            "sample": f"{n_sample}^3 sub-block of the same curvilinear Q{nmech} + PML{npml} "
                      f"set-up, 1 step (5 stages), oracle/wql_oracle.c (gcc -O2, OpenMP "
                      f"{threads} threads over the node loops, no MPI)",
            "seconds": round(el, 3),
            "single_thread": {"value": round(v1, 4), "cores": 1,
                              "sample": f"{n_single}^3 sub-block, 1 step, 1 thread",
                              "seconds": round(el1, 3)}}


def stream_ceiling(gib=4, reps=10):
    """The rate that plain streaming reaches on this GPU, for reading the
    dominant kernel's fraction against what is attainable rather than the
    nominal 8 TB/s: the library's own streaming kernels (k_scale: one read
    and one write per value; k_update: two reads and one write) over two
    `gib` GiB arrays of doubles, timed with events on the default stream,
    which these entry points launch on and synchronize after each call."""
    import torch
    from waveqlab3d_q_gpu_amd import scale_rate_array, update_state_array
    n = gib * 2 ** 30 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.ones(n, dtype=torch.float64, device="cuda")

    def rate(fn, nbytes):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return nbytes / (e0.elapsed_time(e1) / reps / 1e3) / 1e12

    sc = rate(lambda: scale_rate_array(x, 1.0), 16 * n)
    up = rate(lambda: update_state_array(y, x, 0.0), 24 * n)
    del x, y
    torch.cuda.empty_cache()
    return {"k_scale_TBs": round(sc, 3), "k_update_TBs": round(up, 3), "TBs": round(max(sc, up), 3),
            "arrays": f"2 x {gib} GiB of doubles, {reps} launches each"}


def read_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def rel_l2(a, b):
    """||a - b|| / ||b||, scaled by max|b| first so tiny traces do not underflow."""
    s = float(np.abs(b).max())
    if not s > 0:
        return None
    return float(np.linalg.norm(((a - b) / s).ravel()) / np.linalg.norm((b / s).ravel()))


ACCURACY_C3_FILE = "accuracy_config3_385_1000steps.json"
ACCURACY_C5_FILE = "accuracy_config5_fault_2000steps.json"


def _seismogram_summary(s_lo, s_hi):
    """Errors of a receiver set: per receiver (with its peak relative to the
    largest receiver peak), the worst over the receivers with peak >= 1e-2
    and >= 1e-3 of the largest, and the relative L2 of all traces together."""
    peaks = np.abs(s_hi).max(axis=(0, 2))
    top = float(peaks.max())
    per = {int(r): [float("%.3e" % rel_l2(s_lo[:, r], s_hi[:, r])),
                    float("%.3e" % (peaks[r] / top))]
           for r in range(s_hi.shape[1]) if peaks[r] > 0}

    def worst(floor):
        v = [e for e, p in per.values() if p >= floor]
        return (float("%.3e" % max(v)) if v else None), len(v)
    w2, n2 = worst(1e-2)
    w3, n3 = worst(1e-3)
    return {"seismogram_rel_l2_max_peak_ge_1e-2": w2, "receivers_peak_ge_1e-2": n2,
            "seismogram_rel_l2_max_peak_ge_1e-3": w3, "receivers_peak_ge_1e-3": n3,
            "seismogram_rel_l2_all_traces": float("%.3e" % rel_l2(s_lo, s_hi)),
            "seismogram_rel_l2_and_relative_peak_per_receiver": per}


def accuracy_config3(n=385, steps=1000, nmech=8, npml=20):
    """The precision variant at length: the curvilinear Q8 + PML block run
    from rest for ``steps`` steps in FP64, FP32-mixed and FP32-Kahan on this
    GPU (the Ricker source starts at 0 and peaks at 0.6 s; 1000 steps are
    3.0 s, in which P and S reach most of the 64 surface receivers); relative
    L2 of every field at the end and of every receiver seismogram against
    FP64."""
    import torch
    from waveqlab3d_q_gpu_amd import _lib as L
    from waveqlab3d_q_gpu_amd import problems as PB
    prob = PB.curvilinear_q(n=n, h=50.0, nsteps=steps, nmech=nmech, npml=npml, seed=2024, nrec=64)
    res = {}
    for prec, name in ((L.WQL_FP64, "fp64"), (L.WQL_FP32_MIXED, "fp32_mixed"),
                       (L.WQL_FP32_KAHAN, "fp32_kahan")):
        print(f"[accuracy c3] {name}: {n}^3, {steps} steps", file=sys.stderr, flush=True)
        blk = PB.make_device_block(prob, precision=prec, state_mirror=True)
        blk.set_receivers(prob.receivers, steps)
        w0 = time.perf_counter()
        t, done = 0.0, 0
        while done < steps:
            k = min(250, steps - done)
            t = blk.time_step(prob.dt, k, t)
            blk.sync()
            done += k
            print(f"[accuracy c3] {name}: {done} / {steps} steps", file=sys.stderr, flush=True)
        wall = time.perf_counter() - w0
        blk.download(L.WQL_F)
        res[name] = (blk.seismograms(), blk.F.cpu().numpy(), wall)
        blk.close()
        del blk
        torch.cuda.empty_cache()
    s64, F64, _ = res["fp64"]
    out = {"workload": f"{n}^3 curvilinear block (seeded topography, layered + von Karman "
                       f"medium), Q{nmech}, PML {npml} on 4 sides + bottom, free surface, "
                       f"h = 50 m, Ricker f0 = 2 Hz from rest, {steps} steps (dt = "
                       f"{prob.dt:.4e} s, {steps * prob.dt:.2f} s), 64 surface receivers",
           "reference": "FP64 GPU path (parity with the FP64 oracle: <= 1e-12)",
           "fp64_wall_s": round(res["fp64"][2], 2)}
    for name in ("fp32_mixed", "fp32_kahan"):
        s32, F32, wall = res[name]
        out[name] = {
            "field_rel_l2": {c: float("%.3e" % rel_l2(F32[k], F64[k]))
                             for k, c in enumerate(("vx", "vy", "vz", "sxx", "syy", "szz",
                                                    "sxy", "sxz", "syz"))},
            **_seismogram_summary(s32, s64),
            "wall_s": round(wall, 2)}
    out["receiver_rule"] = ("per receiver: relative L2 of its 9 traces over the window; the "
                            "worst is taken over the receivers whose FP64 peak is at least "
                            "1e-2 (or 1e-3) of the largest receiver peak")
    return out


def accuracy_config5(steps=2000, every=10):
    """The rupture variant at length: two curvilinear 201x401x201 blocks
    (h = 100 m), Q8 (Qs 100, Qp 200), PML 20, TPV5 slip-weakening fault, in
    FP64 and FP32-mixed (interface arithmetic FP64 in both) for ``steps``
    steps: on-fault slip rate (the stage-1 fault output every ``every``
    steps), rupture time at the end, and 32 off-fault surface receivers (16
    per block, 2-8 km from the fault) against FP64."""
    import torch
    from waveqlab3d_q_gpu_amd import _lib as L
    from waveqlab3d_q_gpu_amd import problems as PB
    from waveqlab3d_q_gpu_amd.interface import Interface
    tp = PB.fault_two_block(nsteps=steps)
    res = {}
    for prec, name in ((L.WQL_FP64, "fp64"), (L.WQL_FP32_MIXED, "fp32_mixed")):
        print(f"[accuracy c5] {name}: {tp.name}, {steps} steps", file=sys.stderr, flush=True)
        bl = PB.make_device_block(tp.left, precision=prec)
        br = PB.make_device_block(tp.right, precision=prec)
        bl.set_receivers(tp.left.receivers, steps)
        br.set_receivers(tp.right.receivers, steps)
        itf = Interface(bl, br, tp.yf, tp.zf, coupling=tp.coupling)
        svel = []
        t = 0.0
        w0 = time.perf_counter()
        for k in range(0, steps, every):
            itf.set_frames(1)                 # record the first step of the chunk
            t = itf.time_step(tp.dt, min(every, steps - k), t)
            svel.append(itf.frames()["Svel"][0])
        itf.sync()
        wall = time.perf_counter() - w0
        S, DS, trup = itf.state()
        seis = np.concatenate([bl.seismograms(), br.seismograms()], axis=1)
        res[name] = (np.array(svel), trup.copy(), S.copy(), seis, wall)
        itf.close()
        bl.close()
        br.close()
        del bl, br
        torch.cuda.empty_cache()
    v64, tr64, S64, s64, w64 = res["fp64"]
    v32, tr32, S32, s32, w32 = res["fp32_mixed"]
    rup64, rup32 = tr64 < 1e8, tr32 < 1e8
    both = rup64 & rup32
    vmag64 = np.sqrt((v64 ** 2).sum(axis=1))
    vmag32 = np.sqrt((v32 ** 2).sum(axis=1))
    return {
        "workload": f"2 x 201x401x201 curvilinear blocks (seeded topography, planar fault x = "
                    f"0), h = 100 m, Q8 (Qs 100, Qp 200), PML 20, TPV5 slip-weakening fault, "
                    f"{steps} steps (dt = {tp.dt:.4e} s, {steps * tp.dt:.2f} s)",
        "reference": "FP64 GPU path (parity with the FP64 oracle: <= 1e-12 at reduced size)",
        "fault_nodes": int(tr64.size),
        "ruptured_fp64": int(rup64.sum()), "ruptured_fp32_mixed": int(rup32.sum()),
        "rupture_state_differs": int((rup64 != rup32).sum()),
        "rupture_time_max_abs_diff_s": float("%.3e" % np.abs(tr32[both] - tr64[both]).max()),
        "rupture_time_dt": tp.dt,
        "slip_rate_rel_l2": float("%.3e" % rel_l2(v32, v64)),
        "slip_rate_magnitude_rel_l2": float("%.3e" % rel_l2(vmag32, vmag64)),
        "slip_rate_samples": f"stage-1 fault output every {every} steps ({v64.shape[0]} frames)",
        "final_slip_rel_l2": float("%.3e" % rel_l2(S32[:3], S64[:3])),
        "max_slip_m": float(np.abs(S64[:3]).max()),
        **_seismogram_summary(s32, s64),
        "wall_s": {"fp64": round(w64, 1), "fp32_mixed": round(w32, 1)}}


def write_profile(name, data, command):
    path = os.path.join(ROOT, "profiles", name)
    data = dict(data)
    data["command"] = command
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1)
    print(f"[bench] wrote {path}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # >= 50 steady-state steps after warm-up, timed with HIP events
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=385)
    ap.add_argument("--nmech", type=int, default=8)
    ap.add_argument("--npml", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-kahan", action="store_true",
                    help="skip the FP32 state with Kahan-compensated RK update")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the discrete-adjoint cost measurement (129^3 Q4)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the discrete-adjoint cost measurement and print it")
    ap.add_argument("--no-two-block", action="store_true",
                    help="skip the two-block fault workload (2 x 201x401x201, Q8, PML 20)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-mixed vs FP64 field comparison (host copies)")
    # oracle sub-blocks: n^3 on all host threads (~10-30 s of CPU work), and a
    # smaller one on one thread
    ap.add_argument("--cpu-sample", type=int, default=161)
    ap.add_argument("--cpu-sample-single", type=int, default=97)
    ap.add_argument("--long-accuracy", choices=("c3", "c5"), default=None,
                    help="only run the precision variant (c3: 385^3, 1000 steps) or the "
                         "rupture variant (c5: 2 x 201x401x201, 2000 steps) in FP32 against "
                         "FP64 and write its summary under profiles/")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device copy of this many MiB 5 times (scale factor "
                         "for FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X (one block per GPU, no halo "
              "exchange); runs on more than one GPU are out of scope. Use --gpus 1.",
              file=sys.stderr)
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

    import torch  # noqa: F401

    from waveqlab3d_q_gpu_amd import _lib as L
    from waveqlab3d_q_gpu_amd import problems as PB

    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    if args.adjoint_only:
        print(json.dumps({"adjoint_config4": adjoint_config4()}), flush=True)
        return
    if args.long_accuracy:
        cmd = f"python -u bench.py --long-accuracy {args.long_accuracy}"
        if args.long_accuracy == "c3":
            write_profile(ACCURACY_C3_FILE, accuracy_config3(), cmd)
        else:
            write_profile(ACCURACY_C5_FILE, accuracy_config5(), cmd)
        return
    print(f"[bench] building {args.n}^3 curvilinear Q{args.nmech} PML{args.npml} set-up",
          file=sys.stderr, flush=True)
    prob = PB.curvilinear_q(n=args.n, h=50.0, nsteps=args.steps, nmech=args.nmech,
                            npml=args.npml, seed=2024, nrec=64)
    nq, nr, ns = prob.n
    P = nq * nr * ns
    receivers = np.concatenate([prob.receivers, near_source_receivers(prob)])
    nrec_surface = prob.receivers.shape[0]
    nslots = args.steps + args.warmup
    # start so that the Ricker peak falls inside the timed window
    t0 = prob.sources[0].t_init - 0.5 * nslots * prob.dt

    results, finals = {}, {}
    for prec, name in ((L.WQL_FP32_MIXED, "fp32"), (L.WQL_FP32_KAHAN, "fp32_kahan"),
                       (L.WQL_FP64, "fp64")):
        if name == "fp64" and args.no_fp64:
            continue
        if name == "fp32_kahan" and args.no_kahan:
            continue
        print(f"[bench] {name}: creating block", file=sys.stderr, flush=True)
        want_fields = not args.no_accuracy and not args.no_fp64
        blk = PB.make_device_block(prob, precision=prec, state_mirror=want_fields)
        blk.set_receivers(receivers, nslots)
        print(f"[bench] {name}: device bytes {blk.device_bytes() / 2**30:.1f} GiB; timing "
              f"{args.warmup}+{args.steps} steps", file=sys.stderr, flush=True)
        ms, wall, _ = timed_run(blk, prob, args.steps, args.warmup, t0)
        seis = blk.seismograms()
        # per-kernel durations on the block stream over one more call of the
        # same number of steps
        blk.set_timing(True)
        blk.time_step(prob.dt, args.steps, 0.0)
        kt, nst = blk.kernel_times()
        blk.set_timing(False)
        bm = byte_model(prob, prec, steps=args.steps)
        stage_ms = ms / (5 * args.steps)
        kernels = {}
        for k, key in enumerate(("interior", "shell", "memory_update")):
            kms = kt[k] / nst
            kb = bm["kernels"][key]
            kernels[key] = {"ms": round(kms, 4), "bytes": kb,
                            "GBs": round(kb / (kms / 1e3) / 1e9, 1) if kms > 0 else None}
        results[name] = {
            "value": P * 5 * args.steps / (ms / 1e3) / 1e6,
            "ms_per_step": ms / args.steps,
            "wall_ms_per_step": wall * 1e3 / args.steps,
            "stage_ms": stage_ms,
            "stage_bytes": bm["end_to_end"],
            "stage_bytes_all_mechanisms": bm["end_to_end_all_mechanisms"],
            "stage_GBs": bm["end_to_end"] / (stage_ms / 1e3) / 1e9,
            "kernels": kernels,
            "byte_model": bm,
        }
        print(f"[bench] {name}: {results[name]['value']:.1f} Mpt-RK-stage/s, stage "
              f"{stage_ms:.3f} ms, kernels {[kernels[k]['ms'] for k in kernels]}",
              file=sys.stderr, flush=True)
        F = None
        if want_fields:
            # state after warmup + steps + the 2 timing steps
            blk.download(L.WQL_F)
            F = blk.F.cpu().numpy()
        finals[name] = (seis, F)
        blk.close()
        del blk
        torch.cuda.empty_cache()

    # FP32 variants vs FP64: same inputs, same GPU, same time window
    def accuracy(name):
        s32, F32 = finals[name]
        s64, F64 = finals["fp64"]
        peaks = np.abs(s64).max(axis=(0, 2))
        top = float(peaks.max()) if peaks.size else 0.0
        errs = [rel_l2(s32[:, r], s64[:, r]) for r in range(s64.shape[1])
                if top > 0 and peaks[r] >= 1e-3 * top]
        acc = {"seismogram_rel_l2_max": max(errs) if errs else None,
               "seismogram_rel_l2_per_receiver": [float("%.3e" % e) for e in errs],
               "receivers_with_signal": len(errs),
               "receivers": (f"{nrec_surface} seeded surface receivers + 8 receivers 3-4 nodes "
                             "from the source; a receiver counts when its FP64 peak is >= 1e-3 "
                             "of the largest receiver peak in the window"),
               "window": f"{args.warmup + args.steps} steps from rest around the Ricker peak"}
        if F32 is not None:
            acc["field_rel_l2"] = [float("%.3e" % rel_l2(F32[c], F64[c])) for c in range(9)]
        return acc

    r32 = results["fp32"]
    k32 = r32["kernels"]
    dom = max(k32, key=lambda k: k32[k]["ms"])
    traffic = read_traffic(TRAFFIC_FILE)
    roof = {"bound": "hbm", "kernel": dom,
            "scope": "the stage's dominant kernel; the whole RK stage is stage_end_to_end",
            "achieved": k32[dom]["GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(k32[dom]["GBs"] / HBM_PEAK_GBS, 4),
            "traffic": (traffic.get("bytes_per_launch")
                        if traffic and traffic.get("kernel") == dom else None),
            "algorithmic_bytes_per_launch": k32[dom]["bytes"],
            "stage_end_to_end": {"achieved": round(r32["stage_GBs"], 1),
                                 "frac": round(r32["stage_GBs"] / HBM_PEAK_GBS, 4),
                                 "bytes_per_stage": r32["stage_bytes"],
                                 "bytes_per_node": round(r32["stage_bytes"] / P, 1),
                                 "all_mechanisms_model": {
                                     "bytes_per_node": round(r32["stage_bytes_all_mechanisms"] / P, 1),
                                     "frac": round(r32["stage_bytes_all_mechanisms"]
                                                   / (r32["stage_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}}
    ceil = stream_ceiling()
    roof["stream_ceiling"] = ceil
    roof["frac_of_stream_ceiling"] = round(k32[dom]["GBs"] / 1e3 / ceil["TBs"], 4)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_sample, args.cpu_sample_single, args.nmech, args.npml, 2024)
    out = {
        "metric": METRIC, "value": round(r32["value"], 2), "unit": "Mpt-RK-stage/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 stencil/memory + f64 RK accumulation",
        "data": "synthetic (seeded topography, von Karman medium, Ricker source)",
        "config": {"workload": f"{nq}^3 curvilinear block, anelastic Q{args.nmech}, PML "
                               f"{args.npml} on 4 sides + bottom, free surface, h=50 m, "
                               "RK stage = interior stencil + shell + memory-variable kernels",
                   "nodes": P, "precision": "fp32-mixed"},
        # the whole fused stage against HBM peak on the end-to-end byte count
        # (the target's figure); roofline below is the dominant kernel's
        "stage_frac_of_hbm_peak": round(r32["stage_GBs"] / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_fp32": k32,
        "stage_ms_fp32": round(r32["stage_ms"], 4),
    }
    if "fp64" in results:
        r64 = results["fp64"]
        out["fp64"] = {"value": round(r64["value"], 2), "ms_per_step": round(r64["ms_per_step"], 4),
                       "stage_ms": round(r64["stage_ms"], 4),
                       "stage_GBs": round(r64["stage_GBs"], 1),
                       "stage_frac": round(r64["stage_GBs"] / HBM_PEAK_GBS, 4),
                       "kernels": r64["kernels"]}
        out["speedup_fp32_mixed_vs_fp64"] = round(r32["value"] / r64["value"], 3)
    out["bytes_per_node_stage_fp32"] = round(r32["stage_bytes"] / P, 1)
    out["active_mechanisms"] = r32["byte_model"]["active_mechanisms"]
    if "fp64" in finals:
        out["accuracy_fp32_mixed_vs_fp64"] = accuracy("fp32")
    if "fp32_kahan" in results:
        rk = results["fp32_kahan"]
        out["fp32_kahan"] = {"value": round(rk["value"], 2), "ms_per_step": round(rk["ms_per_step"], 4),
                             "stage_ms": round(rk["stage_ms"], 4),
                             "stage_GBs": round(rk["stage_GBs"], 1),
                             "stage_frac": round(rk["stage_GBs"] / HBM_PEAK_GBS, 4),
                             "bytes_per_node_stage": round(rk["stage_bytes"] / P, 1),
                             "kernels": rk["kernels"]}
        if "fp64" in finals:
            out["fp32_kahan"]["accuracy_vs_fp64"] = accuracy("fp32_kahan")
            out["fp32_kahan"]["speedup_vs_fp64"] = round(rk["value"] / results["fp64"]["value"], 3)
    if not args.no_two_block:
        out["two_block_fault"] = two_block_fault(args)
    if not args.no_adjoint:
        out["adjoint_config4"] = adjoint_config4()
    # the long-window accuracy runs (python bench.py --long-accuracy c3 / c5)
    # take minutes each; their committed summaries are carried here
    acc3 = read_traffic(ACCURACY_C3_FILE)
    if acc3:
        out["accuracy_config3_long"] = {
            "file": "profiles/" + ACCURACY_C3_FILE, "workload": acc3.get("workload"),
            **{k: {key: acc3[k].get(key) for key in (
                "seismogram_rel_l2_max_peak_ge_1e-2", "receivers_peak_ge_1e-2",
                "seismogram_rel_l2_max_peak_ge_1e-3", "receivers_peak_ge_1e-3",
                "seismogram_rel_l2_all_traces")}
               | {"field_rel_l2_max": max(acc3[k]["field_rel_l2"].values())}
               for k in ("fp32_mixed", "fp32_kahan") if k in acc3}}
    acc5 = read_traffic(ACCURACY_C5_FILE)
    if acc5:
        out["accuracy_config5_long"] = {
            "file": "profiles/" + ACCURACY_C5_FILE,
            **{k: acc5.get(k) for k in ("workload", "slip_rate_rel_l2",
                                        "rupture_time_max_abs_diff_s", "rupture_state_differs",
                                        "ruptured_fp64", "seismogram_rel_l2_max_peak_ge_1e-2",
                                        "receivers_peak_ge_1e-2", "seismogram_rel_l2_all_traces")}}
    print(json.dumps(out), flush=True)


def two_block_fault(args):
    """The two-block fault workload (two curvilinear 201 x 401 x 201 blocks
    with topography and a planar fault, Q8, PML 20, TPV5 slip-weakening
    interface) in FP32-mixed: whole-domain throughput with the interface
    kernel in every stage, and a check that the rupture nucleated (slip > 0,
    finite fields)."""
    import torch
    from waveqlab3d_q_gpu_amd import _lib as L
    from waveqlab3d_q_gpu_amd import problems as PB
    from waveqlab3d_q_gpu_amd.interface import Interface
    tp = PB.fault_two_block(nsteps=args.steps)
    print(f"[bench] two-block fault {tp.left.n} x 2: creating blocks", file=sys.stderr, flush=True)
    bl = PB.make_device_block(tp.left, precision=L.WQL_FP32_MIXED)
    br = PB.make_device_block(tp.right, precision=L.WQL_FP32_MIXED)
    itf = Interface(bl, br, tp.yf, tp.zf, coupling=tp.coupling)
    t = itf.time_step(tp.dt, args.warmup, 0.0)
    itf.sync()
    torch.cuda.synchronize()
    st = torch.cuda.ExternalStream(bl.stream())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    t = itf.time_step(tp.dt, args.steps, t)
    e1.record(st)
    e1.synchronize()
    itf.sync()
    ms = e0.elapsed_time(e1)
    S, _, _ = itf.state()
    nodes = 2 * int(np.prod(tp.left.n))
    res = {"workload": "2 x 201x401x201 curvilinear blocks (h = 100 m, seeded topography, "
                       "planar fault), anelastic Q8 (Qs 100, Qp 200), PML 20 on the outer "
                       "faces, free surface, TPV5 slip-weakening fault between them; "
                       "FP32-mixed; one stage = interface kernel + both blocks' stages",
           "nodes": nodes, "steps": args.steps, "warmup": args.warmup,
           "value": round(nodes * 5 * args.steps / (ms / 1e3) / 1e6, 2),
           "unit": "Mpt-RK-stage/s", "ms_per_step": round(ms / args.steps, 4),
           "stage_ms": round(ms / (5 * args.steps), 4),
           "max_slip_m": float(np.abs(S[:3]).max()), "finite": bool(np.isfinite(S).all())}
    print(f"[bench] two-block fault: {res['value']} Mpt-RK-stage/s, stage {res['stage_ms']} ms, "
          f"max slip {res['max_slip_m']:.3e}", file=sys.stderr, flush=True)
    itf.close()
    bl.close()
    br.close()
    del bl, br
    torch.cuda.empty_cache()
    return res


def adjoint_byte_model(prob, precision):
    """Algorithmic HBM bytes of one reverse stage (wql_adjoint_stage) per
    node, end to end: the stage input state read (F in the RK state's type,
    eta), the state and rate adjoints read and written (F parts FP64, eta
    parts in the storage type), geometry, material and Q^-1 read, the three
    FP64 gradient planes read and written.  The 18 pointwise adjoint planes
    between k_adj_pre and k_adj_stencil are a hand-off, not counted."""
    es = 8 if precision == 0 else 4
    ss = 4 if precision == 2 else 8
    ne = 6 * active_mechanisms(prob) if prob.nmech else 0
    ngeo = 10 if prob.mesh == 1 else 3
    u = 9 * ss + ne * es
    adj = 2 * (9 * 8 + ne * es)
    return u + 2 * adj + (ngeo + 3 + (2 if prob.nmech else 0)) * es + 3 * 8 * 2


def adjoint_config4(n=129, steps=40, every=7, reps=10):
    """Cost of the discrete adjoint on the differentiation workload (129^3
    curvilinear, Q4, 5 % von Karman, characteristic SAT), FP64 and
    FP32-mixed: a forward step and a reverse step (five wql_adjoint_stage
    calls, i.e. k_adj_grad + k_adj_pre + k_adj_stencil per stage) timed with
    HIP events on the block's stream, the reverse stage against its
    algorithmic bytes, and a whole gradient (forward with checkpoints every
    ``every`` steps, segment recomputation keeping the stage inputs, reverse
    sweep; host clock, the second of two) over ``steps`` steps per step,
    against the forward step."""
    import torch
    from waveqlab3d_q_gpu_amd import _lib as L
    from waveqlab3d_q_gpu_amd import problems as PB
    from waveqlab3d_q_gpu_amd.adjoint import AdjointRunner, receiver_traces
    from waveqlab3d_q_gpu_amd.block import RK_A, RK_B
    prob = PB.differentiation_q4(n, nsteps=steps)
    nodes = int(np.prod(prob.n))
    out = {"workload": f"{n}^3 curvilinear block, h = 100 m, seeded topography, 5 % von "
                       "Karman speeds and density, Q4, characteristic SAT, free surface, "
                       f"Ricker 1.5 Hz, 64 surface receivers; gradient over {steps} steps, "
                       f"checkpoints every {every}",
           "nodes": nodes}
    for prec, name in ((L.WQL_FP64, "fp64"), (L.WQL_FP32_MIXED, "fp32_mixed")):
        print(f"[bench] adjoint {name}: {n}^3 Q4", file=sys.stderr, flush=True)
        blk = PB.make_device_block(prob, precision=prec, state_mirror=True)
        run = AdjointRunner(blk, prob.dt, steps, prob.receivers, every=every)
        h = blk._h
        st = torch.cuda.ExternalStream(blk.stream())
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t = blk.time_step(prob.dt, 3, 0.3)
        blk.sync()
        e0.record(st)
        blk.time_step(prob.dt, reps, t)
        e1.record(st)
        e1.synchronize()
        fwd = e0.elapsed_time(e1) / reps
        L.check(L.lib().wql_adjoint_reset(h))
        for s in reversed(range(5)):
            L.check(L.lib().wql_adjoint_stage(h, RK_A[s], RK_B[s] * prob.dt))
        blk.sync()
        e0.record(st)
        for _ in range(reps):
            for s in reversed(range(5)):
                L.check(L.lib().wql_adjoint_stage(h, RK_A[s], RK_B[s] * prob.dt))
        e1.record(st)
        e1.synchronize()
        rev = e0.elapsed_time(e1) / reps
        M = blk.M
        # the second of two whole gradients is timed (the first call of the
        # process also starts torch's autograd machinery)
        for _ in range(2):
            lam, mu, rho = (M[k].clone().requires_grad_(True) for k in range(3))
            torch.cuda.synchronize()
            w0 = time.perf_counter()
            tr = receiver_traces(rho, lam, mu, run)
            (0.5 * (tr[..., :3] ** 2).sum()).backward()
            torch.cuda.synchronize()
            grad = (time.perf_counter() - w0) * 1e3 / steps
        fin = bool(torch.isfinite(rho.grad).all() and torch.isfinite(mu.grad).all())
        bpn = adjoint_byte_model(prob, prec)
        gbs = bpn * nodes / (rev / 5 / 1e3) / 1e9
        out[name] = {"forward_ms_per_step": round(fwd, 4),
                     "reverse_ms_per_step": round(rev, 4),
                     "reverse_stage_bytes_per_node": bpn,
                     "reverse_stage_GBs": round(gbs, 1),
                     "reverse_stage_frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                     "gradient_ms_per_step": round(grad, 3),
                     "gradient_over_forward": round(grad / fwd, 2),
                     "gradient_finite": fin}
        print(f"[bench] adjoint {name}: forward {fwd:.3f} ms/step, reverse {rev:.3f} ms/step "
              f"({gbs:.0f} GB/s), gradient {grad:.2f} ms/step", file=sys.stderr, flush=True)
        del run, lam, mu, rho, tr
        blk.close()
        del blk
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
