"""The byte model behind bench.py's roofline figures (host logic, CPU only):
per-node bytes of each kernel of the 385^3 curvilinear Q8 + PML20 stage,
counted by hand from the kernels' streams."""
from types import SimpleNamespace

import numpy as np

import bench
from waveqlab3d_q_gpu_amd import problems as PB

N = 385
NODES = N ** 3


def _c3():
    _, weight = PB.q_coefficients(8)
    return SimpleNamespace(n=(N, N, N), mesh=1, nmech=8, pml=(1, 1, 0, 1, 1, 1),
                           npml=(20, 20, 0, 20, 20, 20), weight=weight)


def test_active_mechanisms_and_boxes():
    prob = _c3()
    assert bench.active_mechanisms(prob) == 6          # Q8 weights 2 and 7 are zero
    lo, hi = bench.interior_box(prob)
    assert lo == [20, 6, 20] and hi == [364, 364, 364]
    # five slabs of 20 nodes: x low/high and z low/high over the full faces, y high
    assert bench.slab_nodes(prob) == 5 * 20 * N * N


def test_memory_kernel_bytes_per_node():
    prob = _c3()
    n_int = 345 * 359 * 345
    n_shell = NODES - n_int
    # FP32-mixed: W's 6 strain rates + lambda, mu, Qs^-1, Qp^-1 (4 B each), 6
    # active mechanisms x 6 components x (eta, Deta) read and written (4 B),
    # the six stress components' F and DF read and written (8 B) and their
    # float copy written (4 B); shell nodes also read W's 6 stress rates (the
    # strict interior's are Hooke's law of the strain rates)
    per_node = 6 * 4 + 4 * 4 + 6 * 6 * 4 * 4 + 6 * 4 * 8 + 6 * 4
    assert per_node == 832
    # the first of the five RK stages (A = 0) reads no old rate: the six
    # stress components' DF and the 36 Deta planes; in a long run the fifth
    # stage does not store them either (dead before the next first stage);
    # the model is the five-stage average
    old = 6 * 8 + 6 * 6 * 4
    assert old == 192
    assert bench.byte_model(prob, 1)["kernels"]["memory_update"] == round(
        NODES * per_node + n_shell * 6 * 4 - 2 * NODES * old / 5)
    # a call of one step stores every rate; of 50 steps, 49 fifth stages do not
    assert bench.byte_model(prob, 1, steps=1)["kernels"]["memory_update"] == round(
        NODES * per_node + n_shell * 6 * 4 - NODES * old / 5)
    assert bench.byte_model(prob, 1, steps=50)["kernels"]["memory_update"] == round(
        NODES * per_node + n_shell * 6 * 4 - NODES * old / 5 - 49 / 50 * NODES * old / 5)
    # FP64: everything in 8 B, no float copy
    assert bench.byte_model(prob, 0)["kernels"]["memory_update"] == round(NODES * (
        6 * 8 + 4 * 8 + 6 * 6 * 4 * 8 + 6 * 4 * 8) + n_shell * 6 * 8
        - 2 * NODES * (6 * 8 + 6 * 6 * 8) / 5)
    # FP32-Kahan: F, DF and the compensation in 4 B (F, DF, Fc read + write)
    assert bench.byte_model(prob, 2)["kernels"]["memory_update"] == round(NODES * (
        6 * 4 + 4 * 4 + 6 * 6 * 4 * 4 + 6 * (4 * 4 + 2 * 4)) + n_shell * 6 * 4
        - 2 * NODES * (6 * 4 + 6 * 6 * 4) / 5)


def test_stencil_and_end_to_end_bytes():
    prob = _c3()
    m = bench.byte_model(prob, 1)
    n_int = 345 * 359 * 345
    assert m["interior_nodes"] == n_int
    # FP32-mixed stencil node: F as float (36) + velocity FP64 F read (24) and its
    # float copy written (12) + geometry 10 and material 3 (4 B) + W written
    # (strict interior: the 6 strain rates, 24; shell: the 6 stress rates as
    # well, 48) + velocity DF read and written and F written (3 x 24)
    interior = 36 + 3 * (8 + 4) + 13 * 4 + 6 * 4 + 3 * 24
    shell = 36 + 3 * (8 + 4) + 13 * 4 + 12 * 4 + 3 * 24
    assert (interior, shell) == (220, 244)
    # five-stage averages of a long run: the first stage reads no velocity
    # DF (3 x 8 B) and no PML DQ (9 x 4 B per slab node), the fifth does not
    # store them
    slab = 5 * 20 * N * N
    assert m["kernels"]["interior"] == round(n_int * (interior - 2 * 24 / 5))
    assert m["kernels"]["shell"] == round((NODES - n_int) * (shell - 2 * 24 / 5)
                                          + slab * (36 - 2 * 9 / 5) * 4)
    # end to end: F, DF read + write in FP64, geometry + material, eta/Deta
    # + Q^-1, PML auxiliaries; on average a fifth of the old rates (DF 9 x
    # 8 B, Deta 36 x 4 B, DQ 9 x 4 B) is not read and a fifth not stored
    e2e = NODES * (36 * 8 + 13 * 4 + (24 * 6 + 2) * 4) + slab * 36 * 4
    assert m["end_to_end"] == round(e2e - 2 * (NODES * (9 * 8 + 36 * 4) + slab * 9 * 4) / 5)
    assert np.isclose(e2e / NODES, 961.4, atol=0.1)
    assert np.isclose(m["end_to_end"] / NODES, 871.3, atol=0.1)
    assert m["end_to_end_all_mechanisms"] / NODES > e2e / NODES


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so,
    before it touches a device."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout == ""
