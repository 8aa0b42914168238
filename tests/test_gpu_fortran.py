"""Fortran callers of the library on the GPU:

* the reference's own rk_vector unit test (tests/rk_vector_kernel_test.f90,
  asymmetric ghost bounds -2:3, 3:7, -1:1), unchanged, linked with the HIP
  swap-ins of the rk_vector seam: it prints its PASS line (errors within
  32 eps; the HIP kernels give exactly 0);
* a Fortran host (tests/fortran/kat41_caller.f90) that fills the
  iso_c_binding descriptor from arrays with the reference's ghost bounds
  (-2 .. n+3) and runs the 41^3 reference KAT through wql_time_step: the
  reference's printed max|field| 1.6623128367649574E+02, and the ghost nodes
  untouched."""
import os
import struct
import subprocess

import numpy as np
import pytest

from waveqlab3d_q_gpu_amd import problems as PB

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_rk_vector_unit_test_on_hip_backend():
    exe = os.path.join(ROOT, "oracle", "_ref", "rk_vector_kernel_test_hip")
    if not os.path.exists(exe):
        pytest.skip("built only where the reference sources exist (make -C oracle ref)")
    print("reference unit test rk_vector_kernel_test on the HIP backend", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    assert "PASS: asymmetric RK vector scale/update matches the CPU expression" in out.stdout
    assert "RK vector errors: rate=  0.0000E+00, state=  0.0000E+00" in out.stdout


def _write_kat_input(path, oracle):
    prob = PB.kat_cartesian_41()
    met = oracle.metrics(prob.X)
    nl, nm, nn = PB.face_frames(*met[:3])
    s = prob.sources[0]
    with open(path, "wb") as fh:
        fh.write(struct.pack("<4i6i", *prob.n, prob.nsteps, *prob.bc))
        fh.write(struct.pack("<2d", prob.dt, 1.0 / 0.315949074074074))
        fh.write(struct.pack("<6i", s.kind, s.order, s.alpha, *[int(v) for v in s.near]))
        fh.write(struct.pack("<14d", s.duration, s.t_init, s.mxx, s.myy, s.mzz, s.mxy, s.mxz,
                             s.myz, *s.location, *s.near_phys))
        # (ncomp, nz, ny, nx) C order == Fortran (x, y, z, comp)
        for a in (met[0], met[1], met[2], met[3], prob.M):
            fh.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for f in range(6):
            for arr in (nl[f], nm[f], nn[f]):
                fh.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())
    return prob


def test_fortran_host_runs_reference_kat(oracle, tmp_path):
    exe = os.path.join(ROOT, "tests", "fortran", "kat41_caller")
    assert os.path.exists(exe), "build tests/fortran first (__graft_entry__.build)"
    inp = str(tmp_path / "kat41.bin")
    _write_kat_input(inp, oracle)
    print("Fortran host: 41^3 reference KAT through wql_time_step", flush=True)
    out = subprocess.run([exe, inp], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    vals = {}
    for line in out.stdout.splitlines():
        if "=" in line:
            k, v = line.rsplit("=", 1)
            vals[k.strip()] = v.strip()
    got = float(vals["Elastic final state: max|field|"])
    ref = 1.6623128367649574e+02
    assert abs(got - ref) / ref < 1e-13, got
    assert float(vals["library max_state: max|field|"]) == got
    assert "ghost nodes untouched: T" in out.stdout
