"""The precision variant at its full size and length: 385^3 curvilinear Q8 +
PML 20 block, 1000 steps from rest, FP32-mixed and FP32-Kahan against the
FP64 GPU path (itself at <= 1e-12 of the FP64 oracle): relative L2 of every
field at the end and of every surface receiver seismogram.

Bounds: the seismogram relative L2 of all 64 receivers' traces together
<= 1e-5 (the target's figure; measured 7.3e-7 FP32-mixed, 4.8e-7 Kahan) and
every field <= 1e-5 (measured <= 7.6e-7).  Per receiver, relative to that
receiver's own trace, the error is the FP32 round-off of the stencil
arithmetic where the fields are largest, carried to every receiver alike;
on the weakest receivers it is therefore largest: measured 1.0e-5 (both
variants) over the 57 receivers whose peak is at least 1e-2 of the largest,
bound 2e-5, and 4.3e-5 over all 64 (peaks down to 3e-3 of the largest),
bound 1e-4.  FP32-Kahan compensates the state's accumulation, not the
stencils, so its per-receiver worst equals FP32-mixed's."""
import ctypes

import pytest

from waveqlab3d_q_gpu_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_config3_fp32_variants_1000_steps(capsys):
    import bench
    dev = ctypes.c_int32()
    L.check(L.lib().wql_runtime_init(0, ctypes.byref(dev)))
    # about five minutes: its progress lines go straight to the terminal
    with capsys.disabled():
        out = bench.accuracy_config3(n=385, steps=1000)
    for name in ("fp32_mixed", "fp32_kahan"):
        r = out[name]
        print(f"config 3, {name}: seismogram rel L2 all traces "
              f"{r['seismogram_rel_l2_all_traces']:.3e}; per receiver max "
              f"{r['seismogram_rel_l2_max_peak_ge_1e-2']:.3e} over "
              f"{r['receivers_peak_ge_1e-2']} receivers with peak >= 1e-2 of the largest, "
              f"{r['seismogram_rel_l2_max_peak_ge_1e-3']:.3e} over "
              f"{r['receivers_peak_ge_1e-3']} with >= 1e-3; fields {r['field_rel_l2']}",
              flush=True)
        assert r["seismogram_rel_l2_all_traces"] <= 1e-5
        assert max(r["field_rel_l2"].values()) <= 1e-5
        assert r["receivers_peak_ge_1e-2"] >= 32
        assert r["seismogram_rel_l2_max_peak_ge_1e-2"] <= 2e-5
        assert r["seismogram_rel_l2_max_peak_ge_1e-3"] <= 1e-4
