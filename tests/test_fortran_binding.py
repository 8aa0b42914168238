"""The Fortran side of the drop-in boundary, on the CPU: the bind(C) types of
waveqlab3d_q_gpu_amd/fortran/wql3d_hip_binding.f90, compiled with the
image's flang, have the C layout of include/wql3d.h (every size and field
offset equal to gcc's sizeof / offsetof), and the HIP swap-ins of the
rk_vector seam compile and link with the reference's own unit test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "tests", "fortran")

C_FIELDS = {
    "sizeof_block_desc": "sizeof(wql_block_desc)",
    "bc": "offsetof(wql_block_desc, bc)",
    "arrays_on_device": "offsetof(wql_block_desc, arrays_on_device)",
    "hq": "offsetof(wql_block_desc, hq)",
    "tau": "offsetof(wql_block_desc, tau)",
    "weight": "offsetof(wql_block_desc, weight)",
    "body": "offsetof(wql_block_desc, body)",
    "metricx": "offsetof(wql_block_desc, metricx)",
    "F": "offsetof(wql_block_desc, F)",
    "eta": "offsetof(wql_block_desc, eta)",
    "Deta": "offsetof(wql_block_desc, Deta)",
    "Qs_inv": "offsetof(wql_block_desc, Qs_inv)",
    "pml_box": "offsetof(wql_block_desc, pml_box)",
    "pml_Q": "offsetof(wql_block_desc, pml_Q)",
    "pml_DQ": "offsetof(wql_block_desc, pml_DQ)",
    "face_box": "offsetof(wql_block_desc, face_box)",
    "n_l": "offsetof(wql_block_desc, n_l)",
    "n_n": "offsetof(wql_block_desc, n_n)",
    "sat_form": "offsetof(wql_block_desc, sat_form)",
    "sizeof_source": "sizeof(wql_source)",
    "src_duration": "offsetof(wql_source, duration)",
    "src_location": "offsetof(wql_source, location)",
    "src_near_phys": "offsetof(wql_source, near_phys)",
    "sizeof_iface_desc": "sizeof(wql_iface_desc)",
    "iface_y": "offsetof(wql_iface_desc, y)",
    "iface_Dc": "offsetof(wql_iface_desc, Dc)",
    "iface_W0": "offsetof(wql_iface_desc, W0)",
    "iface_alpha": "offsetof(wql_iface_desc, alpha)",
}


def _flang():
    for p in ("/opt/rocm/bin/amdflang", "/opt/rocm/lib/llvm/bin/flang"):
        if os.path.exists(p):
            return p
    return None


@pytest.fixture(scope="module")
def fortran_built():
    if not _flang():
        pytest.skip("no Fortran compiler in this image")
    from waveqlab3d_q_gpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from waveqlab3d_q_gpu_amd import build
        build.build()
    subprocess.run(["make", "-s", "-C", FDIR], check=True)
    return FDIR


def test_bind_c_layout_matches_c(fortran_built, tmp_path):
    fortran = {}
    out = subprocess.run([os.path.join(FDIR, "desc_layout")], capture_output=True, text=True,
                         check=True).stdout
    for line in out.splitlines():
        k, v = line.split()
        fortran[k] = int(v)
    src = tmp_path / "layout.c"
    body = "".join(f'printf("{k} %zu\\n", (size_t)({expr}));' for k, expr in C_FIELDS.items())
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "wql3d.h"\n'
                   f"int main(void){{{body}return 0;}}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    c = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        k, v = line.split()
        c[k] = int(v)
    assert set(c) == set(fortran)
    assert c == fortran, {k: (c[k], fortran[k]) for k in c if c[k] != fortran[k]}


def test_reference_unit_test_links_with_hip_backend():
    """tests/rk_vector_kernel_test.f90 of the reference, unchanged, linked
    with rk_vector_hip.f90 + rk_vector_test_data_hip.f90 (built by
    make -C oracle ref where the reference sources exist)."""
    from oracle import ref as R
    if not R.available():
        pytest.skip("reference sources absent")
    exe = os.path.join(ROOT, "oracle", "_ref", "rk_vector_kernel_test_hip")
    assert os.path.exists(exe)
    out = subprocess.run(["nm", exe], capture_output=True, text=True, check=True).stdout
    assert "wql_scale_rate_present" in out and "wql_update_state_present" in out
