"""CPU checks of the C ABI: the library loads, exports every entry point
include/wql3d.h declares, the ctypes mirror of the structs matches the C
layout, and descriptor validation reports errors (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wql3d.h")


@pytest.fixture(scope="module")
def L():
    from waveqlab3d_q_gpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from waveqlab3d_q_gpu_amd import build
        build.build()
    _lib.lib()
    return _lib


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wql_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(L):
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(L.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES, n


def test_struct_layout_matches_c(L, tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "wql3d.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(wql_block_desc),"
        " offsetof(wql_block_desc, body), offsetof(wql_block_desc, eta),"
        " offsetof(wql_block_desc, pml_box), offsetof(wql_block_desc, n_n),"
        " sizeof(wql_source), offsetof(wql_source, location));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    D, S = L.BlockDesc, L.Source
    assert out == [ctypes.sizeof(D), D.body.offset, D.eta.offset, D.pml_box.offset,
                   D.n_n.offset, ctypes.sizeof(S), S.location.offset]


def test_descriptor_validation_errors(L):
    from waveqlab3d_q_gpu_amd import problems as PB
    prob = PB.cartesian_bump(n=12, nsteps=1)
    bad = PB.Problem(n=(11, 12, 12), mesh=0, X=prob.X[:, :, :, :11].copy(),
                     M=prob.M[:, :, :, :11].copy(), bc=prob.bc, dt=1.0, nsteps=1)
    met = tuple(np.ones((3,) + bad.X.shape[1:]) for _ in range(3)) + (np.ones(bad.X.shape[1:]),)
    with pytest.raises(L.WqlError) as e:
        PB.make_block(bad, met)
    assert e.value.code == L.WQL_EINVAL and "12" in str(e.value)
    good = PB.Problem(n=(12, 12, 12), mesh=0, X=prob.X, M=prob.M, bc=(5, 1, 1, 1, 1, 1),
                      dt=1.0, nsteps=1)
    met = tuple(np.ones((3, 12, 12, 12)) for _ in range(3)) + (np.ones((12, 12, 12)),)
    with pytest.raises(L.WqlError) as e:
        PB.make_block(good, met)
    assert "bc code" in str(e.value)


def test_capacity_check_env_override(L, monkeypatch):
    monkeypatch.setenv("WQL3D_GPU_MEMORY_BYTES", "1000")
    assert L.lib().wql_capacity_check(10) == L.WQL_OK
    assert L.lib().wql_capacity_check(5000) == L.WQL_ENOMEM
    assert b"RUN-GPU-MEM-002" in L.lib().wql_last_error()


def test_missing_library_fails_loudly(tmp_path):
    """Without the built library the package raises on first use: there is
    no CPU fallback for any entry point (fresh interpreter, library path
    pointed at an empty folder)."""
    code = ("import waveqlab3d_q_gpu_amd._lib as L\n"
            f"L.LIB_PATH = {str(tmp_path / 'libwql3d_hip.so')!r}\n"
            "try:\n"
            "    L.lib()\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    print('loaded')\n")
    r = subprocess.run(["python", "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("raised:") and "is missing" in r.stdout
