"""Output files in the reference's formats (waveqlab3d_q_gpu_amd/output.py):
the ES24.16E3 station lines of write_seismogram (seismogram.f90:525-526) and
the float32 fault files read by python/read_binary.py."""
import numpy as np

from waveqlab3d_q_gpu_amd import output as O


def test_fortran_es_descriptor():
    # ES24.16E3: one leading digit, 16 decimals, 3-digit exponent, width 24
    assert O.fortran_es(1.0) == " 1.0000000000000000E+000"
    assert O.fortran_es(0.0) == " 0.0000000000000000E+000"
    assert O.fortran_es(-0.026665) == "-2.6665000000000001E-002"
    assert O.fortran_es(123456.789) == " 1.2345678900000000E+005"
    assert O.fortran_es(-1.5e-300) == "-1.5000000000000001E-300"
    assert all(len(O.fortran_es(v)) == 24 for v in (1e300, -1e-300, 3.0, -7.25))


def test_station_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    t = np.arange(5) * 0.0108
    seis = rng.normal(size=(5, 2, 9))
    rec = np.array([[21, 11, 11], [3, 4, 5]])
    paths = O.write_station_files(str(tmp_path), "run", 1, rec, t, seis)
    assert [p.rsplit("/", 1)[1] for p in paths] == ["run_21_11_11_block1.dat",
                                                     "run_3_4_5_block1.dat"]
    lines = open(paths[0]).read().splitlines()
    assert len(lines) == 5 and len(lines[0]) == 24 * 4 + 3
    back = O.read_station_file(paths[1])
    assert np.array_equal(back[:, 0], t) and np.array_equal(back[:, 1:], seis[:, 1, :3])


def test_fault_files_match_read_binary_layout(tmp_path):
    """Per step the Fortran array (ny, nz, ncomp) in column order: element
    (y, z, c) at offset y + ny*(z + nz*c) within the step's record."""
    ny, nz, ns = 5, 4, 3
    rng = np.random.default_rng(2)
    frames = {k: rng.normal(size=(ns, nc, nz, ny)) for k, nc in O.FAULT_FILES if k != "state"}
    O.write_fault_files(str(tmp_path), "run", frames)
    raw = np.fromfile(tmp_path / "run_interface.Svel", dtype="<f4")
    assert raw.size == ns * 3 * nz * ny
    step, y, z, c = 2, 3, 1, 2
    off = step * 3 * nz * ny + y + ny * (z + nz * c)
    assert raw[off] == np.float32(frames["Svel"][step, c, z, y])
    back = O.read_fault_file(str(tmp_path / "run_interface.Uface"), 9, ny, nz)
    assert np.allclose(back, frames["Uface"], rtol=1e-7, atol=0)
    assert not np.fromfile(tmp_path / "run_interface.state", dtype="<f4").any()


def test_slice_files_reproduce_the_reference_files(tmp_path):
    """The archived volume slices (test_problems/truth/1/test_rup_cart_fric.
    {Hslice,Vslice}, stored as float32 in tests/golden/tpv5_cart_truth.npz)
    written by write_slice_files are the reference's files byte for byte
    (per step the Fortran array (nq, ns, 9) / (nq, nr, 9)), and read back
    exactly."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "tpv5_cart_truth.npz")
    with np.load(path, allow_pickle=False) as z:
        H, V = z["Hslice"], z["Vslice"]
    assert H.dtype == np.float32 and H.shape == (46, 9, 21, 21)
    paths = O.write_slice_files(str(tmp_path), "run", H, V)
    assert [p.rsplit("/", 1)[1] for p in paths] == ["run.Hslice", "run.Vslice"]
    assert (tmp_path / "run.Hslice").read_bytes() == H.astype("<f4").tobytes()
    assert os.path.getsize(tmp_path / "run.Vslice") == 730296   # the reference file's size
    assert np.array_equal(O.read_slice_file(paths[1], 21, 21), V.astype(np.float64))
    # element (q, s, c) of step n at offset q + nq*(s + ns*c) within the step
    raw = np.fromfile(paths[0], dtype="<f4")
    n, q, s_, c = 7, 20, 3, 5
    assert raw[n * 9 * 441 + q + 21 * (s_ + 21 * c)] == H[n, c, s_, q]


def test_fault_state_file_holds_the_state_frames(tmp_path):
    ny, nz, ns = 4, 3, 2
    rng = np.random.default_rng(3)
    frames = {k: rng.normal(size=(ns, nc, nz, ny)) for k, nc in O.FAULT_FILES}
    O.write_fault_files(str(tmp_path), "run", frames)
    back = O.read_fault_file(str(tmp_path / "run_interface.state"), 1, ny, nz)
    assert np.allclose(back, frames["state"], rtol=1e-7, atol=0)
