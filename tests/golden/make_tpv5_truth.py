"""Writes tpv5_cart_truth.npz and tpv28_curv_truth.npz: the reference's
archived two-block rupture outputs test_problems/truth/{1,2}/
test_rup_{cart,curv}_fric_interface.{S,Svel,Uface,Vface,Uhat,Vhat,trup}
(float32, 46 frames of a 21 x 21 face; layout and checker
python/read_binary.py:7-37, writer fault_output.f90:73-123), the volume
slices test_rup_{cart,curv}_fric.{Hslice,Vslice} (block 1's nine fields on
the planes r = 1 and s = 1, writer slice_output.f90:80-94), and the station
files test_rup_{cart,curv}_fric_{21_11_11,1_11_11}.dat of the same runs.

The files are data of the reference's own tests (test_problems/
test_rup_{cart,curv}_fric.in; serial_test1/2, mpi_test1/2 in
src/CMakeLists.txt); they are stored unchanged, the fault outputs reshaped to
(frame, component, z, y).

    python tests/golden/make_tpv5_truth.py /root/reference
"""
import os
import sys

import numpy as np

N, NT = 21, 46
KEYS = (("S", 4), ("Svel", 3), ("Uface", 9), ("Vface", 9), ("Uhat", 6), ("Vhat", 6), ("trup", 1))
SETS = (("1", "test_rup_cart_fric", "tpv5_cart_truth.npz"),
        ("2", "test_rup_curv_fric", "tpv28_curv_truth.npz"))


def main(ref):
    for set_id, prefix, out_name in SETS:
        tdir = os.path.join(ref, "test_problems", "truth", set_id)
        out = {}
        for key, nc in KEYS:
            a = np.fromfile(os.path.join(tdir, f"{prefix}_interface.{key}"), dtype="<f4")
            if a.size != NT * nc * N * N:
                raise SystemExit(f"{key}: {a.size} values, expected {NT * nc * N * N}")
            out[key] = a.reshape(NT, nc, N, N)
        # volume slices of block 1 (slice_output.f90:80-94): F on the plane
        # r = 1 (.Hslice, per frame 9 planes of (s, q)) and s = 1 (.Vslice,
        # (r, q)), float32, one frame per step at stage 1 like the fault output
        for key in ("Hslice", "Vslice"):
            a = np.fromfile(os.path.join(tdir, f"{prefix}.{key}"), dtype="<f4")
            if a.size != NT * 9 * N * N:
                raise SystemExit(f"{key}: {a.size} values, expected {NT * 9 * N * N}")
            out[key] = a.reshape(NT, 9, N, N)
        # station files: block 1 node (21, 11, 11) on the fault and block 2
        # node (1, 11, 11), one line of t, vx, vy, vz per step
        for label in ("21_11_11", "1_11_11"):
            out["station_" + label] = np.loadtxt(os.path.join(tdir, f"{prefix}_{label}.dat"))
        dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), out_name)
        np.savez_compressed(dst, **out)
        print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
