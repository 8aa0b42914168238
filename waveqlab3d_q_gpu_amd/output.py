"""Output files in the reference's formats.

* Station (seismogram) files: write_seismogram (seismogram.f90:495-541) writes
  one line per step, ``(ES24.16E3,3(1X,ES24.16E3))`` of (t, vx, vy, vz), into
  ``<name>_<i>_<j>_<k>_block<b>.dat`` (seismogram.f90:392-396).
* Fault files: write_fault / write_hats (fault_output.f90:73-123) append one
  float32 record per step, the (ny, nz, ncomp) Fortran array, to
  ``<name>_interface.{Uface,Vface,S,Uhat,Vhat,Svel,trup,state}``
  (fault_output.f90:30-46); python/read_binary.py reads them back.
* Volume slices: write_slice (slice_output.f90:80-94) appends block 1's nine
  fields on the plane r = locationH to ``<name>.Hslice`` (the (nq, ns, 9)
  Fortran array per step) and on s = locationV to ``<name>.Vslice`` ((nq, nr,
  9)), float32.
"""
from __future__ import annotations

import math
import os

import numpy as np

FAULT_FILES = (("Uface", 9), ("Vface", 9), ("S", 4), ("Uhat", 6), ("Vhat", 6), ("Svel", 3),
               ("trup", 1), ("state", 1))


def fortran_es(x: float, digits: int = 16, exp_digits: int = 3, width: int = 24) -> str:
    """Fortran ``ESw.dEe`` edit descriptor (scientific, one leading digit)."""
    x = float(x)
    if not math.isfinite(x):
        return ("NaN" if math.isnan(x) else ("Infinity" if x > 0 else "-Infinity")).rjust(width)
    mant, exp = f"{x:.{digits}E}".split("E")
    e = int(exp)
    s = f"{mant}E{'+' if e >= 0 else '-'}{abs(e):0{exp_digits}d}"
    return s.rjust(width)


def station_filename(name: str, ijk, block: int) -> str:
    i, j, k = (int(v) for v in ijk)
    return f"{name}_{i}_{j}_{k}_block{block}.dat"


def station_lines(t, v) -> list:
    """Lines of one station file: t (nsteps,), v (nsteps, 3)."""
    t = np.asarray(t, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    return [fortran_es(t[n]) + "".join(" " + fortran_es(v[n, c]) for c in range(3))
            for n in range(t.shape[0])]


def write_station_files(directory: str, name: str, block: int, receivers, t, seis) -> list:
    """One file per receiver: ``receivers`` (nrec, 3) 1-based node indices,
    ``t`` (nsteps,) sample times, ``seis`` (nsteps, nrec, 9) as returned by
    Block.seismograms (components 0..2 are vx, vy, vz)."""
    paths = []
    for r, ijk in enumerate(np.asarray(receivers).reshape(-1, 3)):
        path = os.path.join(directory, station_filename(name, ijk, block))
        with open(path, "w") as fh:
            fh.write("\n".join(station_lines(t, seis[:, r, :3])) + "\n")
        paths.append(path)
    return paths


def read_station_file(path: str) -> np.ndarray:
    """(nsteps, 4) of t, vx, vy, vz (comment lines skipped)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.strip() and not line.lstrip().startswith("#"):
                rows.append([float(v.replace("D", "E")) for v in line.split()])
    return np.asarray(rows)


def write_fault_files(directory: str, name: str, frames: dict) -> list:
    """Write the fault output of Interface.frames(): each key an array
    (nsteps, ncomp, nz, ny); a missing key (e.g. the state of a
    slip-weakening fault, whose I%W the reference leaves at zero) is written
    as zeros.  Returns the paths."""
    nsteps, _, nz, ny = frames["S"].shape
    paths = []
    for key, nc in FAULT_FILES:
        a = frames.get(key)
        if a is None:
            a = np.zeros((nsteps, nc, nz, ny))
        if a.shape != (nsteps, nc, nz, ny):
            raise ValueError(f"{key}: shape {a.shape}, expected {(nsteps, nc, nz, ny)}")
        path = os.path.join(directory, f"{name}_interface.{key}")
        # per step the Fortran array (ny, nz, nc) in column order == C (nc, nz, ny)
        np.ascontiguousarray(a, dtype="<f4").tofile(path)
        paths.append(path)
    return paths


def read_fault_file(path: str, ncomp: int, ny: int, nz: int) -> np.ndarray:
    """(nsteps, ncomp, nz, ny) float64 from one fault file."""
    a = np.fromfile(path, dtype="<f4")
    return a.reshape(-1, ncomp, nz, ny).astype(np.float64)


def write_slice_files(directory: str, name: str, hslice=None, vslice=None) -> list:
    """Write ``<name>.Hslice`` from ``hslice`` (nsteps, 9, nz, nx), block 1's
    F on one r plane per step, and ``<name>.Vslice`` from ``vslice``
    (nsteps, 9, ny, nx), F on one s plane; either may be None.  Returns the
    paths."""
    paths = []
    for ext, a in (("Hslice", hslice), ("Vslice", vslice)):
        if a is None:
            continue
        a = np.asarray(a)
        if a.ndim != 4 or a.shape[1] != 9:
            raise ValueError(f"{ext}: shape {a.shape}, expected (nsteps, 9, n2, nx)")
        path = os.path.join(directory, f"{name}.{ext}")
        # per step the Fortran array (nx, n2, 9) in column order == C (9, n2, nx)
        np.ascontiguousarray(a, dtype="<f4").tofile(path)
        paths.append(path)
    return paths


def read_slice_file(path: str, n2: int, nx: int) -> np.ndarray:
    """(nsteps, 9, n2, nx) float64 from one .Hslice (n2 = nz) or .Vslice
    (n2 = ny) file."""
    a = np.fromfile(path, dtype="<f4")
    return a.reshape(-1, 9, n2, nx).astype(np.float64)
