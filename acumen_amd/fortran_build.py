"""Builds the Fortran module `Tomography` (iso_c_binding over libacumen_hip.so)
and the drivers fbp_hip, tv_hip, tv_hip_slice, projections_to_sinogram_hip into
acumen_amd/fortran/bin/.

Usage: python -m acumen_amd.fortran_build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(HERE, "fortran")
BUILD = os.path.join(FDIR, "_build")
BIN = os.path.join(FDIR, "bin")
MODULES = ["acumen_io.F90", "acumen_tomography.F90"]
DRIVERS = ["fbp_hip.F90", "tv_hip.F90", "tv_hip_slice.F90", "projections_to_sinogram_hip.F90"]
FFLAGS = ["-O2"]


def flang() -> str | None:
    for c in (os.environ.get("FC"), shutil.which("flang"), "/opt/rocm/lib/llvm/bin/flang"):
        if c and os.path.exists(c):
            return c
    return None


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("flang failed: " + " ".join(cmd))


def build(verbose: bool = False) -> list[str]:
    fc = flang()
    if fc is None:
        raise RuntimeError("flang not found (needed for the Fortran drivers)")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    for m in MODULES:
        o = os.path.join(BUILD, m.replace(".F90", ".o"))
        _run([fc, *FFLAGS, "-module-dir", BUILD, "-c", os.path.join(FDIR, m), "-o", o], verbose)
        objs.append(o)
    out = []
    for d in DRIVERS:
        exe = os.path.join(BIN, d.replace(".F90", ""))
        _run([fc, *FFLAGS, "-I", BUILD, os.path.join(FDIR, d), *objs, "-L", HERE, "-lacumen_hip",
              "-Wl,-rpath,$ORIGIN/../..", "-o", exe], verbose)
        out.append(exe)
    return out


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
