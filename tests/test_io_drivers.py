"""On-disk contract of the reference drivers and the Fortran drop-in drivers
(host side; no GPU needed)."""
import os
import subprocess

import numpy as np
import pytest

from acumen_amd import io as aio
from acumen_amd import fortran_build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "acumen_amd", "fortran", "bin")


def test_slice_location_matches_reference_formula():
    # tv_hpc.F90:609-610 with nimagef = 10 split over 3 files
    locs = [aio.slice_location(s, 10, 3, 7) for s in range(1, 11)]
    js = [l[0] for l in locs]
    assert js == [1, 1, 1, 2, 2, 2, 3, 3, 3, 3]
    assert [l[1] for l in locs] == [0, 7, 14, 0, 7, 14, 0, 7, 14, 21]


def test_dataset_roundtrip(tmp_path):
    ds = aio.Dataset(str(tmp_path), "exp")
    nimagef, nfiles = 6, 5
    th = np.linspace(0, 180, nfiles).astype(np.float32)
    sino = np.random.default_rng(0).normal(size=(nimagef, nfiles, nimagef + 1)).astype(np.float32)
    aio.write_dataset(ds, th, sino, n_files_sino=2, parm=dict(nouter=3, wpu=2))
    hdr = np.fromfile(ds.path("fileinfo_projections.dat"), np.float32)
    assert list(hdr[:4]) == [5, 6, 2, 1] and list(hdr[10:12]) == [2, 1]
    assert np.array_equal(np.fromfile(ds.path("theta.dat"), np.float32), th)
    for s in range(1, nimagef + 1):
        js, off = aio.slice_location(s, nimagef, 2, (nimagef + 1) * nfiles)
        a = np.fromfile(ds.path(f"sinograms{js}.dat"), np.float32)[off:off + (nimagef + 1) * nfiles]
        assert np.array_equal(a.reshape(nfiles, nimagef + 1), sino[s - 1])
    lines = open(os.path.join(str(tmp_path), "exp", "input_parm.txt")).read().split()
    assert lines[3] == "3" and lines[7] == "2"


def test_angle_window_first_minima():
    th = np.array([-3, 0, 0, 90, 180, 180, 183], np.float32)
    assert aio.angle_window(th) == (1, 4)


def test_fortran_drivers_build_and_link():
    if fortran_build.flang() is None:
        pytest.skip("flang not available")
    if not os.path.exists(os.path.join(ROOT, "acumen_amd", "libacumen_hip.so")):
        pytest.skip("libacumen_hip.so not built")
    exes = fortran_build.build()
    for exe in exes:
        out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
        assert "libacumen_hip.so" in out and "not found" not in out, out


def test_driver_fails_cleanly_without_gpu(tmp_path):
    exe = os.path.join(BIN, "fbp_hip")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    ds = aio.Dataset(str(tmp_path), "exp")
    th = np.linspace(0, 180, 7).astype(np.float32)
    aio.write_dataset(ds, th, np.zeros((4, 7, 5), np.float32), parm=dict(wpu=1))
    env = dict(os.environ, ROOT_CT_DIR=str(tmp_path), EXP_CT_NAME="exp",
               HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode != 0
    assert "acumen:" in r.stdout + r.stderr
