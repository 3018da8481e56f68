"""The reference drivers' on-disk contract.

Directory ``$ROOT_CT_DIR/$EXP_CT_NAME/reconstruction/`` holding (all FP32,
native byte order, first index fastest):

* ``fileinfo_projections.dat`` -- header ``[nfiles, nimagef, n_files_sino,
  n_files_rec]`` as written by tomography/fits_to_hpc.F90:257-266.  fbp_hpc
  and tv_hpc_slice overwrite offsets 2-9 with ``nsub, nsub_t, ninner, nouter,
  ncg, lambda, mu, wpu`` and then read the file counts from offsets 10-11
  (fbp_hpc.F90:220-275), so a dataset meant for every driver carries a
  12-value header whose offsets 2-3 and 10-11 both hold the file counts.
* ``theta.dat`` -- ``nfiles`` angles in degrees.
* ``sinograms{js}.dat`` -- per volume slice ``(nimagef+add_Nx) x nfiles``,
  detector fastest (projections_to_sinogram_hpc.F90:479-487).
* ``fbp{js}.dat``, ``fbp_tv{js}.dat`` (``fbp_sub``/``fbp_sub_tv`` when
  sub-sampled), ``fbp_tv_slice.dat``, ``fbp_sub_slice.dat`` -- ``N x N``
  images, x fastest.
* ``../input_parm.txt`` -- 8 list-directed lines
  ``nsub, nsub_t, ninner, nouter, ncg, lambda, mu, wpu``.

Slice ``s`` (1-based) of ``nimagef`` lives in file ``js = ceil(s / r)`` at
element offset ``(s - floor((js-1) r) - 1) * rowlen`` with
``r = nimagef / nfiles_of_that_kind`` (tv_hpc.F90:609-610, 622-623).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np


@dataclass
class Dataset:
    root: str
    name: str

    @property
    def rec(self) -> str:
        return os.path.join(self.root, self.name, "reconstruction")

    def path(self, fname: str) -> str:
        return os.path.join(self.rec, fname)


def _f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32)


def slice_location(s: int, nimagef: int, nfiles_kind: int, rowlen: int):
    """(js, element offset) of 1-based slice s, as the reference computes it
    in FP32 (tv_hpc.F90:600-610)."""
    r = np.float32(nimagef) / np.float32(nfiles_kind)
    js = int(math.ceil(np.float32(s) / r))
    off = (s - int(math.floor(np.float32(js - 1) * r)) - 1) * rowlen
    return js, off


def write_dataset(ds: Dataset, theta_deg, sinograms, n_files_sino: int = 1,
                  n_files_rec: int = 1, parm=None) -> None:
    """sinograms: array (nslices, nfiles, nimagef+add_Nx) -- one sinogram per
    volume slice, angle-major rows (detector fastest).  nslices must equal
    nimagef (the reference reconstructs one slice per detector row)."""
    os.makedirs(ds.rec, exist_ok=True)
    theta = _f32(theta_deg)
    nfiles = theta.size
    sino = _f32(sinograms)
    nslices, nf, width = sino.shape
    assert nf == nfiles
    nimagef = nslices
    add = 1 if nimagef % 2 == 0 else 0
    assert width == nimagef + add, "sinogram rows are nimagef+add_Nx wide"
    hdr = np.zeros(12, np.float32)
    hdr[:4] = [nfiles, nimagef, n_files_sino, n_files_rec]
    hdr[10:12] = [n_files_sino, n_files_rec]
    hdr.tofile(ds.path("fileinfo_projections.dat"))
    theta.tofile(ds.path("theta.dat"))
    rowlen = width * nfiles
    files = {}
    for s in range(1, nslices + 1):
        js, off = slice_location(s, nimagef, n_files_sino, rowlen)
        files.setdefault(js, []).append((off, sino[s - 1]))
    for js, parts in files.items():
        size = max(o for o, _ in parts) + rowlen
        buf = np.zeros(size, np.float32)
        for o, a in parts:
            buf[o:o + rowlen] = a.ravel()
        buf.tofile(ds.path(f"sinograms{js}.dat"))
    if parm is not None:
        write_input_parm(ds, **parm)


def write_projections(ds: Dataset, theta_deg, projections, n_files_sino: int = 1,
                      n_files_rec: int = 1) -> None:
    """The input of projections_to_sinogram: projections (nfiles, nimage,
    nimage) img(x, y) with x fastest, in projections{js}.dat split as the
    reference reads them (j = ceiling(job / r), offset = (job - floor((j-1) r)
    - 1) nimage^2, r = nfiles / n_files_sino;
    projections_to_sinogram_hpc.F90:391-407), theta.dat and the 4-float
    header [nfiles, nimage, n_files_sino, n_files_rec] (:56-71)."""
    os.makedirs(ds.rec, exist_ok=True)
    theta = _f32(theta_deg)
    proj = _f32(projections)
    nfiles, n, n2 = proj.shape
    assert n == n2 and nfiles == theta.size
    np.array([nfiles, n, n_files_sino, n_files_rec], np.float32).tofile(ds.path("fileinfo_projections.dat"))
    theta.tofile(ds.path("theta.dat"))
    r = np.float32(nfiles) / np.float32(n_files_sino)
    files = {}
    for job in range(1, nfiles + 1):
        j = int(np.ceil(np.float32(job) / r))
        off = (job - int(np.floor(np.float32(j - 1) * r)) - 1) * n * n
        files.setdefault(j, []).append((off, proj[job - 1]))
    for j, parts in files.items():
        size = max(o for o, _ in parts) + n * n
        buf = np.zeros(size, np.float32)
        for o, a in parts:
            buf[o:o + n * n] = a.ravel()
        buf.tofile(ds.path(f"projections{j}.dat"))


def read_sinograms(ds: Dataset, nimage: int, nfiles: int, n_files_sino: int = 1) -> np.ndarray:
    """All slices of sinograms{js}.dat as (nimage, nfiles, Nw)."""
    nw = nimage + (1 if nimage % 2 == 0 else 0)
    out = np.zeros((nimage, nfiles, nw), np.float32)
    cache = {}
    for s in range(1, nimage + 1):
        js, off = slice_location(s, nimage, n_files_sino, nw * nfiles)
        if js not in cache:
            cache[js] = np.fromfile(ds.path(f"sinograms{js}.dat"), np.float32)
        out[s - 1] = cache[js][off:off + nw * nfiles].reshape(nfiles, nw)
    return out


def write_input_parm(ds: Dataset, nsub=1, nsub_t=1, ninner=1, nouter=1, ncg=5,
                     lam=0.1, mu=0.05, wpu=4) -> None:
    with open(os.path.join(ds.root, ds.name, "input_parm.txt"), "w") as fh:
        for v in (nsub, nsub_t, ninner, nouter, ncg, lam, mu, wpu):
            fh.write(f"{v}\n")


def read_images(ds: Dataset, prefix: str, nimagef: int, nimage: int,
                n_files_rec: int = 1) -> np.ndarray:
    """All slices of fbp/fbp_tv{js}.dat as (nimagef, nimage, nimage)."""
    out = np.zeros((nimagef, nimage, nimage), np.float32)
    cache = {}
    for s in range(1, nimagef + 1):
        js, off = slice_location(s, nimagef, n_files_rec, nimage * nimage)
        if js not in cache:
            cache[js] = np.fromfile(ds.path(f"{prefix}{js}.dat"), np.float32)
        out[s - 1] = cache[js][off:off + nimage * nimage].reshape(nimage, nimage)
    return out


def read_image(ds: Dataset, fname: str, nimage: int) -> np.ndarray:
    return np.fromfile(ds.path(fname), np.float32)[:nimage * nimage].reshape(nimage, nimage)


def angle_window(theta_deg):
    """0-based (start, end) of the reference's window: first argmin |theta|
    and first argmin |theta - 180| (tv_hpc.F90:546-547)."""
    th = _f32(theta_deg)
    return int(np.argmin(np.abs(th))), int(np.argmin(np.abs(th - np.float32(180.0))))


# Per-slice done map of a volume run (no counterpart in the reference, whose
# reruns recompute every slice: tv_hpc.F90:583-631 keeps no record of what is
# written).  File ``{prefix}.done`` beside the outputs, ``prefix`` the last
# output kind the run writes (fbp_tv / fbp_sub_tv, or fbp / fbp_sub without
# TV).  Layout, little-endian, the same for acumen_amd/fortran/acumen_io.F90:
#   bytes  0-7   b"ACUDONE1"
#   bytes  8-47  int32 nimagef, nimage, ntheta, nsub, nsub_t, wpu, ninner,
#                nouter, ncg, 0
#   bytes 48-55  float32 mu, lambda
#   bytes 56-63  zero
#   byte  64+s-1 1 once slice s (1-based) has its outputs written, else 0
DONE_MAGIC = b"ACUDONE1"
DONE_HEADER = 64


def done_header(nimagef: int, nimage: int, ntheta: int, nsub: int, nsub_t: int, wpu: int,
                ninner: int, nouter: int, ncg: int, mu: float, lam: float) -> bytes:
    ints = np.array([nimagef, nimage, ntheta, nsub, nsub_t, wpu, ninner, nouter, ncg, 0], "<i4")
    reals = np.array([mu, lam], "<f4")
    h = DONE_MAGIC + ints.tobytes() + reals.tobytes()
    return h + bytes(DONE_HEADER - len(h))


class DoneMap:
    """The done map of one volume run.  ``resume`` keeps the slices an
    earlier run with the same header marked done; otherwise, or when the
    file is missing, short or from a run with other parameters, every slice
    starts pending.  The file is (re)written at once, so a run that stops
    early can be resumed."""

    def __init__(self, path: str, header: bytes, nimagef: int, resume: bool):
        self.path = path
        done = np.zeros(nimagef, np.uint8)
        if resume and os.path.exists(path):
            raw = np.fromfile(path, np.uint8)
            if raw.size == DONE_HEADER + nimagef and raw[:DONE_HEADER].tobytes() == header:
                done = (raw[DONE_HEADER:] == 1).astype(np.uint8)
        self.done = done
        with open(path, "wb") as fh:
            fh.write(header)
            fh.write(done.tobytes())

    def pending(self):
        """1-based slices not yet done, in order."""
        return [int(s) + 1 for s in np.flatnonzero(self.done == 0)]

    def mark(self, slices) -> None:
        """Record slices (1-based) whose outputs are written."""
        with open(self.path, "r+b") as fh:
            for s in slices:
                self.done[s - 1] = 1
                fh.seek(DONE_HEADER + s - 1)
                fh.write(b"\x01")
