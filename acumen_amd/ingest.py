"""FITS ingest: raw neutron radiographs -> normalised projection files, the
restatement of tomography/fits_to_hpc.F90 with the normalisation on the GPU.

Input tree (fits_to_hpc.F90:47-168): ``$ROOT_CT_DIR/$EXP_CT_NAME/`` with
``projections/``, ``OB/`` (open beam) and ``DF/`` (dark field) FITS files,
listed in ``ls`` order (byte order of the names, the C locale).

Output under ``reconstruction/`` (read by projections_to_sinogram_hpc.F90):

* ``theta.dat`` -- nfiles FP32 angles parsed from the projection file
  names (:73-101);
* ``projections{j}.dat`` -- -log(I / max(OB, 1)) of every projection,
  nimage x nimage FP32, split over n_files_sino files (:227-253);
* ``fileinfo_projections.dat`` -- ``[nfiles, nimage, n_files_sino,
  n_files_rec]`` (:257-266).

The open-beam average and the normalisation are the two GPU kernels
(`acu_ingest_mean`, `acu_ingest_normalize`); FITS decoding and file writing
are host I/O.  The dark-field average of :218-231 is computed by the
reference and then unused (its use at :244 is commented out), so it is not
computed here; the DF listing is still reported.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import fits
from ._lib import check

_F = ctypes.POINTER(ctypes.c_float)

NULLVAL = -999.0   # ftgpve's null value (:173)
NAME_LEN = 100     # character(LEN=100) file names (:27)
STRING_LEN = 150   # character(len=150) c_string (:28)


def list_dir(path: str) -> list:
    """`ls path` as the reference reads it: entry names in byte order."""
    return sorted(os.listdir(path), key=lambda s: s.encode("utf-8", "surrogateescape"))


def parse_theta(names, exp_name: str) -> np.ndarray:
    """Angles of :73-101: with L = len_trim(EXP_CT_NAME), the characters at
    1-based positions L+11..L+13 are the hundreds, tens and units and
    L+15..L+17 the tenths, hundredths and thousandths of the angle in
    degrees, accumulated in FP32 as f = ((((( 0 + 100 d1) + 10 d2) + 1 d3)
    + 0.1 d4) + 0.01 d5) + 0.001 d6.  Each digit is read with a list-directed
    integer read into the same variable; a character that is not a digit
    leaves the variable unchanged, so it repeats the previous digit's value
    (before the first file's first digit that variable holds -1, the
    end-of-file status of the listing loop, :63-67).  Names are truncated
    to 100 characters and blank-padded."""
    L = len(exp_name.rstrip(" "))
    c1, c2 = L + 11, L + 15  # 1-based
    weights = (np.float32(100.0), np.float32(10.0), np.float32(1.0), np.float32(0.1), np.float32(0.01),
               np.float32(0.001))
    pos = (c1, c1 + 1, c1 + 2, c2, c2 + 1, c2 + 2)
    out = np.zeros(len(names), np.float32)
    tmp = -1
    for j, nm in enumerate(names):
        s = nm[:NAME_LEN].ljust(STRING_LEN)
        f = np.float32(0.0)
        for w, p in zip(weights, pos):
            ch = s[p - 1] if p - 1 < len(s) else " "
            if "0" <= ch <= "9":
                tmp = ord(ch) - 48
            f = np.float32(f + np.float32(w * np.float32(tmp)))
        out[j] = f
    return out


def file_counts(nimage: int, nfiles: int):
    """(add_Nx, n_files_sino, n_files_rec) of :233-240, FP32 arithmetic left
    to right: ceiling(4.0*(nimage+add_Nx)*nfiles*nimage / 8.0e9) and
    ceiling(4.0*(nimage+add_Nx)*(nimage+add_Nx)*nimage / 8.0e9)."""
    add = 1 if nimage % 2 == 0 else 0
    f = np.float32
    s = f(f(f(f(4.0) * f(nimage + add)) * f(nfiles)) * f(nimage)) / f(8.0e9)
    r = f(f(f(f(4.0) * f(nimage + add)) * f(nimage + add)) * f(nimage)) / f(8.0e9)
    return add, int(math.ceil(s)), int(math.ceil(r))


def writer_layout(nfiles: int, n_files_sino: int, reference_offsets: bool = True):
    """[(j, js, offset_in_images)] of the projection writes of :241-253:
    file j gets projections js = 1 + floor((j-1) r) .. min(nfiles, floor(j r)),
    r = real(nfiles) / real(n_files_sino) in FP32, at image offset
    js - ceiling((j-1) r) - 1.

    The writer's ceiling differs from the reader's floor
    (projections_to_sinogram_hpc.F90:146-147) when (j-1) r is not an integer:
    the first projection of such a file then gets offset -1, which
    MPI_FILE_WRITE_AT rejects with an error the reference does not check,
    so that projection is never written and the file's other projections sit
    one image earlier than the reader expects.  With reference_offsets
    (default) that layout is reproduced (offset -1 entries are returned and
    skipped by the writer); otherwise the reader's floor offsets are used
    so every projection lands where projections_to_sinogram_hpc reads it."""
    f = np.float32
    r = f(nfiles) / f(n_files_sino)
    out = []
    for j in range(1, n_files_sino + 1):
        a = 1 + int(math.floor(f(j - 1) * r))
        b = min(nfiles, int(math.floor(f(j) * r)))
        base = int(math.ceil(f(j - 1) * r)) if reference_offsets else int(math.floor(f(j - 1) * r))
        for js in range(a, b + 1):
            out.append((j, js, js - base - 1))
    return out


def _write_at(path: str, offset: int, data: bytes) -> None:
    """Write data at a byte offset, creating the file but never truncating
    it (MPI_FILE_OPEN with MPI_MODE_CREATE, then MPI_FILE_WRITE_AT)."""
    with open(path, "ab"):
        pass
    with open(path, "r+b") as fh:
        fh.seek(offset)
        fh.write(data)


class Normalizer:
    """The GPU part: open-beam average and -log(I / max(OB, 1))."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._lib = _lib.load()

    def mean(self, imgs) -> np.ndarray:
        a = np.ascontiguousarray(imgs, np.float32)
        if a.ndim < 2 or a.shape[0] < 1:
            raise ValueError("imgs must be (count, ...) with count >= 1")
        out = np.empty(a.shape[1:], np.float32)
        check(self._lib.acu_ingest_mean(self.ctx.handle, a.shape[0], out.size, a.ctypes.data_as(_F),
                                        out.ctypes.data_as(_F)))
        return out

    def normalize(self, imgs, open_beam) -> np.ndarray:
        a = np.ascontiguousarray(imgs, np.float32)
        ob = np.ascontiguousarray(open_beam, np.float32)
        if a.ndim < 2 or a.shape[1:] != ob.shape:
            raise ValueError(f"imgs {a.shape} must be (count,) + open_beam shape {ob.shape}")
        out = np.empty_like(a)
        check(self._lib.acu_ingest_normalize(self.ctx.handle, a.shape[0], ob.size, a.ctypes.data_as(_F),
                                             ob.ctypes.data_as(_F), out.ctypes.data_as(_F)))
        return out


@dataclass
class IngestResult:
    nfiles: int
    nimage: int
    n_files_sino: int
    n_files_rec: int
    n_ob: int
    n_df: int
    theta: np.ndarray
    dropped: list  # projections the reference's writer never writes (offset -1)


def fits_to_hpc(ctx, root: str, exp: str, batch: int = 32, reference_offsets: bool = True,
                log=None) -> IngestResult:
    """The whole of fits_to_hpc.F90 for $ROOT_CT_DIR=root, $EXP_CT_NAME=exp."""
    wdir = os.path.join(root, exp)
    rec = os.path.join(wdir, "reconstruction")
    os.makedirs(rec, exist_ok=True)
    names = list_dir(os.path.join(wdir, "projections"))
    nfiles = len(names)
    if nfiles == 0:
        raise FileNotFoundError(f"no projections under {wdir}/projections")
    theta = parse_theta(names, exp)
    _write_at(os.path.join(rec, "theta.dat"), 0, theta.astype(np.float32).tobytes())
    ob_names = list_dir(os.path.join(wdir, "OB"))
    df_dir = os.path.join(wdir, "DF")
    df_names = list_dir(df_dir) if os.path.isdir(df_dir) else []
    if not ob_names:
        raise FileNotFoundError(f"no open-beam images under {wdir}/OB")
    nimage = fits.image_shape(os.path.join(wdir, "OB", ob_names[0]))[0]
    if nimage < 1:
        raise fits.FitsError("first open-beam image has no NAXIS1")
    gpu = Normalizer(ctx)
    obs = np.stack([fits.read_image(os.path.join(wdir, "OB", nm), nimage, NULLVAL) for nm in ob_names])
    bg = gpu.mean(obs)
    del obs
    _, n_files_sino, n_files_rec = file_counts(nimage, nfiles)
    layout = writer_layout(nfiles, n_files_sino, reference_offsets)
    dropped = [js for _, js, off in layout if off < 0]
    nn = nimage * nimage
    by_file = {}
    for j, js, off in layout:
        by_file.setdefault(j, []).append((js, off))
    for j in range(1, n_files_sino + 1):
        path = os.path.join(rec, f"projections{j}.dat")
        _write_at(path, 0, b"")
        todo = by_file.get(j, [])
        with open(path, "r+b") as fh:
            for k in range(0, len(todo), batch):
                part = todo[k:k + batch]
                imgs = np.stack([fits.read_image(os.path.join(wdir, "projections", names[js - 1]), nimage,
                                                 NULLVAL) for js, _ in part])
                proj = gpu.normalize(imgs, bg)
                for (js, off), p in zip(part, proj):
                    if off < 0:
                        continue
                    fh.seek(off * nn * 4)
                    fh.write(p.tobytes())
                if log:
                    log(f"file {j}: projections {part[0][0]}..{part[-1][0]}")
    info = np.array([nfiles, nimage, n_files_sino, n_files_rec], np.float32)
    _write_at(os.path.join(rec, "fileinfo_projections.dat"), 0, info.tobytes())
    return IngestResult(nfiles, nimage, n_files_sino, n_files_rec, len(ob_names), len(df_names), theta, dropped)


def main(argv=None) -> int:
    import argparse

    from .tomography import Context

    ap = argparse.ArgumentParser(description="FITS radiographs -> normalised projection files (GPU)")
    ap.add_argument("root", nargs="?", default=os.environ.get("ROOT_CT_DIR"))
    ap.add_argument("exp", nargs="?", default=os.environ.get("EXP_CT_NAME"))
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--reader-offsets", action="store_true",
                    help="write every projection where the sinogram step reads it")
    a = ap.parse_args(argv)
    if not a.root or not a.exp:
        ap.error("ROOT_CT_DIR and EXP_CT_NAME (arguments or environment) are required")
    with Context(0) as ctx:
        r = fits_to_hpc(ctx, a.root, a.exp, a.batch, not a.reader_offsets, log=print)
    print(f"{r.nfiles} projections of {r.nimage}^2, {r.n_ob} open-beam, {r.n_df} dark-field images; "
          f"{r.n_files_sino} projection files" + (f"; not written (reference offsets): {r.dropped}"
                                                   if r.dropped else ""))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
