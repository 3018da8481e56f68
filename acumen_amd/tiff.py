"""TIFF export of reconstructions, sinograms and projections: the MATLAB
viewers of the reference (tomography/convert_to_tiff.m,
convert_sub_to_tiff.m, projections_sino_tiff.m) without MATLAB.

Those scripts read an FP32 image from the .dat files into a MATLAB matrix
with `reshape(fread(...), rows, cols)` -- first index fastest, so matrix row
i is the Fortran first index (x for images, the detector for sinograms) --
and save it with `imwrite(matrix, file, 'tiff')`.  For a double matrix,
imwrite writes 8-bit grayscale: values clipped to [0, 1], times 255,
rounded half away from zero; the default TIFF compression for such images
is PackBits.  `matlab_uint8` and `write_tiff` restate exactly that; the
loops restate the scripts' file names and offsets (including
convert_to_tiff.m taking the fbp file index from n_files_sino but its
offset from n_files_rec, convert_to_tiff.m:33-34).

    python -m acumen_amd.tiff ROOT_CT_DIR EXP_CT_NAME [--sub] [--raw]
"""
from __future__ import annotations

import argparse
import math
import os
import struct

import numpy as np


def matlab_uint8(x) -> np.ndarray:
    """imwrite's conversion of a double image: uint8(round(255 * clip(x, 0, 1))),
    rounding half away from zero; NaN becomes 0."""
    a = np.asarray(x, np.float64)
    a = np.where(np.isnan(a), 0.0, a)
    return np.floor(np.clip(a, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def packbits(data: bytes) -> bytes:
    """PackBits run-length encoding (TIFF 6.0, section 9): literal runs of
    1-128 bytes as (n-1, bytes), replicate runs of 2-128 as (1-n, byte)."""
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        j = i + 1
        while j < n and j - i < 128 and data[j] == data[i]:
            j += 1
        if j - i >= 2:
            out.append((257 - (j - i)) & 0xFF)
            out.append(data[i])
            i = j
            continue
        # literal run: up to the next pair of equal bytes, at most 128
        j = i
        while j < n and j - i < 128 and not (j + 1 < n and data[j] == data[j + 1]):
            j += 1
        if j == i:  # a single byte before a repeat
            j = i + 1
        out.append(j - i - 1)
        out.extend(data[i:j])
        i = j
    return bytes(out)


def write_tiff(path: str, img, compression: str = "packbits") -> None:
    """Baseline little-endian 8-bit grayscale TIFF of img (rows, cols) uint8,
    strips of about 8 KB (rows per strip = max(1, 8192 // cols)), 72 dpi."""
    a = np.ascontiguousarray(img, np.uint8)
    if a.ndim != 2:
        raise ValueError("img must be 2-D")
    rows, cols = a.shape
    rps = max(1, 8192 // cols)
    strips = []
    for r0 in range(0, rows, rps):
        raw = a[r0:r0 + rps].tobytes()
        if compression == "packbits":
            strips.append(b"".join(packbits(a[r].tobytes()) for r in range(r0, min(rows, r0 + rps))))
        elif compression == "none":
            strips.append(raw)
        else:
            raise ValueError(compression)
    ns = len(strips)
    comp = 32773 if compression == "packbits" else 1
    # layout: header (8) | strips | offsets[ns] | counts[ns] | resolutions | IFD
    pos = 8
    offs = []
    for s in strips:
        offs.append(pos)
        pos += len(s) + (len(s) & 1)
    off_arr, cnt_arr = pos, pos + 4 * ns
    xres = cnt_arr + 4 * ns
    yres = xres + 8
    ifd = yres + 8
    tags = [
        (254, 4, 1, 0),                      # NewSubfileType
        (256, 4, 1, cols),                   # ImageWidth
        (257, 4, 1, rows),                   # ImageLength
        (258, 3, 1, 8),                      # BitsPerSample
        (259, 3, 1, comp),                   # Compression
        (262, 3, 1, 1),                      # PhotometricInterpretation: BlackIsZero
        (273, 4, ns, offs[0] if ns == 1 else off_arr),   # StripOffsets
        (277, 3, 1, 1),                      # SamplesPerPixel
        (278, 4, 1, rps),                    # RowsPerStrip
        (279, 4, ns, len(strips[0]) if ns == 1 else cnt_arr),  # StripByteCounts
        (282, 5, 1, xres),                   # XResolution
        (283, 5, 1, yres),                   # YResolution
        (284, 3, 1, 1),                      # PlanarConfiguration: chunky
        (296, 3, 1, 2),                      # ResolutionUnit: inch
    ]
    with open(path, "wb") as fh:
        fh.write(b"II" + struct.pack("<HI", 42, ifd))
        for s in strips:
            fh.write(s)
            if len(s) & 1:
                fh.write(b"\0")
        fh.write(struct.pack(f"<{ns}I", *offs))
        fh.write(struct.pack(f"<{ns}I", *[len(s) for s in strips]))
        fh.write(struct.pack("<II", 72, 1) * 2)
        fh.write(struct.pack("<H", len(tags)))
        for tag, typ, count, value in tags:
            if typ == 3 and count == 1:
                fh.write(struct.pack("<HHIHH", tag, typ, count, value, 0))
            else:
                fh.write(struct.pack("<HHII", tag, typ, count, value))
        fh.write(struct.pack("<I", 0))


def _read_f32(path: str, offset: int, count: int) -> np.ndarray:
    with open(path, "rb") as fh:
        fh.seek(4 * offset)
        a = np.fromfile(fh, np.float32, count)
    if a.size != count:
        raise IOError(f"{path}: short read of {count} values at offset {offset}")
    return a


def _header(rec: str):
    h = np.fromfile(os.path.join(rec, "fileinfo_projections.dat"), np.float32, 4)
    return int(h[0]), int(h[1]), int(h[2]), int(h[3])


def export_reconstructions(root: str, exp: str, sub: bool = False) -> list:
    """convert_to_tiff.m (sub: convert_sub_to_tiff.m): every slice of
    fbp{js}.dat and fbp_tv{js}.dat (fbp_sub / fbp_sub_tv) as
    fbp_images/fbp{j}.tiff and fbp_tv_images/fbp_tv{j}.tiff (fbp_sub_images,
    fbp_sub_tv_images) under ROOT_CT_DIR/EXP_CT_NAME."""
    wdir = os.path.join(root, exp)
    rec = os.path.join(wdir, "reconstruction")
    nfiles, nimage, n_files_sino, n_files_rec = _header(rec)
    add = 1 if nimage % 2 == 0 else 0
    w = nimage + add
    rtmp2 = np.float32(nimage) / np.float32(n_files_sino)
    rtmp = np.float32(nimage) / np.float32(n_files_rec)
    a, b = ("fbp_sub", "fbp_sub_tv") if sub else ("fbp", "fbp_tv")
    da, db = os.path.join(wdir, a + "_images"), os.path.join(wdir, b + "_images")
    os.makedirs(da, exist_ok=True)
    os.makedirs(db, exist_ok=True)
    written = []
    for j in range(1, nimage + 1):
        js = int(math.ceil(np.float32(j) / rtmp2))
        off_a = (j - int(math.floor(np.float32(js - 1) * rtmp)) - 1) * w * w     # :33-34
        off_b = (j - int(math.floor(np.float32(js - 1) * rtmp2)) - 1) * w * w    # :40-41
        for prefix, d, off in ((a, da, off_a), (b, db, off_b)):
            img = _read_f32(os.path.join(rec, f"{prefix}{js}.dat"), off, w * w).reshape(w, w)
            # matrix row = Fortran first index (x): the transpose of [y, x]
            p = os.path.join(d, f"{prefix}{j}.tiff")
            write_tiff(p, matlab_uint8(img.T))
            written.append(p)
    return written


def export_projections_sinograms(root: str, exp: str) -> list:
    """projections_sino_tiff.m: projections{js}.dat -> projections_images/
    projections{j}.tiff, sinograms{js}.dat -> sinogram_images/sinogram{j}.tiff."""
    wdir = os.path.join(root, exp)
    rec = os.path.join(wdir, "reconstruction")
    nfiles, nimage, n_files_sino, _ = _header(rec)
    add = 1 if nimage % 2 == 0 else 0
    dp, ds = os.path.join(wdir, "projections_images"), os.path.join(wdir, "sinogram_images")
    os.makedirs(dp, exist_ok=True)
    os.makedirs(ds, exist_ok=True)
    written = []
    rtmp2 = np.float32(nimage) / np.float32(n_files_sino)
    rtmp = np.float32(nfiles) / np.float32(n_files_sino)
    for j in range(1, nfiles + 1):
        js = int(math.ceil(np.float32(j) / rtmp))
        off = (j - int(math.floor(np.float32(js - 1) * rtmp)) - 1) * nimage * nimage
        path = os.path.join(rec, f"projections{js}.dat")
        if not os.path.exists(path):
            break
        img = _read_f32(path, off, nimage * nimage).reshape(nimage, nimage)
        p = os.path.join(dp, f"projections{j}.tiff")
        write_tiff(p, matlab_uint8(img.T))
        written.append(p)
    for j in range(1, nimage + 1):
        js = int(math.ceil(np.float32(j) / rtmp2))
        off = (j - int(math.floor(np.float32(js - 1) * rtmp2)) - 1) * (nimage + add) * nfiles
        sino = _read_f32(os.path.join(rec, f"sinograms{js}.dat"), off, (nimage + add) * nfiles)
        p = os.path.join(ds, f"sinogram{j}.tiff")
        write_tiff(p, matlab_uint8(sino.reshape(nfiles, nimage + add).T))
        written.append(p)
    return written


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="TIFF export of the reconstruction files (MATLAB viewers' format)")
    ap.add_argument("root")
    ap.add_argument("exp")
    ap.add_argument("--sub", action="store_true", help="the sub-sampled reconstructions (fbp_sub*)")
    ap.add_argument("--raw", action="store_true", help="projections and sinograms instead of reconstructions")
    a = ap.parse_args(argv)
    files = export_projections_sinograms(a.root, a.exp) if a.raw else export_reconstructions(a.root, a.exp, a.sub)
    print(f"wrote {len(files)} TIFF files")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
