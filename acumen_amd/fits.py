"""Minimal FITS image I/O for the ingest step (tomography/fits_to_hpc.F90),
which reads its raw projections, open-beam and dark-field images with
cfitsio (linked with -lcfitsio, Read_Me.txt:13; the version is not pinned by
the reference).  cfitsio is not available here, so this module restates the
parts of the FITS standard (version 4.0) and of cfitsio's `ftgpve` that the
reference uses:

* the primary HDU only (ftopen without an extension specifier): header of
  80-character cards in 2880-byte blocks up to END, data from the next
  block boundary, big-endian, NAXIS1 varying fastest;
* `ftgknj(unit, 'NAXIS', 1, 2, naxes, ...)` -> (NAXIS1, NAXIS2);
* `ftgpve(unit, 1, 1, nelem, nullval, buf, ...)`: the first nelem pixels as
  FP32.  cfitsio converts integer pixels p to float as
  (float)(p * BSCALE + BZERO) in double precision when scaling is present,
  else (float)p; pixels equal to BLANK become nullval.  For IEEE pixels it
  replaces values whose exponent bits are all ones (NaN, Inf) by nullval
  and values whose exponent bits are all zero (zero, subnormals) by 0.0,
  before scaling.  The null test is only made when nullval != 0 (the
  reference passes nullval = -999).

`write_fits` writes the same subset (used to make synthetic data sets).
"""
from __future__ import annotations

import math
import os

import numpy as np

BLOCK = 2880
CARD = 80

_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


class FitsError(ValueError):
    pass


def _parse_value(text: str):
    """Value field of a header card (columns 11-80): string, logical, int or
    float (Fortran 'D' exponents accepted); None for an empty value."""
    t = text.strip()
    if not t:
        return None
    if t.startswith("'"):
        out, i = [], 1
        while i < len(t):
            if t[i] == "'":
                if i + 1 < len(t) and t[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(t[i])
            i += 1
        return "".join(out).rstrip()
    t = t.split("/", 1)[0].strip()
    if t in ("T", "F"):
        return t == "T"
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(t.replace("D", "E").replace("d", "e"))
    except ValueError:
        return t


def read_header(fh):
    """(cards dict, data offset in bytes) of the primary HDU of an open
    binary file."""
    cards = {}
    nblocks = 0
    while True:
        block = fh.read(BLOCK)
        if len(block) < BLOCK:
            raise FitsError("truncated FITS header (no END card)")
        nblocks += 1
        text = block.decode("ascii", errors="replace")
        for c in range(0, BLOCK, CARD):
            card = text[c:c + CARD]
            key = card[:8].strip()
            if key == "END":
                return cards, nblocks * BLOCK
            if card[8:10] == "= " and key and key not in cards:
                cards[key] = _parse_value(card[10:])
        if nblocks == 1 and not text.startswith("SIMPLE  ="):
            raise FitsError("not a FITS file (first card is not SIMPLE)")


def image_shape(path: str):
    """(NAXIS1, NAXIS2) of the primary image, as ftgknj(.., 'NAXIS', 1, 2, ..)."""
    with open(path, "rb") as fh:
        cards, _ = read_header(fh)
    return int(cards.get("NAXIS1", 0)), int(cards.get("NAXIS2", 0))


def read_pixels(path: str, nelem: int, nullval: float = -999.0) -> np.ndarray:
    """The first nelem pixels of the primary image as float32, with
    ftgpve's conversion, null and underflow rules (module docstring)."""
    with open(path, "rb") as fh:
        cards, off = read_header(fh)
        bitpix = int(cards.get("BITPIX", 0))
        if bitpix not in _DTYPES:
            raise FitsError(f"{path}: unsupported BITPIX {bitpix}")
        naxis = int(cards.get("NAXIS", 0))
        total = 1 if naxis > 0 else 0
        for k in range(1, naxis + 1):
            total *= int(cards.get(f"NAXIS{k}", 0))
        if nelem > total:
            raise FitsError(f"{path}: {nelem} pixels requested, the image holds {total}")
        fh.seek(off)
        raw = np.fromfile(fh, _DTYPES[bitpix], nelem)
    if raw.size != nelem:
        raise FitsError(f"{path}: data unit shorter than its header says")
    bscale = float(cards.get("BSCALE", 1.0))
    bzero = float(cards.get("BZERO", 0.0))
    scaled = not (bscale == 1.0 and bzero == 0.0)
    check = nullval != 0
    if bitpix > 0:
        null = None
        if check and "BLANK" in cards and isinstance(cards["BLANK"], int):
            null = raw == cards["BLANK"]
        if scaled:
            with np.errstate(over="ignore"):
                out = (raw.astype(np.float64) * bscale + bzero).astype(np.float32)
        else:
            out = raw.astype(np.float32)
        if null is not None:
            out[null] = np.float32(nullval)
        return out
    # IEEE pixels
    if bitpix == -32:
        expo = (raw.view(">u4") >> np.uint32(23)) & np.uint32(0xFF)
        allones, zero = expo == 0xFF, expo == 0
    else:
        expo = (raw.view(">u8") >> np.uint64(52)) & np.uint64(0x7FF)
        allones, zero = expo == 0x7FF, expo == 0
    vals = raw.astype(np.float64)
    if check:
        vals = np.where(zero | allones, 0.0, vals)
    if scaled:
        vals = vals * bscale + bzero
    with np.errstate(over="ignore"):  # C's (float) of a large double: +-inf
        out = vals.astype(np.float32)
    if check:
        out[allones] = np.float32(nullval)
    return out


def read_image(path: str, nimage: int, nullval: float = -999.0) -> np.ndarray:
    """image_buffer(nimage, nimage) of the reference as a C-ordered
    (nimage, nimage) array indexed [y, x]: the first nimage**2 pixels in file
    order (x fastest)."""
    return read_pixels(path, nimage * nimage, nullval).reshape(nimage, nimage)


def _card(key: str, value, comment: str = "") -> str:
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, int):
        v = f"{value:>20d}"
    elif isinstance(value, float):
        v = f"{value:>20.15G}"
        if "." not in v and "E" not in v:
            v = f"{value:>20.1f}"
    else:
        s = "'" + str(value).replace("'", "''").ljust(8) + "'"
        v = f"{s:<20}"
    c = f"{key:<8}= {v}"
    if comment:
        c += f" / {comment}"
    if len(c) > CARD:
        raise FitsError(f"card too long: {c!r}")
    return c.ljust(CARD)


def write_fits(path: str, data, bitpix: int = -32, bzero=None, bscale=None, blank=None) -> None:
    """Primary-HDU image of `data` (C-ordered, the last axis becomes NAXIS1).
    Integer BITPIX stores round(data) (after removing bzero/bscale)."""
    a = np.asarray(data)
    if bitpix not in _DTYPES:
        raise FitsError(f"unsupported BITPIX {bitpix}")
    shape = a.shape[::-1]
    cards = [_card("SIMPLE", True, "conforms to FITS standard"), _card("BITPIX", bitpix),
             _card("NAXIS", len(shape))]
    cards += [_card(f"NAXIS{k + 1}", int(n)) for k, n in enumerate(shape)]
    if bzero is not None:
        cards.append(_card("BZERO", float(bzero)))
    if bscale is not None:
        cards.append(_card("BSCALE", float(bscale)))
    if blank is not None:
        if bitpix < 0:
            raise FitsError("BLANK applies to integer images only")
        cards.append(_card("BLANK", int(blank)))
    cards.append("END".ljust(CARD))
    head = "".join(cards)
    head += " " * ((-len(head)) % BLOCK)
    if bitpix > 0:
        v = a.astype(np.float64)
        if bzero is not None or bscale is not None:
            v = (v - (bzero or 0.0)) / (bscale or 1.0)
        raw = np.rint(v).astype(_DTYPES[bitpix])
    else:
        raw = a.astype(_DTYPES[bitpix])
    body = raw.tobytes()
    body += b"\0" * ((-len(body)) % BLOCK)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(head.encode("ascii"))
        fh.write(body)
    os.replace(tmp, path)


def nblocks(nbytes: int) -> int:
    return int(math.ceil(nbytes / BLOCK))
