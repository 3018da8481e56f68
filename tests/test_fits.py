"""FITS reader: pinned against hand-assembled files (header cards and
big-endian data written byte by byte here, independently of
acumen_amd.fits.write_fits), then write/read round trips."""
import math
import struct

import numpy as np
import pytest

from acumen_amd import fits


def card(text: str) -> bytes:
    assert len(text) <= 80
    return text.ljust(80).encode("ascii")


def assemble(cards, data: bytes) -> bytes:
    head = b"".join(card(c) for c in cards) + card("END")
    head += b" " * ((-len(head)) % 2880)
    return head + data + b"\0" * ((-len(data)) % 2880)


def put(tmp_path, name, blob):
    p = tmp_path / name
    p.write_bytes(blob)
    return str(p)


def test_uint16_camera_image_with_bzero(tmp_path):
    # unsigned 16-bit pixels stored as signed with BZERO = 32768, the usual
    # neutron-camera form: physical = raw + 32768
    phys = [0, 1, 32767, 32768, 65535, 1000]
    raw = struct.pack(">6h", *[v - 32768 for v in phys])
    p = put(tmp_path, "u16.fits", assemble(["SIMPLE  =                    T", "BITPIX  =                   16",
                                            "NAXIS   =                    2", "NAXIS1  =                    3",
                                            "NAXIS2  =                    2", "BZERO   =                32768",
                                            "BSCALE  =                    1"], raw))
    assert fits.image_shape(p) == (3, 2)
    a = fits.read_pixels(p, 6)
    assert a.dtype == np.float32 and a.tolist() == [float(v) for v in phys]


def test_int32_rounds_once_to_float(tmp_path):
    vals = [16777217, -16777219, 2147483647, 7]  # not representable in FP32 (first three)
    p = put(tmp_path, "i32.fits", assemble(["SIMPLE  =                    T", "BITPIX  =                   32",
                                            "NAXIS   =                    1", "NAXIS1  =                    4"],
                                           struct.pack(">4i", *vals)))
    a = fits.read_pixels(p, 4)
    assert a.tolist() == [float(np.float32(v)) for v in vals]
    assert a[0] == 16777216.0 and a[1] == -16777220.0


def test_scaled_bytes_and_blank(tmp_path):
    # physical = raw * 0.5 - 3, computed in double then rounded to FP32; BLANK -> nullval
    raw = bytes([0, 1, 255, 7, 9])
    p = put(tmp_path, "u8.fits", assemble(["SIMPLE  =                    T", "BITPIX  =                    8",
                                           "NAXIS   =                    1", "NAXIS1  =                    5",
                                           "BSCALE  =                  0.5", "BZERO   =                 -3.0",
                                           "BLANK   =                    7"], raw))
    assert fits.read_pixels(p, 5).tolist() == [-3.0, -2.5, 124.5, -999.0, 1.5]
    # nullval = 0 disables the null test: the BLANK pixel is converted like any other
    assert fits.read_pixels(p, 5, nullval=0.0).tolist() == [-3.0, -2.5, 124.5, 0.5, 1.5]


def test_ieee_nulls_and_underflow(tmp_path):
    # NaN and +-Inf -> nullval; zero and subnormals -> 0.0 (cfitsio's fnan rule)
    words = [0x3F800000, 0x7FC00000, 0x7F800000, 0xFF800000, 0x00000001, 0x80000000, 0xC0490FDB, 0x00800000]
    p = put(tmp_path, "f32.fits", assemble(["SIMPLE  =                    T", "BITPIX  =                  -32",
                                            "NAXIS   =                    1", "NAXIS1  =                    8"],
                                           struct.pack(">8I", *words)))
    a = fits.read_pixels(p, 8)
    want = [1.0, -999.0, -999.0, -999.0, 0.0, 0.0, float(np.float32(-3.14159274)), 2.0 ** -126]
    assert a.tolist() == want
    assert not math.copysign(1.0, float(a[5])) < 0  # -0 becomes +0
    b = fits.read_pixels(p, 8, nullval=0.0)  # no checks: values pass through
    assert np.isnan(b[1]) and np.isinf(b[2]) and b[4] == np.float32(1.4e-45)


def test_float64_pixels_round_to_float(tmp_path):
    vals = [1.0 / 3.0, 1e300, 5e-324, -2.5]
    p = put(tmp_path, "f64.fits", assemble(["SIMPLE  =                    T", "BITPIX  =                  -64",
                                            "NAXIS   =                    2", "NAXIS1  =                    2",
                                            "NAXIS2  =                    2"], struct.pack(">4d", *vals)))
    a = fits.read_image(p, 2)
    assert a.shape == (2, 2)
    assert a.ravel().tolist() == [float(np.float32(1 / 3)), float("inf"), 0.0, -2.5]


def test_header_spanning_blocks_and_comments(tmp_path):
    cards = ["SIMPLE  =                    T / standard", "BITPIX  =                  -32 / float",
             "NAXIS   =                    2", "NAXIS1  =                    2", "NAXIS2  =                    1"]
    cards += [f"COMMENT filler {i}" for i in range(40)]  # pushes END into the second block
    cards += ["BZERO   =               1.0D+1 / Fortran exponent", "OBJECT  = 'it''s a / test'"]
    p = put(tmp_path, "long.fits", assemble(cards, struct.pack(">2f", 1.5, -2.0)))
    assert fits.read_pixels(p, 2).tolist() == [11.5, 8.0]
    with open(p, "rb") as fh:
        h, off = fits.read_header(fh)
    assert off == 2 * 2880 and h["OBJECT"] == "it's a / test" and h["BZERO"] == 10.0


def test_errors(tmp_path):
    p = put(tmp_path, "short.fits", assemble(["SIMPLE  =                    T", "BITPIX  =                  -32",
                                              "NAXIS   =                    1", "NAXIS1  =                    2"],
                                             struct.pack(">2f", 1, 2)))
    with pytest.raises(fits.FitsError):
        fits.read_pixels(p, 3)  # more pixels than the image holds
    bad = put(tmp_path, "bad.fits", b"NOTFITS " + b" " * 2872)
    with pytest.raises(fits.FitsError):
        fits.read_pixels(bad, 1)
    trunc = put(tmp_path, "trunc.fits", card("SIMPLE  =                    T") * 3)
    with pytest.raises(fits.FitsError):
        fits.image_shape(trunc)


@pytest.mark.parametrize("bitpix,kw", [(16, dict(bzero=32768.0)), (8, {}), (32, {}), (-32, {}), (-64, {}),
                                       (16, dict(bzero=100.0, bscale=0.25))])
def test_round_trip(tmp_path, bitpix, kw):
    rng = np.random.default_rng(bitpix + 100)
    if bitpix == 8:
        img = rng.integers(0, 256, (5, 7)).astype(np.float64)
    elif bitpix == 16 and "bscale" in kw:
        img = 100.0 + 0.25 * rng.integers(-30000, 30000, (5, 7))
    elif bitpix > 0:
        img = rng.integers(0, 60000, (5, 7)).astype(np.float64)
    else:
        img = rng.normal(size=(5, 7))
    p = str(tmp_path / "rt.fits")
    fits.write_fits(p, img, bitpix=bitpix, **kw)
    assert fits.image_shape(p) == (7, 5)
    got = fits.read_pixels(p, 35).reshape(5, 7)
    np.testing.assert_array_equal(got, img.astype(np.float32))
    assert (tmp_path / "rt.fits").stat().st_size % 2880 == 0
