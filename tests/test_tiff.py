"""TIFF export (the MATLAB viewers of tomography/convert_to_tiff.m,
convert_sub_to_tiff.m, projections_sino_tiff.m): PackBits against the TIFF
6.0 specification's example, imwrite's double -> uint8 conversion, files
decoded by an independent reader (PIL), and the scripts' file loops."""
import os

import numpy as np
import pytest

from acumen_amd import io as aio
from acumen_amd import tiff

Image = pytest.importorskip("PIL.Image")


def test_packbits_specification_example():
    # TIFF 6.0, section 9 (PackBits compression)
    raw = bytes.fromhex("AAAAAA80002AAAAAAAAA80002A22AAAAAAAAAAAAAAAAAAAA")
    assert tiff.packbits(raw) == bytes.fromhex("FEAA0280002AFDAA0380002A22F7AA")


def test_packbits_long_runs_and_literals():
    raw = bytes([7] * 300) + bytes(range(200)) + b"\x01\x01"
    enc = tiff.packbits(raw)
    # decode
    out, i = bytearray(), 0
    while i < len(enc):
        h = enc[i]
        if h < 128:
            out += enc[i + 1:i + 2 + h]
            i += 2 + h
        else:
            out += bytes([enc[i + 1]]) * (257 - h)
            i += 2
    assert bytes(out) == raw


def test_matlab_uint8_conversion():
    x = np.array([-1.0, 0.0, 0.5 / 255, 1.5 / 255, 0.998, 1.0, 2.0, np.nan])
    assert list(tiff.matlab_uint8(x)) == [0, 0, 1, 2, 254, 255, 255, 0]


@pytest.mark.parametrize("compression", ["packbits", "none"])
@pytest.mark.parametrize("shape", [(5, 7), (33, 1000), (300, 2)])
def test_tiff_decodes_with_pil(tmp_path, compression, shape):
    rng = np.random.default_rng(shape[0])
    img = rng.integers(0, 256, size=shape).astype(np.uint8)
    img[: shape[0] // 2, : shape[1] // 2] = 17  # runs
    p = str(tmp_path / "a.tiff")
    tiff.write_tiff(p, img, compression)
    with Image.open(p) as im:
        assert im.mode == "L" and im.size == (shape[1], shape[0])
        assert np.array_equal(np.asarray(im), img)


def test_export_reconstructions_and_raw(tmp_path):
    n, nfiles = 6, 5          # even: images are 7 x 7, sinogram rows 7 wide
    w = n + 1
    ds = aio.Dataset(str(tmp_path), "exp")
    th = np.linspace(0, 180, nfiles).astype(np.float32)
    rng = np.random.default_rng(0)
    sinos = rng.uniform(0, 1, (n, nfiles, w)).astype(np.float32)
    aio.write_dataset(ds, th, sinos, n_files_sino=2, n_files_rec=1)
    # header as fits_to_hpc writes it: [nfiles, nimage, n_files_sino, n_files_rec]
    np.array([nfiles, n, 2, 2], np.float32).tofile(ds.path("fileinfo_projections.dat"))
    imgs = rng.uniform(-0.2, 1.2, (n, w, w)).astype(np.float32)
    # fbp{js}.dat / fbp_tv{js}.dat with 3 slices per file
    for js in (1, 2):
        imgs[3 * (js - 1):3 * js].tofile(ds.path(f"fbp{js}.dat"))
        (imgs[3 * (js - 1):3 * js] * 0.5).tofile(ds.path(f"fbp_tv{js}.dat"))
    files = tiff.export_reconstructions(str(tmp_path), "exp")
    assert len(files) == 2 * n
    for j in (1, 4, 6):
        with Image.open(os.path.join(str(tmp_path), "exp", "fbp_images", f"fbp{j}.tiff")) as im:
            a = np.asarray(im)
        assert np.array_equal(a, tiff.matlab_uint8(imgs[j - 1].T))     # rows = x
        with Image.open(os.path.join(str(tmp_path), "exp", "fbp_tv_images", f"fbp_tv{j}.tiff")) as im:
            assert np.array_equal(np.asarray(im), tiff.matlab_uint8((imgs[j - 1] * 0.5).T))
    raw = tiff.export_projections_sinograms(str(tmp_path), "exp")
    assert len(raw) == n  # no projections files: sinograms only
    with Image.open(os.path.join(str(tmp_path), "exp", "sinogram_images", "sinogram5.tiff")) as im:
        assert np.array_equal(np.asarray(im), tiff.matlab_uint8(sinos[4].T))  # rows = detector
