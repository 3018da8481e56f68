"""FITS ingest host logic (tomography/fits_to_hpc.F90) and the oracle of its
two GPU steps, pinned by known answers."""
import math

import numpy as np
import pytest

from acumen_amd import ingest
from oracle import ref


def name(exp, deg: str, frac: str, idx=1, sep="_"):
    # 10 characters after the experiment name, then ddd, a separator, fff
    return f"{exp}_{idx:04d}_0000{deg}{sep}{frac}.fits"


def test_theta_digits_known_answers():
    exp = "CT_Aluminum_ironcenter_0090"
    names = [name(exp, "000", "000"), name(exp, "123", "456"), name(exp, "359", "999"), name(exp, "007", "5x0")]
    th = ingest.parse_theta(names, exp)
    assert th.dtype == np.float32
    assert th[0] == 0.0
    # FP32 left-to-right sums of FP32 products
    assert th[1] == np.float32(np.float32(np.float32(np.float32(np.float32(100 + 20) + 3) + np.float32(0.1) * 4)
                                          + np.float32(0.01) * 5) + np.float32(0.001) * 6)
    assert abs(float(th[1]) - 123.456) < 1e-4 and abs(float(th[2]) - 359.999) < 1e-4
    # a non-digit repeats the previous digit: "5x0" reads 5, 5, 0
    assert abs(float(th[3]) - 7.550) < 1e-5


def test_theta_carry_and_blank_padding():
    exp = "E"
    # first digit of the first file not a digit: the carried value is -1
    th = ingest.parse_theta([name(exp, "x12", "000")], exp)
    assert th[0] == np.float32(-100 + 10 + 2)
    # a name too short to hold the digits: blanks, every digit carries
    th = ingest.parse_theta([name(exp, "250", "125"), "E_short"], exp)
    assert abs(float(th[0]) - 250.125) < 4e-5  # 1 ulp at 250 is 1.5e-5
    assert abs(float(th[1]) - (555 + 0.555)) < 1e-3
    # trailing blanks of EXP_CT_NAME do not count (len_trim)
    assert ingest.parse_theta([name(exp, "090", "000")], "E   ")[0] == 90.0
    # names are truncated to 100 characters before parsing
    long_exp = "X" * 95
    assert ingest.parse_theta([name(long_exp, "180", "000")], long_exp)[0] == np.float32(-111.111)


def test_listing_order():
    names = ["b.fits", "B.fits", "a_10.fits", "a_9.fits", "a_010.fits"]
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        for n in names:
            open(os.path.join(d, n), "w").close()
        assert ingest.list_dir(d) == ["B.fits", "a_010.fits", "a_10.fits", "a_9.fits", "b.fits"]


@pytest.mark.parametrize("nimage,nfiles", [(2048, 1500), (2049, 1500), (512, 720), (7, 3), (4096, 2000),
                                           (1024, 9001)])
def test_file_counts(nimage, nfiles):
    add, ns, nr = ingest.file_counts(nimage, nfiles)
    assert add == (1 if nimage % 2 == 0 else 0)
    # away from integer boundaries the FP32 formula equals the exact one
    exact_s = 4 * (nimage + add) * nfiles * nimage / 8e9
    exact_r = 4 * (nimage + add) ** 2 * nimage / 8e9
    assert ns == math.ceil(exact_s) and nr == math.ceil(exact_r)
    if (nimage, nfiles) == (2048, 1500):
        assert (ns, nr) == (4, 5)


def reader_file(js, nfiles, nfs):
    # projections_to_sinogram_hpc.F90:143-147: j = ceiling(js / r), offset
    # (js - floor((j-1) r) - 1) images, r = real(nfiles) / real(n_files_sino)
    r = np.float32(nfiles) / np.float32(nfs)
    j = int(math.ceil(np.float32(js) / r))
    return j, js - int(math.floor(np.float32(j - 1) * r)) - 1


@pytest.mark.parametrize("nfiles,nfs", [(10, 3), (12, 3), (1500, 4), (7, 7), (5, 1), (1001, 6)])
def test_writer_layout(nfiles, nfs):
    lay = ingest.writer_layout(nfiles, nfs)
    # every projection is assigned to exactly one file, in order, and that
    # file is the one the sinogram step opens for it
    assert [js for _, js, _ in lay] == list(range(1, nfiles + 1))
    for j, js, off in lay:
        rj, roff = reader_file(js, nfiles, nfs)
        assert j == rj
        # the writer's ceiling offset equals the reader's floor offset
        # exactly when (j-1) r is an integer; otherwise it is one less
        r = np.float32(nfiles) / np.float32(nfs)
        integral = float(np.float32(j - 1) * r).is_integer()
        assert off == (roff if integral else roff - 1)
    dropped = [js for _, js, off in lay if off < 0]
    if nfiles % nfs == 0:
        assert dropped == []
    if (nfiles, nfs) == (10, 3):
        assert dropped == [4, 7]
    fixed = ingest.writer_layout(nfiles, nfs, reference_offsets=False)
    assert all(off == reader_file(js, nfiles, nfs)[1] for _, js, off in fixed)


def test_oracle_mean_accumulates_in_file_order():
    big = np.float32(1e8)
    seq = lambda vals: ref.ingest_mean(np.array(vals, np.float32).reshape(len(vals), 1, 1))[0, 0]
    # FP32 in file order: 1e8 + 1 == 1e8, so the order of the files shows
    assert seq([big, 1.0, -big]) == 0.0
    assert seq([big, -big, 1.0]) == np.float32(1.0) / np.float32(3.0)
    # integer counts (16-bit camera images): every partial sum is exact in
    # FP32, so the mean is the exact mean rounded once
    rng = np.random.default_rng(3)
    c = rng.integers(0, 65536, (13, 9, 11)).astype(np.float32)
    want = (c.astype(np.float64).sum(0) / 13).astype(np.float32)
    np.testing.assert_array_equal(ref.ingest_mean(c), want)
    # and close to the FP64 mean in general
    a = rng.uniform(0, 60000, (13, 9, 11)).astype(np.float32)
    np.testing.assert_allclose(ref.ingest_mean(a), a.astype(np.float64).mean(0), rtol=2e-6)


def test_oracle_normalize_known_answers():
    bg = np.array([[100.0, 0.5, 1.0, 1000.0, 2.0, 5.0]], np.float32)
    img = np.array([[[100.0, 0.5, 1.0, 1000.0 / math.e, 0.0, -999.0]]], np.float32)
    out = ref.ingest_normalize(img, bg)[0, 0]
    assert out[0] == 0.0                       # I = OB
    assert out[1] == np.float32(-math.log(0.5))  # OB < 1 is replaced by 1
    assert out[2] == 0.0
    assert abs(out[3] - 1.0) < 1e-6            # transmission 1/e
    assert out[4] == np.inf                    # no counts
    assert np.isnan(out[5])                    # a null pixel (-999)
    rng = np.random.default_rng(4)
    ob = rng.uniform(0.2, 5e4, (1, 64)).astype(np.float32)
    im = (ob * rng.uniform(0.01, 1.2, (3, 1, 64))).astype(np.float32)
    want = -np.log(im.astype(np.float64) / np.maximum(ob.astype(np.float64), 1.0))
    np.testing.assert_allclose(ref.ingest_normalize(im, ob), want, rtol=1e-6, atol=2e-7)
