"""Projection -> sinogram preprocessing on the GPU against the oracle
(tomography/projections_to_sinogram_hpc.F90): median and correction bit for
bit, the mirror error to FP64 summation order, the centre search to the
oracle's result, and the whole pipeline bit for bit."""
import ctypes

import numpy as np
import pytest

from acumen_amd import _lib, prep
from helpers import HipBuffers
from oracle import ref
from test_oracle_prep import blob_image

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n", [33, 64, 129])
def test_median27_bit_exact(ctx, n):
    say("median27", n)
    rng = np.random.default_rng(n)
    l, c, r = [rng.normal(size=(n, n)).astype(np.float32) for _ in range(3)]
    c[::7, ::5] = l[::7, ::5]  # ties
    hb = HipBuffers()
    try:
        dl, dc, dr, do = hb.upload(l), hb.upload(c), hb.upload(r), hb.alloc(c.nbytes)
        assert _lib.load().acu_prep_median27_dev(ctx.handle, n, dl, dc, dr, do) == 0
        ctx.synchronize()
        got = hb.download(do, (n, n))
    finally:
        hb.close()
    assert np.array_equal(got, ref.prep_median27(l, c, r))


@pytest.mark.parametrize("n", [32, 65])
@pytest.mark.parametrize("theta,shift", [(0.0, 0.0), (0.013, -0.021), (-0.05, 0.0037)])
def test_correct_bit_exact(ctx, n, theta, shift):
    say("correct", n, theta, shift)
    img = blob_image(n, 7) + np.random.default_rng(2).normal(0, 0.01, (n, n)).astype(np.float32)
    add = 1 if n % 2 == 0 else 0
    nw = n + add
    hb = HipBuffers()
    try:
        di, ds = hb.upload(img), hb.alloc(n * 1 * nw * 4)
        assert _lib.load().acu_prep_correct_dev(ctx.handle, n, add, theta, shift, di, 1, 0, 1, ds) == 0
        ctx.synchronize()
        got = hb.download(ds, (n, nw))
    finally:
        hb.close()
    assert np.array_equal(got, ref.prep_correct(img, theta, shift))


def test_mirror_error_matches_oracle(ctx):
    n = 96
    p0 = blob_image(n, 9)
    p180 = p0[:, ::-1].copy() + np.float32(0.01)
    hb = HipBuffers()
    try:
        d0, d1 = hb.upload(p0), hb.upload(p180)
        for th, sh in [(0.0, 0.0), (0.002, 0.01), (-0.01, -0.02)]:
            e = ctypes.c_double()
            assert _lib.load().acu_prep_mirror_error_dev(ctx.handle, n, 1, th, sh, d0, d1, ctypes.byref(e)) == 0
            want = ref.prep_mirror_error(p0, p180, th, sh)
            assert abs(e.value - want) <= 1e-12 * want, (th, sh, e.value, want)
    finally:
        hb.close()


@pytest.mark.parametrize("n,d", [(63, 2), (96, -3)])
def test_find_center_matches_oracle(ctx, n, d):
    say("find_center", n, d)
    p0 = blob_image(n, 11)
    p180 = p0[:, np.clip(n - 1 + 2 * d - np.arange(n), 0, n - 1)].copy()
    th, sh, ev = prep.find_center(ctx, p0, p180)
    rth, rsh = ref.prep_find_center(p0, p180)
    assert ev > 12
    assert th == rth and sh == rsh, ((th, sh), (rth, rsh))


@pytest.mark.parametrize("n,nfiles", [(40, 7), (33, 1), (64, 12)])
def test_sinogram_pipeline_bit_exact(ctx, n, nfiles):
    say("prep pipeline", n, nfiles)
    rng = np.random.default_rng(n + nfiles)
    proj = np.stack([blob_image(n, 20 + j) + rng.normal(0, 0.02, (n, n)).astype(np.float32)
                     for j in range(nfiles)])
    th, sh = 0.0123, -0.0177
    got = prep.sinograms(ctx, proj, th, sh)
    nw = n + (1 if n % 2 == 0 else 0)
    assert got.shape == (n, nfiles, nw)
    for j in range(nfiles):
        med = ref.prep_median27(proj[max(j - 1, 0)], proj[j], proj[min(j + 1, nfiles - 1)])
        want = ref.prep_correct(med, th, sh)
        assert np.array_equal(got[:, j, :], want), j
