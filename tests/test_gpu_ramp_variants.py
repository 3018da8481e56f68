"""The ramp filter's arithmetic variants (FP64 default, FP32 MFMA, BF16 split
into 2 and 3 parts, FFT convolution): each against the FP64 oracle on a
smooth phantom sinogram (heavy cancellation) and on white noise, the FFT
form on both sides of its wrap-around correction, and the FBP each gives."""
import numpy as np
import pytest

from acumen_amd import AcumenError, Batch, _lib, synth
from helpers import HipBuffers, disk_phantom, rel
from oracle import ref

pytestmark = pytest.mark.gpu

V = {"fp64": _lib.RAMP_FP64, "fp32": _lib.RAMP_FP32, "bf16x3": _lib.RAMP_BF16X3,
     "bf16x6": _lib.RAMP_BF16X6, "fft": _lib.RAMP_FFT}
# rel-L2 bounds against the FP64 oracle (FP64 taps), measured on MI355X at
# N = 65..2049 with margin.  On the phantom the reference's FP32 taps alone
# leave 0.5-2.5e-6 (the sum cancels ~N-fold); on noise only the arithmetic
# shows.  bf16x3 (hi + lo, the north star's form) keeps ~16 bits per operand.
PHANTOM = {"fp64": 4e-6, "fp32": 5e-6, "bf16x3": 5e-4, "bf16x6": 5e-6, "fft": 8e-6}
NOISE = {"fp64": 2e-7, "fp32": 2e-6, "bf16x3": 2e-5, "bf16x6": 1e-6, "fft": 5e-7}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("N,T", [(65, 48), (257, 180), (1025, 90)])
def test_ramp_variants_against_fp64(ctx, N, T):
    th = synth.angles(T)
    x = synth.phantom(N, 12, 7 + N)
    hb = HipBuffers()
    try:
        with Batch(ctx, [(N, T, synth.default_wpu(N), th)]) as b:
            s = b.radon(x)
            noise = np.random.default_rng(N).normal(size=(T, N)).astype(np.float32)
            want_s = ref.ramp(s.astype(np.float64), "f64")
            want_n = ref.ramp(noise.astype(np.float64), "f64")
            d_s, d_n, d_o = hb.upload(s), hb.upload(noise), hb.alloc(s.nbytes)
            for name, v in V.items():
                say("ramp", name, N, T)
                b.set_ramp_variant(v)
                assert b.ramp_variant() == v
                b.ramp_dev(0, d_o, d_s)
                ctx.synchronize()
                e_s = rel(hb.download(d_o, (T, N)), want_s)
                b.ramp_dev(0, d_o, d_n)
                ctx.synchronize()
                e_n = rel(hb.download(d_o, (T, N)), want_n)
                print(name, N, T, "phantom %.2e noise %.2e" % (e_s, e_n), flush=True)
                assert e_s < PHANTOM[name], (name, e_s)
                assert e_n < NOISE[name], (name, e_n)
    finally:
        hb.close()


@pytest.mark.parametrize("N", [3, 5, 33, 34, 129, 1537, 2049])
def test_fft_ramp_wraparound(ctx, N):
    # N - 1 a power of two: L = 2N - 2 and two outputs take the correction;
    # otherwise L >= 2N - 1 and the circular convolution is the linear one
    T = 7
    say("fft ramp", N)
    s = np.random.default_rng(N + 1).normal(size=(T, N)).astype(np.float32)
    with Batch(ctx, [(N, T, 1, synth.angles(T))]) as b:
        b.set_ramp_variant(_lib.RAMP_FFT)
        hb = HipBuffers()
        try:
            d_s, d_o = hb.upload(s), hb.alloc(s.nbytes)
            b.ramp_dev(0, d_o, d_s)
            ctx.synchronize()
            got = hb.download(d_o, (T, N))
        finally:
            hb.close()
    want = ref.ramp(s.astype(np.float64), "f64")
    assert rel(got, want) < 5e-7
    # the corrected outputs themselves
    assert np.allclose(got[:, [0, N - 1]], want[:, [0, N - 1]], rtol=1e-5, atol=1e-5 * np.abs(want).max())


def test_fft_ramp_size_limit(ctx):
    with Batch(ctx, [(4099, 2, 4, synth.angles(2))]) as b:
        with pytest.raises(AcumenError):
            b.set_ramp_variant(_lib.RAMP_FFT)
        assert b.ramp_variant() == _lib.RAMP_FP64


def test_fbp_with_each_ramp_variant(ctx):
    N, T = 513, 720
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    x = disk_phantom(N, 20, 77)
    f = ref.radon(x, th, "f64").astype(np.float32)
    ref.set_threads(8)
    try:
        f32 = ref.iradon(f, th, wpu, "f32")
    finally:
        ref.set_threads(1)
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        for name, v in V.items():
            b.set_ramp_variant(v)
            e = rel(b.iradon(f), f32)
            print("fbp", name, "%.2e" % e, flush=True)
            assert e < (2e-3 if name == "bf16x3" else 1e-5), (name, e)
