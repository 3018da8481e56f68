"""Differentiation on the GPU: the exact adjoint of the projector, the loss
and its gradients with respect to the image and the per-angle shifts, and
gradient-based shift registration.  Workload of the gradient checks: 16
slices of N = 513 (512 detectors, odd-padded) with 720 angles, shifts drawn
N(0, 1.5 px), lambda = 0.1, eps = 1e-3."""
import numpy as np
import pytest

from acumen_amd import Batch, synth
from helpers import HipBuffers, fov_mask, rel
from oracle import ref

pytestmark = pytest.mark.gpu

LAM, EPS = 0.1, 1e-3


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("N,T", [(33, 24), (65, 48), (257, 180), (128, 37)])
def test_adjoint_matches_oracle(ctx, N, T):
    say("adjoint", N, T)
    th = synth.angles(T)
    y = np.random.default_rng(N * 7 + T).normal(size=(T, N)).astype(np.float32)
    with Batch(ctx, [(N, T, 4, th)]) as b:
        got = b.radon_adjoint(y)
    want32 = ref.radon_adjoint(y, th, "f32")
    with ref.fp32_angles():
        want64 = ref.radon_adjoint(y.astype(np.float64), th, "f64")
    # the reference's FP32 positions everywhere: FP32 sums of ~4T white-noise
    # terms per pixel (block-summed over 32 angles), the bound the
    # backprojector's white-noise test uses
    assert rel(got, want32) < 5e-6, rel(got, want32)
    # against FP64 inside the field of view: on the border FP32 and FP64
    # place some samples on opposite sides of the bounds test (at 0 and 90
    # degrees the positions sit on integers), which changes which samples
    # reach the border pixels
    # FP32 positions themselves carry ~1 ulp (7.6e-6 px at 128): the
    # attainable accuracy is that of the reference's FP32 arithmetic
    m = fov_mask(N) > 0
    e_ref = rel(want32[m], want64[m])
    assert rel(got[m], want64[m]) <= 1.5 * e_ref + 2e-6, (rel(got[m], want64[m]), e_ref)


def test_adjoint_dot_product_full_grid(ctx):
    N, T = 513, 720
    say("dot product", N, T)
    th = synth.angles(T)
    rng = np.random.default_rng(2024)
    x = rng.normal(size=(N, N)).astype(np.float32)
    y = rng.normal(size=(T, N)).astype(np.float32)
    with Batch(ctx, [(N, T, 16, th)]) as b:
        rx = b.radon(x)
        aty = b.radon_adjoint(y)
    rx64, y64 = rx.astype(np.float64), y.astype(np.float64)
    x64, aty64 = x.astype(np.float64), aty.astype(np.float64)
    lhs, rhs = np.vdot(rx64, y64), np.vdot(x64, aty64)
    # <Rx, y> of random fields is ~sqrt(N T) smaller than the sum of its
    # terms' magnitudes, so the gap is measured against the Cauchy-Schwarz
    # scale |Rx| |y| (and |x| |R^T y|)
    scale = min(np.linalg.norm(rx64) * np.linalg.norm(y64), np.linalg.norm(x64) * np.linalg.norm(aty64))
    print("dot", lhs, rhs, "gap / scale", abs(lhs - rhs) / scale, flush=True)
    assert abs(lhs - rhs) <= 1e-6 * scale


def test_adjoint_dev_matches_host(ctx):
    N, T = 129, 90
    th = synth.angles(T)
    y = np.random.default_rng(9).normal(size=(T, N)).astype(np.float32)
    hb = HipBuffers()
    try:
        with Batch(ctx, [(N, T, 4, th)]) as b:
            want = b.radon_adjoint(y)
            d_y = hb.upload(y)
            d_x = hb.alloc(N * N * 4)
            b.radon_adjoint_dev(0, d_x, d_y)
            ctx.synchronize()
            got = hb.download(d_x, (N, N))
    finally:
        hb.close()
    assert np.array_equal(got, want)


@pytest.fixture(scope="module")
def cfg3():
    """16 slices of 513^2 x 720 angles: phantom images, shifted noisy
    sinograms, a perturbed image at which the gradients are taken."""
    N, T, S = 513, 720, 16
    th = synth.angles(T)
    rng = np.random.default_rng(1003)
    imgs, xs, bs, shifts = [], [], [], []
    ref.set_threads(8)
    try:
        # images supported in the field of view, so that the border samples,
        # where FP32 and FP64 positions can fall on opposite sides of the
        # bounds test, carry zeros
        fov = fov_mask(N, margin=4)
        for i in range(S):
            img = synth.phantom(N, 20, 1003 + i) * fov
            sh = rng.normal(0, 1.5, T).astype(np.float32)
            b = ref.shift(ref.radon(img, th, "f32"), sh, "f32")
            b = b + rng.normal(0, 1e-3, b.shape).astype(np.float32)
            x = img + rng.normal(0, 0.01, img.shape).astype(np.float32) * fov
            imgs.append(img); xs.append(x.astype(np.float32)); bs.append(b.astype(np.float32))
            shifts.append(sh)
    finally:
        ref.set_threads(1)
    return dict(N=N, T=T, S=S, th=th, xs=xs, bs=bs, shifts=shifts, fov=fov > 0)


@pytest.fixture(scope="module")
def cfg3_gpu(ctx, cfg3):
    c = cfg3
    say("loss_grad", c["S"], "slices", c["N"], c["T"])
    with Batch(ctx, [(c["N"], c["T"], 16, c["th"])] * c["S"]) as b:
        for i in range(c["S"]):
            b.set_sino(i, c["bs"][i])
        # perturbed shifts: the gradient is taken away from the optimum
        sh = [s + np.float32(0.3) for s in c["shifts"]]
        loss, gx, gs, hs = b.loss_grad(c["xs"], sh, LAM, EPS, curvature=True)
    return dict(loss=loss, gx=gx, gs=gs, hs=hs, sh=sh)


def test_loss_grad_matches_oracle(cfg3, cfg3_gpu):
    c, g = cfg3, cfg3_gpu
    ref.set_threads(8)
    try:
        m = c["fov"]
        for i in (0, 7, 15):
            say("oracle loss_grad slice", i)
            with ref.fp32_angles():
                L, gx, gs = ref.loss_grad(c["xs"][i], c["th"], g["sh"][i], c["bs"][i], LAM, EPS, "f64")
            L32, gx32, gs32, hs32 = ref.loss_grad(c["xs"][i], c["th"], g["sh"][i], c["bs"][i], LAM, EPS, "f32",
                                                  curvature=True)
            print("slice", i, "loss", g["loss"][i], L, "gx rel", rel(g["gx"][i][m], gx[m]),
                  rel(g["gx"][i], gx32), "gs rel", rel(g["gs"][i], gs), flush=True)
            assert np.allclose(g["loss"][i], L, rtol=1e-5), (g["loss"][i], L)
            assert rel(g["gx"][i][m], gx[m]) < 1e-4
            assert rel(g["gx"][i], gx32) < 1e-4
            assert rel(g["gs"][i], gs) < 1e-4
            assert rel(g["hs"][i], hs32) < 1e-5
    finally:
        ref.set_threads(1)


def test_gradients_match_central_differences(cfg3, cfg3_gpu):
    # FP64 central differences of the oracle loss, h = 1e-4, at 32 random
    # image coordinates and 32 random shift coordinates of one slice
    c, g = cfg3, cfg3_gpu
    i = 3
    N, T, th = c["N"], c["T"], c["th"]
    x = c["xs"][i].astype(np.float64)
    sh = g["sh"][i]
    b = c["bs"][i]
    rng = np.random.default_rng(77)
    h = 1e-4

    def L(xx, ss):
        with ref.fp32_angles():
            return ref.loss_grad(xx, th, ss, b, LAM, EPS, "f64", grad=False)[0][0]

    ref.set_threads(8)
    try:
        say("central differences: image")
        fd, got = [], []
        for yy, xx in zip(rng.integers(8, N - 8, 32), rng.integers(8, N - 8, 32)):
            xp = x.copy(); xp[yy, xx] += h
            xm = x.copy(); xm[yy, xx] -= h
            fd.append((L(xp, sh) - L(xm, sh)) / (2 * h))
            got.append(g["gx"][i][yy, xx])
        fd, got = np.array(fd), np.array(got)
        assert np.linalg.norm(got - fd) <= 1e-3 * np.linalg.norm(fd), (got, fd)
        say("central differences: shifts")
        fd, got = [], []
        for t in rng.choice(T, 32, replace=False):
            sp = sh.copy(); sp[t] = np.float32(sp[t] + np.float32(h))
            sm = sh.copy(); sm[t] = np.float32(sm[t] - np.float32(h))
            if np.floor(-sp[t]) != np.floor(-sm[t]):
                continue  # a kink of the linear interpolation lies in between
            fd.append((L(x, sp) - L(x, sm)) / (float(sp[t]) - float(sm[t])))
            got.append(g["gs"][i][t])
        fd, got = np.array(fd), np.array(got)
        assert len(fd) >= 28
        assert np.linalg.norm(got - fd) <= 1e-3 * np.linalg.norm(fd), (got, fd)
    finally:
        ref.set_threads(1)


def test_loss_grad_dev_matches_host(ctx):
    N, T = 65, 48
    th = synth.angles(T)
    rng = np.random.default_rng(4)
    fov = fov_mask(N, margin=4)
    x = synth.phantom(N, 6, 4) * fov
    sh = rng.normal(0, 1.5, T).astype(np.float32)
    f = ref.shift(ref.radon(x, th, "f32"), sh, "f32")
    hb = HipBuffers()
    try:
        with Batch(ctx, [(N, T, 4, th)]) as b:
            b.set_sino(0, f)
            x1 = (x + np.float32(0.01) * fov).astype(np.float32)
            loss, gx, gs = b.loss_grad([x1], [sh], LAM, EPS)
            import ctypes
            d_x, d_s = hb.upload(x1), hb.upload(sh)
            d_gx, d_gs = hb.alloc(N * N * 4), hb.alloc(T * 4)
            vp = ctypes.c_void_p
            lib = b._lib
            l2 = np.zeros(3, np.float64)
            rc = lib.acu_batch_loss_grad_dev(
                b.handle, (vp * 1)(d_x), (vp * 1)(d_s), LAM, EPS, (vp * 1)(d_gx), (vp * 1)(d_gs), None,
                l2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            assert rc == 0
            assert np.array_equal(hb.download(d_gx, (N, N)), gx[0])
            assert np.array_equal(hb.download(d_gs, (T,)), gs[0])
            assert np.array_equal(l2, loss[0])
            # no shifts, no gradients: the loss alone
            l3, gx3, gs3 = b.loss_grad([x1], None, LAM, EPS, grad_x=False, grad_s=False)
            with ref.fp32_angles():
                L, _, _ = ref.loss_grad(x1, th, None, f, LAM, EPS, "f64", grad=False)
            assert gx3 is None and gs3 is None
            assert np.allclose(l3[0], L, rtol=1e-5)
    finally:
        hb.close()


def test_register_shifts_recovers_centre_of_rotation(ctx):
    N, T, S = 513, 720, 4
    th = synth.angles(T)
    rng = np.random.default_rng(33)
    xs, truth = [], []
    with Batch(ctx, [(N, T, 16, th)] * S) as b:
        for i in range(S):
            x = synth.phantom(N, 20, 500 + i)
            s = rng.normal(0, 1.5, T).astype(np.float32)
            b.set_sino(i, ref.shift(ref.radon(x, th, "f32"), s, "f32"))
            xs.append(x); truth.append(s)
        say("register", S, "slices")
        fit, evals = b.register_shifts(xs, iters=12)
    for s, t in zip(fit, truth):
        err = np.abs(s - t)
        print("registration error: max", err.max(), "rms", np.sqrt(np.mean(err ** 2)), "evals", evals, flush=True)
        assert err.max() < 2e-3, err.max()
    assert evals <= 12
