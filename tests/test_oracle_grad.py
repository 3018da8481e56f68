"""Pins the oracle's differentiation functions (no counterpart in the
reference; built on its projector, tomography/tv_hpc.F90:129-159): the
adjoint by the dot-product test and a known answer, the detector shift by
integer-shift known answers and its transpose, TV_eps by closed forms, and
both gradients by FP64 central differences."""
import numpy as np
import pytest

from oracle import ref


def angles(T):
    return (np.arange(T) * np.float32(180.0 / T)).astype(np.float32)


@pytest.mark.parametrize("N,T", [(33, 24), (20, 7)])
def test_adjoint_dot_product(N, T):
    rng = np.random.default_rng(N + T)
    th = angles(T)
    x = rng.normal(size=(N, N))
    y = rng.normal(size=(T, N))
    a = np.vdot(ref.radon(x, th, "f64"), y)
    b = np.vdot(x, ref.radon_adjoint(y, th, "f64"))
    assert abs(a - b) <= 1e-12 * abs(a)
    a32 = np.vdot(ref.radon(x, th, "f32").astype(np.float64), y)
    b32 = np.vdot(x, ref.radon_adjoint(y.astype(np.float32), th, "f32").astype(np.float64))
    assert abs(a32 - b32) <= 1e-5 * abs(a)


def test_adjoint_known_answer_at_90_degrees():
    # theta = 90: c = 1, s = 0, samples sit on pixels, radon = row sums over
    # x = 2..N-1 for rows y = 2..N-1, divided by N; its transpose spreads
    # sino(y)/N over the same pixels
    N = 17
    y = np.random.default_rng(3).normal(size=(1, N))
    img = ref.radon_adjoint(y, np.array([90.0], np.float32), "f64")
    want = np.zeros((N, N))
    want[1:N - 1, 1:N - 1] = (y[0, 1:N - 1] / N)[:, None]
    assert np.allclose(img, want, rtol=0, atol=1e-15)


def test_adjoint_threads_invariant():
    N, T = 33, 24
    th = angles(T)
    y = np.random.default_rng(5).normal(size=(T, N))
    a = ref.radon_adjoint(y, th, "f64")
    ref.set_threads(4)
    try:
        b = ref.radon_adjoint(y, th, "f64")
    finally:
        ref.set_threads(1)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-15)


def test_shift_integer_known_answers():
    N, T = 9, 3
    y = np.arange(1, T * N + 1, dtype=np.float64).reshape(T, N)
    out = ref.shift(y, np.array([2.0, -1.0, 0.0], np.float32), "f64")
    assert np.array_equal(out[0], np.r_[0, 0, y[0, :-2]])      # out(d) = y(d - 2)
    assert np.array_equal(out[1], np.r_[y[1, 1:], 0])          # out(d) = y(d + 1)
    assert np.array_equal(out[2], y[2])
    half = ref.shift(y, np.array([0.5, -0.5, 0.25], np.float32), "f64")
    assert np.allclose(half[0, 1:], 0.5 * (y[0, :-1] + y[0, 1:]))
    assert np.allclose(half[1, :-1], 0.5 * (y[1, :-1] + y[1, 1:]))
    assert np.allclose(half[2, 1:], 0.25 * y[2, :-1] + 0.75 * y[2, 1:])


def test_shift_adjoint_dot_product():
    rng = np.random.default_rng(11)
    T, N = 16, 41
    sh = rng.normal(0, 3.0, T).astype(np.float32)
    y = rng.normal(size=(T, N))
    z = rng.normal(size=(T, N))
    a = np.vdot(ref.shift(y, sh, "f64"), z)
    b = np.vdot(y, ref.shift_adjoint(z, sh, "f64"))
    assert abs(a - b) <= 1e-13 * abs(a)


def test_tv_closed_forms():
    N, T = 12, 5
    th = angles(T)
    b = np.zeros((T, N))
    eps = 1e-3
    # constant image: every difference is zero, TV = N^2 eps, no TV gradient
    x = np.full((N, N), 2.5)
    loss, gx, _ = ref.loss_grad(x, th, None, b, 0.0, eps)
    assert np.isclose(loss[2], N * N * eps, rtol=1e-12)
    loss1, gx1, _ = ref.loss_grad(x, th, None, b, 0.7, eps)
    assert np.allclose(gx1, gx, rtol=0, atol=1e-12)
    # unit step in x between columns 4 and 5: the periodic difference is
    # nonzero on two columns (the step and the wrap-around)
    x = np.zeros((N, N))
    x[:, 5:] = 1.0
    loss, _, _ = ref.loss_grad(x, th, None, b, 1.0, eps, grad=False)
    want = 2 * N * np.sqrt(1 + eps * eps) + (N * N - 2 * N) * eps
    assert np.isclose(loss[2], want, rtol=1e-12)
    # data term: zero image, loss = |b|^2 / 2
    b = np.random.default_rng(1).normal(size=(T, N))
    loss, _, gs = ref.loss_grad(np.zeros((N, N)), th, np.zeros(T, np.float32), b, 0.0, eps)
    assert np.isclose(loss[1], 0.5 * np.sum(b * b), rtol=1e-12)
    assert np.allclose(gs, 0.0)


def test_gradients_match_central_differences():
    N, T = 33, 24
    rng = np.random.default_rng(7)
    th = angles(T)
    x = rng.normal(size=(N, N))
    sh = rng.normal(0, 1.5, T).astype(np.float32)
    b = rng.normal(size=(T, N)) * 0.05
    lam, eps = 0.1, 1e-3
    _, gx, gs = ref.loss_grad(x, th, sh, b, lam, eps)
    h = 1e-5
    for i, j in rng.integers(0, N, size=(8, 2)):
        xp = x.copy(); xp[i, j] += h
        xm = x.copy(); xm[i, j] -= h
        fd = (ref.loss_grad(xp, th, sh, b, lam, eps, grad=False)[0][0]
              - ref.loss_grad(xm, th, sh, b, lam, eps, grad=False)[0][0]) / (2 * h)
        assert abs(gx[i, j] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, j, gx[i, j], fd)
    for t in rng.integers(0, T, size=8):
        # the loss is piecewise quadratic in s (kinks where -s crosses an
        # integer); step within the piece
        hs = np.float32(1e-3)
        sp = sh.copy(); sp[t] += hs
        sm = sh.copy(); sm[t] -= hs
        if np.floor(-sp[t]) != np.floor(-sm[t]):
            continue
        fd = (ref.loss_grad(x, th, sp, b, lam, eps, grad=False)[0][0]
              - ref.loss_grad(x, th, sm, b, lam, eps, grad=False)[0][0]) / (float(sp[t]) - float(sm[t]))
        assert abs(gs[t] - fd) <= 1e-6 * max(1.0, abs(fd)), (t, gs[t], fd)


def test_shift_curvature_is_the_derivative_of_the_shift_gradient():
    # between kinks of the interpolation the data term is quadratic in s_t
    # and its second derivative is hs_t
    N, T = 33, 24
    rng = np.random.default_rng(8)
    th = angles(T)
    x = rng.normal(size=(N, N))
    sh = rng.normal(0, 1.5, T).astype(np.float32)
    b = rng.normal(size=(T, N)) * 0.05
    _, _, gs, hs = ref.loss_grad(x, th, sh, b, 0.0, 1.0, curvature=True)
    for t in range(T):
        hstep = np.float32(1e-3)
        sp = sh.copy(); sp[t] += hstep
        sm = sh.copy(); sm[t] -= hstep
        if np.floor(-sp[t]) != np.floor(-sm[t]):
            continue
        gp = ref.loss_grad(x, th, sp, b, 0.0, 1.0)[2][t]
        gm = ref.loss_grad(x, th, sm, b, 0.0, 1.0)[2][t]
        fd = (gp - gm) / (float(sp[t]) - float(sm[t]))
        assert abs(hs[t] - fd) <= 1e-7 * max(1.0, abs(fd)), (t, hs[t], fd)


def test_gauss_newton_registration_recovers_shifts():
    # host logic of the registration driven by the oracle's gradient and
    # curvature
    from acumen_amd import synth
    from acumen_amd.registration import gauss_newton_shifts
    N, T = 97, 60
    th = synth.angles(T)
    x = synth.phantom(N, 10, 5).astype(np.float64)
    s_true = np.random.default_rng(3).normal(0, 1.5, T).astype(np.float32)
    b = ref.shift(ref.radon(x, th, "f64"), s_true, "f64")

    def evaluate(sh):
        _, _, gs, hs = ref.loss_grad(x, th, sh[0], b, 0.0, 1.0, curvature=True)
        return [gs], [hs]

    s, evals = gauss_newton_shifts(evaluate, [np.zeros(T, np.float32)], iters=12)
    assert np.max(np.abs(s[0] - s_true)) < 1e-4, np.abs(s[0] - s_true).max()
    assert evals <= 12
