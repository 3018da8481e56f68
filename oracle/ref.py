"""ctypes front-end of the C oracle (liboracle.so) -- test infrastructure only.

Every function restates a part of Acumen's Fortran module ``Tomography``
(tomography/tv_hpc.F90); the C source cites the lines.  Arrays follow the
reference's Fortran layout: an image ``(x, y)`` and a sinogram
``(detector, angle)`` are stored first-index-fastest, which in numpy is a
C-ordered array of shape ``(ny, nx)`` / ``(T, N)``; that is what these
functions take and return.

``prec="f32"`` runs the reference's own FP32 arithmetic (one rounding per
Fortran operation, sequential FP32 sums); ``prec="f64"`` runs the same
algorithm in FP64 and is the accuracy yardstick.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
        for p in ("f32", "f64"):
            ptr = _F if p == "f32" else _D
            real = ctypes.c_float if p == "f32" else ctypes.c_double
            getattr(_lib, f"orc_filter_{p}").argtypes = [ctypes.c_int, ptr]
            getattr(_lib, f"orc_cmask_{p}").argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ptr]
            getattr(_lib, f"orc_ramp_{p}").argtypes = [ctypes.c_int, ctypes.c_int, ptr, ptr, ptr]
            getattr(_lib, f"orc_backproject_{p}").argtypes = [ctypes.c_int, ctypes.c_int, _F, ptr, ptr, ptr]
            getattr(_lib, f"orc_iradon_{p}").argtypes = [ctypes.c_int, ctypes.c_int, _F, ptr, ptr, ptr, ptr]
            getattr(_lib, f"orc_radon_{p}").argtypes = [ctypes.c_int, ctypes.c_int, _F, ptr, ptr]
            getattr(_lib, f"orc_tvopt_{p}").argtypes = [
                ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, _F, real, real,
                ctypes.c_int, ctypes.c_int, ctypes.c_int, ptr, ptr, ptr]
            getattr(_lib, f"orc_radon_adjoint_{p}").argtypes = [ctypes.c_int, ctypes.c_int, _F, ptr, ptr]
            getattr(_lib, f"orc_shift_{p}").argtypes = [ctypes.c_int, ctypes.c_int, _F, ptr, ptr]
            getattr(_lib, f"orc_shift_adjoint_{p}").argtypes = [ctypes.c_int, ctypes.c_int, _F, ptr, ptr]
            getattr(_lib, f"orc_loss_grad_{p}").argtypes = [
                ctypes.c_int, ctypes.c_int, _F, ptr, _F, ptr, ctypes.c_double, ctypes.c_double,
                _D, ptr, ptr, ptr]
        _lib.orc_prep_median27.argtypes = [ctypes.c_int, _F, _F, _F, _F]
        _lib.orc_prep_correct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _F, _F]
        _lib.orc_prep_mirror_error.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _F, _F]
        _lib.orc_prep_mirror_error.restype = ctypes.c_double
        _lib.orc_prep_centroid_shift.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F]
        _lib.orc_prep_centroid_shift.restype = ctypes.c_float
        _lib.orc_ingest_mean.argtypes = [ctypes.c_int, ctypes.c_longlong, _F, _F]
        _lib.orc_ingest_normalize.argtypes = [ctypes.c_int, ctypes.c_longlong, _F, _F, _F]
        _lib.orc_prep_find_center.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F, _F, _F]
        _lib.orc_tvopt_f32d.argtypes = [
            ctypes.c_int, ctypes.c_int, _F, ctypes.c_int, _F, ctypes.c_float, ctypes.c_float,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _F, _F, _F]
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
        _lib.orc_set_fp32_angles.argtypes = [ctypes.c_int]
        _lib.orc_angle_tables_f32.argtypes = [ctypes.c_int, _F, ctypes.c_int, _F, _F]
        _lib.orc_pass_batch_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _F,
                                            ctypes.c_int, _F, _F, _F]
    return _lib


def _dt(prec):
    if prec not in ("f32", "f64"):
        raise ValueError(prec)
    return np.float32 if prec == "f32" else np.float64


def _p(a):
    if a.dtype == np.float32:
        return a.ctypes.data_as(_F)
    return a.ctypes.data_as(_D)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


class fp32_angles:
    """Context manager: FP64 computations use the reference's FP32 per-angle
    sin/cos values, isolating arithmetic error from angle rounding."""

    def __enter__(self):
        lib().orc_set_fp32_angles(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_fp32_angles(0)


def filter_taps(N: int, prec="f32") -> np.ndarray:
    """tv_hpc.F90:53-56 -- 2N+1 ramp-filter taps."""
    h = np.empty(2 * N + 1, _dt(prec))
    getattr(lib(), f"orc_filter_{prec}")(N, _p(h))
    return h


def cmask(N: int, T: int, wpu: int, prec="f32") -> np.ndarray:
    """tv_hpc.F90:58-74 -- (N,N) smooth disk mask times pi/T."""
    m = np.empty((N, N), _dt(prec))
    getattr(lib(), f"orc_cmask_{prec}")(N, T, wpu, _p(m))
    return m


def _theta(theta):
    return np.ascontiguousarray(theta, dtype=np.float32)


def ramp(sino: np.ndarray, prec="f32") -> np.ndarray:
    """tv_hpc.F90:93-105 -- sinogram (T,N) filtered along the detector axis."""
    T, N = sino.shape
    s = np.ascontiguousarray(sino, _dt(prec))
    out = np.empty_like(s)
    h = filter_taps(N, prec)
    getattr(lib(), f"orc_ramp_{prec}")(N, T, _p(h), _p(s), _p(out))
    return out


def backproject(sino_f: np.ndarray, theta, wpu: int, prec="f32") -> np.ndarray:
    """tv_hpc.F90:107-123 -- backprojection of an already filtered sinogram."""
    T, N = sino_f.shape
    th = _theta(theta)
    s = np.ascontiguousarray(sino_f, _dt(prec))
    cm = cmask(N, T, wpu, prec)
    img = np.empty((N, N), _dt(prec))
    getattr(lib(), f"orc_backproject_{prec}")(N, T, th.ctypes.data_as(_F), _p(cm), _p(s), _p(img))
    return img


def iradon(sino: np.ndarray, theta, wpu: int, prec="f32") -> np.ndarray:
    """tv_hpc.F90:84-126 -- ramp filter then backprojection (the FBP operator)."""
    T, N = sino.shape
    th = _theta(theta)
    s = np.ascontiguousarray(sino, _dt(prec))
    h = filter_taps(N, prec)
    cm = cmask(N, T, wpu, prec)
    img = np.empty((N, N), _dt(prec))
    getattr(lib(), f"orc_iradon_{prec}")(N, T, th.ctypes.data_as(_F), _p(h), _p(cm), _p(s), _p(img))
    return img


def radon(img: np.ndarray, theta, prec="f32") -> np.ndarray:
    """tv_hpc.F90:129-159 -- (N,N) image to (T,N) sinogram."""
    N = img.shape[0]
    assert img.shape == (N, N)
    th = _theta(theta)
    T = th.size
    x = np.ascontiguousarray(img, _dt(prec))
    s = np.empty((T, N), _dt(prec))
    getattr(lib(), f"orc_radon_{prec}")(N, T, th.ctypes.data_as(_F), _p(x), _p(s))
    return s


def tvopt(f: np.ndarray, theta, wpu: int, mu: float, lam: float, nouter: int,
          ninner: int = 1, ncg: int = 5, prec="f32"):
    """tv_hpc.F90:162-363 -- split-Bregman TV.  Returns (img_tv, img_fbp, delta).

    ``prec="f32d"`` keeps FP32 operators and arrays but forms the CG inner
    products, alpha, beta and Err_diff in FP64."""
    T, N = f.shape
    th = _theta(theta)
    fi = np.ascontiguousarray(f, np.float32)
    dt = _dt("f32" if prec == "f32d" else prec)
    tv = np.empty((N, N), dt)
    fbp = np.empty((N, N), dt)
    delta = np.zeros(max(nouter, 1), dt)
    getattr(lib(), f"orc_tvopt_{prec}")(N, T, th.ctypes.data_as(_F), wpu, fi.ctypes.data_as(_F),
                                        mu, lam, nouter, ninner, ncg, _p(tv), _p(fbp), _p(delta))
    return tv, fbp, delta[:nouter]


def angle_tables(theta, bp: bool):
    """Per-angle (sin, cos) of the reference's ctheta in FP32."""
    th = _theta(theta)
    sn = np.empty(th.size, np.float32)
    cs = np.empty(th.size, np.float32)
    lib().orc_angle_tables_f32(th.size, th.ctypes.data_as(_F), 1 if bp else 0,
                               sn.ctypes.data_as(_F), cs.ctypes.data_as(_F))
    return sn, cs


def pass_batch_f32(imgs: np.ndarray, theta, wpu: int, threads: int):
    """One radon + one iradon per slice, one slice per thread (CPU baseline)."""
    S, N, _ = imgs.shape
    th = _theta(theta)
    T = th.size
    x = np.ascontiguousarray(imgs, np.float32)
    sinos = np.empty((S, T, N), np.float32)
    out = np.empty_like(x)
    set_threads(threads)
    try:
        lib().orc_pass_batch_f32(S, N, T, th.ctypes.data_as(_F), wpu, x.ctypes.data_as(_F),
                                 sinos.ctypes.data_as(_F), out.ctypes.data_as(_F))
    finally:
        set_threads(1)
    return sinos, out


# ---- differentiation (no counterpart in the reference; built on its radon) ----

def radon_adjoint(sino: np.ndarray, theta, prec="f32") -> np.ndarray:
    """Exact transpose of :func:`radon`: (T,N) sinogram to (N,N) image."""
    T, N = sino.shape
    th = _theta(theta)
    assert th.size == T
    s = np.ascontiguousarray(sino, _dt(prec))
    img = np.empty((N, N), _dt(prec))
    getattr(lib(), f"orc_radon_adjoint_{prec}")(N, T, th.ctypes.data_as(_F), _p(s), _p(img))
    return img


def _shifts(shift, T):
    sh = np.ascontiguousarray(shift, np.float32)
    assert sh.shape == (T,)
    return sh


def shift(sino: np.ndarray, shifts, prec="f32") -> np.ndarray:
    """Each projection row sampled at detector position d - s_t (linear
    interpolation, zero outside the detector)."""
    T, N = sino.shape
    sh = _shifts(shifts, T)
    s = np.ascontiguousarray(sino, _dt(prec))
    out = np.empty_like(s)
    getattr(lib(), f"orc_shift_{prec}")(N, T, sh.ctypes.data_as(_F), _p(s), _p(out))
    return out


def shift_adjoint(sino: np.ndarray, shifts, prec="f32") -> np.ndarray:
    """Transpose of :func:`shift`."""
    T, N = sino.shape
    sh = _shifts(shifts, T)
    s = np.ascontiguousarray(sino, _dt(prec))
    out = np.empty_like(s)
    getattr(lib(), f"orc_shift_adjoint_{prec}")(N, T, sh.ctypes.data_as(_F), _p(s), _p(out))
    return out


def loss_grad(x: np.ndarray, theta, shifts, b: np.ndarray, lam: float, eps: float,
              prec="f64", grad=True, curvature=False):
    """L = 1/2 |shift_s(radon(x)) - b|^2 + lam * TV_eps(x) and its gradients.

    Returns (loss[3] = total, data, TV_eps; dL/dx (N,N); dL/ds (T,)), the
    gradients None when ``grad`` is False, plus the Gauss-Newton curvature
    of the data term in each shift (T,) as a fourth item when
    ``curvature`` is True.  ``shifts`` may be None (zero)."""
    N = x.shape[0]
    th = _theta(theta)
    T = th.size
    dt = _dt(prec)
    xi = np.ascontiguousarray(x, dt)
    bi = np.ascontiguousarray(b, dt)
    assert bi.shape == (T, N)
    loss = np.zeros(3, np.float64)
    gx = np.empty((N, N), dt) if grad else None
    gs = np.empty(T, dt) if grad else None
    hs = np.empty(T, dt) if curvature else None
    sh = None if shifts is None else _shifts(shifts, T)
    getattr(lib(), f"orc_loss_grad_{prec}")(
        N, T, th.ctypes.data_as(_F), _p(xi), None if sh is None else sh.ctypes.data_as(_F),
        _p(bi), float(lam), float(eps), loss.ctypes.data_as(_D),
        None if gx is None else _p(gx), None if gs is None else _p(gs),
        None if hs is None else _p(hs))
    if curvature:
        return loss, gx, gs, hs
    return loss, gx, gs


# ---- projection -> sinogram preprocessing (projections_to_sinogram_hpc.F90) ----

def _img(a):
    return np.ascontiguousarray(a, np.float32)


def prep_median27(l, c, r) -> np.ndarray:
    """3x3x3 median of projection c with neighbours l, r (:420-447)."""
    n = c.shape[0]
    L, C, R = _img(l), _img(c), _img(r)
    out = np.empty_like(C)
    lib().orc_prep_median27(n, L.ctypes.data_as(_F), C.ctypes.data_as(_F), R.ctypes.data_as(_F),
                            out.ctypes.data_as(_F))
    return out


def prep_correct(img, theta: float, shift: float, add_nx=None) -> np.ndarray:
    """Clamped bilinear rotation / shift correction (:449-464): (n, Nw)."""
    n = img.shape[0]
    a = (1 if n % 2 == 0 else 0) if add_nx is None else add_nx
    x = _img(img)
    out = np.empty((n, n + a), np.float32)
    lib().orc_prep_correct(n, a, theta, shift, x.ctypes.data_as(_F), out.ctypes.data_as(_F))
    return out


def prep_mirror_error(p0, p180, theta: float, shift: float, add_nx=None) -> float:
    n = p0.shape[0]
    a = (1 if n % 2 == 0 else 0) if add_nx is None else add_nx
    return lib().orc_prep_mirror_error(n, a, theta, shift, _img(p0).ctypes.data_as(_F),
                                       _img(p180).ctypes.data_as(_F))


def prep_centroid_shift(p0, p180, add_nx=None) -> float:
    n = p0.shape[0]
    a = (1 if n % 2 == 0 else 0) if add_nx is None else add_nx
    return lib().orc_prep_centroid_shift(n, a, _img(p0).ctypes.data_as(_F), _img(p180).ctypes.data_as(_F))


def prep_find_center(p0, p180, add_nx=None):
    """(theta, shift) of the reference's pattern search (:250-364)."""
    n = p0.shape[0]
    a = (1 if n % 2 == 0 else 0) if add_nx is None else add_nx
    th, sh = ctypes.c_float(), ctypes.c_float()
    lib().orc_prep_find_center(n, a, _img(p0).ctypes.data_as(_F), _img(p180).ctypes.data_as(_F),
                               ctypes.byref(th), ctypes.byref(sh))
    return th.value, sh.value


# ---- FITS ingest normalisation (fits_to_hpc.F90) ----


def ingest_mean(imgs) -> np.ndarray:
    """Open-beam average of fits_to_hpc.F90:203-215 over imgs (count, ...)."""
    a = np.ascontiguousarray(imgs, np.float32)
    out = np.empty(a.shape[1:], np.float32)
    lib().orc_ingest_mean(a.shape[0], out.size, a.ctypes.data_as(_F), out.ctypes.data_as(_F))
    return out


def ingest_normalize(imgs, bg) -> np.ndarray:
    """-log(image / max(bg, 1)) of fits_to_hpc.F90:243 for imgs (count, ...)."""
    a = np.ascontiguousarray(imgs, np.float32)
    b = np.ascontiguousarray(bg, np.float32)
    if b.shape != a.shape[1:]:
        raise ValueError("bg must have the shape of one image")
    out = np.empty_like(a)
    lib().orc_ingest_normalize(a.shape[0], b.size, a.ctypes.data_as(_F), b.ctypes.data_as(_F),
                               out.ctypes.data_as(_F))
    return out
