"""Projection -> sinogram preprocessing on the GPU, the restatement of
tomography/projections_to_sinogram_hpc.F90: centre-of-rotation / tilt
search on the 0 and 180 degree projections, 3x3x3 median filter, clamped
bilinear correction and transposition into the sinogram layout.

Arrays are numpy float32 in the reference's layout as C-ordered arrays: a
projection img(x, y) is shape (n, n) indexed [y, x]; the sinogram volume is
(n slices, nfiles angles, Nw detectors), Nw = n + 1 for even n.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

_F = ctypes.POINTER(ctypes.c_float)


def _f32(a, shape):
    out = np.ascontiguousarray(a, np.float32)
    if out.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {out.shape}")
    return out


def angle_window_projection(projections, theta, target: float):
    """The projection at `target` degrees (0 or 180): the first projection
    whose angle is closest to it, or the linear interpolation with its
    neighbour on the far side when the closest angle is not exact
    (projections_to_sinogram_hpc.F90:140-228)."""
    th = np.asarray(theta, np.float32)
    j = int(np.argmin(np.abs(th - np.float32(target))))
    if th[j] == np.float32(target):
        return np.asarray(projections[j], np.float32)
    if th[j] <= target:
        if j + 1 >= th.size:
            return np.asarray(projections[j], np.float32)
        a, b, ta, tb = projections[j], projections[j + 1], th[j], th[j + 1]
    else:
        if j == 0:
            return np.asarray(projections[j], np.float32)
        a, b, ta, tb = projections[j - 1], projections[j], th[j - 1], th[j]
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return ((np.float32(target) - ta) * (b - a) / (tb - ta) + a).astype(np.float32)


def find_center(ctx, p0, p180):
    """(theta, shift, evaluations) of the centre-of-rotation / tilt search."""
    n = np.asarray(p0).shape[0]
    a, b = _f32(p0, (n, n)), _f32(p180, (n, n))
    th, sh, ev = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
    check(_lib.load().acu_prep_find_center(ctx.handle, n, a.ctypes.data_as(_F), b.ctypes.data_as(_F),
                                           ctypes.byref(th), ctypes.byref(sh), ctypes.byref(ev)))
    return th.value, sh.value, ev.value


def sinograms(ctx, projections, theta_corr: float, shift: float):
    """Median-filtered, corrected sinogram volume (n, nfiles, Nw) of the
    projections (nfiles, n, n)."""
    p = np.ascontiguousarray(projections, np.float32)
    nfiles, n, _ = p.shape
    _f32(p[0], (n, n))
    nw = n + (1 if n % 2 == 0 else 0)
    out = np.empty((n, nfiles, nw), np.float32)
    check(_lib.load().acu_prep_sinograms(ctx.handle, n, nfiles, p.ctypes.data_as(_F), float(theta_corr),
                                         float(shift), out.ctypes.data_as(_F)))
    return out
