"""Volume reconstruction streamed through one GPU (acu_volume_reconstruct).

The reference's drivers hand slices to MPI ranks one at a time
(tomography/tv_hpc.F90:583-631, fbp_hpc.F90:340-403); here a volume of
same-shape slices streams through the GPU in batches on two lanes whose
host staging and PCIe transfers overlap each other's kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

_F = ctypes.POINTER(ctypes.c_float)


def reconstruct(ctx, sinos, theta, wpu: int, tv=None, fbp: bool = True, batch: int = 8,
                tune: bool = True):
    """FBP (and, with ``tv = dict(mu, lam, nouter, ninner, ncg)``, split-Bregman
    TV) of every slice of ``sinos`` (S, T, N) float32.

    Returns ``(fbp_images or None, tv_images or None, stats)`` with images
    (S, N, N) and ``stats`` = wall ``ms``, ``slices_per_s`` over the call,
    ``setup_ms`` (lanes, pinned buffers, tuning) and ``stream_slices_per_s``
    after set-up; PCIe and host staging included."""
    s = np.ascontiguousarray(sinos, np.float32)
    if s.ndim != 3:
        raise ValueError("sinos must be (slices, angles, detectors)")
    S, T, N = s.shape
    th = np.ascontiguousarray(theta, np.float32)
    if th.shape != (T,):
        raise ValueError(f"theta must have shape ({T},)")
    out_fbp = np.empty((S, N, N), np.float32) if fbp else None
    out_tv = np.empty((S, N, N), np.float32) if tv is not None else None
    p = dict(mu=0.0, lam=0.0, nouter=0, ninner=1, ncg=5)
    if tv is not None:
        p.update(tv)
    stats = np.zeros(4, np.float64)
    lib = _lib.load()
    check(lib.acu_volume_reconstruct(
        ctx.handle, S, N, T, int(wpu), th.ctypes.data_as(_F), s.ctypes.data_as(_F),
        None if out_fbp is None else out_fbp.ctypes.data_as(_F),
        None if out_tv is None else out_tv.ctypes.data_as(_F),
        float(p["mu"]), float(p["lam"]), int(p["nouter"]), int(p["ninner"]), int(p["ncg"]),
        int(batch), 1 if tune else 0, stats.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out_fbp, out_tv, {"ms": float(stats[0]), "slices_per_s": float(stats[1]),
                             "setup_ms": float(stats[2]), "stream_slices_per_s": float(stats[3])}
