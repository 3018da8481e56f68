"""ctypes front end of oracle/_ref/libacumen_ref.so: the reference's own
Fortran module ``Tomography`` (tomography/tv_hpc.F90:1-390), compiled
unmodified from the reference tree by ``make -C oracle ref``, behind the C
entry points of oracle/ref_shim.F90.

TEST INFRASTRUCTURE ONLY: the test-suite checks the C oracle against it.
The library exists where the reference tree was present at build time;
``available()`` says whether it is here.  The module is global state (one
set-up at a time, as in the reference), so calls are serialised by a lock.

Arrays use the layout of oracle/ref.py: an image (x, y) is a C-ordered
(N, N) array indexed [y, x]; a sinogram (detector, angle) is (T, N).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(_HERE, "_ref", "libacumen_ref.so")
REF_TREE = "/root/reference/tomography"
_lib = None
_lock = threading.Lock()
_F = ctypes.POINTER(ctypes.c_float)


def build() -> bool:
    """Compile the reference module where the reference tree is present.
    Returns whether the library exists afterwards."""
    if os.path.exists(os.path.join(REF_TREE, "tv_hpc.F90")):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)
    return os.path.exists(SO)


def available() -> bool:
    return os.path.exists(SO)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{SO} is not built (make -C oracle ref where the reference tree is present)")
        _lib = ctypes.CDLL(SO)
        c_int, c_float = ctypes.c_int, ctypes.c_float
        _lib.ref_init.argtypes = [c_int, c_int, c_int, _F]
        _lib.ref_tables.argtypes = [_F, _F]
        _lib.ref_radon.argtypes = [_F, _F]
        _lib.ref_iradon.argtypes = [_F, _F]
        _lib.ref_filtered.argtypes = [_F]
        _lib.ref_tvopt.argtypes = [_F, c_float, c_float, c_int, c_int, c_int, _F, _F]
        _lib.ref_get_slice.argtypes = [_F, _F]
        _lib.ref_fin.argtypes = []
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_F)


def _f32(a, shape):
    out = np.ascontiguousarray(a, dtype=np.float32)
    if out.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {out.shape}")
    return out


class Module:
    """One set-up of the module (int_radon with nimage, ntheta, wpu and the
    angles), used inside ``with``: the lock is held until exit."""

    def __init__(self, N: int, theta, wpu: int):
        self.th = np.ascontiguousarray(theta, dtype=np.float32)
        self.N, self.T, self.wpu = int(N), int(self.th.size), int(wpu)

    def __enter__(self):
        _lock.acquire()
        try:
            lib().ref_init(self.N, self.T, self.wpu, _p(self.th))
        except BaseException:
            _lock.release()
            raise
        return self

    def __exit__(self, *exc):
        try:
            lib().ref_fin()
        finally:
            _lock.release()

    def tables(self):
        h = np.empty(2 * self.N + 1, np.float32)
        cm = np.empty((self.N, self.N), np.float32)
        lib().ref_tables(_p(h), _p(cm))
        return h, cm

    def radon(self, img) -> np.ndarray:
        x = _f32(img, (self.N, self.N))
        out = np.empty((self.T, self.N), np.float32)
        lib().ref_radon(_p(x), _p(out))
        return out

    def iradon(self, sino) -> np.ndarray:
        s = _f32(sino, (self.T, self.N))
        out = np.empty((self.N, self.N), np.float32)
        lib().ref_iradon(_p(s), _p(out))
        return out

    def filtered(self) -> np.ndarray:
        """The ramp-filtered sinogram the last iradon left in sinogramCS."""
        out = np.empty((self.T, self.N), np.float32)
        lib().ref_filtered(_p(out))
        return out

    def get_slice(self, sino) -> np.ndarray:
        s = _f32(sino, (self.T, self.N))
        out = np.empty((self.N, self.N), np.float32)
        lib().ref_get_slice(_p(s), _p(out))
        return out

    def tvopt(self, sino, mu, lam, nouter, ninner, ncg):
        """(img_tv, img_fbp) of tvopt."""
        s = _f32(sino, (self.T, self.N))
        tv = np.empty((self.N, self.N), np.float32)
        fbp = np.empty((self.N, self.N), np.float32)
        lib().ref_tvopt(_p(s), float(mu), float(lam), int(nouter), int(ninner), int(ncg), _p(tv), _p(fbp))
        return tv, fbp
