"""acu_prep_mirror_error_noalloc, the entry point that allocates nothing,
through the C ABI: its result equals acu_prep_mirror_error_dev's bit for bit
on two calls in a row (the second on a smaller input, so stale partials of
the first call would show), for odd and even projection widths; a width
whose partials exceed the static scratch is refused before any GPU work."""
import ctypes

import numpy as np
import pytest

from acumen_amd import _lib, synth
from helpers import HipBuffers

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def projections(n, seed):
    rng = np.random.default_rng(seed)
    p0 = synth.phantom(n, 12, seed) + rng.normal(0, 0.01, (n, n)).astype(np.float32)
    p180 = np.ascontiguousarray(p0[:, ::-1]) + rng.normal(0, 0.01, (n, n)).astype(np.float32)
    return p0.astype(np.float32), p180.astype(np.float32)


def test_noalloc_equals_allocating_entry_point(ctx):
    lib = _lib.load()
    lib.acu_prep_mirror_error_noalloc.restype = ctypes.c_int
    lib.acu_prep_mirror_error_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                                  ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    hb = HipBuffers()
    try:
        err_dev = hb.alloc(8)
        # two calls in a row, the second with the smaller input
        for n, th, sh in ((1536, 0.012, 0.0031), (255, -0.004, -0.0102), (96, 0.0, 0.0)):
            add_nx = 1 if n % 2 == 0 else 0
            say("mirror error noalloc", n, add_nx, th, sh)
            p0, p180 = projections(n, n)
            d0, d1 = hb.upload(p0), hb.upload(p180)
            want = ctypes.c_double()
            assert lib.acu_prep_mirror_error_dev(ctx.handle, n, add_nx, th, sh, d0, d1, ctypes.byref(want)) == 0
            assert lib.acu_prep_mirror_error_noalloc(ctx.handle, n, add_nx, th, sh, d0, d1, err_dev) == 0, \
                lib.acu_last_error()
            ctx.synchronize()
            got = hb.download(err_dev, (1,), np.float64)[0]
            assert np.float64(got).view(np.uint64) == np.float64(want.value).view(np.uint64), (n, got, want.value)
            assert got > 0.0
        # partials of n = 4096 (+1 padding column) exceed the scratch: refused, nothing launched
        rc = lib.acu_prep_mirror_error_noalloc(ctx.handle, 4096, 1, 0.0, 0.0, d0, d1, err_dev)
        assert rc == _lib.ACU_ERR_ARG
        assert b"exceed" in lib.acu_last_error()
    finally:
        hb.close()
