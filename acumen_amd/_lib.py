"""ctypes binding of libacumen_hip.so (declared in include/acumen.h).

The library is the product: there is no CPU fallback.  Loading fails with an
ImportError-like exception when the shared library has not been built.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libacumen_hip.so")

ACU_OK = 0
ACU_ERR_ARG = 1
ACU_ERR_HIP = 2
ACU_ERR_STATE = 3

ACU_IMG_TV = 0
ACU_IMG_FBP = 1
ACU_SINO_F = 2

ACU_TV_PLAIN = 0
ACU_TV_BATCHED = 1
ACU_TV_GRAPH = 2

ACU_TV_SCHED_REFERENCE = 0
ACU_TV_SCHED_REUSE = 1
ACU_TV_SCHED_LINEAR = 2

ACU_BP_POS_REFERENCE = 0
ACU_BP_POS_TILE = 1

KIND_RAMP, KIND_BP, KIND_CELLS, KIND_FP, KIND_ADJ, KIND_PREP, KIND_INGEST = 0, 1, 2, 3, 4, 5, 6
RAMP_FP64, RAMP_FP32, RAMP_BF16X3, RAMP_BF16X6, RAMP_FFT = 0, 1, 2, 3, 4

# Every symbol the header declares, with (restype, argtypes).
_c_int, _c_float, _vp = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "acu_last_error": (ctypes.c_char_p, []),
    "acu_ctx_create": (_c_int, [_c_int, ctypes.POINTER(_vp)]),
    "acu_ctx_destroy": (_c_int, [_vp]),
    "acu_ctx_synchronize": (_c_int, [_vp]),
    "acu_plan_create": (_c_int, [_vp, _c_int, _c_int, _c_int, _fp, ctypes.POINTER(_vp)]),
    "acu_batch_create": (_c_int, [_vp, _c_int, _ip, _ip, _ip, ctypes.POINTER(_fp),
                                  ctypes.POINTER(_vp)]),
    "acu_batch_destroy": (_c_int, [_vp]),
    "acu_batch_nslices": (_c_int, [_vp]),
    "acu_batch_set_theta": (_c_int, [_vp, _c_int, _fp]),
    "acu_batch_set_sino": (_c_int, [_vp, _c_int, _fp]),
    "acu_batch_get_image": (_c_int, [_vp, _c_int, _c_int, _fp]),
    "acu_batch_device_ptr": (_vp, [_vp, _c_int, _c_int]),
    "acu_batch_fbp": (_c_int, [_vp]),
    "acu_batch_tvopt": (_c_int, [_vp, _c_float, _c_float, _c_int, _c_int, _c_int, _c_int]),
    "acu_batch_get_delta": (_c_int, [_vp, _c_int, _fp, _c_int]),
    "acu_batch_set_tv_schedule": (_c_int, [_vp, _c_int, _c_int]),
    "acu_batch_tv_schedule": (_c_int, [_vp, _ip]),
    "acu_batch_set_bp_positions": (_c_int, [_vp, _c_int, _c_int]),
    "acu_batch_bp_positions": (_c_int, [_vp, _ip, _ip]),
    "acu_radon_dev": (_c_int, [_vp, _c_int, _vp, _vp]),
    "acu_iradon_dev": (_c_int, [_vp, _c_int, _vp, _vp]),
    "acu_ramp_dev": (_c_int, [_vp, _c_int, _vp, _vp]),
    "acu_backproject_dev": (_c_int, [_vp, _c_int, _vp, _vp]),
    "acu_radon": (_c_int, [_vp, _fp, _fp]),
    "acu_iradon": (_c_int, [_vp, _fp, _fp]),
    "acu_get_slice": (_c_int, [_vp, _fp, _fp]),
    "acu_tvopt": (_c_int, [_vp, _fp, _c_float, _c_float, _c_int, _c_int, _c_int, _fp, _fp, _fp]),
    "acu_batch_autotune": (_c_int, [_vp, _c_int, ctypes.c_char_p, _c_int]),
    "acu_batch_fp_variant": (_c_int, [_vp, _c_int]),
    "acu_batch_set_ramp_variant": (_c_int, [_vp, _c_int]),
    "acu_batch_ramp_variant": (_c_int, [_vp]),
    "acu_batch_set_fp_variant": (_c_int, [_vp, _c_int, _c_int]),
    "acu_batch_set_bp_tile": (_c_int, [_vp, _c_int]),
    "acu_batch_bp_tile": (_c_int, [_vp]),
    "acu_radon_adjoint_dev": (_c_int, [_vp, _c_int, _vp, _vp]),
    "acu_radon_adjoint": (_c_int, [_vp, _fp, _fp]),
    "acu_batch_loss_grad": (_c_int, [_vp, ctypes.POINTER(_fp), ctypes.POINTER(_fp), _c_float, _c_float,
                                     ctypes.POINTER(_fp), ctypes.POINTER(_fp), ctypes.POINTER(_fp),
                                     ctypes.POINTER(ctypes.c_double)]),
    "acu_batch_loss_grad_dev": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _c_float, _c_float,
                                         ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(ctypes.c_double)]),
    "acu_volume_reconstruct": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _fp, _fp, _fp, _fp, _c_float,
                                        _c_float, _c_int, _c_int, _c_int, _c_int, _c_int,
                                        ctypes.POINTER(ctypes.c_double)]),
    "acu_prep_median27_dev": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp]),
    "acu_prep_correct_dev": (_c_int, [_vp, _c_int, _c_int, _c_float, _c_float, _vp, _c_int, _c_int, _c_int,
                                      _vp]),
    "acu_prep_mirror_error_dev": (_c_int, [_vp, _c_int, _c_int, _c_float, _c_float, _vp, _vp,
                                           ctypes.POINTER(ctypes.c_double)]),
    "acu_prep_mirror_error_noalloc": (_c_int, [_vp, _c_int, _c_int, _c_float, _c_float, _vp, _vp, _vp]),
    "acu_prep_find_center": (_c_int, [_vp, _c_int, _fp, _fp, _fp, _fp, ctypes.POINTER(_c_int)]),
    "acu_prep_sinograms": (_c_int, [_vp, _c_int, _c_int, _fp, _c_float, _c_float, _fp]),
    "acu_ingest_mean_dev": (_c_int, [_vp, _c_int, ctypes.c_longlong, _vp, _vp]),
    "acu_ingest_normalize_dev": (_c_int, [_vp, _c_int, ctypes.c_longlong, _vp, _vp, _vp]),
    "acu_ingest_mean": (_c_int, [_vp, _c_int, ctypes.c_longlong, _vp, _vp]),
    "acu_ingest_normalize": (_c_int, [_vp, _c_int, ctypes.c_longlong, _vp, _vp, _vp]),
    "acu_ctx_set_timing": (_c_int, [_vp, _c_int]),
    "acu_ctx_kernel_time": (_c_int, [_vp, _c_int, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_longlong)]),
}

_lib = None


class AcumenError(RuntimeError):
    """A nonzero status from the library, with its message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"acumen error {code}: {msg}")
        self.code = code


def load(path: str | None = None):
    """Load the library (once).  Raises FileNotFoundError if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -m acumen_amd.build` "
            "(the HIP library has no CPU fallback)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != ACU_OK:
        raise AcumenError(rc, load().acu_last_error().decode(errors="replace"))
