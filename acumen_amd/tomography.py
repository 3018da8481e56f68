"""Host-side mirror of Acumen's Fortran module ``Tomography``.

The reference keeps its state in module globals (tomography/tv_hpc.F90:3-7)
and exposes six subroutines; :class:`Tomography` keeps the same names,
argument meanings and calling order on top of the C ABI of
libacumen_hip.so:

    t = Tomography(ctx)                 # the module
    t.nimage, t.ntheta, t.wpu = N, T, w # globals set by the driver
    t.int_radon()                       # tv_hpc.F90:11-78
    t.theta[:] = angles                 # tv_hpc.F90:560-562
    t.f[:] = sinogram                   # tv_hpc.F90:614-618
    img_tv, img_fbp = t.tvopt(jslice)   # tv_hpc.F90:162-363
    img_fbp = t.get_slice()             # fbp_hpc.F90:129-150
    sino = t.radon(img); img = t.iradon(sino)
    t.fin_radon()

Arrays are numpy float32 in the reference's Fortran layout expressed as
C-ordered arrays: an image ``(x, y)`` is shape ``(N, N)`` indexed ``[y, x]``,
a sinogram ``(detector, angle)`` is shape ``(T, N)`` indexed ``[angle, det]``.

:class:`Batch` is the batched form (many slices of their own N, T, wpu and
angles) that replaces the reference's MPI master/worker slice loop.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

_F = ctypes.POINTER(ctypes.c_float)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_F)


def _f32(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and out.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {out.shape}")
    return out


class Context:
    """One device and the HIP stream all work of this context is ordered on."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.acu_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self) -> None:
        check(self._lib.acu_ctx_synchronize(self.handle))

    def set_timing(self, enable: bool) -> None:
        check(self._lib.acu_ctx_set_timing(self.handle, 1 if enable else 0))

    def kernel_time(self, kind: int) -> tuple[float, int]:
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        check(self._lib.acu_ctx_kernel_time(self.handle, kind, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def close(self) -> None:
        if self.handle:
            self._lib.acu_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Slices reconstructed together; a 1-slice batch is a plan."""

    def __init__(self, ctx: Context, slices):
        """slices: iterable of (nimage, ntheta, wpu, theta_deg or None)."""
        self.ctx = ctx
        self._lib = ctx._lib
        slices = list(slices)
        if not slices:
            raise ValueError("empty batch")
        n = len(slices)
        Ns = (ctypes.c_int * n)(*[int(s[0]) for s in slices])
        Ts = (ctypes.c_int * n)(*[int(s[1]) for s in slices])
        Ws = (ctypes.c_int * n)(*[int(s[2]) for s in slices])
        self._thetas = [None if s[3] is None else _f32(s[3], (int(s[1]),)) for s in slices]
        th = (_F * n)(*[(_fptr(t) if t is not None else _F()) for t in self._thetas])
        h = ctypes.c_void_p()
        check(self._lib.acu_batch_create(ctx.handle, n, Ns, Ts, Ws, th, ctypes.byref(h)))
        self.handle = h
        self.shapes = [(int(s[0]), int(s[1])) for s in slices]

    def __len__(self):
        return len(self.shapes)

    def set_theta(self, i: int, theta) -> None:
        t = _f32(theta, (self.shapes[i][1],))
        check(self._lib.acu_batch_set_theta(self.handle, i, _fptr(t)))

    def set_sino(self, i: int, f) -> None:
        N, T = self.shapes[i]
        a = _f32(f, (T, N))
        check(self._lib.acu_batch_set_sino(self.handle, i, _fptr(a)))

    def get_image(self, i: int, which: int = _lib.ACU_IMG_TV) -> np.ndarray:
        N, _ = self.shapes[i]
        out = np.empty((N, N), np.float32)
        check(self._lib.acu_batch_get_image(self.handle, i, which, _fptr(out)))
        return out

    def get_delta(self, i: int, n: int) -> np.ndarray:
        out = np.empty(n, np.float32)
        check(self._lib.acu_batch_get_delta(self.handle, i, _fptr(out), n))
        return out

    def device_ptr(self, i: int, which: int) -> int:
        """Device address of slice i's f, FBP or TV array.  The TV address
        moves when a tvopt runs more outer iterations than any before on
        this batch; ask again after such a call."""
        p = self._lib.acu_batch_device_ptr(self.handle, i, which)
        if not p:
            check(_lib.ACU_ERR_ARG)
        return p

    def fbp(self) -> None:
        check(self._lib.acu_batch_fbp(self.handle))

    def tvopt(self, mu: float, lam: float, nouter: int, ninner: int = 1, ncg: int = 5,
              mode: int = _lib.ACU_TV_GRAPH) -> None:
        check(self._lib.acu_batch_tvopt(self.handle, float(mu), float(lam), int(nouter),
                                        int(ninner), int(ncg), int(mode)))

    def set_tv_schedule(self, sched: int, refresh: int = 0) -> None:
        """Operator schedule of tvopt: _lib.ACU_TV_SCHED_LINEAR (default: 5
        radon + 5 iradon per outer iteration), ACU_TV_SCHED_REUSE (6 + 7,
        bit-identical to the reference's) or ACU_TV_SCHED_REFERENCE (7 + 7);
        refresh > 0 recomputes the linear schedule's carried terms exactly
        every `refresh` outer iterations."""
        check(self._lib.acu_batch_set_tv_schedule(self.handle, int(sched), int(refresh)))

    def tv_schedule(self) -> tuple[int, int]:
        """(schedule, refresh period)."""
        r = ctypes.c_int()
        v = self._lib.acu_batch_tv_schedule(self.handle, ctypes.byref(r))
        if v < 0:
            check(-v)
        return v, r.value

    # single-slice operators on host arrays (slice 0) -- the reference subroutines
    def radon(self, img) -> np.ndarray:
        N, T = self.shapes[0]
        x = _f32(img, (N, N))
        out = np.empty((T, N), np.float32)
        check(self._lib.acu_radon(self.handle, _fptr(out), _fptr(x)))
        return out

    def iradon(self, sino) -> np.ndarray:
        N, T = self.shapes[0]
        s = _f32(sino, (T, N))
        out = np.empty((N, N), np.float32)
        check(self._lib.acu_iradon(self.handle, _fptr(out), _fptr(s)))
        return out

    # ---- tuning ----
    def autotune(self, reps: int = 2) -> dict:
        """Time every forward-projection variant for every distinct slice
        shape and keep the fastest per slice, the ramp-filter arithmetics
        that meet the accuracy bar and the backprojection tile heights on the
        whole batch; returns the report (ms per radon per slice, per batch
        ramp filter and per batch backprojection, and the choices)."""
        import json
        cap = 1536 + 192 * len(self.shapes)
        buf = ctypes.create_string_buffer(cap)
        check(self._lib.acu_batch_autotune(self.handle, int(reps), buf, cap))
        return json.loads(buf.value.decode())

    def set_ramp_variant(self, variant: int) -> None:
        """Ramp-filter arithmetic: _lib.RAMP_FP64 (default), RAMP_FP32, RAMP_BF16X3,
        RAMP_BF16X6, RAMP_FFT."""
        check(self._lib.acu_batch_set_ramp_variant(self.handle, int(variant)))

    def ramp_variant(self) -> int:
        v = self._lib.acu_batch_ramp_variant(self.handle)
        if v < 0:
            check(-v)
        return v

    def fp_variant(self, i: int) -> int:
        """0 = global gathers; 1, 2, 4 = LDS pixel boxes with that many angles
        per lane; 8 = LDS bilinear-cell boxes (16 angles per workgroup)."""
        v = self._lib.acu_batch_fp_variant(self.handle, i)
        if v < 0:
            check(-v)
        return v

    def set_fp_variant(self, i: int, variant: int) -> None:
        check(self._lib.acu_batch_set_fp_variant(self.handle, i, int(variant)))

    def bp_tile(self) -> int:
        """Backprojection tile height of the batch: 16, 32 or 64 rows."""
        v = self._lib.acu_batch_bp_tile(self.handle)
        if v < 0:
            check(-v)
        return v

    def set_bp_tile(self, rows: int) -> None:
        check(self._lib.acu_batch_set_bp_tile(self.handle, int(rows)))

    def set_bp_positions(self, fbp: int, tv: int) -> None:
        """Backprojection positions (_lib.ACU_BP_POS_REFERENCE or
        ACU_BP_POS_TILE) of FBP images and inside tvopt's iteration; both
        default to REFERENCE."""
        check(self._lib.acu_batch_set_bp_positions(self.handle, int(fbp), int(tv)))

    def bp_positions(self) -> tuple[int, int]:
        """(FBP positions, tvopt positions)."""
        f, t = ctypes.c_int(), ctypes.c_int()
        check(self._lib.acu_batch_bp_positions(self.handle, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    # ---- differentiation (no counterpart in the reference; built on radon) ----
    def radon_adjoint(self, sino) -> np.ndarray:
        """Exact transpose of :meth:`radon` (slice 0): (T,N) sinogram to (N,N) image."""
        N, T = self.shapes[0]
        s = _f32(sino, (T, N))
        out = np.empty((N, N), np.float32)
        check(self._lib.acu_radon_adjoint(self.handle, _fptr(out), _fptr(s)))
        return out

    def radon_adjoint_dev(self, i: int, img_ptr: int, sino_ptr: int) -> None:
        check(self._lib.acu_radon_adjoint_dev(self.handle, i, img_ptr, sino_ptr))

    def loss_grad(self, xs, shifts=None, lam: float = 0.1, eps: float = 1e-3,
                  grad_x: bool = True, grad_s: bool = True, curvature: bool = False):
        """L_i = 1/2 |shift_{s_i}(radon(x_i)) - f_i|^2 + lam TV_eps(x_i) for every
        slice, f_i the slice's sinogram (:meth:`set_sino`).

        Returns ``(loss, gx, gs)``: ``loss`` (n, 3) = (L, data term, TV_eps),
        ``gx`` a list of (N,N) image gradients (None if not asked for), ``gs``
        a list of (T,) shift gradients (None if not asked for); with
        ``curvature`` a fourth item, the (T,) Gauss-Newton curvatures of the
        data term in each shift.  ``shifts`` is None (all zero) or a list of
        (T,) arrays / None."""
        n = len(self.shapes)
        if len(xs) != n:
            raise ValueError("one image per slice")
        keep = []
        xp, sp, gxp, gsp, hsp = ((_F * n)() for _ in range(5))
        gx_out, gs_out, hs_out = [], [], []
        for i, (N, T) in enumerate(self.shapes):
            x = _f32(xs[i], (N, N))
            keep.append(x)
            xp[i] = _fptr(x)
            if shifts is not None and shifts[i] is not None:
                sh = np.ascontiguousarray(shifts[i], np.float32)
                if sh.shape != (T,):
                    raise ValueError(f"shifts[{i}] must have shape ({T},)")
                keep.append(sh)
                sp[i] = _fptr(sh)
            for lst, arr, shape in ((gx_out, gxp, (N, N)), (gs_out, gsp, (T,)), (hs_out, hsp, (T,))):
                a = np.empty(shape, np.float32)
                lst.append(a)
                arr[i] = _fptr(a)
        loss = np.zeros((n, 3), np.float64)
        check(self._lib.acu_batch_loss_grad(
            self.handle, xp, sp if shifts is not None else None, float(lam), float(eps),
            gxp if grad_x else None, gsp if grad_s else None, hsp if curvature else None,
            loss.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        out = (loss, gx_out if grad_x else None, gs_out if grad_s else None)
        return out + (hs_out,) if curvature else out

    def register_shifts(self, xs, shifts0=None, iters: int = 10, **kw):
        """Per-angle centre-of-rotation shifts that align each slice's
        sinogram with radon(x_i): Gauss-Newton steps on the data term (see
        :func:`acumen_amd.registration.gauss_newton_shifts`).  Returns the
        fitted (T,) float32 shift arrays and the number of evaluations (one
        batched forward projection each)."""
        from .registration import gauss_newton_shifts
        if shifts0 is None:
            shifts0 = [np.zeros(T, np.float32) for _, T in self.shapes]

        def evaluate(sh):
            _, _, gs, hs = self.loss_grad(xs, sh, lam=0.0, eps=1.0, grad_x=False, grad_s=True,
                                          curvature=True)
            return gs, hs

        return gauss_newton_shifts(evaluate, shifts0, iters=iters, **kw)

    def radon_dev(self, i: int, sino_ptr: int, img_ptr: int) -> None:
        check(self._lib.acu_radon_dev(self.handle, i, sino_ptr, img_ptr))

    def iradon_dev(self, i: int, img_ptr: int, sino_ptr: int) -> None:
        check(self._lib.acu_iradon_dev(self.handle, i, img_ptr, sino_ptr))

    def ramp_dev(self, i: int, out_ptr: int, sino_ptr: int) -> None:
        check(self._lib.acu_ramp_dev(self.handle, i, out_ptr, sino_ptr))

    def backproject_dev(self, i: int, img_ptr: int, filtered_ptr: int) -> None:
        check(self._lib.acu_backproject_dev(self.handle, i, img_ptr, filtered_ptr))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.acu_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tomography:
    """The reference module: globals plus int_radon / iradon / radon / tvopt /
    get_slice / fin_radon (tomography/tv_hpc.F90:1-390)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context()
        # module globals (tv_hpc.F90:3-4); ninnner keeps the reference's spelling
        self.ncg = 5
        self.ninnner = 1
        self.nouter = 1
        self.nimage = 0
        self.ntheta = 0
        self.wpu = 1
        self.add_Nx = 0
        self.mu = 0.0
        self.lambda_ = 0.0
        self.theta: np.ndarray | None = None
        self.f: np.ndarray | None = None
        self.last_delta: np.ndarray | None = None
        self._batch: Batch | None = None

    @property
    def midpoint(self) -> float:
        return float((self.nimage + 1) // 2)

    def int_radon(self) -> None:
        """Allocate the device state for (nimage, ntheta, wpu); theta and f are
        allocated here and filled by the caller afterwards (tv_hpc.F90:559-562)."""
        if self._batch is not None:
            self._batch.close()
        self._batch = Batch(self.ctx, [(self.nimage, self.ntheta, self.wpu, None)])
        self.theta = np.zeros(self.ntheta, np.float32)
        self.f = np.zeros((self.ntheta, self.nimage), np.float32)

    def _sync_theta(self) -> None:
        self._batch.set_theta(0, self.theta)

    def iradon(self, sinogram) -> np.ndarray:
        """iradon(img_rec, sinogram): ramp filter + masked backprojection."""
        self._sync_theta()
        return self._batch.iradon(sinogram)

    def radon(self, img) -> np.ndarray:
        """radon(sinogram, img): ray-driven bilinear projection / nimage."""
        self._sync_theta()
        return self._batch.radon(img)

    def get_slice(self) -> np.ndarray:
        """get_slice(img_rec_fbp) = iradon(f) (fbp_hpc.F90:129-150)."""
        self._sync_theta()
        self._batch.set_sino(0, self.f)
        self._batch.fbp()
        return self._batch.get_image(0, _lib.ACU_IMG_FBP)

    def tvopt(self, jslice: int = 0, mode: int = _lib.ACU_TV_GRAPH):
        """tvopt(img_rec_tv, img_rec_fbp, jslice) with the module's mu,
        lambda, nouter, ninnner, ncg.  Returns (img_rec_tv, img_rec_fbp)."""
        self._sync_theta()
        self._batch.set_sino(0, self.f)
        self._batch.tvopt(self.mu, self.lambda_, self.nouter, self.ninnner, self.ncg, mode)
        self.last_delta = self._batch.get_delta(0, self.nouter) if self.nouter > 0 else None
        return (self._batch.get_image(0, _lib.ACU_IMG_TV),
                self._batch.get_image(0, _lib.ACU_IMG_FBP))

    def fin_radon(self) -> None:
        if self._batch is not None:
            self._batch.close()
            self._batch = None
