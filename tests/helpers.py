"""Shared inputs for the parity tests."""
import numpy as np

from acumen_amd import synth


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fov_mask(N, margin=3):
    m = (N + 1) // 2
    yy, xx = np.mgrid[1:N + 1, 1:N + 1]
    return ((xx - m) ** 2 + (yy - m) ** 2) <= ((N - 1) / 2 - margin) ** 2


def disk_phantom(N, count, seed, margin=3):
    """Seeded ellipse phantom inside the disk inscribed in the field of view
    (the region the reconstruction mask keeps)."""
    return (synth.phantom(N, count, seed) * fov_mask(N, margin)).astype(np.float32)


def interior_noise(N, seed):
    """N(0,1) image that is zero on a 3-pixel border, so no sample of the
    projector sits on the edge of the reference's bounds test."""
    x = np.random.default_rng(seed).normal(size=(N, N)).astype(np.float32)
    x[:3, :] = 0
    x[-3:, :] = 0
    x[:, :3] = 0
    x[:, -3:] = 0
    return x


def noisy_sinogram(N, T, count, seed, I0, radon):
    """Phantom -> radon -> Poisson noise, as in the benchmark configurations."""
    x = synth.phantom(N, count, seed)
    th = synth.angles(T)
    s = radon(x, th)
    return synth.poisson_noise(s, I0, N, seed + 7), th


class HipBuffers:
    """Minimal device-memory helper over libamdhip64 (same HIP runtime as the
    library, so the pointers are valid for the acu_*_dev entry points)."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.ct.c_void_p()
        assert self.hip.hipMalloc(self.ct.byref(p), nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def sign_projections(a, k=64, seed=20261203):
    """k Rademacher projections r_i . a in FP64, r_i in {-1, +1}^a.size drawn
    in order from PCG64(seed).  For two images, the norm of the difference of
    their projections over the norm of one's estimates their rel-L2 distance
    (relative spread ~13% at k = 64), so full-size results can be pinned by
    k numbers instead of the images."""
    a = np.asarray(a, np.float64).ravel()
    rng = np.random.Generator(np.random.PCG64(seed))
    total = a.sum()
    out = np.empty(k)
    for i in range(k):
        plus = rng.integers(0, 2, size=a.size, dtype=np.int8).astype(bool)
        out[i] = 2.0 * a[plus].sum() - total
    return out


def rel_projected(pa, pb):
    """rel-L2 distance of two images estimated from their sign_projections."""
    pa = np.asarray(pa, np.float64)
    pb = np.asarray(pb, np.float64)
    return float(np.linalg.norm(pa - pb) / max(np.linalg.norm(pb), 1e-300))
