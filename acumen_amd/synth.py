"""Seeded synthetic neutron-CT inputs: ellipse phantoms, angle sets, noise.

Layout convention (shared with the whole package): an image ``(x, y)`` in the
reference's Fortran column-major layout is a C-ordered numpy array of shape
``(N, N)`` indexed ``[y-1, x-1]``; a sinogram ``(detector, angle)`` is a
C-ordered array of shape ``(T, N)`` indexed ``[angle, detector]``.

Pixel ``(x, y)`` (1-based) has its centre at coordinates ``(x, y)``; the
rotation centre is the reference's ``midpoint = (N+1)/2`` (integer division,
tomography/tv_hpc.F90:15).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def padded_size(width: int) -> int:
    """Odd internal image side: even detector widths get one padding bin
    (add_Nx, tomography/tv_hpc.F90:551-555)."""
    return width + 1 if width % 2 == 0 else width


def default_wpu(N: int) -> int:
    """Width of the mask's smooth edge in pixels (``Radon_wpu``).

    The reference leaves it to the caller; ``max(4, N // 32)`` keeps the same
    relative taper at every image size."""
    return max(4, N // 32)


def angles(T: int) -> np.ndarray:
    """theta_k = k * 180 / T degrees, k = 0..T-1, stored FP32 (theta.dat)."""
    return (np.arange(T, dtype=np.float64) * (180.0 / T)).astype(np.float32)


@dataclass
class Ellipse:
    cx: float
    cy: float
    a: float
    b: float
    phi: float
    att: float


def random_ellipses(rng: np.random.Generator, N: int, count: int,
                    axes=(5.0, 60.0), att=(0.005, 0.05)) -> list[Ellipse]:
    """Centres uniform in the disk of radius 0.8*(N-1)/2 about the midpoint,
    semi-axes U[axes], rotation U[0, pi), attenuation U[att] per pixel."""
    m = float((N + 1) // 2)
    R = 0.8 * (N - 1) / 2.0
    out = []
    for _ in range(count):
        r = R * np.sqrt(rng.uniform())
        t = rng.uniform(0.0, 2.0 * np.pi)
        out.append(Ellipse(cx=m + r * np.cos(t), cy=m + r * np.sin(t),
                           a=rng.uniform(*axes), b=rng.uniform(*axes),
                           phi=rng.uniform(0.0, np.pi), att=rng.uniform(*att)))
    return out


def render(N: int, ellipses: list[Ellipse]) -> np.ndarray:
    """Sum of ellipse indicator functions sampled at pixel centres (FP32)."""
    img = np.zeros((N, N), np.float64)
    coords = np.arange(1, N + 1, dtype=np.float64)
    for e in ellipses:
        x0 = max(1, int(np.floor(e.cx - max(e.a, e.b))) - 1)
        x1 = min(N, int(np.ceil(e.cx + max(e.a, e.b))) + 1)
        y0 = max(1, int(np.floor(e.cy - max(e.a, e.b))) - 1)
        y1 = min(N, int(np.ceil(e.cy + max(e.a, e.b))) + 1)
        if x0 > x1 or y0 > y1:
            continue
        X = coords[x0 - 1:x1][None, :] - e.cx
        Y = coords[y0 - 1:y1][:, None] - e.cy
        c, s = np.cos(e.phi), np.sin(e.phi)
        u = (X * c + Y * s) / e.a
        v = (-X * s + Y * c) / e.b
        img[y0 - 1:y1, x0 - 1:x1] += np.where(u * u + v * v <= 1.0, e.att, 0.0)
    return img.astype(np.float32)


def phantom(N: int, count: int, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return render(N, random_ellipses(rng, N, count))


def poisson_noise(sino: np.ndarray, I0: float, N: int, seed: int) -> np.ndarray:
    """Beer-Lambert + Poisson counts on the physical line integral N*s
    (the reference's radon divides by N, tomography/tv_hpc.F90:154):
    c ~ Poisson(I0*exp(-N*s)),  s_noisy = -log(max(c,1)/I0)/N, stored FP32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = I0 * np.exp(-N * sino.astype(np.float64))
    c = rng.poisson(lam).astype(np.float64)
    return (-np.log(np.maximum(c, 1.0) / I0) / N).astype(np.float32)


def shift_rows(sino: np.ndarray, shifts: np.ndarray) -> np.ndarray:
    """A centre-of-rotation error per projection: row t of the (T, N)
    sinogram sampled at detector position d - shifts[t] by linear
    interpolation, zero off the detector (FP64 arithmetic, stored FP32)."""
    T, N = sino.shape
    s = sino.astype(np.float64)
    d = np.arange(N, dtype=np.float64)[None, :] - np.asarray(shifts, np.float64)[:, None]
    i0 = np.floor(d).astype(np.int64)
    w = d - i0
    rows = np.arange(T)[:, None]

    def at(i):
        ok = (i >= 0) & (i < N)
        return np.where(ok, s[rows, np.clip(i, 0, N - 1)], 0.0)

    return ((1.0 - w) * at(i0) + w * at(i0 + 1)).astype(np.float32)


MIXED_WIDTHS = (512, 768, 1024, 1536, 2048)


def mixed_batch(n: int = 64, seed: int = 1002):
    """Slices of uneven size for the batching workload: detector width from
    MIXED_WIDTHS and angle count from U{360..1500}, both seeded; per slice
    an ellipse count in 20..40, Poisson I0 ~ U[2e3, 2e4] and a phantom seed.
    Returns a list of dicts (N, T, wpu, theta, count, I0, seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        width = int(rng.choice(MIXED_WIDTHS))
        T = int(rng.integers(360, 1501))
        N = padded_size(width)
        out.append(dict(N=N, T=T, wpu=default_wpu(N), theta=angles(T),
                        count=int(rng.integers(20, 41)), I0=float(rng.uniform(2e3, 2e4)),
                        seed=seed * 1000 + i))
    return out
