"""Writes headline_2049x1500.npz: the reference's own TV of the benchmark's
first slice (2049^2 x 1500 angles, 40-ellipse phantom seed 1001, Poisson
I0 = 5e3 seed 1101, mu = 0.05, lambda = 0.1, 5 CG steps) after 20 and 50
outer iterations, in the reference's FP32 arithmetic (the oracle's FP32
build, bit-identical to the reference's own module, tests/test_oracle_reference.py)
and in FP64, kept as 64 Rademacher projections each (tests/helpers.py
sign_projections), with the same projections of the input sinogram.

The four tvopt runs take ~2.7 hours of 8 CPU threads (FP32: 19 and 47 min,
FP64: 20 and 50 min).  --from DIR reads the same four results saved earlier
by this script's --save DIR (tv_f32_20.npy, tv_f32_50.npy, tv_f64_20.npy,
tv_f64_50.npy and sino.npy) instead of recomputing them.
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from acumen_amd import synth  # noqa: E402
from helpers import sign_projections  # noqa: E402
from oracle import ref  # noqa: E402

N, T = 2049, 1500
NOUTERS = (20, 50)


def headline_sinogram(threads=8):
    ref.set_threads(threads)
    th = synth.angles(T)
    x = synth.phantom(N, 40, 1001)
    return synth.poisson_noise(ref.radon(x, th, "f32"), 5e3, N, 1101)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--from", dest="src", default=None, help="directory of saved results")
    ap.add_argument("--save", default=None, help="directory to save the full results in")
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    sino = headline_sinogram(a.threads)
    res = {}
    if a.src:
        saved = np.load(os.path.join(a.src, "sino.npy"))
        assert np.array_equal(saved, sino), "saved sinogram differs from this script's"
        for p in ("f32", "f64"):
            for n in NOUTERS:
                res[(p, n)] = np.load(os.path.join(a.src, f"tv_{p}_{n}.npy"))
    else:
        for p in ("f32", "f64"):
            for n in NOUTERS:
                tv, _, _ = ref.tvopt(sino, th, wpu, 0.05, 0.1, n, 1, 5, p)
                res[(p, n)] = tv.astype(np.float32) if p == "f32" else tv
                print(p, n, "done", flush=True)
        if a.save:
            os.makedirs(a.save, exist_ok=True)
            np.save(os.path.join(a.save, "sino.npy"), sino)
            for (p, n), v in res.items():
                np.save(os.path.join(a.save, f"tv_{p}_{n}.npy"), v)
    out = {"sino": sign_projections(sino), "N": N, "T": T, "wpu": wpu}
    for (p, n), v in res.items():
        out[f"tv_{p}_{n}"] = sign_projections(v)
    np.savez(os.path.join(HERE, "headline_2049x1500.npz"), **out)
    print("wrote headline_2049x1500.npz", flush=True)


if __name__ == "__main__":
    main()
