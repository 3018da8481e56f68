"""The benchmark's first slice against the reference's own runs.

2049^2 x 1500 angles, 40-ellipse phantom (seed 1001), Poisson I0 = 5e3 (seed
1101), mu = 0.05, lambda = 0.1, 5 CG steps, FBP warm start.  The reference's
TV of this slice after 20 and 50 outer iterations, in its own FP32
arithmetic (the oracle's FP32 build, bit-identical to the reference's
module, tests/test_oracle_reference.py) and in FP64, is kept as 64
Rademacher projections per image (tests/golden/headline_2049x1500.npz,
written by tests/golden/make_headline_projections.py from ~2.7 CPU-hours of
oracle runs); distances below are estimated from the projections (relative
spread of the estimate ~13%).

* The input is regenerated here (the oracle's FP32 radon of the phantom and
  the same noise); its projections equal the fixture's bit for bit.
* 20 outer iterations: every operator schedule is within 3e-5 of FP64
  (measured 5e-6 linear, 1e-5 reference schedule) and within 1.5x the
  reference's own FP32-to-FP64 distance (1.5e-4) of the reference's FP32
  result.
* 50 outer iterations: the iteration amplifies rounding and the reference's
  own FP32 result ends 1.0e-3 from FP64.  The default (linear) schedule,
  whose carried terms are FP64, stays within half that distance of FP64
  (measured 1.4e-4); the reference's operator schedule follows the
  reference's own FP32 trajectory, within half that distance of it
  (measured 1.1e-4).
"""
import os

import numpy as np
import pytest

from acumen_amd import ACU_IMG_TV, ACU_TV_GRAPH, Batch, _lib, synth
from helpers import rel_projected, sign_projections
from oracle import ref

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N, T = 2049, 1500


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def headline():
    g = np.load(os.path.join(GOLD, "headline_2049x1500.npz"))
    assert int(g["N"]) == N and int(g["T"]) == T
    say("headline input: oracle FP32 radon of the 40-ellipse phantom, Poisson I0 = 5e3")
    th = synth.angles(T)
    ref.set_threads(16)
    try:
        sino = synth.poisson_noise(ref.radon(synth.phantom(N, 40, 1001), th, "f32"), 5e3, N, 1101)
    finally:
        ref.set_threads(1)
    return g, sino, th


def test_headline_input_is_the_fixtures(headline):
    g, sino, _ = headline
    assert np.array_equal(sign_projections(sino), g["sino"])


@pytest.mark.parametrize("sched", ["linear", "reference"])
def test_headline_tv_against_the_reference_runs(ctx, headline, sched):
    g, sino, th = headline
    wpu = int(g["wpu"])
    assert wpu == synth.default_wpu(N)
    S = {"linear": _lib.ACU_TV_SCHED_LINEAR, "reference": _lib.ACU_TV_SCHED_REFERENCE}[sched]
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        b.set_tv_schedule(S)
        b.set_sino(0, sino)
        for n in (20, 50):
            say("headline slice", sched, "schedule", n, "outer x 5 CG")
            b.tvopt(0.05, 0.1, n, 1, 5, ACU_TV_GRAPH)
            tv = b.get_image(0, ACU_IMG_TV)
            assert np.all(np.isfinite(tv))
            p = sign_projections(tv)
            e64 = rel_projected(p, g[f"tv_f64_{n}"])
            e32 = rel_projected(p, g[f"tv_f32_{n}"])
            spread = rel_projected(g[f"tv_f32_{n}"], g[f"tv_f64_{n}"])
            print(f"{sched} {n} outer: vs FP64 {e64:.2e}, vs reference FP32 {e32:.2e}; "
                  f"reference FP32 vs FP64 {spread:.2e}", flush=True)
            if n == 20:
                assert e64 < 3e-5
                assert e32 < 1.5 * spread
            elif sched == "linear":
                assert e64 < 0.5 * spread
            else:
                assert e32 < 0.5 * spread
