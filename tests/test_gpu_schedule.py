"""Operator schedules of tvopt on the GPU (acu_batch_set_tv_schedule).

The reference recomputes iradon(fk), radon(uk), iradon(radon(uk)) and
radon(uk) again every outer iteration (tomography/tv_hpc.F90:215,228-229,
339): 7 radon + 7 iradon.  REUSE skips the radon of :228 (bit-identical);
LINEAR carries the three terms by linearity in FP64 (5 + 5).  Checked here:
REUSE equals REFERENCE bit for bit; every schedule against the reference's
own FP32 outputs (tests/golden/ref_257x180.npz: 257^2 x 180, 20 outer) and
the FP64 oracle; at 513^2 x 360 after 20 and 50 outer iterations against
the reference's FP32 arithmetic (the oracle's FP32 build, bit-identical to
the reference) and FP64; LINEAR against REFERENCE at the production size
(2049^2 x 1500, 20 x 5); the refresh option and graph replay; and the state
rules of the delta log and of new sinograms or angles.
"""
import os

import numpy as np
import pytest

from acumen_amd import ACU_IMG_FBP, ACU_IMG_TV, ACU_TV_BATCHED, ACU_TV_GRAPH, ACU_TV_PLAIN, AcumenError, Batch, _lib, synth
from helpers import disk_phantom, rel
from oracle import ref

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCHEDS = {"reference": _lib.ACU_TV_SCHED_REFERENCE, "reuse": _lib.ACU_TV_SCHED_REUSE,
          "linear": _lib.ACU_TV_SCHED_LINEAR}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    ref.set_threads(16)
    yield
    ref.set_threads(1)


def run(ctx, N, T, th, wpu, f, sched, nouter, ninner=1, ncg=5, lam=0.1, mode=ACU_TV_GRAPH, refresh=0):
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        b.set_tv_schedule(sched, refresh)
        assert b.tv_schedule() == (sched, refresh)
        b.set_sino(0, f)
        b.tvopt(0.05, lam, nouter, ninner, ncg, mode)
        return b.get_image(0, ACU_IMG_TV), b.get_image(0, ACU_IMG_FBP), b.get_delta(0, nouter)


def test_default_schedule_is_linear(ctx):
    with Batch(ctx, [(33, 8, 4, synth.angles(8))]) as b:
        if os.environ.get("ACUMEN_TV_SCHEDULE") is None:
            assert b.tv_schedule() == (_lib.ACU_TV_SCHED_LINEAR, int(os.environ.get("ACUMEN_TV_REFRESH", "0")))
        with pytest.raises(AcumenError):
            b.set_tv_schedule(7)
        with pytest.raises(AcumenError):
            b.set_tv_schedule(_lib.ACU_TV_SCHED_LINEAR, -1)


@pytest.mark.parametrize("mode", [ACU_TV_PLAIN, ACU_TV_GRAPH])
def test_reuse_equals_reference_bitwise(ctx, mode):
    N, T = 97, 64
    say("reuse vs reference schedule", N, T, mode)
    th = synth.angles(T)
    f = ref.radon(disk_phantom(N, 6, 9), th, "f64").astype(np.float32)
    a = run(ctx, N, T, th, 4, f, _lib.ACU_TV_SCHED_REFERENCE, 4, 2, 4, mode=mode)
    b = run(ctx, N, T, th, 4, f, _lib.ACU_TV_SCHED_REUSE, 4, 2, 4, mode=mode)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("name", sorted(SCHEDS))
def test_schedules_against_reference_golden_vectors(ctx, name):
    """257^2 x 180, 10-ellipse phantom, Poisson I0 = 1e4, 20 outer x 5 CG:
    the reference's own FP32 FBP and TV (tests/golden/ref_257x180.npz)."""
    g = np.load(os.path.join(GOLD, "ref_257x180.npz"))
    th, wpu, f = g["theta"], int(g["wpu"]), g["sino"]
    say("schedule", name, "against the reference's 257x180 vectors")
    tv, fbp, _ = run(ctx, 257, 180, th, wpu, f, SCHEDS[name], int(g["nouter"]))
    tv64, fbp64, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, int(g["nouter"]), 1, 5, "f64")
    e_ref, e64 = rel(tv, g["tv"]), rel(tv, tv64)
    print(f"{name}: TV vs reference FP32 {e_ref:.2e}, vs FP64 {e64:.2e}; FBP vs reference {rel(fbp, g['fbp']):.2e}",
          flush=True)
    assert rel(fbp, g["fbp"]) < 1e-5 and rel(fbp, fbp64) < 1e-5
    assert e_ref < 1e-4 and e64 < 1e-4


@pytest.fixture(scope="module")
def case513():
    N, T = 513, 360
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    x = synth.phantom(N, 10, 1000)
    f = synth.poisson_noise(ref.radon(x, th, "f32"), 1e4, N, 1007)
    return N, T, th, wpu, f


def test_20_outer_against_reference_fp32_and_fp64(ctx, case513):
    """513^2 x 360, 20 outer x 5 CG: before the iteration's sensitivity to
    rounding grows (below), every schedule stays within 3e-5 of the
    reference's FP32 arithmetic (the oracle's FP32 build reproduces the
    reference bit for bit, tests/test_oracle_reference.py) and within the
    1e-4 bar of FP64; the linear schedule's carried terms add no more than the
    GPU's own rounding does."""
    N, T, th, wpu, f = case513
    tv32, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 20, 1, 5, "f32")
    tv64, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 20, 1, 5, "f64")
    out = {}
    for name in ("reference", "linear"):
        say("schedule", name, N, T, "20 outer")
        out[name] = run(ctx, N, T, th, wpu, f, SCHEDS[name], 20)[0]
        e32, e64 = rel(out[name], tv32), rel(out[name], tv64)
        print(f"20 outer, {name}: vs reference FP32 {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
        assert e32 < 3e-5 and e64 < 1e-4
    assert rel(out["linear"], out["reference"]) < 3e-5


def test_50_outer_within_the_references_own_spread(ctx, case513):
    """513^2 x 360, 50 outer x 5 CG.  Past ~25 outer iterations the
    split-Bregman iteration at mu = 0.05, lambda = 0.1 amplifies rounding
    differences of ~1e-7 to ~1e-4 (its CG runs on a non-symmetric operator
    and the shrink never activates, so the Bregman terms only accumulate):
    here the reference's own FP32 result is ~9e-5 from its FP64 result.  The
    GPU result (default schedule) is held to the 1e-4 bar against the
    reference's FP32 result, and against FP64 to 1.5 times the reference's
    own distance from FP64."""
    N, T, th, wpu, f = case513
    tv32, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 50, 1, 5, "f32")
    tv64, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 50, 1, 5, "f64")
    say("default schedule", N, T, "50 outer")
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        b.set_sino(0, f)
        b.tvopt(0.05, 0.1, 50, 1, 5, ACU_TV_GRAPH)
        tv = b.get_image(0, ACU_IMG_TV)
    spread = rel(tv32, tv64)
    e32, e64 = rel(tv, tv32), rel(tv, tv64)
    print(f"50 outer: GPU vs reference FP32 {e32:.2e}, vs FP64 {e64:.2e}; reference FP32 vs FP64 {spread:.2e}",
          flush=True)
    assert e32 < 1e-4
    assert e64 < max(1e-4, 1.5 * spread)


def test_tv_backprojection_positions(ctx, case513):
    """Backprojection positions of FBP images and inside tvopt's iteration
    (acu_batch_set_bp_positions; both the reference's by default).
    513^2 x 360, 20 outer x 5 CG, all four combinations: the FBP image
    depends on the FBP positions only (bit for bit); every TV is within the
    1e-4 bar of the reference's FP32 arithmetic and of FP64; a fresh batch
    set to the reference's positions everywhere reproduces that TV bit for
    bit (switching leaves no state behind)."""
    N, T, th, wpu, f = case513
    REF, TILE = _lib.ACU_BP_POS_REFERENCE, _lib.ACU_BP_POS_TILE
    tv32, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 20, 1, 5, "f32")
    tv64, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 20, 1, 5, "f64")
    out = {}
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        if os.environ.get("ACUMEN_BP_POSITIONS") is None and os.environ.get("ACUMEN_TV_BP_POSITIONS") is None:
            assert b.bp_positions() == (REF, REF)
        with pytest.raises(AcumenError):
            b.set_bp_positions(REF, 2)
        b.set_sino(0, f)
        for fp in (REF, TILE):
            for tp in (REF, TILE):
                say("bp positions fbp", fp, "tv", tp, N, T, "20 outer")
                b.set_bp_positions(fp, tp)
                assert b.bp_positions() == (fp, tp)
                with pytest.raises(AcumenError):
                    b.get_image(0, ACU_IMG_TV)  # invalidated by the setting
                b.tvopt(0.05, 0.1, 20, 1, 5, ACU_TV_GRAPH)
                out[(fp, tp)] = (b.get_image(0, ACU_IMG_TV), b.get_image(0, ACU_IMG_FBP))
                e32, e64 = rel(out[(fp, tp)][0], tv32), rel(out[(fp, tp)][0], tv64)
                print(f"positions fbp {fp} tv {tp}: TV vs reference FP32 {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
                assert e32 < 1e-4 and e64 < 1e-4
    for tp in (REF, TILE):
        assert np.array_equal(out[(REF, tp)][1], out[(REF, REF)][1])
        assert np.array_equal(out[(TILE, tp)][1], out[(TILE, TILE)][1])
    assert not np.array_equal(out[(REF, REF)][1], out[(TILE, TILE)][1])
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        b.set_sino(0, f)
        b.set_bp_positions(REF, REF)
        b.tvopt(0.05, 0.1, 20, 1, 5, ACU_TV_GRAPH)
        assert np.array_equal(b.get_image(0, ACU_IMG_TV), out[(REF, REF)][0])


def test_linear_refresh_and_graph(ctx):
    """refresh = 3 recomputes the carried terms every 3 outer iterations; the
    graph path (two captured bodies) equals the batched path bit for bit, and
    both stay within the bar of the FP64 oracle."""
    N, T, nouter = 129, 96, 7
    say("linear refresh 3", N, T, nouter)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    f = ref.radon(disk_phantom(N, 6, 5), th, "f64").astype(np.float32)
    g = run(ctx, N, T, th, wpu, f, _lib.ACU_TV_SCHED_LINEAR, nouter, refresh=3)
    b = run(ctx, N, T, th, wpu, f, _lib.ACU_TV_SCHED_LINEAR, nouter, refresh=3, mode=ACU_TV_BATCHED)
    p = run(ctx, N, T, th, wpu, f, _lib.ACU_TV_SCHED_LINEAR, nouter, refresh=3, mode=ACU_TV_PLAIN)
    for x in (b, p):
        assert np.array_equal(g[0].view(np.uint32), x[0].view(np.uint32))
        assert np.array_equal(g[2], x[2])
    tv64, _, d64 = ref.tvopt(f, th, wpu, 0.05, 0.1, nouter, 1, 5, "f64")
    assert rel(g[0], tv64) < 1e-4
    assert np.allclose(g[2], d64, rtol=1e-3)


def test_full_size_linear_against_reference_schedule(ctx):
    """2049^2 x 1500, 40 ellipses, I0 = 5e3, 20 x 5: the linear schedule
    against the reference's operator schedule on the GPU, before the
    iteration's amplification of rounding sets in."""
    N, T = 2049, 1500
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    say("full-size linear vs reference schedule", N, T, "20 x 5")
    with Batch(ctx, [(N, T, wpu, th)]) as p:
        f = synth.poisson_noise(p.radon(synth.phantom(N, 40, 1001)), 5e3, N, 1101)
    lin = run(ctx, N, T, th, wpu, f, _lib.ACU_TV_SCHED_LINEAR, 20)
    refs = run(ctx, N, T, th, wpu, f, _lib.ACU_TV_SCHED_REFERENCE, 20)
    e = rel(lin[0], refs[0])
    print(f"full size, 20 outer: linear vs reference schedule {e:.2e}", flush=True)
    assert np.all(np.isfinite(lin[0]))
    assert e < 3e-5
    assert np.array_equal(lin[1], refs[1])  # the FBP warm start does not depend on the schedule


def test_delta_log_and_state_rules(ctx):
    """get_delta accepts at most the last tvopt's nouter entries, even after a
    longer earlier run; a new sinogram or new angles invalidate the slice's
    images (ACU_ERR_STATE) until the next FBP / tvopt."""
    N, T = 65, 48
    say("delta log and state rules", N, T)
    th = synth.angles(T)
    f = ref.radon(disk_phantom(N, 4, 3), th, "f64").astype(np.float32)
    with Batch(ctx, [(N, T, 4, th)]) as b:
        b.set_sino(0, f)
        b.tvopt(0.05, 0.1, 6, 1, 3, ACU_TV_GRAPH)
        d6 = b.get_delta(0, 6)
        b.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_GRAPH)
        assert np.array_equal(b.get_delta(0, 2), d6[:2])
        with pytest.raises(AcumenError) as e:
            b.get_delta(0, 3)
        assert e.value.code == _lib.ACU_ERR_ARG
        b.get_image(0, ACU_IMG_TV)
        b.set_sino(0, f)
        for which in (ACU_IMG_TV, ACU_IMG_FBP):
            with pytest.raises(AcumenError) as e:
                b.get_image(0, which)
            assert e.value.code == _lib.ACU_ERR_STATE
        with pytest.raises(AcumenError):
            b.get_delta(0, 1)
        b.fbp()
        b.get_image(0, ACU_IMG_FBP)
        b.tvopt(0.05, 0.1, 1, 1, 3, ACU_TV_GRAPH)
        b.get_image(0, ACU_IMG_TV)
        b.set_theta(0, th)
        with pytest.raises(AcumenError) as e:
            b.get_image(0, ACU_IMG_TV)
        assert e.value.code == _lib.ACU_ERR_STATE
