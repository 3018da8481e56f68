"""Forward-projection variants and the autotuner on the GPU: every variant
matches the oracle, a batch whose slices use different variants (one launch
per variant group) equals single-slice plans bit for bit, and autotuning
keeps results within the parity bar."""
import numpy as np
import pytest

from acumen_amd import ACU_IMG_TV, ACU_TV_GRAPH, ACU_TV_PLAIN, AcumenError, Batch, _lib, synth
from helpers import disk_phantom, rel
from oracle import ref

pytestmark = pytest.mark.gpu

FBP_TOL, TV_TOL = 1e-5, 1e-4
VARIANTS = (0, 1, 2, 4, 8)  # gather, LDS boxes with 1/2/4 angles per lane, cell boxes


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("N,T", [(257, 180), (1025, 60)])
def test_every_variant_matches_oracle(ctx, N, T):
    th = synth.angles(T)
    x = disk_phantom(N, 10, 3 * N + T)
    want = ref.radon(x, th, "f32")
    with Batch(ctx, [(N, T, synth.default_wpu(N), th)]) as b:
        for v in VARIANTS:
            say("variant", v, N, T)
            b.set_fp_variant(0, v)
            assert b.fp_variant(0) == v
            got = b.radon(x)
            assert rel(got, want) < FBP_TOL, (v, rel(got, want))


@pytest.mark.parametrize("N,T", [(257, 1500), (513, 1500), (129, 720)])
def test_cell_boxes_equal_raw_boxes(ctx, N, T):
    """The cell-box projector (8) and the raw-box projector with 4 angles per
    lane (4) share blocks, chunks, boxes and arithmetic: where every box fits
    both LDS layouts (fine angular steps keep the 16-angle drift small) the
    sinograms are identical bit for bit."""
    say("cell16 vs lds4", N, T)
    th = synth.angles(T)
    x = disk_phantom(N, 12, 7 * N + T)
    with Batch(ctx, [(N, T, synth.default_wpu(N), th)]) as b:
        b.set_fp_variant(0, 4)
        raw = b.radon(x)
        b.set_fp_variant(0, 8)
        cell = b.radon(x)
    assert np.array_equal(raw.view(np.uint32), cell.view(np.uint32))


def test_mixed_variant_batch_equals_single_plans(ctx):
    shapes = [(65, 48), (97, 60), (129, 37), (65, 100), (161, 24), (97, 33)]
    variants = [0, 1, 2, 4, 1, 8]
    say("mixed variants", shapes, variants)
    rng = np.random.default_rng(11)
    items = []
    for i, (N, T) in enumerate(shapes):
        th = np.sort(rng.uniform(0, 180, T)).astype(np.float32)
        f = ref.radon(disk_phantom(N, 5, 40 + i), th, "f64").astype(np.float32)
        items.append((N, T, synth.default_wpu(N), th, f))
    with Batch(ctx, [(N, T, w, th) for N, T, w, th, _ in items]) as b:
        for i, it in enumerate(items):
            b.set_sino(i, it[4])
            b.set_fp_variant(i, variants[i])
        b.tvopt(0.05, 0.1, 3, 1, 4, ACU_TV_GRAPH)
        batched = [b.get_image(i, ACU_IMG_TV) for i in range(len(items))]
    for i, (N, T, w, th, f) in enumerate(items):
        with Batch(ctx, [(N, T, w, th)]) as p:
            p.set_fp_variant(0, variants[i])
            p.set_sino(0, f)
            p.tvopt(0.05, 0.1, 3, 1, 4, ACU_TV_PLAIN)
            assert np.array_equal(p.get_image(0, ACU_IMG_TV), batched[i]), i
        tv64, _, _ = ref.tvopt(f, th, w, 0.05, 0.1, 3, 1, 4, "f64")
        assert rel(batched[i], tv64) < TV_TOL


def test_autotune_report_and_results(ctx):
    shapes = [(257, 180), (513, 90), (257, 180), (129, 400)]
    say("autotune", shapes)
    items = []
    for i, (N, T) in enumerate(shapes):
        th = synth.angles(T)
        f = ref.radon(disk_phantom(N, 6, 60 + i), th, "f64").astype(np.float32)
        items.append((N, T, synth.default_wpu(N), th, f))
    with Batch(ctx, [(N, T, w, th) for N, T, w, th, _ in items]) as b:
        for i, it in enumerate(items):
            b.set_sino(i, it[4])
        rep = b.autotune(reps=1)
        assert [s["slice"] for s in rep["slices"]] == list(range(len(shapes)))
        names = {0: "gather", 1: "lds1", 2: "lds2", 4: "lds4", 8: "cell16"}
        for i, s in enumerate(rep["slices"]):
            assert (s["N"], s["T"]) == shapes[i]
            assert set(s["ms"]) == set(names.values())
            assert s["choice"] == min(s["ms"], key=s["ms"].get)
            assert names[b.fp_variant(i)] == s["choice"]
        # identical shapes are timed once and get the same choice
        assert rep["slices"][0]["ms"] == rep["slices"][2]["ms"]
        # ramp filter: fastest among the arithmetics that meet the bar
        rmp = rep["ramp"]
        assert set(rmp["ms"]) == {"fp64", "fp32", "bf16x6", "fft"}
        assert rmp["choice"] == min(rmp["ms"], key=rmp["ms"].get)
        rnames = {_lib.RAMP_FP64: "fp64", _lib.RAMP_FP32: "fp32", _lib.RAMP_BF16X6: "bf16x6",
                  _lib.RAMP_FFT: "fft"}
        assert rnames[b.ramp_variant()] == rmp["choice"]
        b.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_GRAPH)
        for i, (N, T, w, th, f) in enumerate(items):
            tv64, _, _ = ref.tvopt(f, th, w, 0.05, 0.1, 2, 1, 3, "f64")
            assert rel(b.get_image(i, ACU_IMG_TV), tv64) < TV_TOL
        with pytest.raises(AcumenError):
            b.set_fp_variant(0, 3)
