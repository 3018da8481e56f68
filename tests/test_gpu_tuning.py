"""Forward-projection variants, backprojection tile heights and the
autotuner on the GPU: every variant and tile height matches the oracle, a
batch whose slices use different variants (one launch per variant group)
equals single-slice plans bit for bit, and autotuning keeps results within
the parity bar."""
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


@pytest.mark.parametrize("N,T", [(65, 48), (257, 180), (1025, 60), (129, 7), (2049, 40)])
def test_every_bp_tile_matches_oracle(ctx, N, T):
    """iradon of a noisy sinogram at tile heights 16, 32, 64 against the FP64
    oracle: within 1e-5, or no further than the reference's own FP32
    arithmetic is where the noise takes that past 1e-5 (1025 x 60: 4.9e-5).
    The heights differ only in where FP32 positions are rounded (relative to
    the tile origin): they agree with each other within 1e-5."""
    th = synth.angles(T)
    rng = np.random.default_rng(N + T)
    s = (ref.radon(disk_phantom(N, 8, N + 2 * T), th, "f64")
         + rng.normal(0.0, 0.05, (T, N))).astype(np.float32)
    w = synth.default_wpu(N)
    want = ref.iradon(s, th, w, "f64")
    bar = max(FBP_TOL, rel(ref.iradon(s, th, w, "f32"), want))
    got = {}
    with Batch(ctx, [(N, T, w, th)]) as b:
        assert b.bp_tile() == 32
        for ty in (16, 32, 64):
            say("bp tile", ty, N, T)
            b.set_bp_tile(ty)
            assert b.bp_tile() == ty
            got[ty] = b.iradon(s)
            say("tile", ty, "vs FP64:", rel(got[ty], want), "bar:", bar)
            assert rel(got[ty], want) < bar, (ty, rel(got[ty], want), bar)
        with pytest.raises(AcumenError):
            b.set_bp_tile(48)
        assert b.bp_tile() == 64
    for ty in (16, 64):
        say("tile", ty, "vs 32:", rel(got[ty], got[32]))
        assert rel(got[ty], got[32]) < FBP_TOL, (ty, rel(got[ty], got[32]))


@pytest.mark.parametrize("ty", [16, 64])
def test_bp_tile_batch_equals_single_plans(ctx, ty):
    """A batch at one tile height (graph-replayed TV) equals single-slice
    plans at the same height bit for bit."""
    shapes = [(65, 48), (129, 37), (97, 100)]
    say("bp tile batch", ty, shapes)
    rng = np.random.default_rng(ty)
    items = []
    for i, (N, T) in enumerate(shapes):
        th = np.sort(rng.uniform(0, 180, T)).astype(np.float32)
        f = ref.radon(disk_phantom(N, 5, 80 + i), th, "f64").astype(np.float32)
        items.append((N, T, synth.default_wpu(N), th, f))
    with Batch(ctx, [(N, T, w, th) for N, T, w, th, _ in items]) as b:
        for i, it in enumerate(items):
            b.set_sino(i, it[4])
        b.tvopt(0.05, 0.1, 1, 1, 2, ACU_TV_GRAPH)   # a captured graph at the default height
        b.set_bp_tile(ty)                            # is rebuilt at the new one
        b.tvopt(0.05, 0.1, 3, 1, 4, ACU_TV_GRAPH)
        batched = [b.get_image(i, ACU_IMG_TV) for i in range(len(items))]
    for i, (N, T, w, th, f) in enumerate(items):
        with Batch(ctx, [(N, T, w, th)]) as p:
            p.set_bp_tile(ty)
            p.set_sino(0, f)
            p.tvopt(0.05, 0.1, 3, 1, 4, ACU_TV_PLAIN)
            assert np.array_equal(p.get_image(0, ACU_IMG_TV), batched[i]), i
        tv64, _, _ = ref.tvopt(f, th, w, 0.05, 0.1, 3, 1, 4, "f64")
        assert rel(batched[i], tv64) < TV_TOL


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
        # backprojection tile height: fastest over the batch
        bpr = rep["bp"]
        assert set(bpr["ms"]) == {"16", "32", "64"}
        assert str(bpr["choice"]) == min(bpr["ms"], key=bpr["ms"].get)
        assert b.bp_tile() == bpr["choice"]
        b.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_GRAPH)
        for i, (N, T, w, th, f) in enumerate(items):
            tv64, _, _ = ref.tvopt(f, th, w, 0.05, 0.1, 2, 1, 3, "f64")
            assert rel(b.get_image(i, ACU_IMG_TV), tv64) < TV_TOL
        with pytest.raises(AcumenError):
            b.set_fp_variant(0, 3)


@pytest.mark.parametrize("variant", [1, 2, 4])
def test_dispatch_order_changes_no_bit(ctx, monkeypatch, variant):
    """The LDS projector's dispatch orders (the work-balanced permutation of
    a one-slice or mixed launch, the natural order of balanced launches)
    only reorder workgroups: a mixed batch's TV and a single slice's radon
    equal, bit for bit, those of batches created with ACUMEN_FPL_ORDER=rb
    (every launch in ray-block order)."""
    shapes = [(513, 360), (257, 180), (769, 240), (513, 360)]
    say("dispatch order", shapes, variant)
    rng = np.random.default_rng(23)
    items = []
    for i, (N, T) in enumerate(shapes):
        th = np.sort(rng.uniform(0, 180, T)).astype(np.float32)
        f = ref.radon(disk_phantom(N, 5, 60 + i), th, "f64").astype(np.float32)
        items.append((N, T, synth.default_wpu(N), th, f))

    def run():
        with Batch(ctx, [(N, T, w, th) for N, T, w, th, _ in items]) as b:
            for i, it in enumerate(items):
                b.set_sino(i, it[4])
                b.set_fp_variant(i, variant)
            b.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_GRAPH)
            tv = [b.get_image(i, ACU_IMG_TV) for i in range(len(items))]
        N, T, w, th, _ = items[2]
        with Batch(ctx, [(N, T, w, th)]) as p:
            p.set_fp_variant(0, variant)
            sino = p.radon(disk_phantom(N, 7, 99))
        return tv, sino

    monkeypatch.delenv("ACUMEN_FPL_ORDER", raising=False)
    tv_bal, sino_bal = run()
    monkeypatch.setenv("ACUMEN_FPL_ORDER", "rb")
    tv_rb, sino_rb = run()
    for i in range(len(items)):
        assert np.array_equal(tv_bal[i].view(np.uint32), tv_rb[i].view(np.uint32)), i
    assert np.array_equal(sino_bal.view(np.uint32), sino_rb.view(np.uint32))
