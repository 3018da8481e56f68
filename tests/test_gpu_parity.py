"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Tolerances (relative L2):
  * radon / ramp / backprojection / iradon / FBP: <= 1e-5 against both the
    FP64 oracle and the reference's FP32 arithmetic;
  * split-Bregman TV after a fixed iteration count: <= 1e-4 against FP64.
Runs on one MI355X; every case prints its shape before the GPU work starts.
"""
import numpy as np
import pytest

from acumen_amd import ACU_IMG_FBP, ACU_IMG_TV, ACU_TV_BATCHED, ACU_TV_GRAPH, ACU_TV_PLAIN
from acumen_amd import Batch, Tomography, _lib, synth
from helpers import HipBuffers, disk_phantom, interior_noise, rel
from oracle import ref

pytestmark = pytest.mark.gpu

FBP_TOL = 1e-5
TV_TOL = 1e-4


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    ref.set_threads(16)
    yield
    ref.set_threads(1)


def say(*a):
    print("[case]", *a, flush=True)


def plan(ctx, N, T, th, wpu=None):
    return Batch(ctx, [(N, T, synth.default_wpu(N) if wpu is None else wpu, th)])


@pytest.mark.parametrize("N,T", [(65, 48), (257, 180), (513, 720)])
def test_radon_matches_oracle(ctx, N, T):
    say("radon", N, T)
    th = synth.angles(T)
    x = disk_phantom(N, 10, N)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
    assert rel(s, ref.radon(x, th, "f64")) < FBP_TOL
    assert rel(s, ref.radon(x, th, "f32")) < FBP_TOL


@pytest.mark.parametrize("N,T", [(1025, 60), (2049, 150)])
def test_radon_boxes_too_large_for_lds(ctx, N, T):
    # few angles: an angle block drifts farther than the LDS box allows, so
    # chunks away from the centre read the global bilinear cells
    say("radon global-cells path", N, T)
    th = synth.angles(T)
    x = disk_phantom(N, 12, N + T)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
    assert rel(s, ref.radon(x, th, "f32")) < FBP_TOL
    with ref.fp32_angles():
        assert rel(s, ref.radon(x, th, "f64")) < FBP_TOL


def test_radon_random_angles_and_image(ctx):
    N, T = 129, 97
    say("radon random", N, T)
    th = np.random.default_rng(1).uniform(0, 180, T).astype(np.float32)
    x = interior_noise(N, 2)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
    assert rel(s, ref.radon(x, th, "f64")) < FBP_TOL


@pytest.mark.parametrize("N,T", [(65, 48), (257, 180), (1025, 333)])
def test_ramp_and_backprojection_match_oracle(ctx, N, T):
    say("ramp+bp", N, T)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    S = np.random.default_rng(N + T).normal(size=(T, N)).astype(np.float32)
    hb = HipBuffers()
    with plan(ctx, N, T, th) as b:
        d_in = hb.upload(S)
        d_f = hb.alloc(S.nbytes)
        d_img = hb.alloc(N * N * 4)
        b.ramp_dev(0, d_f, d_in)
        b.backproject_dev(0, d_img, d_f)
        ctx.synchronize()
        filt = hb.download(d_f, (T, N))
        img = hb.download(d_img, (N, N))
    hb.close()
    f64 = ref.ramp(S, "f64")
    assert rel(filt, f64) < 2e-7  # FP64 accumulation: only the output rounding remains
    assert rel(filt, ref.ramp(S, "f32")) < 1e-6
    # The backprojector computes positions as the reference does (FP32, rounded
    # at magnitude N), so on white noise it reproduces the reference's FP32
    # backprojection of the same filtered sinogram to FP32 rounding of the
    # sums, and sits at the reference's own distance from FP64 (2.5e-6 /
    # 1.0e-5 / 4.1e-5 here, on the reference's angle tables).
    bp32 = ref.backproject(filt, th, wpu, "f32")
    with ref.fp32_angles():
        bp64 = ref.backproject(filt.astype(np.float64), th, wpu, "f64")
    assert rel(img, bp32) < 1e-6
    assert rel(img, bp64) < 1.1 * rel(bp32, bp64) + 1e-6


@pytest.mark.parametrize("N,T", [(65, 48), (257, 180)])
def test_iradon_matches_oracle(ctx, N, T):
    say("iradon", N, T)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    s = ref.radon(disk_phantom(N, 10, 3), th, "f64").astype(np.float32)
    with plan(ctx, N, T, th) as b:
        img = b.iradon(s)
    assert rel(img, ref.iradon(s, th, wpu, "f64")) < FBP_TOL
    assert rel(img, ref.iradon(s, th, wpu, "f32")) < FBP_TOL


def cfg0_sinogram():
    """One 256x256 slice of 10 seeded ellipses, 180 angles, I0 = 1e4."""
    N, T = 257, 180
    th = synth.angles(T)
    x = synth.phantom(N, 10, 1000)
    s = ref.radon(x, th, "f64")
    return synth.poisson_noise(s, 1e4, N, 1000), th


def test_fbp_small_config_through_module_mirror(ctx):
    say("fbp cfg0 via Tomography")
    f, th = cfg0_sinogram()
    t = Tomography(ctx)
    t.nimage, t.ntheta, t.wpu = 257, 180, synth.default_wpu(257)
    t.int_radon()
    t.theta[:] = th
    t.f[:] = f
    img = t.get_slice()
    t.fin_radon()
    assert rel(img, ref.iradon(f, th, t.wpu, "f64")) < FBP_TOL
    assert rel(img, ref.iradon(f, th, t.wpu, "f32")) < FBP_TOL


def test_full_size_fbp_and_radon(ctx):
    """2049^2 x 1500 angles (the production slice): one radon and one FBP
    against the oracle (multi-threaded, tens of seconds).

    The FBP bar of 1e-5 is met against the reference's own FP32 arithmetic.
    Against FP64 the reference itself is 1.83e-5 away at this size: its FP32
    filter taps are rounded, and the ramp filter's cancellation (the result
    is ~N times smaller than h(0)*s) amplifies that rounding, which the
    backprojection amplifies again.  The GPU uses the reference's taps, so it
    sits at the same distance; the test bounds it by 2.5e-5."""
    N, T = 2049, 1500
    say("full-size radon+fbp", N, T)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    x = disk_phantom(N, 40, 1001)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
        img = b.iradon(s)
    assert rel(s, ref.radon(x, th, "f64")) < FBP_TOL
    assert rel(s, ref.radon(x, th, "f32")) < FBP_TOL
    assert rel(img, ref.iradon(s, th, wpu, "f32")) < FBP_TOL
    assert rel(img, ref.iradon(s, th, wpu, "f64")) < 2.5e-5


def test_full_size_tv_against_fp64(ctx):
    """The production slice (2049^2 x 1500, 40-ellipse phantom, Poisson
    I0 = 5e3): split-Bregman TV, 1 outer x 2 CG iterations, against the FP64
    oracle -- with the library defaults, with the autotuner's choices at
    this shape (LDS projector with 4 angles per lane, FFT ramp filter), with
    the cell-box projector and with backprojection tiles of 16 and 64 rows."""
    N, T = 2049, 1500
    say("full-size TV", N, T)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    x = synth.phantom(N, 40, 1001)
    results = {}
    with plan(ctx, N, T, th) as b:
        f = synth.poisson_noise(b.radon(x), 5e3, N, 1101)
        b.set_sino(0, f)
        b.tvopt(0.05, 0.1, 1, 1, 2, ACU_TV_GRAPH)
        results["default"] = b.get_image(0, ACU_IMG_TV)
        b.set_fp_variant(0, 4)
        b.set_ramp_variant(_lib.RAMP_FFT)
        b.tvopt(0.05, 0.1, 1, 1, 2, ACU_TV_GRAPH)
        results["tuned"] = b.get_image(0, ACU_IMG_TV)
        b.set_fp_variant(0, 8)  # cell-box projector
        b.tvopt(0.05, 0.1, 1, 1, 2, ACU_TV_GRAPH)
        results["cell16"] = b.get_image(0, ACU_IMG_TV)
        b.set_fp_variant(0, 4)
        for ty in (16, 64):  # backprojection tile heights
            b.set_bp_tile(ty)
            b.tvopt(0.05, 0.1, 1, 1, 2, ACU_TV_GRAPH)
            results[f"bp_tile{ty}"] = b.get_image(0, ACU_IMG_TV)
    ref.set_threads(16)
    try:
        tv64, _, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 1, 1, 2, "f64")
    finally:
        ref.set_threads(1)
    for k, img in results.items():
        e = rel(img, tv64)
        print("full-size TV", k, "rel-L2 vs FP64 %.2e" % e, flush=True)
        assert e < TV_TOL, (k, e)


@pytest.mark.parametrize("N,T,nouter", [(65, 48, 2), (129, 96, 5)])
def test_tvopt_matches_oracle(ctx, N, T, nouter):
    say("tvopt", N, T, nouter)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    f = ref.radon(disk_phantom(N, 6, 5), th, "f64").astype(np.float32)
    t = Tomography(ctx)
    t.nimage, t.ntheta, t.wpu = N, T, wpu
    t.mu, t.lambda_, t.nouter, t.ninnner, t.ncg = 0.05, 0.1, nouter, 1, 5
    t.int_radon()
    t.theta[:] = th
    t.f[:] = f
    tv, fbp = t.tvopt(1)
    delta = t.last_delta
    t.fin_radon()
    tv64, fbp64, d64 = ref.tvopt(f, th, wpu, 0.05, 0.1, nouter, 1, 5, "f64")
    assert rel(fbp, fbp64) < FBP_TOL
    assert rel(tv, tv64) < TV_TOL
    assert np.allclose(delta, d64, rtol=1e-3)


@pytest.mark.parametrize("mode", [ACU_TV_PLAIN, ACU_TV_GRAPH])
def test_tvopt_active_shrink_and_two_inner(ctx, mode):
    """lambda = 100 (tomo_run.pbs:21): the soft-threshold 1/lambda = 0.01 is
    active, unlike at lambda = 0.1; with two inner iterations per outer one.
    Against the FP64 oracle, and the shrink visibly changes the result."""
    N, T = 97, 72
    say("tvopt lambda=100 ninner=2", N, T, mode)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    f = ref.radon(disk_phantom(N, 8, 21), th, "f64").astype(np.float32)
    with plan(ctx, N, T, th) as b:
        b.set_sino(0, f)
        b.tvopt(0.05, 100.0, 3, 2, 4, mode)
        tv = b.get_image(0, ACU_IMG_TV)
        b.tvopt(0.05, 1e-3, 3, 2, 4, mode)
        tv_weak = b.get_image(0, ACU_IMG_TV)
    tv64, _, _ = ref.tvopt(f, th, wpu, 0.05, 100.0, 3, 2, 4, "f64")
    assert rel(tv, tv64) < TV_TOL, rel(tv, tv64)
    assert rel(tv, tv_weak) > 1e-3  # the shrink changed u


def test_tvopt_small_config(ctx):
    """cfg0: 256^2 (N=257), 180 angles, I0=1e4, mu=0.05, lambda=0.1, 20 x 5."""
    say("tvopt cfg0 20x5")
    f, th = cfg0_sinogram()
    wpu = synth.default_wpu(257)
    with plan(ctx, 257, 180, th) as b:
        b.set_sino(0, f)
        b.tvopt(0.05, 0.1, 20, 1, 5, ACU_TV_GRAPH)
        tv = b.get_image(0, ACU_IMG_TV)
        fbp = b.get_image(0, ACU_IMG_FBP)
    tv64, fbp64, _ = ref.tvopt(f, th, wpu, 0.05, 0.1, 20, 1, 5, "f64")
    assert rel(fbp, fbp64) < FBP_TOL
    assert rel(tv, tv64) < TV_TOL


def test_tv_modes_are_bit_identical(ctx):
    N, T = 97, 64
    say("tv modes", N, T)
    th = synth.angles(T)
    f = ref.radon(disk_phantom(N, 6, 9), th, "f64").astype(np.float32)
    outs = []
    for mode in (ACU_TV_PLAIN, ACU_TV_BATCHED, ACU_TV_GRAPH, ACU_TV_GRAPH):
        with plan(ctx, N, T, th) as b:
            b.set_sino(0, f)
            b.tvopt(0.05, 0.1, 4, 1, 5, mode)
            outs.append((b.get_image(0, ACU_IMG_TV), b.get_delta(0, 4)))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.array_equal(o[1], outs[0][1])


def test_batch_of_uneven_slices_equals_single_plans(ctx):
    shapes = [(65, 48), (97, 60), (129, 37), (65, 100)]
    say("batch uneven", shapes)
    rng = np.random.default_rng(7)
    items = []
    for i, (N, T) in enumerate(shapes):
        th = np.sort(rng.uniform(0, 180, T)).astype(np.float32)
        f = ref.radon(disk_phantom(N, 5, 20 + i), th, "f64").astype(np.float32)
        items.append((N, T, synth.default_wpu(N), th, f))
    with Batch(ctx, [(N, T, w, th) for N, T, w, th, _ in items]) as b:
        for i, it in enumerate(items):
            b.set_sino(i, it[4])
        b.tvopt(0.05, 0.1, 3, 2, 4, ACU_TV_GRAPH)
        batched = [b.get_image(i, ACU_IMG_TV) for i in range(len(items))]
        fbps = [b.get_image(i, ACU_IMG_FBP) for i in range(len(items))]
    for i, (N, T, w, th, f) in enumerate(items):
        with plan(ctx, N, T, th, w) as p:
            p.set_sino(0, f)
            p.tvopt(0.05, 0.1, 3, 2, 4, ACU_TV_PLAIN)
            assert np.array_equal(p.get_image(0, ACU_IMG_TV), batched[i])
            assert np.array_equal(p.get_image(0, ACU_IMG_FBP), fbps[i])
        tv64, _, _ = ref.tvopt(f, th, w, 0.05, 0.1, 3, 2, 4, "f64")
        assert rel(batched[i], tv64) < TV_TOL


def test_zero_sinogram_gives_zero_not_nan(ctx):
    """The reference divides 0/0 in its CG step size (tv_hpc.F90:277) on an
    all-zero slice; this library takes alpha = 0 there (documented deviation)."""
    N, T = 33, 16
    say("zero sinogram", N, T)
    th = synth.angles(T)
    with plan(ctx, N, T, th) as b:
        b.set_sino(0, np.zeros((T, N), np.float32))
        b.tvopt(0.05, 0.1, 3, 1, 5, ACU_TV_GRAPH)
        tv = b.get_image(0, ACU_IMG_TV)
    assert np.all(tv == 0.0)


@pytest.mark.parametrize("N,T,wpu", [(3, 1, 1), (5, 2, 2), (33, 1, 2), (31, 7, 2)])
def test_tiny_and_degenerate_shapes(ctx, N, T, wpu):
    """Special angles (exactly 0, 90, 180 degrees) and images that are nonzero
    up to the border.  With wpu >= 2 the mask keeps every backprojected pixel
    away from the detector edge, so the backprojection has no bounds-test
    ambiguity; with wpu = 1 it reaches the edge, where the reference's FP32
    rounding decides (N = 3 has no unmasked pixel)."""
    say("tiny", N, T, wpu)
    th = np.array([0.0, 90.0, 45.0, 180.0, 135.0, 10.0, 170.0][:T], np.float32)
    x = np.random.default_rng(N).normal(size=(N, N)).astype(np.float32)
    with plan(ctx, N, T, th, wpu) as b:
        s = b.radon(x)
        img = b.iradon(s)
    # angles of exactly 0/90/180 put samples on the edge of the bounds test,
    # where only the FP32 evaluation order decides; the GPU's sample
    # positions are the FP32 reference's, so it matches that one
    assert rel(s, ref.radon(x, th, "f32")) < 1e-6
    assert rel(img, ref.iradon(s, th, wpu, "f64")) < 1e-5


@pytest.mark.parametrize("N,T", [(64, 48), (256, 180), (130, 37)])
def test_even_image_side(ctx, N, T):
    """Even N (the drivers pad detectors to odd widths, but the module takes
    any N): the midpoint (N+1)/2 is then N/2 by integer division, off the
    image centre.  radon, FBP and TV against the oracle as for odd N."""
    say("even N", N, T)
    th = synth.angles(T) if T != 37 else np.sort(np.random.default_rng(T).uniform(0, 180, T)).astype(np.float32)
    wpu = synth.default_wpu(N)
    x = disk_phantom(N, 6, N + T)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
        img = b.iradon(s)
        b.set_sino(0, s)
        b.tvopt(0.05, 0.1, 2, 1, 4, ACU_TV_GRAPH)
        tv = b.get_image(0, ACU_IMG_TV)
    assert rel(s, ref.radon(x, th, "f64")) < FBP_TOL
    assert rel(s, ref.radon(x, th, "f32")) < FBP_TOL
    assert rel(img, ref.iradon(s, th, wpu, "f64")) < FBP_TOL
    tv64, _, _ = ref.tvopt(s, th, wpu, 0.05, 0.1, 2, 1, 4, "f64")
    assert rel(tv, tv64) < TV_TOL


@pytest.mark.parametrize("N,T", [(129, 97), (257, 400)])
def test_unsorted_repeated_and_out_of_range_angles(ctx, N, T):
    """The module takes any angle list: unsorted, with repeats, below 0 and
    above 180 degrees (the projector's angle blocks and box sizes follow the
    table as given)."""
    say("angle set", N, T)
    rng = np.random.default_rng(N * T)
    th = rng.uniform(-60.0, 240.0, T).astype(np.float32)
    th[T // 3:T // 3 + 5] = th[0]          # repeats
    th[-3:] = (-90.0, 270.0, 360.0)          # special angles beyond the usual range
    wpu = synth.default_wpu(N)
    x = disk_phantom(N, 6, N + 5)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
        img = b.iradon(s)
        b.set_sino(0, s)
        b.tvopt(0.05, 0.1, 1, 1, 3, ACU_TV_GRAPH)
        tv = b.get_image(0, ACU_IMG_TV)
    assert rel(s, ref.radon(x, th, "f64")) < FBP_TOL
    assert rel(img, ref.iradon(s, th, wpu, "f64")) < FBP_TOL
    tv64, _, _ = ref.tvopt(s, th, wpu, 0.05, 0.1, 1, 1, 3, "f64")
    assert rel(tv, tv64) < TV_TOL


@pytest.mark.parametrize("N,T", [(4097, 24), (8193, 8)])
def test_large_images(ctx, N, T):
    """Image sides beyond the production size: 4097 (the largest the FFT
    filter takes) and 8193 (Toeplitz filter only), radon and FBP against the
    oracle, with the library's defaults and with the FFT filter where it
    applies.  The reference's FP32 taps, positions and sums put its own FBP
    5.5e-5 (4097) and 7.7e-5 (8193) from FP64; the GPU computes positions as
    the reference does, so its FBP is within 1e-5 of the reference's and no
    more than 10% further from FP64."""
    say("large", N, T)
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    x = disk_phantom(N, 12, N)
    with plan(ctx, N, T, th) as b:
        s = b.radon(x)
        img = b.iradon(s)
        fft = None
        if N <= 4097:
            b.set_ramp_variant(_lib.RAMP_FFT)
            fft = b.iradon(s)
    ref.set_threads(16)
    try:
        s64 = ref.radon(x, th, "f64")
        img64 = ref.iradon(s, th, wpu, "f64")
        img32 = ref.iradon(s, th, wpu, "f32")
        bar = max(FBP_TOL, 1.1 * rel(img32, img64))
    finally:
        ref.set_threads(1)
    assert rel(s, s64) < FBP_TOL
    say("FBP vs FP64", rel(img, img64), "vs reference FP32", rel(img, img32), "bar", bar)
    for got in (img, fft):
        if got is not None:
            assert rel(got, img32) < FBP_TOL
            assert rel(got, img64) <= bar, rel(got, img64)


def test_extreme_mixed_batch_equals_single_plans(ctx):
    """Slices of N = 3, 65 and 2049 (1, 1000 and 40 angles) in one batch:
    graph-replayed TV equals the single-slice plans bit for bit."""
    shapes = [(3, 1, 1), (2049, 40, synth.default_wpu(2049)), (65, 1000, synth.default_wpu(65))]
    say("extreme mix", shapes)
    items = []
    for i, (N, T, w) in enumerate(shapes):
        th = synth.angles(T)
        f = ref.radon(disk_phantom(N, 4, 70 + i) if N > 8 else np.ones((N, N), np.float32), th, "f64")
        items.append((N, T, w, th, f.astype(np.float32)))
    with Batch(ctx, [(N, T, w, th) for N, T, w, th, _ in items]) as b:
        for i, it in enumerate(items):
            b.set_sino(i, it[4])
        b.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_GRAPH)
        batched = [b.get_image(i, ACU_IMG_TV) for i in range(len(items))]
    for i, (N, T, w, th, f) in enumerate(items):
        with plan(ctx, N, T, th, w) as p:
            p.set_sino(0, f)
            p.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_PLAIN)
            assert np.array_equal(p.get_image(0, ACU_IMG_TV), batched[i]), N


def test_contexts_on_two_threads_equal_sequential_runs():
    """One context per host thread: two threads running batches at once give
    the bits the same calls give one after another."""
    import threading
    from acumen_amd import Context
    cases = [(257, 180, 11), (129, 360, 12)]
    sinos = {(N, T, seed): [ref.radon(disk_phantom(N, 5, seed + i), synth.angles(T), "f64").astype(np.float32)
                            for i in range(2)] for N, T, seed in cases}

    def run(N, T, seed, out, k):
        th = synth.angles(T)
        with Context(0) as c, Batch(c, [(N, T, synth.default_wpu(N), th)] * 2) as b:
            for i in range(2):
                b.set_sino(i, sinos[(N, T, seed)][i])
            b.tvopt(0.05, 0.1, 2, 1, 3, ACU_TV_GRAPH)
            out[k] = [b.get_image(i, ACU_IMG_TV) for i in range(2)] + [b.get_image(i, ACU_IMG_FBP) for i in range(2)]

    say("two threads")
    seq = [None, None]
    for k, cse in enumerate(cases):
        run(*cse, seq, k)
    par = [None, None]
    threads = [threading.Thread(target=run, args=(*cse, par, k)) for k, cse in enumerate(cases)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert all(p is not None for p in par), "a thread did not finish"
    for a, b in zip(seq, par):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_rejects_bad_arguments(ctx):
    from acumen_amd import AcumenError
    with pytest.raises(AcumenError):
        Batch(ctx, [(64, 10, 0, None)])  # wpu = 0 divides by zero in the mask
    with pytest.raises(AcumenError):
        Batch(ctx, [(2, 10, 1, None)])
    with plan(ctx, 33, 10, None) as b:
        with pytest.raises(AcumenError):
            b.radon(np.zeros((33, 33), np.float32))  # theta never set
        b.set_theta(0, synth.angles(10))
        with pytest.raises(AcumenError):
            b.get_image(0, ACU_IMG_TV)  # nothing computed yet
        with pytest.raises(AcumenError):
            b.tvopt(0.05, 0.1, 1, 1, 61)
