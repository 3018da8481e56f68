#!/usr/bin/env python3
"""Headline benchmark: split-Bregman TV slices/s at 2048^2 on one MI355X.

Workload (one slice): 2048 detector bins (odd-padded to N = 2049) x 1500
angles on [0, 180), a 2048^2 phantom of 40 seeded ellipses, Poisson noise at
I0 = 5e3, Ram-Lak FBP warm start then split-Bregman TV with mu = 0.05,
lambda = 0.1, 50 outer x 1 inner x 5 CG iterations -- FP32 operators, FP64
CG scalars.  A step is that whole reconstruction for a batch of `--batch`
slices (default 8, ~7.3 s) resident in HBM, run by the batched path (every
kernel launch covers all slices, CG scalars stay on the device).  `value`
is slices/s over the timed steps, with HIP events bracketing every operator
launch on the library's stream (they give the roofline's kernel times); the
warm-up steps replay each outer iteration from a HIP graph instead and
report their time beside it.  After the timed steps, each within what is
left of --budget-s: FBP slices/s of the batched path; the plain per-slice
GPU path (one slice at a time, each reduction returned to the host before
the next kernel, as the reference is structured) with the same operators
as the batched path (`batched_vs_plain`: batching alone) and with the
reference's 7 radon + 7 iradon per outer iteration; the CPU baseline.

roofline: the forward projector dominates TV time; its algorithmic work is
14 FP32 flops per ray sample (2 coordinates, 2 fractions, 3 lerps, 1
accumulate) over N*N*T samples per slice.  It runs on the FP32 VALU with
its bilinear gathers from LDS, whose bank conflicts bound it, so `bound` is
"lds" (the `lds` object gives its LDS read rate), `frac` is against the
published FP32 vector peak (157.3 TFLOP/s, packed issue) and
`frac_nonpacked` against the 78.6 TFLOP/s of one FMA per lane and clock,
the rate of code built without packed FP32.  `traffic` is the HBM bytes of
a launch from the rocprofv3 counter summary under profiles/.
`roofline_backprojection` gives the same figures for the backprojector
(6 flop per pixel-angle update over N*N*T updates, compulsory bytes
4(N*T + 2*N*N) per slice) and its LDS read rate: 8 bytes per update over
the updates of the tiles that hold mask pixels (the others only store
zeros).

cpu_baseline: the oracle's FP32 restatement of the reference's loops (the
reference's own arithmetic, bit for bit) times tvopt on one slice of the
workload with nouter = 0 and nouter = 1 on the host cores (at most 16);
a slice costs t(0) + 50 (t(1) - t(0)).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N_DET = 2048
T_ANG = 1500
MU, LAM, NOUTER, NINNER, NCG = 0.05, 0.1, 50, 1, 5
FP_FLOPS_PER_SAMPLE = 14
BP_FLOPS_PER_UPDATE = 6  # coordinate (2), fraction, lerp (2), accumulate
GRAD_LAM, GRAD_EPS = 0.1, 1e-3  # differentiation workload: TV weight and smoothing
FP32_PEAK_TFLOPS = 157.3       # MI355X's published FP32 vector peak (packed FP32 issue)
FP32_NONPACKED_TFLOPS = 78.6   # one FP32 FMA per lane and clock: the rate of non-packed code
LDS_BYTES_PER_SAMPLE = 16  # the 2x2 cell of a bilinear sample
N_CU, CLOCK_GHZ = 256, 2.4


T_START = time.perf_counter()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_inputs(ctx, batch, N, T, seed0):
    """Phantoms -> GPU radon -> Poisson noise (host), per slice."""
    import numpy as np

    from acumen_amd import Batch, synth
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    sinos = []
    with Batch(ctx, [(N, T, wpu, th)]) as p:
        for i in range(batch):
            x = synth.phantom(N, 40, seed0 + i)
            s = p.radon(x)
            sinos.append(synth.poisson_noise(s, 5e3, N, seed0 + 100 + i))
    return th, wpu, sinos


def fp_kernel_name(variant=None):
    """The forward-projection kernel of a projector variant (the autotuner's
    choice, else ACUMEN_FP_PATH)."""
    if variant is None:
        variant = "gather" if os.environ.get("ACUMEN_FP_PATH") == "gather" else "lds4"
    return {"gather": "fp_kernel", "cell16": "fplc_kernel"}.get(variant, "fpl_kernel")


def traffic_from_profile(name=None, key="hbm_bytes_per_launch_per_slice"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 counter
    summary under profiles/, if one is present (default: the forward
    projector at 2049 x 1500, per slice)."""
    name = name or f"{fp_kernel_name()}_traffic_2049x1500.txt"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def bp_active_fraction(N, wpu, ty):
    """Fraction of the N*N pixels the backprojector computes: its 64 x ty
    tiles that hold a pixel of nonzero mask (tomography/tv_hpc.F90:58-74,
    nonzero where 1 - 2(r - rad_n)/wpu > 0)."""
    import numpy as np
    half = (np.float32(N) - 1) / 2
    d = (np.arange(1, N + 1, dtype=np.float32) - half) - 1
    r = np.sqrt(d[None, :] ** 2 + d[:, None] ** 2)
    rad_n = np.float32((N - 1) / 2) - np.float32(wpu)
    on = (1 - (2 * (r - rad_n)) / np.float32(wpu)) > 0
    nty, ntx = -(-N // ty), -(-N // 64)
    pad = np.zeros((nty * ty, ntx * 64), bool)
    pad[:N, :N] = on
    tiles = pad.reshape(nty, ty, ntx, 64).any(axis=(1, 3))
    return float(tiles.sum()) * 64 * ty / (N * N)


def cpu_threads(requested: int = 0) -> int:
    """Threads for a CPU leg: the host cores this process may run on, at most
    16 (one GPU's share of a shared host), unless a count is requested."""
    if requested > 0:
        return requested
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline_tv(ctx, sino, th, wpu, threads):
    """CPU baseline of the headline: the oracle's FP32 restatement of the
    reference's loops (bit for bit the reference's own arithmetic) runs tvopt
    on one 2049^2 x 1500 slice with nouter = 0 (the FBP warm start) and
    nouter = 1 (one complete outer iteration: 7 radon, 7 iradon and the
    elementwise updates, tomography/tv_hpc.F90:205-353), its operators split
    over `threads` host threads (each output element stays one sequential
    sum).  A slice of the headline costs t(0) + NOUTER * (t(1) - t(0)): the
    only extrapolation is in the outer-iteration count, in which the
    reference's work is linear.  Also the checker: the GPU's FBP and one
    outer iteration of the same slice against the oracle's results -- the
    reference's FP32 arithmetic, and FP32 operators with FP64 CG scalars
    (the GPU's precision policy: the reference's sequential FP32 dot
    products over N*N pixels move its TV ~2e-4 from FP64 at N = 2049 after
    one outer iteration, which FP64 scalars do not)."""
    import numpy as np

    from acumen_amd import ACU_IMG_FBP, ACU_IMG_TV, ACU_TV_PLAIN, Batch
    from oracle import ref

    N = sino.shape[1]
    ref.set_threads(threads)
    try:
        t = time.perf_counter()
        _, fbp0, _ = ref.tvopt(sino, th, wpu, MU, LAM, 0, NINNER, NCG, "f32")
        t_fbp = time.perf_counter() - t
        t = time.perf_counter()
        tv1, _, _ = ref.tvopt(sino, th, wpu, MU, LAM, 1, NINNER, NCG, "f32")
        t_one = time.perf_counter() - t
        tv1d, _, _ = ref.tvopt(sino, th, wpu, MU, LAM, 1, NINNER, NCG, "f32d")
    finally:
        ref.set_threads(1)
    per_slice = t_fbp + NOUTER * (t_one - t_fbp)
    with Batch(ctx, [(N, sino.shape[0], wpu, th)]) as one:
        one.set_sino(0, sino)
        one.tvopt(MU, LAM, 1, NINNER, NCG, ACU_TV_PLAIN)
        g_tv, g_fbp = one.get_image(0, ACU_IMG_TV), one.get_image(0, ACU_IMG_FBP)

    def rel(a, b):
        return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64)))

    try:
        nproc = os.cpu_count()
    except Exception:
        nproc = None
    return {"value": round(1.0 / per_slice, 6), "unit": "slices/s", "cores": threads, "kind": "port",
            "sample": (f"oracle FP32 tvopt of one {N}^2 x {sino.shape[0]} slice on {threads} threads "
                       f"(os.cpu_count() = {nproc}): nouter=0 (FBP) {t_fbp:.1f} s, nouter=1 {t_one:.1f} s; "
                       f"slice = t(0) + {NOUTER} x (t(1) - t(0)) = {per_slice:.0f} s"),
            "fbp_s": round(t_fbp, 2), "outer_iteration_s": round(t_one - t_fbp, 2),
            "gpu_vs_cpu_fbp_rel_l2": float(f"{rel(g_fbp, fbp0):.3e}"),
            "gpu_vs_cpu_tv_1_outer_rel_l2": float(f"{rel(g_tv, tv1):.3e}"),
            "gpu_vs_fp32_operators_fp64_scalars_tv_1_outer_rel_l2": float(f"{rel(g_tv, tv1d):.3e}"),
            "cpu_fp32_vs_fp64_scalars_tv_1_outer_rel_l2": float(f"{rel(tv1, tv1d):.3e}")}


def headline_accuracy(ctx, batch, th, wpu, threads):
    """Checker of the headline's accuracy (part of the CPU baseline leg): the
    timed batch's configuration (schedule, autotuned kernels) runs tvopt to 20
    and 50 outer iterations on the slice whose reference results
    tests/golden/headline_2049x1500.npz keeps -- the reference's own FP32
    arithmetic (the oracle's FP32 build) and FP64, as 64 Rademacher
    projections per image.  The slice's sinogram is the oracle's FP32 radon
    of the 40-ellipse phantom (seed 1001) with Poisson noise (I0 = 5e3, seed
    1101), regenerated here and matched to the fixture; rel-L2 distances
    are estimated from the projections (~13% spread)."""
    import numpy as np

    from acumen_amd import ACU_IMG_TV, ACU_TV_GRAPH, Batch, synth
    from oracle import ref
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import rel_projected, sign_projections

    g = np.load(os.path.join(ROOT, "tests", "golden", "headline_2049x1500.npz"))
    N, T = int(g["N"]), int(g["T"])
    ref.set_threads(threads)
    try:
        sino = synth.poisson_noise(ref.radon(synth.phantom(N, 40, 1001), th, "f32"), 5e3, N, 1101)
    finally:
        ref.set_threads(1)
    if not np.array_equal(sign_projections(sino), g["sino"]):
        return {"error": "the regenerated input differs from the fixture's"}
    out = {}
    sched, refresh = batch.tv_schedule()
    with Batch(ctx, [(N, T, wpu, th)]) as one:
        one.set_ramp_variant(batch.ramp_variant())
        one.set_fp_variant(0, batch.fp_variant(0))
        one.set_bp_tile(batch.bp_tile())
        one.set_tv_schedule(sched, refresh)
        one.set_sino(0, sino)
        for n in (20, 50):
            one.tvopt(MU, LAM, n, NINNER, NCG, ACU_TV_GRAPH)
            p = sign_projections(one.get_image(0, ACU_IMG_TV))
            out[f"tv_{n}_outer"] = {
                "vs_fp64": float(f"{rel_projected(p, g[f'tv_f64_{n}']):.3e}"),
                "vs_reference_fp32": float(f"{rel_projected(p, g[f'tv_f32_{n}']):.3e}"),
                "reference_fp32_vs_fp64": float(f"{rel_projected(g[f'tv_f32_{n}'], g[f'tv_f64_{n}']):.3e}")}
    return out


def run_mixed(args) -> int:
    """64 slices, widths from {512, 768, 1024, 1536, 2048} and U{360..1500}
    angles (seeded), 20-40 ellipses, I0 ~ U[2e3, 2e4], TV as the headline:
    the batched path on all slices in one batch against the plain per-slice
    path on the same slices.  slices/s over the timed steps."""
    import numpy as np

    from acumen_amd import ACU_IMG_TV, ACU_TV_BATCHED, ACU_TV_GRAPH, ACU_TV_PLAIN, Batch, Context, _lib, synth

    specs = synth.mixed_batch(64, 1002)
    ctx = Context(0)
    log(f"inputs: {len(specs)} mixed slices")
    sinos = []
    for sp in specs:
        with Batch(ctx, [(sp["N"], sp["T"], sp["wpu"], sp["theta"])]) as p:
            s = p.radon(synth.phantom(sp["N"], sp["count"], sp["seed"]))
        sinos.append(synth.poisson_noise(s, sp["I0"], sp["N"], sp["seed"] + 500))
    shapes = [(sp["N"], sp["T"], sp["wpu"], sp["theta"]) for sp in specs]
    batch = Batch(ctx, shapes)
    for i, f in enumerate(sinos):
        batch.set_sino(i, f)
    ctx.synchronize()
    tune_s, choices = None, None
    if not args.no_tune:
        log("autotune: forward-projection variant per slice shape")
        t = time.perf_counter()
        rep = batch.autotune(reps=2)
        tune_s = time.perf_counter() - t
        choices = {}
        for sl in rep["slices"]:
            choices[sl["choice"]] = choices.get(sl["choice"], 0) + 1
        choices["ramp"] = rep["ramp"]["choice"]
        choices["bp_tile"] = rep["bp"]["choice"]
    for w in range(args.warmup):
        log(f"warmup step {w + 1}/{args.warmup}")
        batch.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_GRAPH)
        ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_GRAPH)
        ctx.synchronize()
        log(f"step {k + 1}/{args.steps} done at {time.perf_counter() - t0:.2f} s")
    elapsed = time.perf_counter() - t0
    value = len(specs) * args.steps / elapsed
    graph_imgs = [batch.get_image(i, ACU_IMG_TV).view(np.uint32).copy() for i in range(len(specs))]
    assert all(np.all(np.isfinite(g.view(np.float32))) for g in graph_imgs)

    def same_as_graph() -> bool:
        return all(np.array_equal(batch.get_image(i, ACU_IMG_TV).view(np.uint32), g)
                   for i, g in enumerate(graph_imgs))

    # kernel breakdown of one batched step (events; graph replay is off while timing)
    ctx.set_timing(True)
    batch.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_BATCHED)
    ctx.synchronize()
    kt = {name: ctx.kernel_time(kind) for name, kind in
          (("ramp", _lib.KIND_RAMP), ("bp", _lib.KIND_BP), ("cells", _lib.KIND_CELLS),
           ("fp", _lib.KIND_FP))}
    ctx.set_timing(False)
    # full-size property: the three modes give the same bits on every slice
    modes_equal = {"batched": same_as_graph()}
    plain_rate = None
    if not args.no_plain:
        log("plain per-slice path: all slices, one after another")
        ctx.synchronize()
        t2 = time.perf_counter()
        batch.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_PLAIN)
        ctx.synchronize()
        plain_rate = len(specs) / (time.perf_counter() - t2)
        modes_equal["plain"] = same_as_graph()
    batch.close()
    ctx.close()
    widths = {}
    for sp in specs:
        widths[sp["N"] - 1] = widths.get(sp["N"] - 1, 0) + 1
    out = {
        "metric": "TV-SB slices/s, 64 mixed-size slices in one batch, batched vs plain",
        "value": round(value, 5),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "split-Bregman TV, 64 slices, widths from {512,768,1024,1536,2048} and "
                               "U{360..1500} angles (seed 1002), 20-40 ellipses, I0~U[2e3,2e4], "
                               f"mu={MU}, lambda={LAM}, {args.nouter} outer x {NCG} CG, graph-replayed "
                               "batched path",
                   "slices_per_step": len(specs), "width_counts": widths,
                   "total_N2T": float(sum(sp["N"] ** 2 * sp["T"] for sp in specs))},
        "plain_slices_per_s": None if plain_rate is None else round(plain_rate, 5),
        "batched_vs_plain": None if plain_rate is None else round(value / plain_rate, 3),
        "tv_equals_graph_bitwise": modes_equal,
        "kernel_ms_per_step": {k: round(v[0], 2) for k, v in kt.items()},
        "autotune_s": None if tune_s is None else round(tune_s, 2),
        "fp_variant_counts": choices,
    }
    print(json.dumps(out), flush=True)
    return 0


def volume_cpu_baseline(sinos, th, wpu, samples, fbp, variants, threads):
    """CPU baseline of the volume workload: the oracle's FP32 FBP (the
    reference's loops) of the sampled slices on `threads` host threads,
    timed; as the checker, the streamed FBP and every ramp arithmetic's FBP
    of those slices against the oracle's FP64 FBP."""
    import numpy as np

    from oracle import ref

    ref.set_threads(threads)
    rel = lambda a, w: float(np.linalg.norm(a - w) / np.linalg.norm(w))
    acc, wall = {}, 0.0
    try:
        for k in samples:
            f64 = ref.iradon(sinos[k].astype(np.float64), th, wpu, "f64")
            t0 = time.time()
            f32 = ref.iradon(sinos[k], th, wpu, "f32")
            wall += time.time() - t0
            row = {"stream_vs_f64": rel(fbp[k], f64), "stream_vs_ref_f32": rel(fbp[k], f32),
                   "ref_f32_vs_f64": rel(f32, f64)}
            for (kk, name), img in variants.items():
                if kk == k:
                    row[name + "_vs_f64"] = rel(img, f64)
            acc[str(k)] = {a: float(f"{b:.3e}") for a, b in row.items()}
    finally:
        ref.set_threads(1)
    cpu = {"value": round(len(samples) / wall, 4), "unit": "slices/s", "cores": threads, "kind": "port",
           "sample": f"FBP (ramp filter + backprojection) of {len(samples)} sampled slices by the oracle's "
                     f"FP32 loops on {threads} threads, {wall:.1f} s"}
    return cpu, acc


def run_volume(args) -> int:
    """A 2048^2 x 256-slice volume (2048 detectors -> N = 2049, 1500 angles,
    Poisson I0 = 5e3; 16 distinct phantoms cycled) streamed from host memory
    through the GPU: FBP of every slice, FBP + 30-iteration TV of the first
    --tv-slices, rates PCIe and host staging included.  FBP of 4 sampled
    slices is compared with the FP64 and FP32 oracles for every ramp-filter
    arithmetic."""
    import numpy as np

    from acumen_amd import Batch, Context, _lib, synth, volume

    N, T = synth.padded_size(N_DET), T_ANG
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    ctx = Context(0)
    distinct = 16
    log(f"inputs: {distinct} distinct slices of {N}x{N}, {T} angles, cycled to {args.slices}")
    base = []
    with Batch(ctx, [(N, T, wpu, th)]) as p:
        for i in range(distinct):
            s = p.radon(synth.phantom(N, 40, 4000 + i))
            base.append(synth.poisson_noise(s, 5e3, N, 4100 + i))
    sinos = np.stack([base[i % distinct] for i in range(args.slices)])
    log("stream: FBP of every slice")
    _ = volume.reconstruct(ctx, sinos[:min(8, args.slices)], th, wpu, tv=None, batch=8, tune=True)  # warm-up
    fbp, _, st_fbp = volume.reconstruct(ctx, sinos, th, wpu, tv=None, batch=16, tune=True)
    log(f"stream: FBP + TV(30) of {args.tv_slices} slices")
    tvp = dict(mu=MU, lam=LAM, nouter=30, ninner=NINNER, ncg=NCG)
    _, tv, st_tv = volume.reconstruct(ctx, sinos[:args.tv_slices], th, wpu, tv=tvp, batch=8, tune=True)
    assert np.all(np.isfinite(tv))
    # FBP of 4 sampled slices with every ramp-filter arithmetic (the product)
    names = {_lib.RAMP_FP64: "fp64", _lib.RAMP_FP32: "fp32", _lib.RAMP_BF16X3: "bf16x3",
             _lib.RAMP_BF16X6: "bf16x6", _lib.RAMP_FFT: "fft"}
    samples = [0, args.slices // 3, (2 * args.slices) // 3, args.slices - 1]
    variants = {}
    with Batch(ctx, [(N, T, wpu, th)]) as b:
        for k in samples:
            for v, name in names.items():
                b.set_ramp_variant(v)
                variants[(k, name)] = b.iradon(sinos[k])
    cpu, acc = None, None
    if not args.no_cpu:
        log("cpu baseline: the oracle's FBP of 4 sampled slices; accuracy against its FP64 form")
        cpu, acc = volume_cpu_baseline(sinos, th, wpu, samples, fbp, variants, cpu_threads(args.cpu_threads))
    ctx.close()
    out = {
        "metric": "FBP slices/s streamed (PCIe incl.), 2048^2 x 1500 x 256-slice volume; FBP+TV(30) slices/s",
        "value": round(st_fbp["slices_per_s"], 3),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(st_fbp["ms"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"volume of {args.slices} slices of {N_DET}^2 (N={N}) x {T} angles, "
                               "I0=5e3, streamed from host memory in batches of 16 on two lanes, "
                               "autotuned", "slices": args.slices, "tv_slices": args.tv_slices},
        "fbp_stream_slices_per_s": round(st_fbp["slices_per_s"], 3),
        "fbp_stream_after_setup_slices_per_s": round(st_fbp["stream_slices_per_s"], 3),
        "setup_ms": round(st_fbp["setup_ms"], 1),
        "tv30_stream_slices_per_s": round(st_tv["slices_per_s"], 4),
        "fbp_accuracy_sampled": acc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def prep_cpu_baseline(proj, sino, th, sh, ks):
    """CPU baseline of the preprocessing workload: the oracle's 3x3x3 median
    and correction of projections ks (one host thread), timed; as the
    checker, the GPU's sinogram rows of those projections bit for bit."""
    import numpy as np

    from oracle import ref

    exact, wall = True, 0.0
    for k in ks:
        t0 = time.time()
        row = ref.prep_correct(ref.prep_median27(proj[k - 1], proj[k], proj[k + 1]), th, sh)
        wall += time.time() - t0
        exact = exact and bool(np.array_equal(sino[:, k, :], row))
    n = proj.shape[1]
    cpu = {"value": round(len(ks) / wall, 3), "unit": "projections/s", "cores": 1, "kind": "port",
           "sample": f"median + correction of {len(ks)} projections of {n}^2 by the oracle, one thread, "
                     f"{wall:.1f} s"}
    return cpu, exact


def run_prep(args) -> int:
    """Projection -> sinogram preprocessing at 2048^2 projections: the
    centre-of-rotation / tilt search on the 0 and 180 degree projections,
    and the median filter + correction of 64 projections into sinogram rows.
    value: the kernels' rate (their HIP-event time); the host -> GPU -> host
    pipeline rate, PCIe and pageable copies included, is reported beside it.  cpu_baseline: the oracle's
    median + correction of a sample of projections on one host thread, which
    also checks the GPU's sinogram rows of that sample bit for bit."""
    import numpy as np

    from acumen_amd import Context, _lib, prep, synth

    n, count = N_DET, 64
    ctx = Context(0)
    rng = np.random.default_rng(3000)
    base = synth.phantom(n, 40, 3001)  # stands in for one detector image
    proj = np.stack([base + rng.normal(0, 0.01, base.shape).astype(np.float32) for _ in range(count)])
    # 180 degree projection: the mirror image about a column 3.4 px right of the centre
    x = np.arange(n, dtype=np.float64)
    src = (n - 1) + 2 * 3.4 - x
    i0 = np.clip(np.floor(src).astype(int), 0, n - 1)
    i1 = np.clip(i0 + 1, 0, n - 1)
    w = (src - np.floor(src)).astype(np.float32)
    p180 = (proj[0][:, i0] * (1 - w) + proj[0][:, i1] * w).astype(np.float32)
    log("centre search at", n)
    t = time.perf_counter()
    th, sh, evals = prep.find_center(ctx, proj[0], p180)
    t_search = time.perf_counter() - t
    log("median + correction of", count, "projections")
    prep.sinograms(ctx, proj[:4], th, sh)  # warm-up
    ctx.set_timing(True)
    t = time.perf_counter()
    sino = prep.sinograms(ctx, proj, th, sh)
    t_pipe = time.perf_counter() - t
    k_ms, _ = ctx.kernel_time(_lib.KIND_PREP)
    ctx.set_timing(False)
    ctx.close()
    # compulsory HBM bytes per projection: read it, write its median, read the
    # median, write the corrected sinogram rows
    prep_bytes = 4 * 4.0 * n * n * count
    cpu, exact = None, None
    if not args.no_cpu:
        log("cpu baseline: the oracle's median + correction of 4 projections")
        cpu, exact = prep_cpu_baseline(proj, sino, th, sh, [1, count // 3, count // 2, count - 2])
    out = {
        "metric": "projections/s through median filter + correction into sinogram rows, 2048^2 "
                  "(kernels on resident data; the host-to-host pipeline rate beside it)",
        "value": round(count / (k_ms * 1e-3), 1),
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(k_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{count} projections of {n}^2, axis 3.4 px off centre", "n": n,
                   "projections": count},
        "pipeline_projections_per_s_pcie": round(count / t_pipe, 2),
        "pipeline_ms": round(t_pipe * 1e3, 2),
        "kernels_compulsory_hbm_GBps": round(prep_bytes / (k_ms * 1e-3) / 1e9, 1),
        "center_search_s": round(t_search, 3),
        "center_search_evaluations": evals,
        "found_shift_px": round(sh * n, 4),
        "found_theta_rad": th,
        "bit_exact_vs_oracle_sample": exact,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_ingest(args) -> int:
    """FITS ingest normalisation, -log(I / max(OB, 1)), of 256 projections
    of 2048^2 resident in HBM (the GPU part of tomography/fits_to_hpc.F90).
    A step is one launch over the whole batch.  Algorithmic traffic: each
    projection read and written once (8 B per pixel) plus the open beam
    read once per launch; the launch duration comes from HIP events on the
    library's stream.  The host-buffer form (PCIe included) and the oracle
    on the host cores are reported beside it."""
    import numpy as np
    import torch

    from acumen_amd import Context, _lib, ingest

    n, count = N_DET, 256
    npix = n * n
    ctx = Context(0)
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4000)
    ob = (torch.rand((n, n), generator=gen, device=dev) * 4000.0 + 18000.0).floor()
    trans = torch.rand((count, n, n), generator=gen, device=dev) * 0.9 + 0.05
    imgs = (ob * trans).floor()
    del trans
    out = torch.empty_like(imgs)
    torch.cuda.synchronize()

    def launch():
        _lib.check(lib.acu_ingest_normalize_dev(ctx.handle, count, npix, imgs.data_ptr(), ob.data_ptr(),
                                                out.data_ptr()))

    log("ingest normalisation:", count, "projections of", n, "^2 in HBM")
    for _ in range(args.warmup):
        launch()
    ctx.synchronize()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        launch()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    k_ms, k_n = ctx.kernel_time(_lib.KIND_INGEST)
    ctx.set_timing(False)
    avg_s = k_ms / max(k_n, 1) / 1e3
    alg_bytes = 8.0 * npix * count + 4.0 * npix
    got16 = out[:16].cpu().numpy()
    # host-buffer form (PCIe both ways), 32 projections
    h_imgs, h_ob = imgs[:32].cpu().numpy(), ob.cpu().numpy()
    norm = ingest.Normalizer(ctx)
    norm.normalize(h_imgs[:2], h_ob)
    t = time.perf_counter()
    norm.normalize(h_imgs, h_ob)
    pcie = 32 / (time.perf_counter() - t)
    cpu, ulps = None, None
    if not args.no_cpu:
        log("cpu baseline: oracle normalisation, one thread")
        from oracle import ref
        done, t = 0, time.perf_counter()
        while done < 256 and time.perf_counter() - t < 10.0:
            want = ref.ingest_normalize(h_imgs[done % 32:done % 32 + 16], h_ob)
            if done == 0:  # the checker: the GPU's first 16 projections, in ulps
                gi = got16.view(np.int32).astype(np.int64)
                ulps = int(np.abs(gi - want.view(np.int32).astype(np.int64)).max())
            done += 16
        el = time.perf_counter() - t
        cpu = {"value": round(done / el, 2), "unit": "projections/s", "cores": 1, "kind": "port",
               "sample": f"{done} projections of {n}^2 through the oracle's FP32 loop (C, logf), "
                         f"one thread, {el:.1f} s"}
    ctx.close()
    traffic = traffic_from_profile("ingest_normalize_traffic_2048x256.txt", "hbm_bytes_per_launch")
    res = {
        "metric": "projections/s through -log(I / max(OB, 1)), 2048^2, inputs in HBM",
        "value": round(count * args.steps / wall, 1),
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / max(args.steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ingest normalisation of {count} projections of {n}^2 (open beam ~2e4 counts)",
                   "n": n, "projections": count},
        "roofline": {"bound": "hbm", "achieved": round(alg_bytes / avg_s / 1e9, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(alg_bytes / avg_s / 1e9 / 8000.0, 4), "traffic": traffic},
        "kernel_ms": round(avg_s * 1e3, 4),
        "pcie_incl_projections_per_s": round(pcie, 1),
        "max_ulps_vs_oracle_sample": ulps,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    return 0


class DeviceBuffers:
    """Device memory from libamdhip64, the HIP runtime the library uses, for
    the acu_*_dev entry points."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.ct.c_void_p()
        if self.hip.hipMalloc(self.ct.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptrs.append(p.value)
        return p.value

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy failed")
        return p

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def grad_cpu_baseline(x, th, sh, b, threads):
    """The oracle's FP32 loss + gradients of one slice on `threads` threads."""
    from oracle import ref
    ref.set_threads(threads)
    try:
        t = time.perf_counter()
        ref.loss_grad(x, th, sh, b, GRAD_LAM, GRAD_EPS, "f32")
        wall = time.perf_counter() - t
    finally:
        ref.set_threads(1)
    return {"value": round(1.0 / wall, 4), "unit": "slice evaluations/s", "cores": threads, "kind": "port",
            "sample": f"one slice's loss, dL/dx and dL/ds by the oracle's FP32 loops on {threads} threads, "
                      f"{wall:.2f} s"}


def run_grad(args) -> int:
    """Differentiation: 16 slices of 512^2 (N = 513) x 720 angles whose
    sinograms carry per-angle centre-of-rotation shifts drawn N(0, 1.5 px),
    applied by linear interpolation along the detector.  value: slice
    evaluations/s of L = 1/2 |shift_s(radon(x)) - b|^2 + lambda TV_eps(x)
    with dL/dx and dL/ds (lambda = 0.1, eps = 1e-3) for all 16 slices in one
    batched call, images, shifts and gradients resident in device memory.
    Also: the same with host buffers (PCIe included), and Gauss-Newton
    registration of the shifts from zero (evaluations, time, error against
    the drawn shifts).  cpu_baseline: the oracle's FP32 evaluation of one
    slice."""
    import ctypes

    import numpy as np

    from acumen_amd import Batch, Context, synth

    N, T, S = synth.padded_size(512), 720, 16
    th = synth.angles(T)
    wpu = synth.default_wpu(N)
    rng = np.random.default_rng(1003)
    ctx = Context(0)
    log(f"inputs: {S} slices of {N}x{N}, {T} angles, shifts N(0, 1.5 px)")
    xs, bs, truth = [], [], []
    with Batch(ctx, [(N, T, wpu, th)]) as p:
        for i in range(S):
            x = synth.phantom(N, 20, 1300 + i)
            sh = rng.normal(0.0, 1.5, T).astype(np.float32)
            xs.append(x)
            truth.append(sh)
            bs.append(synth.shift_rows(p.radon(x), sh))
    b = Batch(ctx, [(N, T, wpu, th)] * S)
    for i in range(S):
        b.set_sino(i, bs[i])
    sh0 = [t + np.float32(0.3) for t in truth]  # evaluated away from the optimum
    steps = max(args.steps, 5)
    log("loss + gradients, host buffers")
    b.loss_grad(xs, sh0, GRAD_LAM, GRAD_EPS)
    t = time.perf_counter()
    for _ in range(steps):
        loss_h, gx_h, gs_h = b.loss_grad(xs, sh0, GRAD_LAM, GRAD_EPS)
    host_rate = S * steps / (time.perf_counter() - t)
    log("loss + gradients, device buffers")
    db = DeviceBuffers()
    vp = ctypes.c_void_p
    try:
        d_x = (vp * S)(*[db.upload(x) for x in xs])
        d_s = (vp * S)(*[db.upload(s) for s in sh0])
        d_gx = (vp * S)(*[db.alloc(N * N * 4) for _ in range(S)])
        d_gs = (vp * S)(*[db.alloc(T * 4) for _ in range(S)])
        loss = np.zeros((S, 3), np.float64)
        lp = loss.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

        def dev_eval():
            rc = b._lib.acu_batch_loss_grad_dev(b.handle, d_x, d_s, GRAD_LAM, GRAD_EPS, d_gx, d_gs, None, lp)
            if rc != 0:
                raise RuntimeError(b._lib.acu_last_error().decode())

        for _ in range(args.warmup):
            dev_eval()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            dev_eval()
        ctx.synchronize()
        elapsed = time.perf_counter() - t
        dev_equal = bool(np.array_equal(loss, loss_h))
    finally:
        db.close()
    log("registration from zero shifts")
    t = time.perf_counter()
    fit, evals = b.register_shifts(xs, iters=12)
    t_reg = time.perf_counter() - t
    err = np.concatenate([f - s for f, s in zip(fit, truth)])
    b.close()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline: the oracle's evaluation of one slice")
        cpu = grad_cpu_baseline(xs[0], th, sh0[0], bs[0], cpu_threads(args.cpu_threads))
    ctx.close()
    out = {
        "metric": "slice evaluations/s of the TV-regularised loss with gradients in the image and the "
                  "per-angle shifts, 512^2 x 720, 16 slices per call",
        "value": round(S * steps / elapsed, 2),
        "unit": "slice evaluations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{S} slices of 512^2 (N={N}) x {T} angles, 20-ellipse phantoms, shifts "
                               f"N(0, 1.5 px) (seed 1003), lambda={GRAD_LAM}, eps={GRAD_EPS}",
                   "slices": S, "nimage": N, "ntheta": T},
        "host_buffers_slice_evaluations_per_s": round(host_rate, 2),
        "device_loss_equals_host_loss": dev_equal,
        "registration": {"evaluations": evals, "seconds": round(t_reg, 3),
                         "max_error_px": float(f"{np.abs(err).max():.3e}"),
                         "rms_error_px": float(f"{np.sqrt(np.mean(err ** 2)):.3e}")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_production(args) -> int:
    """The headline: split-Bregman TV of --batch resident 2049^2 x 1500
    slices per step through the batched path, with per-launch HIP events on
    the library's stream over the timed steps (they give the roofline
    kernels' average launch durations).  After the timed steps, each within
    what remains of --budget-s: FBP slices/s, the plain per-slice path with
    the same operators (batching alone) and with the reference's operators,
    and the CPU baseline."""
    import numpy as np

    from acumen_amd import ACU_IMG_FBP, ACU_IMG_TV, ACU_TV_BATCHED, ACU_TV_GRAPH, ACU_TV_PLAIN, Batch, Context, _lib, synth

    N = synth.padded_size(N_DET)
    T = T_ANG
    B = args.batch
    ctx = Context(0)
    log(f"inputs: {B} slices of {N}x{N}, {T} angles")
    th, wpu, sinos = make_inputs(ctx, B, N, T, 1001)
    batch = Batch(ctx, [(N, T, wpu, th)] * B)
    for i, f in enumerate(sinos):
        batch.set_sino(i, f)
    ctx.synchronize()
    sched, refresh = batch.tv_schedule()
    tuning = None
    if not args.no_tune:
        log("autotune: projector variant per slice shape, ramp-filter arithmetic, backprojection tile")
        t = time.perf_counter()
        rep = batch.autotune(reps=2)
        tuning = {"seconds": round(time.perf_counter() - t, 2),
                  "fp_variant": rep["slices"][0]["choice"], "ramp": rep["ramp"]["choice"],
                  "bp_tile": rep["bp"]["choice"],
                  "ramp_ms_per_batch": {k: round(v, 4) for k, v in rep["ramp"]["ms"].items()},
                  "fp_ms_per_slice": {k: round(v, 4) for k, v in rep["slices"][0]["ms"].items()},
                  "bp_ms_per_batch": {k: round(v, 4) for k, v in rep["bp"]["ms"].items()}}

    # warm-up steps replay each outer iteration from a HIP graph (no events)
    graph_ms = []
    for w in range(args.warmup):
        t = time.perf_counter()
        batch.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_GRAPH)
        ctx.synchronize()
        graph_ms.append((time.perf_counter() - t) * 1e3)
        log(f"warmup step {w + 1}/{args.warmup} (graph): {graph_ms[-1]:.0f} ms")

    # timed steps: every launch bracketed by HIP events on the library's stream
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_BATCHED)
        ctx.synchronize()
        log(f"step {k + 1}/{args.steps} done at {time.perf_counter() - t0:.2f} s")
    elapsed = time.perf_counter() - t0
    kt = {name: ctx.kernel_time(kind) for name, kind in
          (("ramp", _lib.KIND_RAMP), ("bp", _lib.KIND_BP), ("cells", _lib.KIND_CELLS),
           ("fp", _lib.KIND_FP))}
    ctx.set_timing(False)
    tv_img = batch.get_image(0, ACU_IMG_TV)
    assert np.all(np.isfinite(tv_img)), "non-finite TV result"
    value = B * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3

    def remaining():
        return args.budget_s - (time.perf_counter() - T_START)

    legs = {}

    def leg(name, cost_s, fn):
        if remaining() < cost_s:
            legs[name] = f"skipped: {remaining():.0f} s of the {args.budget_s:.0f} s budget left, needs ~{cost_s:.0f} s"
            log(legs[name])
            return None
        try:
            return fn()
        except Exception as e:  # a supplementary leg never loses the headline
            legs[name] = f"failed: {type(e).__name__}: {e}"
            log(legs[name])
            return None

    # FBP slices/s of the batched path
    def fbp_leg():
        reps = 5
        batch.fbp()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            batch.fbp()
        ctx.synchronize()
        return B * reps / (time.perf_counter() - t)

    fbp_rate = leg("fbp", 5, fbp_leg)
    per_slice_s = ms_per_step / 1e3 / B

    # the plain per-slice path: one slice at a time, every reduction back on the host
    def plain_leg(schedule):
        with Batch(ctx, [(N, T, wpu, th)]) as one:
            one.set_ramp_variant(batch.ramp_variant())
            one.set_fp_variant(0, batch.fp_variant(0))
            one.set_bp_tile(batch.bp_tile())
            one.set_tv_schedule(schedule, refresh if schedule == sched else 0)
            one.set_sino(0, sinos[0])
            ctx.synchronize()
            t = time.perf_counter()
            one.tvopt(MU, LAM, args.nouter, NINNER, NCG, ACU_TV_PLAIN)
            ctx.synchronize()
            rate = 1.0 / (time.perf_counter() - t)
            img = one.get_image(0, ACU_IMG_TV)
        return rate, img

    plain = None if args.no_plain else leg("plain", 4 * per_slice_s + 5, lambda: plain_leg(sched))
    plain_rate = plain_equal = None
    if plain is not None:
        plain_rate = plain[0]
        # full-size property: the batched path's slice 0 is the plain path's, bit for bit
        plain_equal = bool(np.array_equal(plain[1].view(np.uint32), tv_img.view(np.uint32)))
    ref_struct = None if args.no_plain else leg("plain_reference_operators", 5 * per_slice_s + 5,
                                                lambda: plain_leg(_lib.ACU_TV_SCHED_REFERENCE))
    ref_rate = ref_dist = None
    if ref_struct is not None:
        ref_rate = ref_struct[0]
        ref_dist = float(np.linalg.norm(ref_struct[1].astype(np.float64) - tv_img)
                         / np.linalg.norm(ref_struct[1].astype(np.float64)))

    cpu = None
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        threads = cpu_threads(args.cpu_threads)
        cpu = leg("cpu", 75, lambda: cpu_baseline_tv(ctx, sinos[0], th, wpu, threads))
        if cpu is not None:
            acc = leg("accuracy", 20, lambda: headline_accuracy(ctx, batch, th, wpu, threads))
            if acc is not None:
                cpu["headline_slice_vs_reference_runs"] = acc

    batch.close()
    ctx.close()

    fp_ms, fp_n = kt["fp"]
    fp_avg_ms = fp_ms / max(fp_n, 1)
    samples_per_launch = float(N) * N * T * B
    achieved = FP_FLOPS_PER_SAMPLE * samples_per_launch / (fp_avg_ms * 1e-3) / 1e12
    fp_name = fp_kernel_name(tuning.get("fp_variant") if tuning else None)
    traffic_file = f"{fp_name}_traffic_2049x1500.txt"
    traffic = traffic_from_profile(traffic_file)
    lds_bytes = LDS_BYTES_PER_SAMPLE * samples_per_launch / (fp_avg_ms * 1e-3)
    roofline = {
        "bound": "lds",
        "kernel": f"{fp_name} (ray-driven bilinear projection on the FP32 VALU, gathers from LDS)",
        "limiters": "bank conflicts of the bilinear ds_read2 gathers (LDS busy ~52% of the cycles, 70% "
                    "of it in conflicts) beside 1.05 VALU wave-instructions per CU and cycle "
                    "(SQ counters under profiles/)",
        "achieved": round(achieved, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
        "peak_nonpacked": FP32_NONPACKED_TFLOPS,
        "frac_nonpacked": round(achieved / FP32_NONPACKED_TFLOPS, 4),
        "traffic": None if traffic is None else traffic * B,
        "traffic_source": None if traffic is None else
        f"profiles/{traffic_file}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per 2049^2 x 1500 slice-launch, x {B}",
        "traffic_over_algorithmic": None if traffic is None else
        round(traffic / (4.0 * (N * N + N * T)), 1),
        "avg_launch_ms": round(fp_avg_ms, 4),
        "launches": fp_n,
        "work_per_launch": f"{FP_FLOPS_PER_SAMPLE} flop x {N}*{N}*{T} samples x {B} slices",
        "lds": {"bytes_per_sample": LDS_BYTES_PER_SAMPLE, "achieved_TBps": round(lds_bytes / 1e12, 2),
                "B_per_clk_per_cu": round(lds_bytes / (N_CU * CLOCK_GHZ * 1e9), 1),
                "frac_of_128_B_per_clk": round(lds_bytes / (N_CU * CLOCK_GHZ * 1e9) / 128.0, 3)},
    }

    bp_ms, bp_n = kt["bp"]
    bp_avg_s = bp_ms / max(bp_n, 1) * 1e-3
    bp_achieved = BP_FLOPS_PER_UPDATE * samples_per_launch / bp_avg_s / 1e12
    bp_ty = tuning["bp_tile"] if tuning else 32
    bp_active = bp_active_fraction(N, wpu, bp_ty)
    bp_lds = 8.0 * samples_per_launch * bp_active / bp_avg_s
    roofline_bp = {
        "bound": "valu+lds",
        "kernel": "bp_kernel (pixel-driven linear-interpolation backprojection on the FP32 VALU)",
        "achieved": round(bp_achieved, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(bp_achieved / FP32_PEAK_TFLOPS, 4),
        "frac_nonpacked": round(bp_achieved / FP32_NONPACKED_TFLOPS, 4),
        "avg_launch_ms": round(bp_avg_s * 1e3, 4),
        "launches": bp_n,
        "work_per_launch": f"{BP_FLOPS_PER_UPDATE} flop x {N}*{N}*{T} updates x {B} slices",
        "compulsory_hbm_GBps": round(4.0 * (N * T + 2.0 * N * N) * B / bp_avg_s / 1e9, 1),
        # 8 bytes of staged sinogram pairs per update, over the tiles that hold mask pixels
        "lds": {"bytes_per_update": 8, "tile_rows": bp_ty, "computed_fraction": round(bp_active, 4),
                "achieved_TBps": round(bp_lds / 1e12, 2),
                "B_per_clk_per_cu": round(bp_lds / (N_CU * CLOCK_GHZ * 1e9), 1)},
    }
    names = {_lib.ACU_TV_SCHED_REFERENCE: "reference (7 radon + 7 iradon per outer iteration)",
             _lib.ACU_TV_SCHED_REUSE: "reuse (6 radon + 7 iradon)",
             _lib.ACU_TV_SCHED_LINEAR: "linear (5 radon + 5 iradon)"}
    ops = {_lib.ACU_TV_SCHED_REFERENCE: (7, 7), _lib.ACU_TV_SCHED_REUSE: (6, 7), _lib.ACU_TV_SCHED_LINEAR: (5, 5)}
    out = {
        "metric": "TV-SB slices/s batched vs plain @2048^2 (rel-L2<=1e-4 vs FP64); FBP slices/s",
        "value": round(value, 5),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"split-Bregman TV, {B} slices of {N_DET}^2 (N={N}) x {T} angles per step, "
                               f"40-ellipse phantoms, I0=5e3, mu={MU}, lambda={LAM}, {args.nouter} outer x "
                               f"{NCG} CG, FBP warm start, batched path, inputs resident in HBM",
                   "slices_per_step": B, "nimage": N, "ntheta": T, "nouter": args.nouter,
                   "ncg": NCG, "wpu": wpu, "operator_schedule": names[sched],
                   "radon_iradon_per_outer": list(ops[sched]), "refresh": refresh,
                   "timed_mode": "batched launches, HIP events around every operator launch"},
        "graph_mode_ms_per_step": None if len(graph_ms) < 2 else round(float(np.median(graph_ms[1:])), 2),
        "plain_slices_per_s": None if plain_rate is None else round(plain_rate, 5),
        "batched_vs_plain": None if plain_rate is None else round(value / plain_rate, 3),
        "tv_batched_equals_plain_bitwise": plain_equal,
        "plain_reference_operators_slices_per_s": None if ref_rate is None else round(ref_rate, 5),
        "batched_vs_plain_reference_operators": None if ref_rate is None else round(value / ref_rate, 3),
        "tv_rel_l2_vs_reference_operators": None if ref_dist is None else float(f"{ref_dist:.3e}"),
        "fbp_slices_per_s": None if fbp_rate is None else round(fbp_rate, 3),
        "kernel_avg_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items()},
        "roofline": roofline,
        "roofline_backprojection": roofline_bp,
        "cpu_baseline": cpu,
        "autotune": tuning,
        "legs": legs or None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: production 5; ingest 50; others 2)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=8, help="slices per step (production workload)")
    ap.add_argument("--budget-s", type=float, default=400.0,
                    help="production: wall-clock budget from start; the legs after the timed steps "
                         "(FBP, plain paths, CPU) run only while it lasts (the timed work never shrinks)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-slice leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (default: the host cores this process may use, at most 16)")
    ap.add_argument("--nouter", type=int, default=NOUTER,
                    help="outer iterations (default: the workload's 50; fewer only for profiling runs)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--workload", choices=("production", "mixed", "volume", "prep", "ingest", "grad"),
                    default="production",
                    help="production: 2048^2 x 1500 slices (the headline); mixed: 64 slices of "
                         "uneven width and angle count in one batch; volume: a 2048^2 x 1500 x 256 "
                         "volume streamed through FBP (+ TV on --tv-slices slices); prep: projection -> "
                         "sinogram preprocessing; ingest: FITS ingest normalisation; grad: loss and "
                         "gradients for shift registration")
    ap.add_argument("--slices", type=int, default=256, help="volume workload: slices streamed through FBP")
    ap.add_argument("--tv-slices", type=int, default=16,
                    help="volume workload: slices streamed through FBP + 30-iteration TV")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"ingest": 50, "production": 5}.get(args.workload, 2)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(slices are independent: several GPUs would each take a share of the slices)",
              file=sys.stderr)
        return 2
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py: multi-process launch is not supported (one GPU)", file=sys.stderr)
        return 2
    runners = {"mixed": run_mixed, "volume": run_volume, "prep": run_prep, "ingest": run_ingest,
               "grad": run_grad, "production": run_production}
    return runners[args.workload](args)


if __name__ == "__main__":
    sys.exit(main())
