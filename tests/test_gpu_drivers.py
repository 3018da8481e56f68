"""The Fortran drop-in drivers (fbp_hip, tv_hip, tv_hip_slice) on a synthetic
volume: same results as the C ABI path, bit for bit, and parity with the
oracle."""
import os
import subprocess

import numpy as np
import pytest

from acumen_amd import ACU_IMG_FBP, ACU_IMG_TV, ACU_TV_GRAPH, Batch, synth
from acumen_amd import io as aio
from helpers import disk_phantom, rel
from oracle import ref

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "acumen_amd", "fortran", "bin")
NIMAGEF, NFILES = 32, 37          # 32 slices, 37 angles 0..180 step 5 -> N = 33
MU, LAM, NOUTER, NCG, WPU = 0.05, 0.1, 2, 3, 4


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def volume(tmp_path_factory):
    root = tmp_path_factory.mktemp("ct")
    ds = aio.Dataset(str(root), "synthetic")
    th = np.arange(NFILES, dtype=np.float32) * np.float32(5.0)
    N = NIMAGEF + 1
    sinos = np.stack([ref.radon(disk_phantom(N, 4, 300 + s), th, "f64").astype(np.float32)
                      for s in range(NIMAGEF)])
    aio.write_dataset(ds, th, sinos, parm=dict(nsub=1, nsub_t=1, ninner=1, nouter=NOUTER, ncg=NCG,
                                               lam=LAM, mu=MU, wpu=WPU))
    return ds, th, sinos


def run(exe, ds, extra=None):
    env = dict(os.environ, ROOT_CT_DIR=ds.root, EXP_CT_NAME=ds.name)
    env.update(extra or {})
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (python -m acumen_amd.fortran_build)")
    say("driver", exe)
    r = subprocess.run([path], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_fbp_driver(ctx, volume):
    ds, th, sinos = volume
    N = NIMAGEF + 1
    out = run("fbp_hip", ds, {"ACUMEN_BATCH": "10"})
    assert "Reconstruction with Resolution:" in out
    imgs = aio.read_images(ds, "fbp", NIMAGEF, N)
    with Batch(ctx, [(N, NFILES, WPU, th)]) as b:
        for s in (0, 13, NIMAGEF - 1):
            b.set_sino(0, sinos[s])
            b.fbp()
            assert np.array_equal(imgs[s], b.get_image(0, ACU_IMG_FBP))
            assert rel(imgs[s], ref.iradon(sinos[s], th, WPU, "f32")) < 1e-5


def test_tv_slice_driver(ctx, volume):
    ds, th, sinos = volume
    N = NIMAGEF + 1
    run("tv_hip_slice", ds)
    tv = aio.read_image(ds, "fbp_tv_slice.dat", N)
    fbp = aio.read_image(ds, "fbp_sub_slice.dat", N)
    s = N // 2 - 1  # job_id = nimage/2 (1-based)
    with Batch(ctx, [(N, NFILES, WPU, th)]) as b:
        b.set_sino(0, sinos[s])
        b.tvopt(MU, LAM, NOUTER, 1, NCG, ACU_TV_GRAPH)
        assert np.array_equal(tv, b.get_image(0, ACU_IMG_TV))
        assert np.array_equal(fbp, b.get_image(0, ACU_IMG_FBP))
    tv64, _, _ = ref.tvopt(sinos[s], th, WPU, MU, LAM, NOUTER, 1, NCG, "f64")
    assert rel(tv, tv64) < 1e-4


def tv_env(nsub, nsub_t):
    return {"Sub_Sample_Radial": str(nsub), "Sub_Sample_Theta": str(nsub_t), "Ninner_Loops": "1",
            "Nouter_Loops": str(NOUTER), "Opt_ncg": str(NCG), "Opt_lambda": str(LAM),
            "Opt_mu": str(MU), "Radon_wpu": str(WPU), "ACUMEN_BATCH": "12"}


@pytest.mark.parametrize("nsub,nsub_t", [(1, 1), (2, 2)])
def test_tv_volume_driver(ctx, volume, nsub, nsub_t):
    ds, th, sinos = volume
    run("tv_hip", ds, tv_env(nsub, nsub_t))
    width = NIMAGEF + 1
    N = (width - 1) // nsub + 1
    T = NFILES // nsub_t
    thw = th[::nsub_t][:T]
    pre = "fbp" if nsub + nsub_t == 2 else "fbp_sub"
    tvs = aio.read_images(ds, pre + "_tv", NIMAGEF, N)
    fbps = aio.read_images(ds, pre, NIMAGEF, N)
    with Batch(ctx, [(N, T, WPU, thw)]) as b:
        for s in (0, 17, NIMAGEF - 1):
            f = sinos[s][::nsub_t][:T, ::nsub][:, :N]
            b.set_sino(0, f)
            b.tvopt(MU, LAM, NOUTER, 1, NCG, ACU_TV_GRAPH)
            assert np.array_equal(tvs[s], b.get_image(0, ACU_IMG_TV))
            assert np.array_equal(fbps[s], b.get_image(0, ACU_IMG_FBP))


@pytest.mark.parametrize("nsub,nsub_t", [(1, 1), (2, 2)])
def test_streamed_dataset_equals_batch_path(ctx, tmp_path, nsub, nsub_t):
    """acumen_amd.volume.reconstruct_dataset (threaded file I/O around the
    two-lane streaming engine) writes the same FBP / TV files, bit for bit,
    as the C ABI's batch path produces for each slice -- in a dataset split
    over 3 sinogram files and 2 reconstruction files."""
    from acumen_amd import volume
    say("streamed dataset", nsub, nsub_t)
    th = np.arange(NFILES, dtype=np.float32) * np.float32(5.0)
    N0 = NIMAGEF + 1
    sinos = np.stack([ref.radon(disk_phantom(N0, 4, 500 + s), th, "f64").astype(np.float32)
                      for s in range(NIMAGEF)])
    ds = aio.Dataset(str(tmp_path), "stream")
    aio.write_dataset(ds, th, sinos, n_files_sino=3, n_files_rec=2)
    tv = dict(mu=MU, lam=LAM, nouter=NOUTER, ninner=1, ncg=NCG)
    st = volume.reconstruct_dataset(ctx, ds.root, ds.name, nsub=nsub, nsub_t=nsub_t, wpu=WPU, tv=tv,
                                    batch=5, tune=False, threads=4)
    N = (N0 - 1) // nsub + 1
    T = NFILES // nsub_t
    assert (st["slices"], st["nimage"], st["ntheta"]) == (NIMAGEF, N, T)
    pre = "fbp" if nsub + nsub_t == 2 else "fbp_sub"
    nn = N * N
    thw = th[::nsub_t][:T]
    with Batch(ctx, [(N, T, WPU, thw)]) as b:
        for s in (0, 11, 16, NIMAGEF - 1):
            js, off = aio.slice_location(s + 1, NIMAGEF, 2, nn)
            got_fbp = np.fromfile(ds.path(f"{pre}{js}.dat"), np.float32, nn, offset=4 * off).reshape(N, N)
            got_tv = np.fromfile(ds.path(f"{pre}_tv{js}.dat"), np.float32, nn, offset=4 * off).reshape(N, N)
            b.set_sino(0, sinos[s][::nsub_t][:T, ::nsub][:, :N])
            b.tvopt(MU, LAM, NOUTER, 1, NCG, ACU_TV_GRAPH)
            assert np.array_equal(got_tv, b.get_image(0, ACU_IMG_TV)), s
            assert np.array_equal(got_fbp, b.get_image(0, ACU_IMG_FBP)), s


def _sentinel(ds, prefix, s, nimagef, nn, n_files_rec=1):
    """Overwrite 1-based slice s of {prefix}{js}.dat with -7."""
    js, off = aio.slice_location(s, nimagef, n_files_rec, nn)
    with open(ds.path(f"{prefix}{js}.dat"), "r+b") as fh:
        fh.seek(4 * off)
        fh.write(np.full(nn, -7.0, np.float32).tobytes())


def test_tv_volume_driver_resume(ctx, volume):
    """tv_hip with ACUMEN_MAX_BATCHES=1 stops after one batch and marks its
    slices in fbp_tv.done; ACUMEN_RESUME=1 then reconstructs only the other
    slices (a sentinel written into a done slice survives), and the volume
    equals an uninterrupted run's bit for bit."""
    ds, th, sinos = volume
    N = NIMAGEF + 1
    nn = N * N
    env = tv_env(1, 1)  # batches of 12: 12 + 12 + 8 slices
    run("tv_hip", ds, env)
    full_tv = aio.read_images(ds, "fbp_tv", NIMAGEF, N)
    full_fbp = aio.read_images(ds, "fbp", NIMAGEF, N)
    assert np.array_equal(np.fromfile(ds.path("fbp_tv.done"), np.uint8)[64:], np.ones(NIMAGEF, np.uint8))
    for f in ("fbp1.dat", "fbp_tv1.dat"):
        os.remove(ds.path(f))
    out = run("tv_hip", ds, dict(env, ACUMEN_MAX_BATCHES="1"))
    assert "Stopped after" in out
    done = np.fromfile(ds.path("fbp_tv.done"), np.uint8)[64:]
    assert list(done) == [1] * 12 + [0] * (NIMAGEF - 12)
    _sentinel(ds, "fbp_tv", 4, NIMAGEF, nn)
    out = run("tv_hip", ds, dict(env, ACUMEN_RESUME="1"))
    assert "Resuming:" in out
    assert np.array_equal(np.fromfile(ds.path("fbp_tv.done"), np.uint8)[64:], np.ones(NIMAGEF, np.uint8))
    got_tv = aio.read_images(ds, "fbp_tv", NIMAGEF, N)
    assert (got_tv[3] == -7.0).all()
    got_tv[3] = full_tv[3]
    assert np.array_equal(got_tv, full_tv)
    assert np.array_equal(aio.read_images(ds, "fbp", NIMAGEF, N), full_fbp)
    # a run with other parameters does not take the marks: everything is recomputed
    out = run("tv_hip", ds, dict(env, ACUMEN_RESUME="1", Opt_ncg=str(NCG + 1)))
    assert "Resuming:" not in out


def test_streamed_dataset_resume(ctx, tmp_path):
    """reconstruct_dataset in chunks of 10 with max_chunks=2 reconstructs 20
    slices and marks them done; resume=True reconstructs the other 12 only
    (a sentinel written into a done slice survives), and the files equal an
    uninterrupted run's bit for bit."""
    from acumen_amd import volume
    say("streamed dataset resume")
    th = np.arange(NFILES, dtype=np.float32) * np.float32(5.0)
    N = NIMAGEF + 1
    nn = N * N
    sinos = np.stack([ref.radon(disk_phantom(N, 4, 700 + s), th, "f64").astype(np.float32)
                      for s in range(NIMAGEF)])
    ds = aio.Dataset(str(tmp_path), "resume")
    aio.write_dataset(ds, th, sinos, n_files_sino=3, n_files_rec=2)
    tv = dict(mu=MU, lam=LAM, nouter=NOUTER, ninner=1, ncg=NCG)
    kw = dict(wpu=WPU, tv=tv, batch=5, tune=False, threads=4)
    st = volume.reconstruct_dataset(ctx, ds.root, ds.name, **kw)
    assert (st["slices"], st["pending"], st["skipped"]) == (NIMAGEF, 0, 0)
    full_tv = aio.read_images(ds, "fbp_tv", NIMAGEF, N, n_files_rec=2)
    full_fbp = aio.read_images(ds, "fbp", NIMAGEF, N, n_files_rec=2)
    for f in ("fbp1.dat", "fbp2.dat", "fbp_tv1.dat", "fbp_tv2.dat"):
        os.remove(ds.path(f))
    st = volume.reconstruct_dataset(ctx, ds.root, ds.name, chunk=10, max_chunks=2, **kw)
    assert (st["slices"], st["pending"], st["skipped"]) == (20, NIMAGEF - 20, 0)
    _sentinel(ds, "fbp_tv", 5, NIMAGEF, nn, n_files_rec=2)
    st = volume.reconstruct_dataset(ctx, ds.root, ds.name, chunk=10, resume=True, **kw)
    assert (st["slices"], st["pending"], st["skipped"]) == (NIMAGEF - 20, 0, 20)
    got_tv = aio.read_images(ds, "fbp_tv", NIMAGEF, N, n_files_rec=2)
    assert (got_tv[4] == -7.0).all()
    got_tv[4] = full_tv[4]
    assert np.array_equal(got_tv, full_tv)
    assert np.array_equal(aio.read_images(ds, "fbp", NIMAGEF, N, n_files_rec=2), full_fbp)
    # nothing left: a resumed run reconstructs nothing
    st = volume.reconstruct_dataset(ctx, ds.root, ds.name, resume=True, **kw)
    assert (st["slices"], st["pending"], st["skipped"]) == (0, 0, NIMAGEF)


def test_projections_to_sinogram_driver(ctx, tmp_path):
    # angles without an exact 0 / 180 (both end projections interpolated),
    # two projection and sinogram files, even n (Nw = n + 1)
    from acumen_amd import prep
    from test_oracle_prep import blob_image
    n, nfiles, nfs = 32, 25, 2
    th = np.linspace(-1.3, 181.2, nfiles).astype(np.float32)
    rng = np.random.default_rng(4)
    proj = np.stack([blob_image(n, 50 + j) + rng.normal(0, 0.01, (n, n)).astype(np.float32)
                     for j in range(nfiles)])
    ds = aio.Dataset(str(tmp_path), "raw")
    aio.write_projections(ds, th, proj, n_files_sino=nfs)
    out = run("projections_to_sinogram_hip", ds)
    assert "Best theta:" in out
    got = aio.read_sinograms(ds, n, nfiles, n_files_sino=nfs)
    p0 = prep.angle_window_projection(proj, th, 0.0)
    p180 = prep.angle_window_projection(proj, th, 180.0)
    t, s, _ = prep.find_center(ctx, p0, p180)
    assert np.array_equal(got, prep.sinograms(ctx, proj, t, s))
