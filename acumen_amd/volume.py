"""Volume reconstruction streamed through one GPU (acu_volume_reconstruct).

The reference's drivers hand slices to MPI ranks one at a time
(tomography/tv_hpc.F90:583-631, fbp_hpc.F90:340-403); here a volume of
same-shape slices streams through the GPU in batches on two lanes whose
host staging and PCIe transfers overlap each other's kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

_F = ctypes.POINTER(ctypes.c_float)


def reconstruct(ctx, sinos, theta, wpu: int, tv=None, fbp: bool = True, batch: int = 8,
                tune: bool = True):
    """FBP (and, with ``tv = dict(mu, lam, nouter, ninner, ncg)``, split-Bregman
    TV) of every slice of ``sinos`` (S, T, N) float32.

    Returns ``(fbp_images or None, tv_images or None, stats)`` with images
    (S, N, N) and ``stats`` = wall ``ms``, ``slices_per_s`` over the call,
    ``setup_ms`` (lanes, pinned buffers, tuning) and ``stream_slices_per_s``
    after set-up; PCIe and host staging included."""
    s = np.ascontiguousarray(sinos, np.float32)
    if s.ndim != 3:
        raise ValueError("sinos must be (slices, angles, detectors)")
    S, T, N = s.shape
    th = np.ascontiguousarray(theta, np.float32)
    if th.shape != (T,):
        raise ValueError(f"theta must have shape ({T},)")
    out_fbp = np.empty((S, N, N), np.float32) if fbp else None
    out_tv = np.empty((S, N, N), np.float32) if tv is not None else None
    p = dict(mu=0.0, lam=0.0, nouter=0, ninner=1, ncg=5)
    if tv is not None:
        p.update(tv)
    stats = np.zeros(4, np.float64)
    lib = _lib.load()
    check(lib.acu_volume_reconstruct(
        ctx.handle, S, N, T, int(wpu), th.ctypes.data_as(_F), s.ctypes.data_as(_F),
        None if out_fbp is None else out_fbp.ctypes.data_as(_F),
        None if out_tv is None else out_tv.ctypes.data_as(_F),
        float(p["mu"]), float(p["lam"]), int(p["nouter"]), int(p["ninner"]), int(p["ncg"]),
        int(batch), 1 if tune else 0, stats.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out_fbp, out_tv, {"ms": float(stats[0]), "slices_per_s": float(stats[1]),
                             "setup_ms": float(stats[2]), "stream_slices_per_s": float(stats[3])}


def _write_at(path: str, offset_elems: int, a: np.ndarray) -> None:
    """FP32 values at an element offset, creating the file without
    truncating it (the drivers' write_reals_at)."""
    with open(path, "ab"):
        pass
    with open(path, "r+b") as fh:
        fh.seek(4 * offset_elems)
        fh.write(np.ascontiguousarray(a, np.float32).tobytes())


def reconstruct_dataset(ctx, root: str, exp: str, *, nsub: int = 1, nsub_t: int = 1, wpu: int = 64,
                        tv=None, batch: int = 16, tune: bool = True, threads: int = 8, log=None,
                        resume: bool = False, chunk: int = 0, max_chunks: int = 0):
    """Every slice of the dataset $ROOT_CT_DIR/$EXP_CT_NAME streamed through
    the GPU with the file contract of tomography/tv_hpc.F90:392-652 (and of
    acumen_amd/fortran/tv_hip.F90): header offsets 0-3 of
    fileinfo_projections.dat, theta.dat, the angle window, radial (nsub) and
    angular (nsub_t) sub-sampling, odd padding, sinograms{js}.dat read and
    fbp{js}.dat / fbp_tv{js}.dat (fbp_sub / fbp_sub_tv when sub-sampling)
    written at the reference's 64-bit element offsets.  Without ``tv`` only
    the FBP files are written.

    Slices go in chunks of ``chunk`` (0: all at once): a pool of ``threads``
    host threads reads a chunk's sinograms, ``reconstruct`` runs it (two
    lanes overlapping PCIe, staging and kernels), the pool writes its images
    and the chunk's slices are marked in the done map ``{fbp_tv|fbp}.done``
    (``io.DoneMap``).  ``resume`` skips the slices an earlier run with the
    same parameters marked done; ``max_chunks`` (0: no limit) stops after
    that many chunks, for runs split over several jobs.  Returns the
    summed ``reconstruct`` statistics of this call plus the read and write
    seconds, the slices reconstructed and the slices still pending."""
    import math
    import time
    from concurrent.futures import ThreadPoolExecutor

    from . import io as aio

    ds = aio.Dataset(root, exp)
    hdr = np.fromfile(ds.path("fileinfo_projections.dat"), np.float32, 4)
    nfiles, nimagef, n_files_sino, n_files_rec = (int(v) for v in hdr)
    theta = np.fromfile(ds.path("theta.dat"), np.float32, nfiles)
    js0, je0 = aio.angle_window(theta)  # 0-based
    ntheta = int(math.floor(np.float32(je0 - js0 + 1) / np.float32(nsub_t)))
    add = 1 if nimagef % 2 == 0 else 0
    width = nimagef + add
    nimage = int(math.floor(np.float32(width - 1) / np.float32(nsub))) + 1
    if ntheta < 1:
        raise ValueError("the angle window holds no projection")
    th = theta[js0 + np.arange(ntheta) * nsub_t]
    rowlen = width * nfiles
    sub = nsub + nsub_t != 2
    names = ("fbp_sub", "fbp_sub_tv") if sub else ("fbp", "fbp_tv")
    nn = nimage * nimage
    if tv is None:
        dh = aio.done_header(nimagef, nimage, ntheta, nsub, nsub_t, wpu, 0, 0, 0, 0.0, 0.0)
    else:
        q = dict(ninner=1, ncg=5)
        q.update(tv)
        dh = aio.done_header(nimagef, nimage, ntheta, nsub, nsub_t, wpu, q["ninner"], q["nouter"], q["ncg"],
                             q["mu"], q["lam"])
    done = aio.DoneMap(ds.path(names[1 if tv is not None else 0] + ".done"), dh, nimagef, resume)
    todo = done.pending()
    step = max(1, len(todo) if chunk <= 0 else chunk)

    def read(into, i, job):
        js, off = aio.slice_location(job, nimagef, n_files_sino, rowlen)
        a = np.fromfile(ds.path(f"sinograms{js}.dat"), np.float32, rowlen, offset=4 * off)
        if a.size != rowlen:
            raise IOError(f"sinograms{js}.dat: short read for slice {job}")
        a = a.reshape(nfiles, width)
        into[i] = a[js0::nsub_t][:ntheta, ::nsub][:, :nimage]

    def write(fbp, tvi, i, job):
        js, off = aio.slice_location(job, nimagef, n_files_rec, nn)
        _write_at(ds.path(f"{names[0]}{js}.dat"), off, fbp[i])
        if tvi is not None:
            _write_at(ds.path(f"{names[1]}{js}.dat"), off, tvi[i])

    stats = dict(ms=0.0, setup_ms=0.0)
    t_read = t_write = 0.0
    nrec = 0
    with ThreadPoolExecutor(max(1, threads)) as ex:
        for k, c0 in enumerate(range(0, len(todo), step)):
            if max_chunks and k >= max_chunks:
                break
            jobs = todo[c0:c0 + step]
            sinos = np.empty((len(jobs), ntheta, nimage), np.float32)
            t0 = time.perf_counter()
            list(ex.map(lambda a: read(sinos, *a), enumerate(jobs)))
            t_read += time.perf_counter() - t0
            if log:
                log(f"{len(jobs)} slices of {nimage}^2 x {ntheta} angles read ({nrec + len(jobs)} of "
                    f"{len(todo)} pending)")
            fbp, tvi, st = reconstruct(ctx, sinos, th, wpu, tv=tv, fbp=True, batch=batch, tune=tune)
            stats["ms"] += st["ms"]
            stats["setup_ms"] += st["setup_ms"]
            t0 = time.perf_counter()
            # slices of one file are written by one thread (no two writers on a file)
            by_file = {}
            for i, job in enumerate(jobs):
                by_file.setdefault(aio.slice_location(job, nimagef, n_files_rec, nn)[0], []).append((i, job))
            list(ex.map(lambda items: [write(fbp, tvi, i, job) for i, job in items], by_file.values()))
            t_write += time.perf_counter() - t0
            done.mark(jobs)
            nrec += len(jobs)
    sec = stats["ms"] / 1e3
    stream = sec - stats["setup_ms"] / 1e3
    stats.update(slices_per_s=nrec / sec if sec > 0 else 0.0,
                 stream_slices_per_s=nrec / stream if stream > 0 else 0.0,
                 read_s=t_read, write_s=t_write, nimage=nimage, ntheta=ntheta, slices=nrec,
                 pending=len(todo) - nrec, skipped=nimagef - len(todo))
    return stats


def main(argv=None) -> int:
    """python -m acumen_amd.volume: the reference's TV driver environment
    (ROOT_CT_DIR, EXP_CT_NAME, Sub_Sample_Radial, Sub_Sample_Theta,
    Ninner_Loops, Nouter_Loops, Opt_ncg, Opt_lambda, Opt_mu, Radon_wpu);
    Nouter_Loops = 0 writes the FBP files only."""
    import argparse
    import os

    from .tomography import Context

    ap = argparse.ArgumentParser(description="Stream a dataset's slices through FBP (+ TV) on one GPU")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--chunk", type=int, default=0,
                    help="slices per chunk; each chunk's slices are marked in the done map once written")
    ap.add_argument("--resume", action="store_true", help="skip slices an earlier run marked done")
    ap.add_argument("--max-chunks", type=int, default=0, help="stop after this many chunks (0: no limit)")
    a = ap.parse_args(argv)
    env = os.environ
    nouter = int(env.get("Nouter_Loops", "0"))
    tv = None
    if nouter > 0:
        tv = dict(mu=float(env["Opt_mu"]), lam=float(env["Opt_lambda"]), nouter=nouter,
                  ninner=int(env.get("Ninner_Loops", "1")), ncg=int(env.get("Opt_ncg", "5")))
    with Context(0) as ctx:
        st = reconstruct_dataset(ctx, env["ROOT_CT_DIR"], env["EXP_CT_NAME"],
                                 nsub=int(env.get("Sub_Sample_Radial", "1")),
                                 nsub_t=int(env.get("Sub_Sample_Theta", "1")),
                                 wpu=int(env.get("Radon_wpu", "64")), tv=tv, batch=a.batch,
                                 tune=not a.no_tune, threads=a.threads, log=print, resume=a.resume,
                                 chunk=a.chunk, max_chunks=a.max_chunks)
    print(f"{st['slices']} slices of {st['nimage']}^2 x {st['ntheta']} angles: read {st['read_s']:.2f} s, "
          f"reconstruction {st['ms'] / 1e3:.2f} s ({st['slices_per_s']:.1f} slices/s), write {st['write_s']:.2f} s; "
          f"{st['skipped']} done earlier, {st['pending']} pending")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
