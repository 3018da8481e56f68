"""The per-slice done map of a volume run (acumen_amd/io.py DoneMap and the
Fortran done_map_open / done_map_mark of acumen_amd/fortran/acumen_io.F90):
what a resumed run keeps, what resets it, and that both sides read and write
the same file."""
import os
import subprocess

import numpy as np
import pytest

from acumen_amd import fortran_build
from acumen_amd import io as aio

HDR = dict(nimagef=9, nimage=33, ntheta=37, nsub=1, nsub_t=1, wpu=4, ninner=1, nouter=2, ncg=3,
           mu=0.05, lam=0.1)


def header(**kw):
    h = dict(HDR)
    h.update(kw)
    return aio.done_header(**h)


def test_done_map_resume_and_reset(tmp_path):
    path = str(tmp_path / "fbp_tv.done")
    m = aio.DoneMap(path, header(), 9, resume=True)  # no file yet: all pending
    assert m.pending() == list(range(1, 10))
    assert os.path.getsize(path) == aio.DONE_HEADER + 9
    m.mark([2, 5])
    raw = np.fromfile(path, np.uint8)
    assert raw[:8].tobytes() == b"ACUDONE1"
    assert list(np.flatnonzero(raw[aio.DONE_HEADER:]) + 1) == [2, 5]
    assert aio.DoneMap(path, header(), 9, resume=True).pending() == [1, 3, 4, 6, 7, 8, 9]
    # other parameters: the earlier marks do not apply, and the file is reset
    assert aio.DoneMap(path, header(nouter=3), 9, resume=True).pending() == list(range(1, 10))
    assert aio.DoneMap(path, header(), 9, resume=True).pending() == list(range(1, 10))
    m = aio.DoneMap(path, header(), 9, resume=False)
    m.mark([9])
    assert aio.DoneMap(path, header(mu=0.06), 9, resume=True).pending() == list(range(1, 10))
    m = aio.DoneMap(path, header(), 9, resume=False)
    m.mark([1, 9])
    # without resume every slice is pending again and the marks are cleared
    assert aio.DoneMap(path, header(), 9, resume=False).pending() == list(range(1, 10))
    assert aio.DoneMap(path, header(), 9, resume=True).pending() == list(range(1, 10))
    # a file of another length (another volume) is ignored
    m = aio.DoneMap(path, header(), 9, resume=False)
    m.mark([3])
    assert aio.DoneMap(path, header(nimagef=8), 8, resume=True).pending() == list(range(1, 9))


FPROG = """
program done_probe
  use acumen_io
  implicit none
  integer(kind=1) :: done(9)
  character(len=256) :: path, mode
  integer :: k
  call get_command_argument(1, path)
  call get_command_argument(2, mode)
  if (trim(mode) == "fresh") then
    call done_map_open(trim(path), [9, 33, 37, 1, 1, 4, 1, 2, 3, 0], [0.05, 0.1], .false., done)
    call done_map_mark(trim(path), 2)
    call done_map_mark(trim(path), 5)
  else
    call done_map_open(trim(path), [9, 33, 37, 1, 1, 4, 1, 2, 3, 0], [0.05, 0.1], .true., done)
  end if
  write(*, '(9i2)') (int(done(k)), k = 1, 9)
end program done_probe
"""


def test_done_map_fortran_and_python_share_the_file(tmp_path):
    fc = fortran_build.flang()
    if fc is None:
        pytest.skip("flang not available")
    src = tmp_path / "done_probe.F90"
    src.write_text(FPROG)
    io_o = str(tmp_path / "acumen_io.o")
    subprocess.run([fc, "-O1", "-module-dir", str(tmp_path), "-c",
                    os.path.join(fortran_build.FDIR, "acumen_io.F90"), "-o", io_o], check=True)
    exe = str(tmp_path / "done_probe")
    subprocess.run([fc, "-O1", "-I", str(tmp_path), str(src), io_o, "-o", exe], check=True)
    path = str(tmp_path / "fbp_tv.done")

    def fortran(mode):
        out = subprocess.run([exe, path, mode], check=True, capture_output=True, text=True).stdout
        return [int(v) for v in out.split()]

    assert fortran("fresh") == [0, 0, 0, 0, 0, 0, 0, 0, 0]  # the state before the marks
    m = aio.DoneMap(path, header(), 9, resume=True)       # Python reads Fortran's header and marks
    assert m.pending() == [1, 3, 4, 6, 7, 8, 9]
    m.mark([7])
    assert fortran("resume") == [0, 1, 0, 0, 1, 0, 1, 0, 0]  # Fortran reads Python's marks
    # the bytes written by either side are the same
    a = open(path, "rb").read()
    aio.DoneMap(path, header(), 9, resume=True)
    assert open(path, "rb").read() == a


def test_reconstruct_dataset_chunks_and_resume_host_logic(tmp_path, monkeypatch):
    """The chunking, done-map and resume logic of volume.reconstruct_dataset
    with the GPU reconstruction replaced by a deterministic stand-in (each
    image a function of its own sinogram only): a run stopped after one
    chunk and resumed writes the same files as one uninterrupted run, and the
    resumed run reads only the pending slices."""
    from acumen_amd import volume
    reads = []

    def fake_reconstruct(ctx, sinos, theta, wpu, tv=None, fbp=True, batch=16, tune=True):
        reads.append(len(sinos))
        f = np.stack([np.outer(s.sum(0), s[0]) for s in sinos]).astype(np.float32)
        return f, (None if tv is None else f * np.float32(2)), dict(ms=1.0, setup_ms=0.5)

    monkeypatch.setattr(volume, "reconstruct", fake_reconstruct)
    nimagef, nfiles = 12, 7
    th = np.arange(nfiles, dtype=np.float32) * np.float32(30.0)
    rng = np.random.default_rng(3)
    sinos = rng.random((nimagef, nfiles, nimagef + 1), dtype=np.float32)
    ds = aio.Dataset(str(tmp_path), "fake")
    aio.write_dataset(ds, th, sinos, n_files_sino=2, n_files_rec=3)
    tv = dict(mu=0.05, lam=0.1, nouter=2, ninner=1, ncg=3)
    st = volume.reconstruct_dataset(None, ds.root, ds.name, wpu=4, tv=tv, threads=3)
    assert (st["slices"], st["pending"], st["skipped"]) == (nimagef, 0, 0) and reads == [nimagef]
    N = nimagef + 1
    full = [aio.read_images(ds, p, nimagef, N, n_files_rec=3) for p in ("fbp", "fbp_tv")]
    for js in (1, 2, 3):
        for p in ("fbp", "fbp_tv"):
            os.remove(ds.path(f"{p}{js}.dat"))
    reads.clear()
    st = volume.reconstruct_dataset(None, ds.root, ds.name, wpu=4, tv=tv, threads=3, chunk=5, max_chunks=1)
    assert (st["slices"], st["pending"]) == (5, nimagef - 5) and reads == [5]
    st = volume.reconstruct_dataset(None, ds.root, ds.name, wpu=4, tv=tv, threads=3, chunk=5, resume=True)
    assert (st["slices"], st["pending"], st["skipped"]) == (nimagef - 5, 0, 5) and reads == [5, 5, 2]
    for p, want in zip(("fbp", "fbp_tv"), full):
        assert np.array_equal(aio.read_images(ds, p, nimagef, N, n_files_rec=3), want)
    # an FBP-only run keeps its own map (fbp.done) and does not take the TV run's marks
    reads.clear()
    st = volume.reconstruct_dataset(None, ds.root, ds.name, wpu=4, threads=3, resume=True)
    assert st["skipped"] == 0 and reads == [nimagef]
    assert os.path.exists(ds.path("fbp.done")) and os.path.exists(ds.path("fbp_tv.done"))
