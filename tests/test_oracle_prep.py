"""Pins the oracle of the projection -> sinogram preprocessing
(tomography/projections_to_sinogram_hpc.F90): the median against numpy's,
the correction by known answers (identity, integer shifts, exactness on a
linear ramp), the mirror error and the centre search by mirrored
projections with a known axis."""
import numpy as np
import pytest

from oracle import ref


def blob_image(n, seed):
    """Smooth positive image, zero in a margin of n/6."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    img = np.zeros((n, n))
    for _ in range(6):
        cx, cy = rng.uniform(n / 3, 2 * n / 3, 2)
        r = rng.uniform(n / 14, n / 8)
        img += rng.uniform(0.5, 1.0) * np.clip(1 - ((x - cx) ** 2 + (y - cy) ** 2) / r ** 2, 0, None)
    return img.astype(np.float32)


@pytest.mark.parametrize("n", [9, 16, 31])
def test_median27_is_numpy_median(n):
    """The pinned median is the intended 27-neighbour one.  The reference's
    code stores its register slots at 3*jxs+jys+2 (+9, +18), which run from
    -2 to 24 in an array of 27 (projections_to_sinogram_hpc.F90:433-435
    against :134): 3 neighbours go out of bounds and 3 stale slots are sorted
    in their place, an undefined result the oracle does not reproduce."""
    rng = np.random.default_rng(n)
    l, c, r = [rng.integers(0, 5, size=(n, n)).astype(np.float32) for _ in range(3)]  # many ties
    got = ref.prep_median27(l, c, r)
    st = np.pad(np.stack([l, c, r]), ((0, 0), (1, 1), (1, 1)), mode="edge")
    want = np.empty_like(c)
    for y in range(n):
        for x in range(n):
            want[y, x] = np.median(st[:, y:y + 3, x:x + 3])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [8, 9])
def test_correct_identity_and_integer_shift(n):
    img = np.random.default_rng(1).normal(size=(n, n)).astype(np.float32)
    nw = n + (1 if n % 2 == 0 else 0)
    out = ref.prep_correct(img, 0.0, 0.0)
    assert out.shape == (n, nw)
    # xr = jx, yr = jy; at jx or jy = n the clamped tap pair has zero weights
    assert np.array_equal(out[:n - 1, :n - 1], img[:n - 1, :n - 1])
    assert not out[n - 1].any() and not out[:, n - 1:].any()
    for k in (1, 2, -2):
        out = ref.prep_correct(img, 0.0, np.float32(k) / np.float32(n))
        for jx in range(1, n):          # 1-based output column
            src = jx - k                # xr = jx - k, clamped into [1, n]
            if src >= n:
                assert not out[:n - 1, jx - 1].any()
            else:
                assert np.array_equal(out[:n - 1, jx - 1], img[:n - 1, max(src, 1) - 1]), (k, jx)


def test_correct_reproduces_a_linear_ramp():
    n = 64
    y, x = np.mgrid[1:n + 1, 1:n + 1].astype(np.float64)
    a, b, c0 = 0.3, -0.7, 50.0
    img = (a * x + b * y + c0).astype(np.float32)
    th, sh = 0.02, 0.013
    out = ref.prep_correct(img, th, sh)
    nw = n + 1
    mx, my, S = (nw - 1) / 2, (n - 1) / 2, n
    jy, jx = np.mgrid[1:n + 1, 1:nw + 1].astype(np.float64)
    u = jx - mx - S * np.float32(sh)
    xr = np.cos(th) * u - np.sin(th) * (jy - my) + mx
    yr = np.sin(th) * u + np.cos(th) * (jy - my) + my
    inside = (xr > 1.5) & (xr < n - 1.5) & (yr > 1.5) & (yr < n - 1.5)
    want = a * xr + b * yr + c0
    assert inside.sum() > 0.6 * n * n
    assert np.abs(out[inside] - want[inside]).max() < 2e-4 * c0


def test_mirror_error_and_centre_of_mirrored_projections():
    n = 63  # odd: Nw = n
    p0 = blob_image(n, 3)
    p180 = p0[:, ::-1].copy()  # p180(x) = p0(n + 1 - x)
    assert ref.prep_mirror_error(p0, p180, 0.0, 0.0) == 0.0
    assert ref.prep_mirror_error(p0, p180, 0.0, 0.05) > 0.0
    th, sh = ref.prep_find_center(p0, p180)
    assert abs(th) < 1e-4 and abs(sh * n) < 0.02


@pytest.mark.parametrize("d", [2, -3])
def test_centre_search_recovers_a_shifted_axis(d):
    # axis d pixels right of the centre: p180(x) = p0(n + 1 + 2d - x), which
    # the resampled pair matches at shift = -d / n
    n = 63
    p0 = blob_image(n, 5)
    idx = np.clip(n - 1 + 2 * d - np.arange(n), 0, n - 1)   # 1-based n + 1 + 2d - x
    p180 = p0[:, idx].copy()
    assert abs(ref.prep_centroid_shift(p0, p180) * n + d) < 0.5
    th, sh = ref.prep_find_center(p0, p180)
    assert abs(sh * n + d) < 0.02 and abs(th) < 1e-3, (th, sh * n)
