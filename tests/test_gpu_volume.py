"""Volume streaming (acu_volume_reconstruct): two lanes, a partial tail
batch, FBP-only and FBP+TV, bit-identical to single-slice plans with the
same variants."""
import numpy as np
import pytest

from acumen_amd import ACU_IMG_FBP, ACU_IMG_TV, ACU_TV_GRAPH, AcumenError, Batch, synth, volume
from helpers import disk_phantom, rel
from oracle import ref

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def make_volume(S, N, T):
    th = synth.angles(T)
    sinos = np.stack([ref.radon(disk_phantom(N, 5, 900 + i), th, "f32") for i in range(S)])
    return th, sinos


@pytest.mark.parametrize("S,batch", [(7, 3), (4, 4), (1, 8)])
def test_volume_fbp_tv_equal_single_plans(ctx, S, batch):
    N, T, wpu = 97, 60, 4
    say("volume", S, "slices, batch", batch)
    th, sinos = make_volume(S, N, T)
    tvp = dict(mu=0.05, lam=0.1, nouter=3, ninner=1, ncg=4)
    fbp, tv, st = volume.reconstruct(ctx, sinos, th, wpu, tv=tvp, batch=batch, tune=False)
    assert st["ms"] > 0 and st["slices_per_s"] > 0
    for i in range(S):
        with Batch(ctx, [(N, T, wpu, th)]) as b:
            b.set_sino(0, sinos[i])
            b.tvopt(0.05, 0.1, 3, 1, 4, ACU_TV_GRAPH)
            assert np.array_equal(fbp[i], b.get_image(0, ACU_IMG_FBP)), i
            assert np.array_equal(tv[i], b.get_image(0, ACU_IMG_TV)), i
    tv64, _, _ = ref.tvopt(sinos[S - 1], th, wpu, 0.05, 0.1, 3, 1, 4, "f64")
    assert rel(tv[S - 1], tv64) < 1e-4


def test_volume_fbp_only_with_tuning(ctx):
    S, N, T, wpu = 5, 129, 90, 4
    say("volume fbp only", S)
    th, sinos = make_volume(S, N, T)
    fbp, tv, _ = volume.reconstruct(ctx, sinos, th, wpu, tv=None, batch=2, tune=True)
    assert tv is None
    for i in (0, S - 1):
        assert rel(fbp[i], ref.iradon(sinos[i], th, wpu, "f32")) < 1e-5


def test_volume_rejects_tv_without_iterations(ctx):
    th, sinos = make_volume(2, 33, 8)
    with pytest.raises(AcumenError):
        volume.reconstruct(ctx, sinos, th, 2, tv=dict(mu=0.05, lam=0.1, nouter=0), batch=2)
