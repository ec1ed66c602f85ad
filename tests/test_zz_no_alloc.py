"""gdem_scene_download_noalloc: the same result as gdem_scene_download with
no device allocation, through the library's fixed staging array, on two
calls in a row (the second with a smaller scene, so stale entries of the
first would show), and a clear error beyond the staging size."""
import numpy as np
import pytest

from gpudem_amd import GDEM_FP64, GDEM_MIXED, GdemError, Scene, engine, scenes

pytestmark = pytest.mark.gpu


def make(ctx, spec, precision):
    m, k, e = engine.material(spec.r, spec.density, spec.young, spec.cor)
    return Scene(ctx, spec.x, spec.r, m, k, e, spec.verts, dict(spec.params, precision=precision))


def same(a, b):
    for key in ("x", "v", "angv", "active"):
        assert np.array_equal(a[key], b[key]), key


@pytest.mark.parametrize("precision", [GDEM_FP64, GDEM_MIXED])
def test_download_noalloc_matches_download(ctx, precision):
    big = make(ctx, scenes.box_scene(71, 20000), precision)
    small = make(ctx, scenes.box_scene(72, 3000), precision)
    print(f"box scenes 20000 and 3000, precision {precision}", flush=True)
    big.step(400)
    small.step(300)
    same(big.download_noalloc(), big.download())
    # second call, smaller scene: every entry it returns is its own
    got = dict(x=np.full((3000, 3), np.nan), v=np.full((3000, 3), np.nan), angv=np.full((3000, 3), np.nan),
               active=np.full(3000, -7, np.int32))
    small.download_noalloc(got)
    same(got, small.download())
    # the staging keeps no state a later call depends on
    big.step(100)
    same(big.download_noalloc(), big.download())
    big.close()
    small.close()


def test_download_noalloc_rejects_oversized_scene(ctx):
    sc = make(ctx, scenes.box_scene(73, (1 << 18) + 1), GDEM_MIXED)
    with pytest.raises(GdemError) as err:
        sc.download_noalloc()
    assert err.value.code == -1
    sc.close()
