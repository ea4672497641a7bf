"""GPU LightTracer (srt_render_light) against the oracle's LightTracer: per
pixel within 1e-4 relative (plus a floor of 1e-6 of the image's largest
value for near-empty pixels) at equal seeds, on the Cornell box and on a
diffuse floor under a quad light; the splats are float atomic adds, so the
sum order, and the last bits, vary between runs and chunkings."""
import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O
from test_oracle_light import lit_floor

pytestmark = pytest.mark.gpu


def close(a, b):
    floor = 1e-6 * max(np.abs(a).max(), np.abs(b).max())
    d = np.abs(a - b)
    return np.all(d <= 1e-4 * np.maximum(np.abs(a), np.abs(b)) + floor), d.max()


@pytest.mark.parametrize("scene,W,H,spp,depth", [("cornell", 48, 36, 2, 5), ("floor", 40, 30, 4, 3)])
def test_matches_oracle(ctx, scene, W, H, spp, depth):
    s = scenes.cornell() if scene == "cornell" else lit_floor()
    h = ctx.upload(s)
    img = ctx.render_light(h, s.camera, W, H, spp=spp, max_depth=depth, rr_depth=2, seed=7)
    ctx.release(h)
    sc = O.Scene(s)
    ref = O.render_light(sc, s.camera, W, H, spp=spp, max_depth=depth, rr_depth=2, seed=7, bvh=O.Bvh(sc))
    ok, dmax = close(img[:, :3], ref[:, :3])
    assert ok, dmax
    assert np.all(img[:, 3] == 1.0)
    assert img[:, :3].max() > 0


def test_chunking_and_reruns(ctx):
    s = scenes.cornell()
    h = ctx.upload(s)
    a = ctx.render_light(h, s.camera, 32, 24, spp=3, max_depth=4, seed=2)
    ctx.set_option("path_pool", 1000)
    b = ctx.render_light(h, s.camera, 32, 24, spp=3, max_depth=4, seed=2)
    ctx.set_option("path_pool", 1 << 27)
    ctx.release(h)
    ok, dmax = close(a[:, :3], b[:, :3])
    assert ok, dmax


def test_scene_without_lights_is_an_error(ctx):
    b = scenes._Builder()
    b.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), 0, (0, 0, 1))
    s = b.finish([scenes.material()], with_normals=False)
    s.camera = dict(position=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=30.0)
    h = ctx.upload(s)
    with pytest.raises(Exception):
        ctx.render_light(h, s.camera, 8, 8)
    ctx.release(h)
