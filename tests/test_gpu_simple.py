"""GPU Simple NPR pass (srt_render_simple) against the oracle's Simple pass:
every colouring, bit for bit, on the Cornell box from outside and from inside
its box, and on the 1M-triangle scene; its camera rays are the path tracer's,
so its first hits are the path tracer's first hits."""
import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cornell_pair():
    s = scenes.cornell()
    sc = O.Scene(s)
    return s, sc, O.Bvh(sc)


@pytest.mark.parametrize("type", list(O.SIMPLE))
def test_cornell_matches_oracle(ctx, cornell_pair, type):
    s, sc, ob = cornell_pair
    W, H, spp, seed = 64, 48, 4, 3
    h = ctx.upload(s)
    img = ctx.render_simple(h, s.camera, W, H, type=type, spp=spp, seed=seed)
    ctx.release(h)
    ref, _ = O.render_simple(sc, s.camera, W, H, type=type, spp=spp, seed=seed, bvh=ob)
    assert np.array_equal(img, ref), (type, np.abs(img - ref).max())
    assert np.all(img[:, 3] == 1.0)


def test_camera_inside_the_scene_box(ctx, cornell_pair):
    s, sc, ob = cornell_pair
    cam = dict(position=(0.3, 1.5, 0.9), target=(-0.4, 0.6, -0.8), up=(0.0, 1.0, 0.0), fov=60.0)
    h = ctx.upload(s)
    img = ctx.render_simple(h, cam, 40, 30, type="Depth", spp=2, seed=9)
    ctx.release(h)
    ref, (dmin, _) = O.render_simple(sc, cam, 40, 30, type="Depth", spp=2, seed=9, bvh=ob)
    assert dmin == 0.0
    assert np.array_equal(img, ref)


def test_spheres_matches_oracle_and_path_tracer_first_hits(ctx):
    s = scenes.spheres()
    W, H, seed = 96, 54, 7
    h = ctx.upload(s)
    ctx.build_bvh(h)
    img = {t: ctx.render_simple(h, s.camera, W, H, type=t, spp=2, seed=seed) for t in ("Normal", "Depth")}
    # sample 0's camera rays of the path tracer: the same first hits as the Simple pass
    o, d = O.primary_rays(s.camera, W, H, seed, 0)
    t, u, v, prim = ctx.trace(h, o, d)
    ctx.release(h)
    sc = O.Scene(s)
    ob = O.Bvh(sc)
    for typ, im in img.items():
        ref, _ = O.render_simple(sc, s.camera, W, H, type=typ, spp=2, seed=seed, bvh=ob)
        assert np.array_equal(im, ref), typ
    one, _ = O.render_simple(sc, s.camera, W, H, type="Normal", spp=1, seed=seed, bvh=ob)
    miss = prim < 0
    assert np.array_equal(miss, np.all(one[:, :3] == 1.0, axis=1))


def test_rejects_bad_arguments(ctx):
    s = scenes.cornell()
    h = ctx.upload(s)
    with pytest.raises(Exception):
        ctx.render_simple(h, s.camera, 8, 8, type=5)
    with pytest.raises(Exception):
        ctx.render_simple(h, s.camera, 8, 8, spp=0)
    ctx.release(h)
