"""GPU traversal parity: first-hit primitive ids, t and barycentrics bit-exact
against the oracle's traversal (itself pinned to brute force on the CPU), and
shadow-ray visibility equal to the oracle's any-hit answer."""
import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _first_hits(ctx, s, W, H, seed=1):
    h = ctx.upload(s)
    ctx.build_bvh(h)
    o, d = O.primary_rays(s.camera, W, H, seed)
    t, u, v, prim = ctx.trace(h, o, d)
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    t2, u2, v2, p2 = O.trace(osc, o, d, bvh=ob)
    ctx.release(h)
    return (t, u, v, prim), (t2, u2, v2, p2)


def test_cornell_first_hit_bit_exact(ctx):
    s = scenes.cornell()
    (t, u, v, p), (t2, u2, v2, p2) = _first_hits(ctx, s, 256, 256)
    assert (p >= 0).mean() > 0.3
    np.testing.assert_array_equal(p, p2)
    hit = p >= 0
    assert np.array_equal(t[hit].view(np.uint32), t2[hit].view(np.uint32))
    assert np.array_equal(u[hit].view(np.uint32), u2[hit].view(np.uint32))
    assert np.array_equal(v[hit].view(np.uint32), v2[hit].view(np.uint32))


def test_spheres_first_hit_bit_exact(ctx):
    s = scenes.spheres()
    (t, u, v, p), (t2, u2, v2, p2) = _first_hits(ctx, s, 480, 270)
    assert (p >= 0).mean() > 0.2
    np.testing.assert_array_equal(p, p2)
    hit = p >= 0
    assert np.array_equal(t[hit].view(np.uint32), t2[hit].view(np.uint32))
    assert np.array_equal(u[hit].view(np.uint32), u2[hit].view(np.uint32))


def test_random_rays_closest_and_any(ctx):
    rng = np.random.default_rng(3)
    s = scenes.cornell()
    h = ctx.upload(s)
    ctx.build_bvh(h)
    n = 20000
    o = rng.uniform([-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmax = rng.uniform(0.05, 3.0, size=n).astype(np.float32)
    t, u, v, p = ctx.trace(h, o, d, tmax=tmax)
    osc = O.Scene(s)
    t2, u2, v2, p2 = O.trace(osc, o, d, tmax=tmax)          # brute force
    np.testing.assert_array_equal(p, p2)
    vis = ctx.trace(h, o, d, tmax=tmax, any_hit=True)[3]
    vis2 = O.trace(osc, o, d, tmax=tmax, mode="any")
    np.testing.assert_array_equal(vis, vis2)
    assert 0.05 < vis.mean() < 0.95
    ctx.release(h)


def test_watertight_and_axis_rays_match_brute_force(ctx):
    """Rays through vertices / edges of a closed mesh and rays lying in box
    planes with zero direction components (the slab NaN case)."""
    uv, ut = scenes.icosphere_f4()
    v = uv.astype(np.float32)
    s = scenes.SceneData(positions=v, indices=ut.astype(np.uint32), mesh_first=np.array([0, 320], np.int32),
                         mesh_material=np.array([0], np.int32),
                         materials=np.asarray([scenes.material()], scenes.MAT_DTYPE))
    h = ctx.upload(s)
    ctx.build_bvh(h)
    targets = np.concatenate([v, 0.5 * (v[ut[:, 0]] + v[ut[:, 1]])]).astype(np.float32)
    o = (targets * 3.0).astype(np.float32)
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t, u, vv, p = ctx.trace(h, o, d)
    assert np.all(p >= 0)
    osc = O.Scene(s)
    np.testing.assert_array_equal(p, O.trace(osc, o, d)[3])
    ctx.release(h)


@pytest.mark.parametrize("name", ["cornell", "spheres", "tube"])
def test_scene_upload_round_trip(ctx, name):
    """srt_scene_upload keeps the vertices bit for bit and builds the light and
    area CDFs exactly as the oracle does (scene.cu:240-287, mesh.cpp:3-18)."""
    s = getattr(scenes, name)()
    h = ctx.upload(s)
    nv = s.positions.shape[0]
    has_n = s.normals is not None
    has_t = getattr(s, "tangents", None) is not None
    p, n, t = ctx.download_vertices(h, nv, normals=has_n, tangents=has_t)
    assert np.array_equal(p.view(np.uint32), np.ascontiguousarray(s.positions, np.float32).view(np.uint32))
    if has_n:
        assert np.array_equal(n.view(np.uint32), np.ascontiguousarray(s.normals, np.float32).view(np.uint32))
    if has_t:
        assert np.array_equal(t.view(np.uint32), np.ascontiguousarray(s.tangents, np.float32).view(np.uint32))
    osc = O.Scene(s)
    got = ctx.lights(h, s.num_triangles, len(s.mesh_material))
    ref = osc.lights()
    assert len(got[0]) == len(ref[0]) > 0
    for a, b in zip(got[:4], ref[:4]):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    assert got[4] == ref[4]
    ctx.release(h)
