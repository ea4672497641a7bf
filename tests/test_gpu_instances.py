"""GPU two-level (instanced) scenes against the oracle (scene.cu:225-288,
hit_info.h:58-62, gmesh.h:77-78, optix_ray_tracer.cu:337-359 and 443-463):

- first hits (t, u, v, primitive, instance) and any hits equal the oracle's
  bit for bit: both transform the world ray by the same float inverse and
  order ties (t, primitive, instance);
- renders (object-space shading transformed by the instance, instanced light
  sampling) equal the oracle within 1e-4 per pixel;
- every mesh as one identity instance renders the flattened scene's image bit
  for bit;
- new instance transforms (a top-level refit, no rebuild) render what the
  oracle renders with the moved instances, and what a fresh build renders;
- the Simple pass (every colouring, the depth range from the instanced scene
  box of scene.cu:299-307) equals the oracle's bit for bit, and the
  LightTracer's splats equal the oracle's within 1e-4 (float atomic adds);
- skinning and the binary-tree download refuse an instanced scene.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle as O
from srt_amd import abi, scenes

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-6)


def _scene(kind):
    if kind == "cornell":
        return scenes.cornell_instanced()
    return scenes.spheres_instanced(num_spheres=200, num_meshes=5, seed=4)


def _rays(s, n, seed):
    """Half the rays aim at instance centres (jittered), half go anywhere from
    points around them."""
    rng = np.random.default_rng(seed)
    c = s.inst_xf[:, :, 3].astype(np.float64)
    lo, hi = c.min(0) - 1.0, c.max(0) + 1.0
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    aim = rng.uniform(size=n) < 0.5
    tgt = c[rng.integers(0, c.shape[0], aim.sum())] + rng.normal(scale=0.2, size=(aim.sum(), 3))
    d[aim] = tgt - o[aim]
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("kind", ["cornell", "spheres"])
def test_first_and_any_hits_equal_oracle(ctx, kind):
    _say(f"[inst] hits {kind}")
    s = _scene(kind)
    h = ctx.upload(s)
    ctx.build_bvh(h)
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    o, d = _rays(s, 20000, 1)
    po, pd = O.primary_rays(s.camera, 64, 64, 3)
    o = np.concatenate([o, po])
    d = np.concatenate([d, pd])
    t, u, v, p, i = ctx.trace_inst(h, o, d)
    t0, u0, v0, p0, i0 = O.trace_inst(osc, o, d, bvh=ob)
    assert np.array_equal(p, p0), f"{np.count_nonzero(p != p0)} primitives differ"
    assert np.array_equal(i, i0)
    hit = p0 >= 0
    assert hit.mean() > 0.2
    for a, b in ((t, t0), (u, u0), (v, v0)):
        assert np.array_equal(a[hit].view(np.uint32), b[hit].view(np.uint32))
    # any hit below a random fraction of the closest distance, and unbounded
    rng = np.random.default_rng(2)
    tm = np.where(hit, t0 * rng.uniform(0.0, 1.5, size=t0.shape), 1e16).astype(np.float32)
    a = ctx.trace_inst(h, o, d, tmax=tm, any_hit=True)
    a0 = O.trace_inst(osc, o, d, tmax=tm, mode="any", bvh=ob)
    assert np.array_equal(a, a0)
    assert np.array_equal(ctx.trace_inst(h, o, d, any_hit=True), hit.astype(np.int32))
    ctx.release(h)


def test_single_instance_scene(ctx):
    """One instance (the top-level root is a node with one child) of a box."""
    _say("[inst] single instance")
    b = scenes._Builder()
    b.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0)
    s = b.finish([scenes.material()], with_normals=False)
    s.inst_mesh = np.array([0], np.int32)
    s.inst_xf = scenes.affine(scenes._rotation((1, 2, 3), 0.7), (0.25, -0.5, 1.0), 1.5)[None]
    h = ctx.upload(s)
    ctx.build_bvh(h)
    osc = O.Scene(s)
    rng = np.random.default_rng(5)
    o = rng.uniform(-4, 4, size=(4096, 3)).astype(np.float32)
    d = (rng.normal(scale=0.5, size=(4096, 3)) + np.array([0.25, -0.5, 1.0]) - o).astype(np.float32)
    t, u, v, p, i = ctx.trace_inst(h, o, d)
    t0, u0, v0, p0, i0 = O.trace_inst(osc, o, d, bvh=O.Bvh(osc))
    assert np.array_equal(p, p0) and np.array_equal(i, i0)
    assert (p0 >= 0).mean() > 0.2
    assert np.array_equal(t.view(np.uint32), t0.view(np.uint32))
    ctx.release(h)


@pytest.mark.parametrize("kind,W,H,spp,depth", [("cornell", 64, 64, 8, 5), ("spheres", 48, 32, 4, 6)])
def test_render_equals_oracle(ctx, kind, W, H, spp, depth):
    _say(f"[inst] render {kind}")
    s = _scene(kind)
    h = ctx.upload(s)
    ctx.build_bvh(h)
    img = ctx.render(h, s.camera, W, H, spp, depth, 4, 1, 1, 7)
    osc = O.Scene(s)
    ref, _ = O.render(osc, s.camera, W, H, spp, depth, 4, 1, 1, 7, bvh=O.Bvh(osc))
    ctx.release(h)
    r = _rel(img[:, :3], ref[:, :3])
    assert r.max() <= TOL, f"max rel {r.max()} at {np.unravel_index(r.argmax(), r.shape)}"
    assert img[:, :3].mean() > 1e-3


def test_identity_instances_render_flattened_image(ctx):
    _say("[inst] identity instances")
    s = scenes.cornell()
    si = scenes.identity_instanced(s)
    h0 = ctx.upload(s)
    h1 = ctx.upload(si)
    a = ctx.render(h0, s.camera, 64, 64, 8, 5, 4, 1, 1, 9)
    b = ctx.render(h1, s.camera, 64, 64, 8, 5, 4, 1, 1, 9)
    ctx.release(h0)
    ctx.release(h1)
    assert np.array_equal(a, b)


def test_update_instances_refit(ctx):
    """Node animation: the boxes turn and move; the refit top-level tree
    renders the oracle's image of the moved instances and a fresh build's."""
    _say("[inst] update instances")
    s = scenes.cornell_instanced()
    h = ctx.upload(s)
    ctx.build_bvh(h)
    ctx.render(h, s.camera, 32, 32, 2, 5, 4, 1, 1, 1)
    moved = scenes.cornell_instanced(box_angles=(1.1, 0.6))
    xf = moved.inst_xf.copy()
    xf[6, :, 3] += np.array([0.1, 0.0, -0.2], np.float32)
    xf[7, :, 3] += np.array([-0.05, 0.1, 0.1], np.float32)
    moved.inst_xf = xf
    ctx.update_instances(h, xf)
    img = ctx.render(h, s.camera, 64, 64, 8, 5, 4, 1, 1, 13)
    osc = O.Scene(moved)
    ref, _ = O.render(osc, s.camera, 64, 64, 8, 5, 4, 1, 1, 13, bvh=O.Bvh(osc))
    assert _rel(img[:, :3], ref[:, :3]).max() <= TOL
    h2 = ctx.upload(moved)
    ctx.build_bvh(h2)
    fresh = ctx.render(h2, s.camera, 64, 64, 8, 5, 4, 1, 1, 13)
    ctx.release(h2)
    ctx.release(h)
    assert np.array_equal(img, fresh)


@pytest.mark.parametrize("kind", ["cornell", "spheres"])
@pytest.mark.parametrize("type", list(O.SIMPLE))
def test_simple_pass_equals_oracle(ctx, kind, type):
    _say(f"[inst] simple {kind} {type}")
    s = _scene(kind)
    h = ctx.upload(s)
    img = ctx.render_simple(h, s.camera, 48, 40, type=type, spp=2, seed=4)
    ctx.release(h)
    osc = O.Scene(s)
    ref, _ = O.render_simple(osc, s.camera, 48, 40, type=type, spp=2, seed=4, bvh=O.Bvh(osc))
    assert np.array_equal(img, ref), np.abs(img - ref).max()


def test_light_tracer_equals_oracle(ctx):
    _say("[inst] light tracer")
    s = scenes.cornell_instanced()
    h = ctx.upload(s)
    img = ctx.render_light(h, s.camera, 48, 36, spp=2, max_depth=5, rr_depth=2, seed=7)
    ctx.release(h)
    osc = O.Scene(s)
    ref = O.render_light(osc, s.camera, 48, 36, spp=2, max_depth=5, rr_depth=2, seed=7, bvh=O.Bvh(osc))
    floor = 1e-6 * max(np.abs(img[:, :3]).max(), np.abs(ref[:, :3]).max())
    d = np.abs(img[:, :3] - ref[:, :3])
    assert np.all(d <= 1e-4 * np.maximum(np.abs(img[:, :3]), np.abs(ref[:, :3])) + floor), d.max()
    assert img[:, :3].max() > 0


def test_instanced_scene_refusals(ctx):
    _say("[inst] refusals")
    s = scenes.cornell_instanced()
    h = ctx.upload(s)
    ctx.build_bvh(h)
    with pytest.raises(abi.SrtError):
        ctx.bvh_download(h)
    nv = s.positions.shape[0]
    with pytest.raises(abi.SrtError):
        ctx.set_skin(h, np.zeros((nv, 4), np.int32), np.full((nv, 4), 0.25, np.float32))
    with pytest.raises(abi.SrtError):
        ctx.update_instances(h, np.zeros((s.inst_xf.shape[0], 3, 4), np.float32))   # singular
    ctx.release(h)


def test_upload_refuses_a_mesh_without_triangles(ctx):
    _say("[inst] empty mesh")
    b = scenes._Builder()
    b.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0)
    s = b.finish([scenes.material()], with_normals=False)
    s.mesh_first = np.array([0, 12, 12], np.int32)
    s.mesh_material = np.array([0, 0], np.int32)
    s.inst_mesh = np.array([0, 1], np.int32)
    s.inst_xf = np.stack([scenes.affine(), scenes.affine(None, (3.0, 0.0, 0.0))])
    with pytest.raises(abi.SrtError):
        ctx.upload(s)


def test_serving_instanced_scenes_mix_equals_alone():
    """Requests on instanced and flattened scenes through the serving loop
    render the same bits as alone."""
    _say("[inst] serving")
    c = abi.Context(0)
    try:
        ci = scenes.cornell_instanced()
        cf = scenes.cornell()
        si = _scene("spheres")
        hi = c.upload(ci)
        hf = c.upload(cf)
        hs = c.upload(si)
        for h in (hi, hf, hs):
            c.build_bvh(h)
        reqs = [(hi, ci.camera, 48, 40, 4, 5, 21), (hf, cf.camera, 32, 32, 8, 5, 22), (hs, si.camera, 40, 30, 3, 6, 23),
                (hi, ci.camera, 24, 24, 16, 4, 24)]
        alone = [c.render(h, cam, W, H, spp, d, 4, 1, 1, seed) for (h, cam, W, H, spp, d, seed) in reqs]
        c.serve_start(5000)
        try:
            tickets = [c.submit(h, cam, W, H, spp, d, 4, 1, 1, seed) for (h, cam, W, H, spp, d, seed) in reqs]
            for k, (t, img) in enumerate(tickets):
                res = c.wait(t)
                assert res["status"] == 2, (k, res)
                assert np.array_equal(img, alone[k]), k
        finally:
            c.serve_stop()
    finally:
        c.close()


def test_mirrored_and_sheared_instances():
    """Transforms with a reflection (negative determinant: the winding flips)
    and with shear: first hits and renders equal the oracle's."""
    _say("[inst] mirror and shear")
    s = scenes.cornell_instanced()
    xf = s.inst_xf.copy()
    mirror = np.diag([-1.0, 1.0, 1.0]) @ (xf[6, :, :3].astype(np.float64))
    xf[6, :, :3] = mirror.astype(np.float32)
    shear = np.array([[1.0, 0.35, 0.0], [0.0, 1.0, 0.0], [0.2, 0.0, 1.0]]) @ xf[7, :, :3].astype(np.float64)
    xf[7, :, :3] = shear.astype(np.float32)
    s.inst_xf = xf
    c = abi.Context(0)
    try:
        h = c.upload(s)
        c.build_bvh(h)
        osc = O.Scene(s)
        ob = O.Bvh(osc)
        o, d = _rays(s, 20000, 3)
        t, u, v, p, i = c.trace_inst(h, o, d)
        t0, u0, v0, p0, i0 = O.trace_inst(osc, o, d, bvh=ob)
        assert np.array_equal(p, p0) and np.array_equal(i, i0)
        hit = p0 >= 0
        assert np.isin(i0[hit], [6, 7]).any()
        assert np.array_equal(t[hit].view(np.uint32), t0[hit].view(np.uint32))
        img = c.render(h, s.camera, 48, 48, 8, 5, 4, 1, 1, 5)
        ref, _ = O.render(osc, s.camera, 48, 48, 8, 5, 4, 1, 1, 5, bvh=ob)
        assert _rel(img[:, :3], ref[:, :3]).max() <= TOL
    finally:
        c.close()


def test_identity_instanced_million_triangle_field_equals_flat():
    """The 1M-triangle field with each of its 3129 meshes as one identity
    instance (3129 per-mesh trees under the top-level tree) renders the flat
    field's image bit for bit."""
    _say("[inst] 1M identity instances")
    c = abi.Context(0)
    try:
        s = scenes.spheres()
        si = scenes.identity_instanced(s)
        hf = c.upload(s)
        hi = c.upload(si)
        c.build_bvh(hf)
        c.build_bvh(hi)
        a = c.render(hf, s.camera, 320, 180, 4, 8, 4, 1, 1, 3)
        b = c.render(hi, s.camera, 320, 180, 4, 8, 4, 1, 1, 3)
        assert np.array_equal(a, b)
        assert a[:, :3].mean() > 1e-3
    finally:
        c.close()
