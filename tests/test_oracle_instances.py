"""Oracle pins of the two-level (instanced) scene form (scene.cu:225-288,
hit_info.h:58-62, gmesh.h:77-78, optix_ray_tracer.cu:337-359).

- Every mesh as one identity instance: hits and renders equal the flattened
  scene's bit for bit (apply_point / apply_vector with the identity are exact).
- Exactly representable transforms (dyadic translations; the cyclic axis
  permutation, a rotation by 120 degrees about (1,1,1)) applied to dyadic
  geometry: the instanced scene's first hits (t, u, v, primitive) equal those
  of the flattened scene whose vertices were moved by the same transform, bit
  for bit (every rounding step is exact or permuted identically).
- Closest hits over the per-mesh BVHs equal brute force over all instances
  and triangles, (t, primitive, instance) lexicographic.
- A scaled light instance: samples lie on the transformed quad, the normal is
  apply_vector of the object-space normal without renormalising (its length
  is the scale), and the pdf uses the mesh's object-space area.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from srt_amd import scenes


def _render(s, W=32, H=32, spp=4, depth=5, seed=3, bvh=True):
    osc = O.Scene(s)
    ob = O.Bvh(osc) if bvh else None
    img, _ = O.render(osc, s.camera, W, H, spp, depth, 4, 1, 1, seed, bvh=ob)
    return img


def test_identity_instances_equal_flattened():
    s = scenes.cornell()
    si = scenes.identity_instanced(s)
    assert np.array_equal(_render(s), _render(si))
    o, d = O.primary_rays(s.camera, 48, 48, 5)
    osc, osi = O.Scene(s), O.Scene(si)
    t0, u0, v0, p0 = O.trace(osc, o, d, bvh=O.Bvh(osc))
    t1, u1, v1, p1, i1 = O.trace_inst(osi, o, d, bvh=O.Bvh(osi))
    assert np.array_equal(p0, p1) and np.array_equal(t0.view(np.uint32), t1.view(np.uint32))
    hit = p0 >= 0
    assert np.array_equal(i1[hit], np.searchsorted(s.mesh_first, p0[hit], side="right") - 1)
    assert np.all(i1[~hit] == -1)


def _dyadic_boxes():
    """Two boxes and a floor with coordinates that are multiples of 1/8, lit by a quad."""
    grey = scenes.material(base_color=(0.7, 0.7, 0.7), roughness=1.0)
    red = scenes.material(base_color=(0.7, 0.2, 0.1), roughness=0.5, metallic=0.5)
    light = scenes.material(base_color=(0.7, 0.7, 0.7), emission=(12.0, 12.0, 12.0))
    b = scenes._Builder()
    b.quad((-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2), 0, (0, 1, 0))
    b.box((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5), 1)
    b.box((-0.25, 0.0, -0.25), (0.25, 0.5, 0.25), 0)
    b.quad((-0.5, 3, -0.5), (0.5, 3, -0.5), (0.5, 3, 0.5), (-0.5, 3, 0.5), 2, (0, -1, 0))
    s = b.finish([grey, red, light], with_normals=False, env=(0.1, 0.1, 0.1))
    s.camera = dict(position=(0.5, 1.5, 6.0), target=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
    return s


def _moved(s, which, M):
    """Flattened copy with mesh `which`'s vertices moved by the 3x4 M (float64, exact here)."""
    out = scenes.SceneData(positions=s.positions.copy(), indices=s.indices, mesh_first=s.mesh_first,
                           mesh_material=s.mesh_material, materials=s.materials, env=s.env, camera=s.camera)
    vid = np.unique(s.indices[s.mesh_first[which]:s.mesh_first[which + 1]])
    p = s.positions[vid].astype(np.float64)
    out.positions[vid] = (p @ M[:, :3].T.astype(np.float64) + M[:, 3]).astype(np.float32)
    return out


@pytest.mark.parametrize("kind", ["translate", "cyclic"])
def test_exact_transforms_equal_moved_geometry(kind):
    s = _dyadic_boxes()
    n = s.mesh_material.shape[0]
    xf = np.tile(np.eye(4, dtype=np.float32)[:3], (n, 1, 1))
    if kind == "translate":
        M = scenes.affine(t=(0.75, 0.0, -0.5))
    else:
        # a rotation by 120 degrees about (1, 1, 1) permutes the axes (x -> y -> z -> x)
        M = scenes.affine(R=[[0, 0, 1], [1, 0, 0], [0, 1, 0]], t=(0.0, 1.0, 0.0))
    xf[1] = M
    si = scenes.SceneData(positions=s.positions, indices=s.indices, mesh_first=s.mesh_first,
                          mesh_material=s.mesh_material, materials=s.materials, env=s.env, camera=s.camera,
                          inst_mesh=np.arange(n, dtype=np.int32), inst_xf=xf)
    flat = _moved(s, 1, M)
    o, d = O.primary_rays(s.camera, 64, 64, 9)
    osi, osf = O.Scene(si), O.Scene(flat)
    t1, u1, v1, p1, i1 = O.trace_inst(osi, o, d, bvh=O.Bvh(osi))
    t0, u0, v0, p0 = O.trace(osf, o, d, bvh=O.Bvh(osf))
    assert np.array_equal(p0, p1)
    for a, b in ((t0, t1), (u0, u1), (v0, v1)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (i1 == 1).sum() > 100
    # hit positions are transformed after interpolation (hit_info.h:58), which
    # rounds differently from interpolating moved vertices: the images agree
    # in the mean, not bit for bit
    a, b = _render(si, 32, 32, 32), _render(flat, 32, 32, 32)
    assert abs(a[:, :3].mean() / b[:, :3].mean() - 1) < 0.03


def test_per_mesh_bvh_equals_brute_force():
    s = scenes.spheres_instanced(num_spheres=200, num_meshes=5, seed=4)
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    rng = np.random.default_rng(2)
    n = 3000
    o = rng.uniform(-60, 60, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    aim = rng.uniform(size=n) < 0.5                  # half the rays aim at an instance's centre
    d[aim] = s.inst_xf[rng.integers(0, 200, aim.sum()), :, 3] - o[aim]
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    a = O.trace_inst(osc, o, d, bvh=ob)
    b = O.trace_inst(osc, o, d, bvh=None)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert 0.05 < (a[3] >= 0).mean() < 0.95
    tm = rng.uniform(1, 100, n).astype(np.float32)
    assert np.array_equal(O.trace_inst(osc, o, d, tmax=tm, mode="any", bvh=ob),
                          O.trace_inst(osc, o, d, tmax=tm, mode="any", bvh=None))


def test_scaled_light_instance_sampling():
    s = _dyadic_boxes()
    n = s.mesh_material.shape[0]
    xf = np.tile(np.eye(4, dtype=np.float32)[:3], (n, 1, 1))
    xf[3] = scenes.affine(scale=2.0, t=(0.0, -3.0, 0.0))        # light quad doubled, moved to y = 3
    si = scenes.SceneData(positions=s.positions, indices=s.indices, mesh_first=s.mesh_first,
                          mesh_material=s.mesh_material, materials=s.materials, env=s.env, camera=s.camera,
                          inst_mesh=np.arange(n, dtype=np.int32), inst_xf=xf)
    osc = O.Scene(si)
    cdf = np.zeros(4, np.float32); lm = np.zeros(4, np.int32); wsum = np.zeros(1, np.float32)
    nl = O.lib().orc_scene_lights(osc.h, O._p(cdf), O._p(lm), None, None, O._p(wsum))
    assert nl == 1
    inten = float(s.materials[2]["intensity"])
    assert wsum[0] == pytest.approx(inten * 1.0, rel=1e-6)        # object-space area 1 (not 4)
    rnd = np.random.default_rng(0).uniform(size=(500, 2)).astype(np.float32)
    out = np.zeros((500, 8), np.float32)
    O.lib().orc_light_sample(osc.h, 500, O._p(rnd), O._p(out))
    pos, nrm, pdf = out[:, :3], out[:, 3:6], out[:, 6]
    assert np.allclose(pos[:, 1], 3.0) and np.all(np.abs(pos[:, 0]) <= 1.0) and np.all(np.abs(pos[:, 2]) <= 1.0)
    assert np.allclose(nrm, [0.0, -2.0, 0.0])                     # apply_vector, not renormalised
    assert np.allclose(pdf, inten / wsum[0])
    assert np.abs(pos[:, 0]).max() > 0.9                          # covers the doubled quad


def test_singular_instance_rejected():
    s = _dyadic_boxes()
    n = s.mesh_material.shape[0]
    xf = np.tile(np.eye(4, dtype=np.float32)[:3], (n, 1, 1))
    xf[0, 1, 1] = 0.0
    with pytest.raises(ValueError):
        O.Scene(scenes.SceneData(positions=s.positions, indices=s.indices, mesh_first=s.mesh_first,
                                 mesh_material=s.mesh_material, materials=s.materials,
                                 inst_mesh=np.arange(n, dtype=np.int32), inst_xf=xf))


@pytest.mark.parametrize("type", ["Depth", "Normal", "BaseColor", "Ambient", "FaceOrientation"])
def test_identity_instances_simple_pass_equals_flattened(type):
    """The Simple pass over identity instances: the scene box (the union of
    the instances' transformed mesh boxes) is the vertex box, and every hit
    colour equals the flattened scene's."""
    s = scenes.cornell()
    si = scenes.identity_instanced(s)
    osc, osi = O.Scene(s), O.Scene(si)
    a, ra = O.render_simple(osc, s.camera, 32, 32, type, 2, 5, bvh=O.Bvh(osc))
    b, rb = O.render_simple(osi, s.camera, 32, 32, type, 2, 5, bvh=O.Bvh(osi))
    assert ra == rb
    assert np.array_equal(a, b)


def test_identity_instances_light_tracer_equals_flattened():
    s = scenes.cornell()
    si = scenes.identity_instanced(s)
    osc, osi = O.Scene(s), O.Scene(si)
    a = O.render_light(osc, s.camera, 32, 32, 2, 5, 4, 3, bvh=O.Bvh(osc))
    b = O.render_light(osi, s.camera, 32, 32, 2, 5, 4, 3, bvh=O.Bvh(osi))
    assert np.array_equal(a, b)
    assert a[:, :3].sum() > 0


def test_instanced_scene_box_and_depth_range():
    """Depth range of an instanced scene: a unit cube scaled by 2 about the
    origin and moved by (3, 0, 0), seen from (3, 0, 10): box [2, 4] x [-1, 1] x
    [-1, 1], so the range is |(1, 1, 9)| .. |(1, 1, 11)| (dyadic transform:
    every corner is exact)."""
    b = scenes._Builder()
    b.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0)
    s = b.finish([scenes.material()], with_normals=False)
    s.inst_mesh = np.array([0], np.int32)
    s.inst_xf = scenes.affine(None, (3.0, 0.0, 0.0), 2.0)[None]
    cam = dict(position=(3.0, 0.0, 10.0), target=(3.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=30.0)
    osc = O.Scene(s)
    img, (lo, hi) = O.render_simple(osc, cam, 16, 16, "Depth", 1, 0, bvh=O.Bvh(osc))
    assert lo == np.float32(np.sqrt(np.float32(83.0)))
    assert hi == np.float32(np.sqrt(np.float32(123.0)))
    # a pixel next to the centre sees the front face at distance 9 / cos(angle < 0.05)
    d = (img.reshape(16, 16, 4)[8, 8, 0] - 0.2) / 0.8 * (hi - lo) + lo
    assert 9.0 - 1e-4 <= d <= 9.0 / np.cos(0.05)


def test_mesh_without_triangles_is_refused():
    """Every mesh holds triangles (srt_scene_upload refuses an empty one; so
    does the oracle's scene)."""
    b = scenes._Builder()
    b.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0)
    s = b.finish([scenes.material()], with_normals=False)
    s.mesh_first = np.array([0, 12, 12], np.int32)
    s.mesh_material = np.array([0, 0], np.int32)
    with pytest.raises(ValueError):
        O.Scene(s)


def test_coplanar_faces_closest_hit_is_order_independent():
    """A mirrored and a sheared box standing on the floor (bottom faces
    coplanar with it), rays from anywhere including inside the boxes: the
    per-mesh trees give brute force's (t, primitive, instance) minimum.  The
    watertight test's t of a face can lie below its own box's computed entry
    distance by more than the slab rounding; the traversal's cull margin
    (best distance x (1 + 2^-12)) keeps the order from deciding."""
    s = scenes.cornell_instanced()
    xf = s.inst_xf.copy()
    xf[6, :, :3] = (np.diag([-1.0, 1.0, 1.0]) @ xf[6, :, :3].astype(np.float64)).astype(np.float32)
    xf[7, :, :3] = (np.array([[1.0, 0.35, 0.0], [0.0, 1.0, 0.0], [0.2, 0.0, 1.0]]) @
                    xf[7, :, :3].astype(np.float64)).astype(np.float32)
    s.inst_xf = xf
    osc = O.Scene(s)
    rng = np.random.default_rng(3)
    c = s.inst_xf[:, :, 3].astype(np.float64)
    n = 20000
    o = rng.uniform(c.min(0) - 1.0, c.max(0) + 1.0, size=(n, 3))
    d = rng.normal(size=(n, 3))
    aim = rng.uniform(size=n) < 0.5
    d[aim] = c[rng.integers(0, c.shape[0], aim.sum())] + rng.normal(scale=0.2, size=(aim.sum(), 3)) - o[aim]
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    a = O.trace_inst(osc, o, d, bvh=O.Bvh(osc))
    b = O.trace_inst(osc, o, d, bvh=None)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
