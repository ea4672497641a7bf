"""GPU wavefront integrator parity with the host execution of the reference
integrator (oracle, float): per-pixel radiance within 1e-4 relative at equal
seeds (observed: identical bits), independence from queue chunking, and
post-processing parity."""
import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-4   # per-pixel relative tolerance at equal seeds


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-6)


def _render_pair(ctx, s, W, H, spp, depth, seed=1, pool=None, threads=0):
    h = ctx.upload(s)
    ctx.build_bvh(h)
    if pool:
        ctx.set_option("path_pool", pool)
    img = ctx.render(h, s.camera, W, H, spp, depth, 4, 1, 1, seed)
    if pool:
        ctx.set_option("path_pool", 1 << 27)
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    ref, _ = O.render(osc, s.camera, W, H, spp, depth, 4, 1, 1, seed, bvh=ob, threads=threads)
    ctx.release(h)
    return img, ref


def test_cornell_full_config(ctx):
    s = scenes.cornell()
    img, ref = _render_pair(ctx, s, 256, 256, 16, 5)
    r = _rel(img[:, :3], ref[:, :3])
    assert r.max() <= TOL, f"max rel {r.max()} at {np.unravel_index(r.argmax(), r.shape)}"
    assert np.all(img[:, 3] == 1.0)
    assert img[:, :3].mean() > 0.01


def test_cornell_deep_paths(ctx):
    s = scenes.cornell()
    img, ref = _render_pair(ctx, s, 64, 64, 8, 16, seed=5)
    assert _rel(img[:, :3], ref[:, :3]).max() <= TOL


def test_chunking_independent(ctx):
    s = scenes.cornell()
    h = ctx.upload(s)
    a = ctx.render(h, s.camera, 64, 48, 8, 5, 4, 1, 1, 3)
    ctx.set_option("path_pool", 1000)   # chunk boundaries inside sample planes
    b = ctx.render(h, s.camera, 64, 48, 8, 5, 4, 1, 1, 3)
    ctx.set_option("path_pool", 1 << 27)
    ctx.release(h)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("pool", [64, 65, 100, 2047, 2049])
def test_queue_partition_shapes(ctx, pool):
    """Chunks of 64 to 2049 paths leave most of the 32 queue partitions empty or
    ragged (partition capacity = 64 * ceil(ceil(n / 64) / 32)): the image and
    the queue totals equal the single-chunk render's."""
    s = scenes.cornell()
    h = ctx.upload(s)
    a = ctx.render(h, s.camera, 40, 30, 4, 5, 4, 1, 1, 11)
    ca = ctx.last_counts()
    ctx.set_option("path_pool", pool)
    b = ctx.render(h, s.camera, 40, 30, 4, 5, 4, 1, 1, 11)
    cb = ctx.last_counts()
    ctx.set_option("path_pool", 1 << 27)
    ctx.release(h)
    assert np.array_equal(a, b)
    assert ca == cb


@pytest.mark.parametrize("W,H,spp", [(1, 1, 1), (3, 2, 5), (7, 9, 1)])
def test_tiny_images(ctx, W, H, spp):
    """Fewer paths than one 64-entry turn of a partition: every other partition
    is empty in every queue."""
    s = scenes.cornell()
    img, ref = _render_pair(ctx, s, W, H, spp, 5, seed=21)
    assert np.array_equal(img, ref)


@pytest.mark.parametrize("depth", [0, 1])
def test_shallow_paths(ctx, depth):
    """max_depth 0 (no bounce: the radiance is cleared, not shaded) and 1 (the
    first bounce alone, whose stages compute the camera rays) equal the oracle,
    also after a deeper render has left radiance in the same path slots."""
    s = scenes.cornell()
    _render_pair(ctx, s, 24, 16, 3, 5, seed=4)
    img, ref = _render_pair(ctx, s, 24, 16, 3, depth, seed=4)
    assert np.array_equal(img, ref)
    if depth == 0:
        assert np.all(img[:, :3] == 0.0)


@pytest.mark.parametrize("packet", [1, 16, 64])
def test_packet_traversal_equals_per_ray(ctx, packet):
    """Camera rays traversed as 64-ray packets (option trace_packet: lanes with
    postponed leaves that make a wave test them) give the per-ray traversal's
    image bit for bit, also for partial last waves and scene edges."""
    for s, W, H, spp, depth in ((scenes.spheres(), 96, 54, 2, 3), (scenes.cornell(), 37, 23, 3, 4)):
        h = ctx.upload(s)
        ctx.build_bvh(h)
        try:
            ctx.set_option("trace_packet", 0)
            ref = ctx.render(h, s.camera, W, H, spp, depth, 4, 1, 1, 5)
            ctx.set_option("trace_packet", packet)
            img = ctx.render(h, s.camera, W, H, spp, depth, 4, 1, 1, 5)
        finally:
            ctx.set_option("trace_packet", 0)
            ctx.release(h)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    with pytest.raises(Exception):
        ctx.set_option("trace_packet", 65)


def test_spheres_small(ctx):
    s = scenes.spheres()
    img, ref = _render_pair(ctx, s, 192, 108, 2, 8)
    assert _rel(img[:, :3], ref[:, :3]).max() <= TOL


def test_tube_small(ctx):
    s = scenes.tube()
    img, ref = _render_pair(ctx, s, 128, 72, 2, 6)
    assert _rel(img[:, :3], ref[:, :3]).max() <= TOL


def test_nee_mis_switches(ctx):
    s = scenes.cornell()
    h = ctx.upload(s)
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    for nee, mis in ((0, 0), (1, 0), (0, 1)):
        img = ctx.render(h, s.camera, 48, 48, 4, 6, 2, nee, mis, 9)
        ref, _ = O.render(osc, s.camera, 48, 48, 4, 6, 2, nee, mis, 9, bvh=ob)
        assert _rel(img[:, :3], ref[:, :3]).max() <= TOL, (nee, mis)
    ctx.release(h)


@pytest.mark.parametrize("tm", ["None", "Clamp", "Reinhard", "Uncharted2", "ACES"])
def test_tonemap_matches_oracle(ctx, tm):
    rng = np.random.default_rng(1)
    px = np.concatenate([rng.uniform(0, 4, size=(4000, 4)), rng.uniform(0, 1e-3, size=(100, 4))]).astype(np.float32)
    for gamma in (True, False):
        for ex in (0.0, 1.5):
            a = ctx.tonemap(px, tm, ex, gamma)
            b = O.tonemap(px, tm, ex, gamma)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (tm, gamma, ex)


def test_accumulate_matches_oracle(ctx):
    rng = np.random.default_rng(2)
    acc = np.zeros((1000, 4), np.float32)
    acc2 = acc.copy()
    for f in range(5):
        px = rng.uniform(size=(1000, 4)).astype(np.float32)
        acc, out = ctx.accumulate(acc, px, f)
        acc2, out2 = O.accumulate(acc2, px, f)
        assert np.array_equal(acc, acc2) and np.array_equal(out, out2)


def test_converges_to_fp64_golden(ctx):
    """GPU float render at 256 spp against the committed FP64 4096-spp image
    (MC noise bound: the oracle's own float render measures RMSE 0.017)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cornell_64x64_fp64_4096spp.npy"))
    s = scenes.cornell()
    h = ctx.upload(s)
    img = ctx.render(h, s.camera, 64, 64, 256, 5, 4, 1, 1, 1000)
    ctx.release(h)
    rmse = np.sqrt(np.mean((img[:, :3].astype(np.float64) - g) ** 2))
    assert rmse < 0.03, rmse
    assert abs(img[:, :3].mean() / g.mean() - 1) < 0.01


def test_nan_samples_dropped_like_the_oracle(ctx):
    """R12: samples whose radiance is NaN/Inf are dropped but keep the 1/spp
    weight; the GPU frame equals the oracle's on a scene with a NaN material."""
    from test_oracle_closed_forms import nan_cornell
    s = nan_cornell()
    img, ref = _render_pair(ctx, s, 48, 48, 4, 5, seed=3)
    assert np.isfinite(img).all()
    assert np.array_equal(img, ref)
    assert (img[:, :3] == 0).all(axis=1).sum() > 10


@pytest.mark.parametrize("name,W,H,spp,depth", [("cornell", 64, 64, 8, 5), ("spheres", 192, 108, 2, 8),
                                                ("tube", 128, 72, 2, 6)])
def test_queue_counts_equal_megakernel_traces(ctx, name, W, H, spp, depth):
    """The wavefront queues hold exactly the megakernel's rays: segments traced
    = the oracle's closest-hit traces, shadow rays = its shadow tests, for the
    same seeded paths (no entry lost or duplicated by compaction)."""
    s = getattr(scenes, name)()
    h = ctx.upload(s)
    ctx.build_bvh(h)
    ctx.render(h, s.camera, W, H, spp, depth, 4, 1, 1, 2)
    seg, sh = ctx.last_counts()
    ctx.release(h)
    osc = O.Scene(s)
    O.render(osc, s.camera, W, H, spp, depth, 4, 1, 1, 2, bvh=O.Bvh(osc))
    traces, shadows = O.last_counts()
    assert seg == traces
    assert sh == shadows
    assert seg >= W * H * spp


@pytest.mark.parametrize("frame", [17, 42])
def test_skinned_tube_posed_frame(ctx, frame):
    """Config 2 on a posed frame: linear blend skinning of the 200k-triangle
    tube with frame `frame`'s bone matrices, BVH refit, then a render at
    reduced resolution (depth 6, 2 spp), per pixel within 1e-4 of the oracle's
    render of the oracle-skinned vertices (scene.cu:336-395, then
    optix_ray_tracer.cu:443-463, then path_tracer.cu:50-169)."""
    s = scenes.tube()
    h = ctx.upload(s)
    ctx.build_bvh(h)
    ctx.set_skin(h, s.skin["bone_ids"], s.skin["bone_weights"])
    M = s.skin["matrices"](frame)
    ctx.skin_refit(h, M)
    W, H = 160, 90
    img = ctx.render(h, s.camera, W, H, 2, 6, 4, 1, 1, 100 + frame)
    ctx.release(h)
    p, n, t = O.skin(s.positions, s.normals, s.tangents, s.skin["bone_ids"], s.skin["bone_weights"], M)
    posed = scenes.SceneData(positions=p, indices=s.indices, mesh_first=s.mesh_first, mesh_material=s.mesh_material,
                             materials=s.materials, normals=n, tangents=t, env=s.env, camera=s.camera)
    osc = O.Scene(posed)
    ob = O.Bvh(osc)
    ref, _ = O.render(osc, s.camera, W, H, 2, 6, 4, 1, 1, 100 + frame, bvh=ob)
    r = _rel(img[:, :3], ref[:, :3])
    assert r.max() <= TOL, f"max rel {r.max()} at {np.unravel_index(r.argmax(), r.shape)}"
    # the pose moved the tube: the image differs from the rest pose's
    rest, _ = O.render(O.Scene(s), s.camera, W, H, 2, 6, 4, 1, 1, 100 + frame)
    assert np.mean(np.any(np.abs(rest[:, :3] - ref[:, :3]) > 1e-3, axis=1)) > 0.02
