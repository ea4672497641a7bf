"""GPU BVH build / refit / skinning parity: node bytes, Morton order and
triangle arrays identical to the oracle's build of the same scene."""
import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _soup(n, seed, spread=10.0, size=0.5):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, size=(n, 3))
    v = (c[:, None, :] + rng.normal(scale=size, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    mats = np.asarray([scenes.material()], scenes.MAT_DTYPE)
    return scenes.SceneData(positions=v, indices=idx, mesh_first=np.array([0, n], np.int32),
                            mesh_material=np.array([0], np.int32), materials=mats)


def _compare(ctx, s, max_leaf=4, cluster_size=0):
    h = ctx.upload(s)
    ctx.build_bvh(h, max_leaf=max_leaf, cluster_size=cluster_size)
    nodes, order, tris, info = ctx.bvh_download(h)
    ob = O.Bvh(O.Scene(s), max_leaf=max_leaf, cluster_size=cluster_size)
    n2, o2, t2, _ = ob.arrays()
    oi = ob.info()
    ctx.release(h)
    assert info["root"] == oi["root"]
    assert info["max_depth"] == oi["max_depth"]
    assert info["clusters"] == oi["clusters"]
    np.testing.assert_array_equal(order, o2)
    assert np.array_equal(tris.view(np.uint32), t2.view(np.uint32))
    assert nodes.shape == n2.shape
    bad = np.nonzero((nodes != n2).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} nodes differ, first {bad[:5]}"
    return info


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 65, 1000])
def test_small_soups(ctx, n):
    _compare(ctx, _soup(n, seed=n))


@pytest.mark.parametrize("cluster", [8, 16, 100])
def test_top_level_sah(ctx, cluster):
    info = _compare(ctx, _soup(5000, seed=9), cluster_size=cluster)
    assert info["clusters"] > 1


@pytest.mark.parametrize("leaf", [1, 2, 8])
def test_leaf_sizes(ctx, leaf):
    _compare(ctx, _soup(3000, seed=4), max_leaf=leaf, cluster_size=64)


def test_cluster_size_below_leaf_size_is_raised(ctx):
    info = _compare(ctx, _soup(2000, seed=11), max_leaf=8, cluster_size=2)
    assert info["cluster_size"] == 8


def test_cornell(ctx):
    _compare(ctx, scenes.cornell())


def test_rebuilds_with_kept_scratch(ctx):
    """Builds take the device scratch chunks kept by earlier builds (larger and
    smaller scenes, the same scene again): every tree equals the oracle's, and
    with keep_build_scratch 0 the chunks are freed."""
    _compare(ctx, _soup(40000, seed=21))
    _compare(ctx, _soup(3000, seed=22))
    s = _soup(20000, seed=23)
    h = ctx.upload(s)
    ctx.build_bvh(h)
    first = ctx.bvh_download(h)
    for _ in range(2):
        ctx.build_bvh(h)
        again = ctx.bvh_download(h)
        for a, b in zip(first[:3], again[:3]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    kept = ctx.lib.srt_hbm_used(ctx.h)
    ctx.set_option("keep_build_scratch", 0)
    try:
        assert ctx.lib.srt_hbm_used(ctx.h) < kept
        ctx.build_bvh(h)
        assert ctx.lib.srt_hbm_used(ctx.h) < kept
        again = ctx.bvh_download(h)
        for a, b in zip(first[:3], again[:3]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    finally:
        ctx.set_option("keep_build_scratch", 1)
        ctx.release(h)
    _compare(ctx, _soup(5000, seed=24), cluster_size=16)


def test_spheres_1m(ctx):
    info = _compare(ctx, scenes.spheres(), max_leaf=2)     # the library's defaults
    assert info["triangles"] == 1000008
    assert info["clusters"] > 100


def test_refit_identity_and_skin(ctx):
    s = scenes.tube()
    h = ctx.upload(s)
    ctx.build_bvh(h)
    nodes0, order0, tris0, _ = ctx.bvh_download(h)
    ctx.set_skin(h, s.skin["bone_ids"], s.skin["bone_weights"])
    ident = np.tile(np.eye(4, dtype=np.float32)[:3], (s.skin["num_bones"], 1, 1))
    ctx.skin_refit(h, ident)
    nodes1, order1, tris1, _ = ctx.bvh_download(h)
    # identity bones: weights sum to 1 in float only approximately, so compare to the oracle
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    n0, _, _, _ = ob.arrays()
    assert np.array_equal(nodes0, n0)
    for frame in (0, 17, 42):
        M = s.skin["matrices"](frame)
        ctx.skin_refit(h, M)
        p, n, t = ctx.download_vertices(h, s.positions.shape[0], normals=True, tangents=True)
        p2, n2, t2 = O.skin(s.positions, s.normals, s.tangents, s.skin["bone_ids"], s.skin["bone_weights"], M)
        assert np.array_equal(p.view(np.uint32), p2.view(np.uint32))
        assert np.array_equal(n.view(np.uint32), n2.view(np.uint32))
        assert np.array_equal(t.view(np.uint32), t2.view(np.uint32))
        osc.set_vertices(p2, n2, t2)
        ob.refit()
        nodes, order, tris, _ = ctx.bvh_download(h)
        nr, orr, tr, _ = ob.arrays()
        np.testing.assert_array_equal(order, orr)
        assert np.array_equal(tris.view(np.uint32), tr.view(np.uint32))
        assert np.array_equal(nodes, nr)
    ctx.release(h)
