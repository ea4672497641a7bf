"""Oracle pins: BVH build, traversal and refit.

- The Karras radix tree equals a top-down radix tree built by splitting each
  sorted key range at its highest differing bit (keys = code << 32 | index),
  on small inputs with many duplicate codes.
- Structure: every triangle is in exactly one leaf, leaves hold at most
  max_leaf triangles, every child box is exactly the union of its triangles'
  vertex bounds (also after a refit), reported depth is the real depth.
- Closest hit over the BVH equals brute force over all triangles: primitive,
  t, u, v bit for bit (lexicographic (t, primitive) rule); any-hit equals
  brute-force any-hit.
- The binned-SAH top rebuild lowers the SAH cost of the tree on clustered input.
- Watertight test: rays through shared edges and vertices of a closed mesh
  never fall through.
"""
import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O


def soup(n, seed, spread=10.0, size=0.5, dup=False):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, size=(n, 3))
    if dup:
        c[n // 2:] = c[: n - n // 2]   # duplicate centroids -> equal Morton codes
    v = (c[:, None, :] + rng.normal(scale=size, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    if dup:
        v = v.reshape(n, 3, 3)
        v[n // 2:] = v[: n - n // 2]
        v = v.reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    mats = np.asarray([scenes.material()], scenes.MAT_DTYPE)
    return scenes.SceneData(positions=v, indices=idx, mesh_first=np.array([0, n], np.int32),
                            mesh_material=np.array([0], np.int32), materials=mats)


def topdown(keys, lo, hi, out):
    """Children (as ('n', split) / ('l', pos)) of the radix-tree node covering [lo, hi]."""
    if lo == hi:
        return ("l", lo)
    x = keys[lo] ^ keys[hi]
    bit = x.bit_length() - 1
    split = lo
    while split + 1 <= hi and ((keys[split + 1] >> bit) & 1) == ((keys[lo] >> bit) & 1):
        split += 1
    return split


def check_karras(codes, kl, kr):
    n = len(codes)
    keys = [(int(c) << 32) | i for i, c in enumerate(codes)]
    # node index: internal node covering [lo, hi] has index == lo or hi (Karras); traverse from root 0
    stack = [(0, 0, n - 1)]
    seen = 0
    while stack:
        node, lo, hi = stack.pop()
        seen += 1
        g = topdown(keys, lo, hi, None)
        left = -(g + 1) if g == lo else g
        right = -(g + 2) if g + 1 == hi else g + 1
        assert kl[node] == left and kr[node] == right, (node, lo, hi)
        if g != lo:
            stack.append((g, lo, g))
        if g + 1 != hi:
            stack.append((g + 1, g + 1, hi))
    assert seen == n - 1


@pytest.mark.parametrize("n,dup", [(2, False), (9, False), (40, True), (300, True), (1000, False)])
def test_karras_equals_topdown_radix_tree(n, dup):
    s = soup(n, seed=n, dup=dup)
    b = O.Bvh(O.Scene(s))
    _, _, _, codes = b.arrays()
    kl, kr = b.karras()
    assert np.all(np.diff(codes.astype(np.float64)) >= 0)
    if dup:
        assert np.any(np.diff(codes) == 0)
    check_karras(codes, kl, kr)


def walk(nodes, root, n):
    """Yield (node, side, ref, box lo, box hi) for every child slot; returns leaves and depth."""
    f = nodes.view(np.float32)
    i32 = nodes.view(np.int32)
    leaves, depth = [], 0
    if root < 0:
        return [root], 0, []
    stack = [(root, 1)]
    slots = []
    while stack:
        x, d = stack.pop()
        depth = max(depth, d)
        for side in (0, 1):
            if side == 0:
                lo = (f[x, 0], f[x, 2], f[x, 8]); hi = (f[x, 1], f[x, 3], f[x, 9]); ref = i32[x, 12]
            else:
                lo = (f[x, 4], f[x, 6], f[x, 10]); hi = (f[x, 5], f[x, 7], f[x, 11]); ref = i32[x, 13]
            slots.append((x, side, int(ref), np.array(lo), np.array(hi)))
            if ref < 0:
                leaves.append(int(ref))
            else:
                stack.append((int(ref), d + 1))
    return leaves, depth, slots


def leaf_range(ref):
    u = ref & 0xFFFFFFFF
    return (u & 0x7FFFFFFF) >> 3, (u & 7) + 1


def subtree_tris(nodes, ref):
    if ref < 0:
        f, c = leaf_range(ref)
        return list(range(f, f + c))
    i32 = nodes.view(np.int32)
    return subtree_tris(nodes, int(i32[ref, 12])) + subtree_tris(nodes, int(i32[ref, 13]))


def check_structure(b, max_leaf):
    nodes, order, tris, _ = b.arrays()
    info = b.info()
    n = len(order)
    leaves, depth, slots = walk(nodes, info["root"], n)
    cover = np.zeros(n, int)
    for r in leaves:
        f, c = leaf_range(r)
        assert c <= max_leaf
        cover[f:f + c] += 1
    assert np.all(cover == 1)
    assert sorted(order.tolist()) == list(range(n))
    assert depth == info["max_depth"]
    tv = tris.reshape(n, 3, 4)[:, :, :3]
    for x, side, ref, lo, hi in slots:
        t = subtree_tris(nodes, ref)
        pts = tv[t].reshape(-1, 3)
        assert np.array_equal(pts.min(0), lo.astype(np.float32)) and np.array_equal(pts.max(0), hi.astype(np.float32))


@pytest.mark.parametrize("n,leaf,cluster", [(1, 4, 0), (2, 4, 0), (5, 1, 0), (200, 4, 0), (3000, 4, 32),
                                            (3000, 8, 100), (3000, 2, 0)])
def test_structure(n, leaf, cluster):
    s = soup(n, seed=n + leaf)
    b = O.Bvh(O.Scene(s), max_leaf=leaf, cluster_size=cluster)
    check_structure(b, leaf)
    if cluster and n > cluster:
        assert b.info()["clusters"] > 1


def random_rays(n, rng, lo=-12, hi=12):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a few axis-aligned directions (zero components)
    d[: n // 20] = np.eye(3, dtype=np.float32)[rng.integers(0, 3, n // 20)] * rng.choice([-1, 1], size=(n // 20, 1))
    return o, d


@pytest.mark.parametrize("n,leaf,cluster", [(7, 4, 0), (500, 4, 16), (2000, 8, 0), (2000, 1, 64)])
def test_closest_equals_brute_force(n, leaf, cluster):
    rng = np.random.default_rng(n)
    s = soup(n, seed=3 * n, size=1.5)
    osc = O.Scene(s)
    b = O.Bvh(osc, max_leaf=leaf, cluster_size=cluster)
    o, d = random_rays(4000, rng)
    tmax = rng.uniform(1, 40, size=4000).astype(np.float32)
    a = O.trace(osc, o, d, tmax=tmax, bvh=b)
    r = O.trace(osc, o, d, tmax=tmax)
    assert np.array_equal(a[3], r[3])
    assert (a[3] >= 0).mean() > (0.05 if n >= 500 else 0.0)
    for k in range(3):
        assert np.array_equal(a[k].view(np.uint32), r[k].view(np.uint32))
    assert np.array_equal(O.trace(osc, o, d, tmax=tmax, bvh=b, mode="any"), O.trace(osc, o, d, tmax=tmax, mode="any"))


def test_refit_moves_boxes_exactly():
    rng = np.random.default_rng(5)
    s = soup(1500, seed=5)
    osc = O.Scene(s)
    b = O.Bvh(osc, cluster_size=64)
    n0, _, _, _ = b.arrays()
    b.refit()                                   # unchanged vertices: same bytes
    n1, _, _, _ = b.arrays()
    assert np.array_equal(n0, n1)
    moved = s.positions + rng.normal(scale=0.3, size=s.positions.shape).astype(np.float32)
    osc.set_vertices(moved)
    b.refit()
    check_structure(b, 4)
    o, d = random_rays(3000, rng)
    a = O.trace(osc, o, d, bvh=b)
    r = O.trace(osc, o, d)
    assert np.array_equal(a[3], r[3])


def sah_cost(nodes, root, tris):
    """float64 SAH cost: sum of child-box areas (internal) + area * count (leaves)."""
    f = nodes.view(np.float32).astype(np.float64)
    i32 = nodes.view(np.int32)
    total = 0.0
    stack = [root]
    while stack:
        x = stack.pop()
        for lo, hi, ref in (((f[x, 0], f[x, 2], f[x, 8]), (f[x, 1], f[x, 3], f[x, 9]), i32[x, 12]),
                            ((f[x, 4], f[x, 6], f[x, 10]), (f[x, 5], f[x, 7], f[x, 11]), i32[x, 13])):
            e = np.subtract(hi, lo)
            a = 2 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0])
            if ref < 0:
                total += a * leaf_range(int(ref))[1]
            else:
                total += a
                stack.append(int(ref))
    return total


def test_top_level_sah_lowers_cost():
    s = scenes.spheres(num_spheres=200, seed=3)
    osc = O.Scene(s)
    plain = O.Bvh(osc, cluster_size=10 ** 9)      # no rebuild: pure LBVH + collapse
    ref = O.Bvh(osc)                              # automatic clusters
    assert plain.info()["clusters"] == 1 and ref.info()["clusters"] > 1
    c0 = sah_cost(plain.arrays()[0], plain.info()["root"], None)
    c1 = sah_cost(ref.arrays()[0], ref.info()["root"], None)
    assert c1 < c0, (c1, c0)


def test_watertight_closed_mesh():
    """Rays aimed exactly at vertices and edge midpoints of a closed icosphere always hit."""
    uv, ut = scenes.icosphere_f4()
    v = uv.astype(np.float32)
    s = scenes.SceneData(positions=v, indices=ut.astype(np.uint32), mesh_first=np.array([0, 320], np.int32),
                         mesh_material=np.array([0], np.int32),
                         materials=np.asarray([scenes.material()], scenes.MAT_DTYPE))
    osc = O.Scene(s)
    b = O.Bvh(osc)
    targets = np.concatenate([v, 0.5 * (v[ut[:, 0]] + v[ut[:, 1]]), 0.5 * (v[ut[:, 1]] + v[ut[:, 2]])]).astype(np.float32)
    rng = np.random.default_rng(1)
    o = np.tile(np.float32([0.013, -0.021, 0.007]), (len(targets), 1))   # from inside
    d = targets - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    prim = O.trace(osc, o, d, bvh=b)[3]
    assert np.all(prim >= 0)
    o2 = (targets * 3.0).astype(np.float32)                            # from outside, towards the centre
    d2 = targets - o2
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    assert np.all(O.trace(osc, o2, d2, bvh=b)[3] >= 0)


def test_coplanar_faces_closest_equals_brute_force():
    """Rays from inside the Cornell boxes toward the floor: each box's bottom
    face is coplanar with the floor, so two triangles are hit at nearly the
    same computed distance.  The tree traversal (culling at the best distance
    x (1 + 2^-12)) returns brute force's (t, primitive) minimum, also against
    the wall that a box touches."""
    s = scenes.cornell()
    osc = O.Scene(s)
    ob = O.Bvh(osc)
    rng = np.random.default_rng(12)
    n = 20000
    mf = s.mesh_first
    o = []
    for m in (5, 6):                              # the two boxes
        P = s.positions[s.indices[mf[m]:mf[m + 1]].ravel()]
        lo, hi = P.min(0), P.max(0)
        o.append(rng.uniform(lo, hi, size=(n // 2, 3)))
    o = np.concatenate(o).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[:, 1] = -np.abs(d[:, 1]) * rng.uniform(0.02, 1.0, size=n)     # downward, many at grazing angles
    d = d.astype(np.float32)
    a = O.trace(osc, o, d, bvh=ob)
    b = O.trace(osc, o, d, bvh=None)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_cluster_size_holds_at_least_a_leaf():
    """A cluster size below the leaf size is raised to it (a collapsed leaf
    never spans the cluster frontier), so the build is the clamped one."""
    s = soup(2000, seed=11, size=1.5)
    osc = O.Scene(s)
    a = O.Bvh(osc, max_leaf=8, cluster_size=2)
    b = O.Bvh(osc, max_leaf=8, cluster_size=8)
    assert a.info() == b.info() and a.info()["cluster_size"] == 8
    for x, y in zip(a.arrays(), b.arrays()):
        assert np.array_equal(x, y)
