"""Closed-form pins of the oracle's linear blend skinning (scene.cu:362-392).

Linear blend skinning (Magnenat-Thalmann et al. 1988; the reference's
Scene::update_gscene) maps a vertex v with (bone id, weight) pairs to
sum_j w_j B_j [v; 1], and normals and tangents to sum_j w_j B_j [n; 0] (no
inverse transpose and no renormalisation, Transform::apply_vector,
transform.h:47-53).  A bone id of -1 ends the list; -1 in the first slot
keeps the original vertex; ids at or above the bone count are skipped.

The expectations here come from homogeneous 4x4 matrix products in float64
(numpy), not from the oracle's per-component loop:
- one bone, weight 1, a rigid transform with small-integer entries (axis
  permutations with signs, integer translations): exact in float32, so the
  oracle must return it bit for bit;
- one bone, weight 1, a general rotation + translation: R v + t and R n
  within float32 rounding;
- identity bones leave every vertex unchanged for weights that sum to 1
  (exactly for weights 1/2, 1/2, 0, 0; within rounding for four quarters);
- weights (w, 1 - w) on two bones give the matrix blend w A + (1 - w) B;
- the list rules (-1 in slot 0, -1 after k bones, ids >= the bone count).
"""
import numpy as np

from oracle import pyoracle as O

RNG = np.random.default_rng(2026)


def _verts(n):
    p = RNG.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    nn = RNG.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    t = RNG.normal(size=(n, 3))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    return p, nn.astype(np.float32), t.astype(np.float32)


def _rot(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K      # Rodrigues


def _bone(R, t):
    m = np.zeros((3, 4))
    m[:, :3] = R
    m[:, 3] = t
    return m.astype(np.float32)


def _apply(m34, v, w_h):
    """[B; 0 0 0 1] [v; w_h] in float64, first three components."""
    M = np.eye(4)
    M[:3, :] = m34.astype(np.float64)
    vh = np.concatenate([v.astype(np.float64), np.full((v.shape[0], 1), w_h)], axis=1)
    return (vh @ M.T)[:, :3]


def _ids(n, *cols):
    ids = np.full((n, 4), -1, np.int32)
    for k, c in enumerate(cols):
        ids[:, k] = c
    return ids


def test_single_rigid_bone_exact():
    n = 64
    p, nn, t = _verts(n)
    p = np.round(p * 4) / 4                       # quarter steps: exact products and sums
    perm = np.array([[0, 0, -1], [1, 0, 0], [0, -1, 0]], np.float64)
    bones = np.stack([_bone(np.eye(3), [0, 0, 0]), _bone(perm, [2, -3, 5])])
    ids = _ids(n, 1)
    w = np.zeros((n, 4), np.float32); w[:, 0] = 1
    P, N, T = O.skin(p, nn, t, ids, w, bones)
    assert np.array_equal(P, _apply(bones[1], p, 1.0).astype(np.float32))
    assert np.array_equal(N, _apply(bones[1], nn, 0.0).astype(np.float32))   # vectors: no translation
    assert np.array_equal(T, _apply(bones[1], t, 0.0).astype(np.float32))


def test_single_general_bone():
    n = 200
    p, nn, t = _verts(n)
    B = _bone(_rot([0.3, -1.0, 0.4], 0.83), [1.5, -0.25, 4.0])
    ids = _ids(n, 0)
    w = np.zeros((n, 4), np.float32); w[:, 0] = 1
    P, N, T = O.skin(p, nn, t, ids, w, B[None])
    np.testing.assert_allclose(P, _apply(B, p, 1.0), rtol=0, atol=4e-6 * 8)
    np.testing.assert_allclose(N, _apply(B, nn, 0.0), rtol=0, atol=4e-7 * 4)
    np.testing.assert_allclose(T, _apply(B, t, 0.0), rtol=0, atol=4e-7 * 4)
    # rigid: lengths of normals and tangents are kept (no renormalisation needed)
    np.testing.assert_allclose(np.linalg.norm(N, axis=1), 1.0, atol=2e-6)


def test_identity_bones_keep_vertices():
    n = 128
    p, nn, t = _verts(n)
    I = _bone(np.eye(3), [0, 0, 0])
    bones = np.stack([I, I, I, I])
    ids = _ids(n, 0, 1, 2, 3)
    w = np.tile(np.array([0.5, 0.5, 0.0, 0.0], np.float32), (n, 1))
    P, N, T = O.skin(p, nn, t, ids, w, bones)
    # 0.5 v + 0.5 v + 0 v + 0 v is exact for every float v (no overflow)
    assert np.array_equal(P, p) and np.array_equal(N, nn) and np.array_equal(T, t)
    # four equal weights: v within float rounding
    w = np.tile(np.array([0.25, 0.25, 0.25, 0.25], np.float32), (n, 1))
    P, N, T = O.skin(p, nn, t, ids, w, bones)
    np.testing.assert_allclose(P, p, rtol=2e-7, atol=0)
    np.testing.assert_allclose(N, nn, rtol=2e-7, atol=1e-30)


def test_two_bone_blend_is_matrix_blend():
    n = 300
    p, nn, t = _verts(n)
    A = _bone(_rot([1, 0.2, 0], 0.4), [0.5, 1.0, -2.0])
    B = _bone(_rot([0, 1, 1], -1.1), [-1.0, 0.0, 3.0])
    wa = RNG.uniform(0, 1, n).astype(np.float32)
    w = np.zeros((n, 4), np.float32)
    w[:, 0] = wa
    w[:, 1] = 1 - wa
    ids = _ids(n, 0, 1)
    P, N, T = O.skin(p, nn, t, ids, w, np.stack([A, B]))
    wa64 = wa.astype(np.float64)[:, None]
    wb64 = w[:, 1].astype(np.float64)[:, None]
    expP = wa64 * _apply(A, p, 1.0) + wb64 * _apply(B, p, 1.0)
    expN = wa64 * _apply(A, nn, 0.0) + wb64 * _apply(B, nn, 0.0)
    np.testing.assert_allclose(P, expP, rtol=0, atol=5e-6 * 8)
    np.testing.assert_allclose(N, expN, rtol=0, atol=1e-6 * 4)
    # the blend of two rotations shortens normals (LBS does not renormalise)
    assert np.min(np.linalg.norm(N, axis=1)) < 0.99


def test_list_rules():
    n = 6
    p, nn, t = _verts(n)
    A = _bone(_rot([0, 0, 1], 0.7), [1, 2, 3])
    B = _bone(_rot([1, 0, 0], -0.3), [-2, 0, 1])
    bones = np.stack([A, B])
    ids = np.array([[-1, 0, 1, 1],        # no bone: the original vertex
                    [0, -1, 1, 1],        # the list ends after bone 0
                    [5, 0, -1, 0],        # id 5 >= 2 bones: skipped
                    [0, 1, -1, -1],
                    [-1, -1, -1, -1],
                    [1, 7, 0, -1]], np.int32)
    w = np.array([[0.3, 0.3, 0.2, 0.2],
                  [1.0, 0.5, 0.5, 0.5],
                  [0.9, 1.0, 0.0, 0.0],
                  [0.25, 0.75, 0.0, 0.0],
                  [1.0, 0.0, 0.0, 0.0],
                  [0.5, 0.3, 0.5, 0.0]], np.float32)
    P, N, T = O.skin(p, nn, t, ids, w, bones)
    assert np.array_equal(P[0], p[0]) and np.array_equal(N[0], nn[0]) and np.array_equal(T[0], t[0])
    assert np.array_equal(P[4], p[4])
    np.testing.assert_allclose(P[1], _apply(A, p[1:2], 1.0)[0], atol=3e-5)
    np.testing.assert_allclose(P[2], _apply(A, p[2:3], 1.0)[0], atol=3e-5)
    np.testing.assert_allclose(P[3], 0.25 * _apply(A, p[3:4], 1.0)[0] + 0.75 * _apply(B, p[3:4], 1.0)[0], atol=3e-5)
    np.testing.assert_allclose(P[5], 0.5 * _apply(B, p[5:6], 1.0)[0] + 0.5 * _apply(A, p[5:6], 1.0)[0], atol=3e-5)
    np.testing.assert_allclose(N[5], 0.5 * _apply(B, nn[5:6], 0.0)[0] + 0.5 * _apply(A, nn[5:6], 0.0)[0], atol=3e-6)


def test_missing_normals_and_tangents():
    n = 16
    p, _, _ = _verts(n)
    B = _bone(_rot([1, 1, 0], 0.5), [0, 1, 0])
    ids = _ids(n, 0)
    w = np.zeros((n, 4), np.float32); w[:, 0] = 1
    P, N, T = O.skin(p, None, None, ids, w, B[None])
    assert N is None and T is None
    np.testing.assert_allclose(P, _apply(B, p, 1.0), atol=3e-5)
