"""The oracle's Simple NPR pass (shader/npr/simple.cu:42-97) against closed
forms: the scene-box distances of aabb.h:24-41 (including the reference's
per-axis min_distance, which is not the Euclidean distance to the box), the
five colourings on a single square whose normal and distances are known, and
white misses."""
import numpy as np

from srt_amd import scenes
from oracle import pyoracle as O

FACE_FRONT = np.array([0.34, 0.73, 0.76])
FACE_BACK = np.array([1.0, 0.85, 0.0])


def square(facing=(0, 0, 1), color=(0.2, 0.5, 0.8)):
    """Square [-1, 1]^2 at z = 0 whose geometric normal points along `facing`,
    and a hidden one at z = -5 behind it, so the scene box has depth."""
    b = scenes._Builder()
    b.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), 0, facing)
    b.quad((-1, -1, -5), (1, -1, -5), (1, 1, -5), (-1, 1, -5), 0, (0, 0, 1))
    s = b.finish([scenes.material(base_color=color)], with_normals=False, name="square")
    # 20 degrees at distance 3: the view is 1.06 wide at z = 0, inside the square
    s.camera = dict(position=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=20.0)
    return s


def render(s, type, spp=1, W=24, H=16, seed=4, cam=None):
    sc = O.Scene(s)
    return O.render_simple(sc, cam or s.camera, W, H, type=type, spp=spp, seed=seed, bvh=O.Bvh(sc))


def test_box_distances_follow_aabb_h():
    b = scenes._Builder()
    b.box((0, 0, 0), (1, 1, 1), 0)
    s = b.finish([scenes.material()], with_normals=False)
    sc = O.Scene(s)
    # outside: per axis the nearer of |p - lo|, |p - hi| (0.5 on x and y although p is inside those slabs)
    cam = dict(position=(0.5, 0.5, 3.0), target=(0.5, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=30.0)
    _, (dmin, dmax) = O.render_simple(sc, cam, 4, 4, type="Depth", spp=1, bvh=O.Bvh(sc))
    assert dmin == np.float32(np.sqrt(np.float32(4.5)))
    assert dmax == np.float32(np.sqrt(np.float32(9.5)))
    # inside the box: min_distance is 0
    cam = dict(position=(0.5, 0.5, 0.5), target=(0.5, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=30.0)
    _, (dmin, dmax) = O.render_simple(sc, cam, 4, 4, type="Depth", spp=1, bvh=O.Bvh(sc))
    assert dmin == 0.0 and dmax == np.float32(np.sqrt(np.float32(0.75)))


def test_normal_and_face_orientation_of_a_square():
    img, _ = render(square((0, 0, 1)), "Normal", spp=16)
    assert np.array_equal(img[:, :3], np.tile(np.float32([0.5, 0.5, 1.0]), (img.shape[0], 1)))
    assert np.all(img[:, 3] == 1.0)
    front, _ = render(square((0, 0, 1)), "FaceOrientation", spp=16)
    back, _ = render(square((0, 0, -1)), "FaceOrientation", spp=16)
    assert np.allclose(front[:, :3], FACE_FRONT, rtol=0, atol=1e-6)
    assert np.allclose(back[:, :3], FACE_BACK, rtol=0, atol=1e-6)
    # flipping the winding flips the reported normal
    img, _ = render(square((0, 0, -1)), "Normal", spp=16)
    assert np.array_equal(img[:, :3], np.tile(np.float32([0.5, 0.5, 0.0]), (img.shape[0], 1)))


def test_misses_are_white():
    s = square()
    away = dict(position=(0.0, 0.0, 3.0), target=(0.0, 0.0, 6.0), up=(0.0, 1.0, 0.0), fov=20.0)
    for t in O.SIMPLE:
        img, _ = render(s, t, spp=16, cam=away)
        assert np.array_equal(img, np.ones_like(img)), t


def test_base_color_ambient_depth_closed_forms():
    s = square(color=(0.2, 0.5, 0.8))
    W, H, seed = 24, 16, 4
    o, d = O.primary_rays(s.camera, W, H, seed, 0)
    o, d = o.astype(np.float64), d.astype(np.float64)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    col = np.float32([0.2, 0.5, 0.8]).astype(np.float64)
    img, _ = render(s, "BaseColor", W=W, H=H, seed=seed)
    assert np.array_equal(img[:, :3], np.tile(np.float32([0.2, 0.5, 0.8]), (W * H, 1)))
    img, _ = render(s, "Ambient", W=W, H=H, seed=seed)
    expect = col[None, :] * (np.abs(dn[:, 2:3]) * 0.5 + 0.5)
    assert np.abs(img[:, :3] - expect).max() < 1e-6
    img, (dmin, dmax) = render(s, "Depth", W=W, H=H, seed=seed)
    hit = o - (o[:, 2:3] / d[:, 2:3]) * d                # on z = 0
    dist = np.linalg.norm(hit - o, axis=1)
    ratio = (dist - dmin) / (dmax - dmin)
    assert np.abs(img[:, 0] - (ratio * 0.8 + 0.2)).max() < 2e-6
    assert np.all(img[:, 0] == img[:, 1]) and np.all(img[:, 0] == img[:, 2])
    # box [-1, 1]^2 x [-5, 0] seen from z = 3: min_distance sqrt(1 + 1 + 9), max_distance sqrt(1 + 1 + 64)
    assert dmin == np.float32(np.sqrt(np.float32(11.0))) and dmax == np.float32(np.sqrt(np.float32(66.0)))
    # the per-axis min_distance exceeds the true nearest distance (3): near pixels fall below 0.2
    assert img[:, 0].max() < 0.2
