"""Closed forms that pin the oracle's camera (R3), tone operators
(tonemapping.cu:14-67), running mean (accumulate.cu:22-26), NaN guard (R12)
and trace accounting, independently of the code they check."""
import math

import numpy as np
import pytest

from srt_amd import scenes
from oracle import pyoracle as O


CAM = dict(position=(0.3, 1.2, 5.0), target=(-0.2, 0.9, 0.4), up=(0.0, 1.0, 0.0), fov=37.0)


def test_camera_basis_closed_form():
    W, H = 320, 200
    b = O.camera_basis(CAM, W, H).astype(np.float64)
    pos, X, Y, Z, nw, nh = b[0:3], b[3:6], b[6:9], b[9:12], b[12], b[13]
    z = np.subtract(CAM["target"], CAM["position"])
    z /= np.linalg.norm(z)
    np.testing.assert_allclose(Z, z, atol=1e-6)
    x = np.cross(CAM["up"], z)
    x /= np.linalg.norm(x)
    np.testing.assert_allclose(X, x, atol=1e-6)               # X = normalize(cross(up, Z))
    np.testing.assert_allclose(Y, np.cross(z, x), atol=1e-6)  # Y = cross(Z, X)
    assert abs(nh - 2.0 * math.tan(math.radians(CAM["fov"]) / 2)) < 1e-6    # vertical fov
    assert abs(nw - nh * W / H) < 1e-5
    np.testing.assert_allclose(pos, CAM["position"], atol=0)


def test_primary_rays_fall_in_their_pixel():
    """Every primary ray, projected back onto the image plane, lands inside
    its own pixel (row 0 at the bottom), and the vertical half-angle of the
    frustum is fov/2."""
    W, H = 64, 40
    b = O.camera_basis(CAM, W, H).astype(np.float64)
    X, Y, Z, nw, nh = b[3:6], b[6:9], b[9:12], b[12], b[13]
    o, d = O.primary_rays(CAM, W, H, seed=9)
    d = d.astype(np.float64)
    assert np.allclose(o, CAM["position"])
    dx, dy, dz = d @ X, d @ Y, d @ Z
    xx = dx / dz / nw + 0.5
    yy = dy / dz / nh + 0.5
    ix = np.arange(W * H) % W
    iy = np.arange(W * H) // W
    eps = 1e-5
    assert np.all(xx >= ix / W - eps) and np.all(xx <= (ix + 1) / W + eps)
    assert np.all(yy >= iy / H - eps) and np.all(yy <= (iy + 1) / H + eps)
    half = np.degrees(np.arctan(np.abs(dy / dz)))
    assert half.max() <= CAM["fov"] / 2 + 1e-3
    assert half.max() > CAM["fov"] / 2 * 0.97                # the top row reaches the frustum edge


def _tonemap_ref(c, kind, exposure, gamma):
    """float64 restatement of the five tone operators."""
    c = c * 2.0 ** exposure
    if kind == "Clamp":
        c = np.clip(c, 0.0, 1.0)
    elif kind == "Reinhard":
        L = c @ np.array([0.2126, 0.7152, 0.0722])
        c = c / (1.0 + L)[:, None]
    elif kind == "Uncharted2":
        A, B, C, D, E, F, Wt = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30, 11.2
        f = lambda x: (x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F) - E / F
        c = f(c) / f(Wt)
    elif kind == "ACES":
        c = c * 0.6
        c = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0.0, 1.0)
    if gamma:
        c = np.power(c, 1.0 / 2.2)
    return c


@pytest.mark.parametrize("kind", ["None", "Clamp", "Reinhard", "Uncharted2", "ACES"])
@pytest.mark.parametrize("exposure,gamma", [(0.0, True), (1.5, False), (-2.0, True)])
def test_tonemap_closed_forms(kind, exposure, gamma):
    rng = np.random.default_rng(5)
    c = np.concatenate([rng.uniform(0.0, 4.0, size=(500, 3)), [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [11.2, 0.5, 30.0]]])
    rgba = np.concatenate([c, np.ones((c.shape[0], 1))], axis=1).astype(np.float32)
    got = O.tonemap(rgba, kind, exposure, gamma).astype(np.float64)
    ref = _tonemap_ref(c.astype(np.float32).astype(np.float64), kind, exposure, False)
    lin = np.power(got[:, :3], 2.2) if gamma else got[:, :3]   # compare in linear space
    # Uncharted2 subtracts E/F = 0.0667 from a ratio of the same size: float32
    # cancellation leaves ~1e-8 absolute error near 0
    np.testing.assert_allclose(lin, ref, rtol=4e-6, atol=5e-8)
    assert np.all(got[:, 3] == 1.0)


def test_running_mean_closed_form():
    rng = np.random.default_rng(2)
    frames = rng.uniform(0, 2, size=(7, 50, 4)).astype(np.float32)
    acc = np.zeros((50, 4), np.float32)
    for n, f in enumerate(frames):
        acc, out = O.accumulate(acc, f, n)
        np.testing.assert_allclose(out[:, :3], frames[: n + 1, :, :3].astype(np.float64).mean(0), rtol=2e-6)


def test_nan_samples_are_dropped_with_their_weight():
    """A material whose base colour is NaN poisons every path that scatters
    off it; those samples are dropped but keep the 1/spp weight (R12), so
    pixels stay finite, and pixels that see only that surface directly read 0."""
    s = nan_cornell()
    osc = O.Scene(s)
    W = H = 48
    spp = 4
    img, _ = O.render(osc, s.camera, W, H, spp, 5, 4, 1, 1, 3)
    assert np.isfinite(img).all()
    # pixels whose every sample's camera ray hits the poisoned mesh first
    direct = np.ones(W * H, bool)
    for smp in range(spp):
        o, d = O.primary_rays(s.camera, W, H, 3, smp)
        p = O.trace(osc, o, d)[3]
        direct &= (p >= s.mesh_first[0]) & (p < s.mesh_first[1])
    assert direct.sum() > 10
    assert np.all(img[direct, :3] == 0.0)
    base = scenes.cornell()
    ref, _ = O.render(O.Scene(base), base.camera, W, H, spp, 5, 4, 1, 1, 3)
    assert np.all(ref[direct, :3].sum(1) > 0)      # the same pixels are lit without the NaN


def nan_cornell():
    """Cornell box whose first mesh has a material of its own with NaN base colour."""
    s = scenes.cornell()
    mats = np.concatenate([s.materials, s.materials[int(s.mesh_material[0])][None]])
    mats[-1]["base_color"] = (np.nan, np.nan, np.nan)
    s.materials = mats
    s.mesh_material = s.mesh_material.copy()
    s.mesh_material[0] = len(mats) - 1
    return s


def test_trace_counts_bound_the_paths():
    s = scenes.cornell()
    img, _ = O.render(O.Scene(s), s.camera, 32, 32, 4, 5, 4, 1, 1, 1)
    traces, shadows = O.last_counts()
    paths = 32 * 32 * 4
    assert paths <= traces <= 5 * paths          # one trace per path at least, max_depth at most
    assert 0 < shadows <= traces
