"""The oracle's LightTracer (shader/integrator/light_tracer.cu:57-176):
camera.h get_xy inverts the camera rays (every pixel's own ray lands back in
it), and the depth-1 image of a diffuse floor under a quad light equals a
float64 restatement built from the oracle's separately pinned parts (light
sampling, PCG streams, closest/any hit): the splat weight
aspect * beta * f * cos / t^2 / pw per particle, with beta = Le cos / pdf and
pw = cos / pi of the emitted direction."""
import numpy as np

from srt_amd import scenes
from oracle import pyoracle as O


def lit_floor():
    b = scenes._Builder()
    b.quad((-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2), 0, (0, 1, 0))
    b.quad((-0.5, 2, -0.5), (0.5, 2, -0.5), (0.5, 2, 0.5), (-0.5, 2, 0.5), 1, (0, -1, 0))
    mats = [scenes.material(type=scenes.BXDF_DIFFUSE, base_color=(0.6, 0.5, 0.4)),
            scenes.material(type=scenes.BXDF_DIFFUSE, base_color=(0.0, 0.0, 0.0), emission=(3.0, 2.0, 1.0))]
    s = b.finish(mats, with_normals=False, name="lit_floor")
    s.camera = dict(position=(0.0, 3.0, 4.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
    return s


def test_camera_pixel_inverts_camera_rays():
    s = scenes.cornell()
    W, H = 40, 30
    _, d = O.primary_rays(s.camera, W, H, 5, 0)
    px = O.camera_pixel(s.camera, W, H, -d)
    assert np.mean(px == np.arange(W * H)) > 0.999
    # a ray through a pixel's far corner may land in the neighbour: never farther
    bad = np.nonzero(px != np.arange(W * H))[0]
    assert np.all(np.abs(px[bad] % W - bad % W) <= 1) and np.all(np.abs(px[bad] // W - bad // W) <= 1)
    # get_xy has no side test (the perspective divide cancels the sign): the
    # reversed direction lands in the same pixel, so the reference splats a
    # point behind the camera where its mirror image through the eye would be
    assert np.array_equal(O.camera_pixel(s.camera, W, H, d), px)


def test_depth_one_image_matches_float64_restatement():
    s = lit_floor()
    sc = O.Scene(s)
    ob = O.Bvh(sc)
    W, H, seed = 20, 15, 3
    img = O.render_light(sc, s.camera, W, H, spp=1, max_depth=1, rr_depth=4, seed=seed, bvh=ob)
    cam = np.asarray(s.camera["position"], np.float64)
    le = (s.materials[1]["emission_color"].astype(np.float32) * s.materials[1]["intensity"]).astype(np.float64)
    rho = s.materials[0]["base_color"].astype(np.float64)
    ref = np.zeros((W * H, 3))
    n = W * H
    fl = np.stack([O.path_floats(i, seed, 0, 4) for i in range(n)])
    ls = sc.light_sample(fl[:, :2])
    pos, nrm, pdf = ls[:, :3].astype(np.float64), ls[:, 3:6].astype(np.float64), ls[:, 6].astype(np.float64)
    z = 1 - 2 * fl[:, 2].astype(np.float64)
    r = np.sqrt(1 - z * z)
    phi = 2 * np.pi * fl[:, 3].astype(np.float64)
    wi = nrm + np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    cos_t = np.sum(wi * nrm, 1)
    pw = cos_t / np.pi
    beta = le[None, :] * (cos_t / pdf)[:, None]
    t, _, _, prim = O.trace(sc, pos, wi, bvh=ob)
    floor = (prim == 0) | (prim == 1)
    p = pos + t[:, None] * wi
    sd = cam[None, :] - p
    dist = np.linalg.norm(sd, axis=1)
    sd /= dist[:, None]
    px = O.camera_pixel(s.camera, W, H, sd)
    vis = O.trace(sc, p, sd, tmax=(dist - 1e-3).astype(np.float32), mode="any", bvh=ob)[3] == 0
    ok = floor & (px >= 0) & vis
    cos_o = sd[:, 1]                                       # floor normal (0, 1, 0)
    f = np.where(cos_o > 0, 1.0, 0.0)[:, None] * rho[None, :] / np.pi
    L = (W / H) * beta * f * cos_o[:, None] / (dist ** 2)[:, None] / pw[:, None]
    np.add.at(ref, px[ok], L[ok])
    assert ok.sum() > 50
    scale = np.abs(ref).max()
    assert np.abs(img[:, :3] - ref).max() <= 2e-5 * scale
    assert np.all(img[:, 3] == 1.0)
