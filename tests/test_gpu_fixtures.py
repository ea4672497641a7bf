"""The reference's example config and Cornell OBJ (tests/fixtures, copied
unchanged from example/) run through srtracer as written, and its CLI section
overrides (src/main/main.cu:44-70: -r, -p and -s each replace the whole
render, passes or scene section) take effect.  With the overrides, the image
srtracer writes equals the C-ABI render of the same loaded scene bit for bit
(film orientation: flip x and y, texture.cu:219-230)."""
import json
import os
import subprocess

import numpy as np
import pytest

from srt_amd import scenes
from test_gpu_host_cli import SRTRACER, film_to_image, read_pfm
from test_host_fixtures import FIX, OBJ, host_load

pytestmark = pytest.mark.gpu
EXAMPLE = os.path.join(FIX, "example.json")


def _srtracer(*args):
    r = subprocess.run([SRTRACER, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def _fixture_scene(env=(0.0, 0.0, 0.0)):
    pos, idx, mf, mm, mats, _ = host_load(OBJ)
    return scenes.SceneData(positions=pos, indices=idx, mesh_first=mf, mesh_material=mm, materials=mats, env=env)


def test_example_json_runs_as_written():
    """Interactive renderer, 1920x1080, PathTracer -> Accumulate -> Denoise ->
    ToneMapping on the OBJ Cornell box; without a window the interactive
    loop runs its frame cap (16 frames) and exits normally."""
    r = _srtracer("-c", EXAMPLE)
    assert "ERROR" not in r.stdout


def test_render_and_passes_overrides(ctx, tmp_path):
    W, H = 80, 60
    out = tmp_path / "override.pfm"
    (tmp_path / "render.json").write_text(json.dumps({"type": "image", "resolution": [W, H], "output": str(out)}))
    (tmp_path / "passes.json").write_text(json.dumps(
        [{"name": "PathTracer", "params": {"samples_per_pixel": 16, "max_depth": 5, "rr_depth": 4, "seed": 7}}]))
    _srtracer("-c", EXAMPLE, "-r", str(tmp_path / "render.json"), "-p", str(tmp_path / "passes.json"))
    img = read_pfm(out)
    cfg = json.load(open(EXAMPLE))
    cam = cfg["scene"]["camera"]
    h = ctx.upload(_fixture_scene())
    ctx.build_bvh(h)
    ref = ctx.render(h, cam, W, H, 16, 5, 4, 1, 1, 7)
    ctx.release(h)
    assert np.array_equal(img, film_to_image(ref, W, H))
    assert 0.01 < img.mean() < 10.0


def test_scene_override(ctx, tmp_path):
    W, H = 64, 64
    out = tmp_path / "scene_override.pfm"
    cam = {"type": "perspective", "position": [0.6, 1.2, 4.0], "target": [0, 0.8, 0], "up": [0, 1, 0], "fov": 45}
    (tmp_path / "render.json").write_text(json.dumps({"type": "image", "resolution": [W, H], "output": str(out)}))
    (tmp_path / "passes.json").write_text(json.dumps(
        [{"name": "PathTracer", "params": {"samples_per_pixel": 8, "max_depth": 4, "seed": 2}}]))
    (tmp_path / "scene.json").write_text(json.dumps(
        {"model": OBJ, "camera": cam, "environment": {"type": "constant", "color": [0.1, 0.2, 0.3]}}))
    _srtracer("-c", EXAMPLE, "-r", str(tmp_path / "render.json"), "-p", str(tmp_path / "passes.json"),
              "-s", str(tmp_path / "scene.json"))
    img = read_pfm(out)
    h = ctx.upload(_fixture_scene(env=(0.1, 0.2, 0.3)))
    ctx.build_bvh(h)
    ref = ctx.render(h, cam, W, H, 8, 4, 4, 1, 1, 2)
    ctx.release(h)
    assert np.array_equal(img, film_to_image(ref, W, H))
