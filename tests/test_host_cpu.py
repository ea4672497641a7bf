"""CPU-side checks of the product: the C-ABI library loads and exports every
function include/srt_hip.h declares; without a GPU srt_create reports
SRT_E_NO_DEVICE instead of falling back to anything; the procedural workloads
have the sizes and properties the benchmark configurations state."""
import ctypes
import os
import re

import numpy as np
import pytest

from srt_amd import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "srt_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(srt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(abi.EXPORTED) == names


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    lib = abi.load_library()
    h = ctypes.c_void_p()
    assert lib.srt_create(0, 0, ctypes.byref(h)) == -2
    assert not h.value
    with pytest.raises(abi.SrtError):
        abi.Context(0)


def test_cornell_scene():
    s = scenes.cornell()
    assert s.num_triangles == 36
    assert s.materials[s.mesh_material[-1]]["intensity"] == pytest.approx(np.sqrt(17 ** 2 + 12 ** 2 + 4 ** 2), rel=1e-6)
    light = s.positions[s.indices[s.mesh_first[-2]:s.mesh_first[-1]].reshape(-1)]
    assert np.allclose(light[:, 1], 1.98)
    assert s.render == dict(width=256, height=256, spp=16, max_depth=5, rr_depth=4, seed=1, use_nee=1, use_mis=1)
    assert s.camera["fov"] == 40.0


def test_icosphere_is_closed_and_outward():
    v, t = scenes.icosphere_f4()
    assert v.shape == (162, 3) and t.shape == (320, 3)
    edges = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)                       # closed 2-manifold
    assert 162 - counts.size + 320 == 2              # Euler characteristic of a sphere
    n = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    assert np.all(np.einsum("ij,ij->i", n, v[t].mean(1)) > 0)


def test_spheres_scene_is_1m_triangles_and_deterministic():
    a = scenes.spheres()
    b = scenes.spheres()
    assert a.num_triangles == 3125 * 320 + 8
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.materials, b.materials)
    m = a.materials[:3125]
    assert 0.25 < (m["metallic"] == 1).mean() < 0.35
    assert 0.15 < (m["clearcoat"] == 1).mean() < 0.25
    assert m["roughness"].min() >= 0.05 and m["specular"].min() == 0.5
    c = a.positions[: 3125 * 162].reshape(3125, 162, 3).mean(1)
    assert np.abs(c).max() <= 50.0 + 1e-3
    assert a.render["width"] == 1920 and a.render["spp"] == 64 and a.render["max_depth"] == 8


def test_tube_scene_and_bones():
    s = scenes.tube()
    assert s.num_triangles == 200000 + 4
    w = s.skin["bone_weights"][: 1001 * 100]
    assert np.allclose(w.sum(1), 1.0, atol=1e-6)
    ids = s.skin["bone_ids"]
    assert ids[: 1001 * 100].min() >= 0 and ids[: 1001 * 100].max() < 64
    assert np.all(ids[1001 * 100:] == -1)
    M0 = scenes.bone_matrices(0)
    assert M0.shape == (64, 3, 4)
    # rotation blocks are orthonormal
    R = M0[:, :, :3].astype(np.float64)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
