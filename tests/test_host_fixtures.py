"""The reference's own example fixtures, as written: example/example.json and
example/cornell_box_obj/cornell_box.{obj,mtl} (public-domain scene data,
copied unchanged under tests/fixtures/).

The host loader (srt_amd/host/scene_io.cpp load_obj, used by srtracer) must
map them as the reference's assimp import does (src/importer/assimp.cu:17-79):
one mesh per (group, material) run of faces in first-appearance order,
polygons fan-triangulated, a Disney BxDF with base colour Kd (Ks if Kd is
black), emission Ke -> intensity |Ke| and colour Ke/|Ke|, Ni -> ior, and the
importer's defaults roughness 0.5, specular 0.5, clearcoatGloss max(0, 0.001).
The expectation is built by a separate reader of the OBJ/MTL text in this
file.  No GPU is used."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from srt_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "fixtures")
OBJ = os.path.join(FIX, "cornell_box_obj", "cornell_box.obj")


def read_obj_expected(path):
    """(positions, triangles, meshes [(group, material, first tri, n tris)], mtl dict) from the text."""
    mtl = {}
    cur = None
    verts, meshes, tris = [], [], []
    group, usemtl = "default", ""
    keys, faces = [], {}
    for raw in open(path):
        line = raw.split("#", 1)[0].split()
        if not line:
            continue
        tag, rest = line[0], line[1:]
        if tag == "mtllib":
            for mraw in open(os.path.join(os.path.dirname(path), rest[0])):
                ml = mraw.split("#", 1)[0].split()
                if not ml:
                    continue
                if ml[0] == "newmtl":
                    cur = ml[1]
                    mtl[cur] = {"Kd": [0.0] * 3, "Ks": [0.0] * 3, "Ke": [0.0] * 3, "Ni": 1.5}
                elif cur and ml[0] in ("Kd", "Ks", "Ke"):
                    mtl[cur][ml[0]] = [float(x) for x in ml[1:4]]
                elif cur and ml[0] == "Ni":
                    mtl[cur]["Ni"] = float(ml[1])
        elif tag == "v":
            verts.append([float(x) for x in rest[:3]])
        elif tag == "g":
            group = rest[0]
        elif tag == "usemtl":
            usemtl = rest[0]
        elif tag == "f":
            idx = [int(t.split("/")[0]) for t in rest]
            idx = [i - 1 if i > 0 else len(verts) + i for i in idx]
            k = (group, usemtl)
            if k not in faces:
                keys.append(k)
                faces[k] = []
            for j in range(1, len(idx) - 1):
                faces[k].append((idx[0], idx[j], idx[j + 1]))
    pos = np.asarray(verts, np.float32)
    out_tris, first = [], 0
    for k in keys:
        meshes.append((k[0], k[1], first, len(faces[k])))
        first += len(faces[k])
        out_tris += faces[k]
    return pos, np.asarray(out_tris, np.int64), meshes, mtl


def host_load(path):
    lib = C.CDLL(os.path.join(ROOT, "srt_amd", "libsrt_host.so"))
    f = lib.srt_host_load_model
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_void_p] + [C.c_void_p] * 5 + [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    counts = np.zeros(4, np.int32)
    err = C.create_string_buffer(512)
    assert f(path.encode(), counts.ctypes.data, None, None, None, None, None, None, 0, err, 512) == 0, err.value
    nv, nf, nm, nk = (int(x) for x in counts)
    pos = np.zeros((nv, 3), np.float32)
    idx = np.zeros((nf, 3), np.uint32)
    mf = np.zeros(nm + 1, np.int32)
    mm = np.zeros(nm, np.int32)
    mats = np.zeros(nk, scenes.MAT_DTYPE)
    names = C.create_string_buffer(4096)
    assert f(path.encode(), counts.ctypes.data, pos.ctypes.data, idx.ctypes.data, mf.ctypes.data, mm.ctypes.data,
             mats.ctypes.data, names, 4096, err, 512) == 0, err.value
    name_list = names.raw.split(b"\0")[:nm]
    return pos, idx, mf, mm, mats, [n.decode() for n in name_list]


def test_fixtures_are_the_reference_files():
    cfg = json.load(open(os.path.join(FIX, "example.json")))
    assert cfg["scene"]["model"] == "cornell_box_obj/cornell_box.obj"
    assert [p["name"] for p in cfg["passes"]] == ["PathTracer", "Accumulate", "Denoise", "ToneMapping"]
    assert cfg["scene"]["camera"]["fov"] == 30 and cfg["render"]["type"] == "interactive"


def test_obj_maps_like_the_reference_importer():
    pos_e, tri_e, meshes_e, mtl = read_obj_expected(OBJ)
    pos, idx, mf, mm, mats, names = host_load(OBJ)
    assert len(tri_e) == 36 and idx.shape[0] == 36            # 18 quads
    assert len(mf) - 1 == len(meshes_e)
    # the group/material runs: walls, then the short box's faces filed under
    # group leftWall with material shortBox, the tall box's under group
    # shortBox with material tallBox, the light
    assert [(g, m) for g, m, _, _ in meshes_e] == [
        ("floor", "floor"), ("ceiling", "ceiling"), ("backWall", "backWall"), ("rightWall", "rightWall"),
        ("leftWall", "leftWall"), ("leftWall", "shortBox"), ("shortBox", "tallBox"), ("light", "light")]
    assert names == [g for g, _, _, _ in meshes_e]
    for k, (g, m, first, n) in enumerate(meshes_e):
        assert mf[k] == first and mf[k + 1] == first + n
        tri_world = pos[idx[first:first + n].astype(np.int64)]
        exp_world = pos_e[tri_e[first:first + n]]
        assert np.array_equal(tri_world, exp_world), g       # same triangles, same vertex order
        mat = mats[mm[k]]
        rec = mtl[m]
        ke = np.asarray(rec["Ke"], np.float32)
        inten = np.float32(np.sqrt(np.float32(ke @ ke)))
        assert mat["type"] == scenes.BXDF_DISNEY
        assert mat["intensity"] == pytest.approx(float(inten), rel=1e-7)
        if inten > 0:
            np.testing.assert_allclose(mat["emission_color"], ke / inten, rtol=1e-6)
        else:
            assert np.all(mat["emission_color"] == 0)
        kd = np.asarray(rec["Kd"], np.float32)
        base = kd if kd @ kd > 0 else np.asarray(rec["Ks"], np.float32)
        np.testing.assert_array_equal(mat["base_color"], base)
        assert mat["ior"] == np.float32(rec["Ni"])
        assert (mat["roughness"], mat["specular"], mat["metallic"], mat["spec_trans"]) == (0.5, 0.5, 0.0, 0.0)
        assert mat["clearcoat_gloss"] == np.float32(0.001)
    light = mats[mm[-1]]
    assert light["intensity"] == pytest.approx(np.sqrt(17 ** 2 + 12 ** 2 + 4 ** 2), rel=1e-6)   # 21.19
    assert mats[mm[4]]["ior"] == np.float32(1.5) and mats[mm[0]]["ior"] == np.float32(1.0)
    # the light quad of the fixture: y = 1.98, x in [-0.24, 0.23], z in [-0.22, 0.16]
    lv = pos[idx[mf[-2]:mf[-1]].reshape(-1)]
    assert np.allclose(lv[:, 1], 1.98) and lv[:, 0].min() == np.float32(-0.24) and lv[:, 2].max() == np.float32(0.16)


def test_missing_model_is_an_error():
    lib = C.CDLL(os.path.join(ROOT, "srt_amd", "libsrt_host.so"))
    f = lib.srt_host_load_model
    f.restype = C.c_int
    counts = np.zeros(4, np.int32)
    err = C.create_string_buffer(256)
    assert f(b"/nonexistent/box.obj", counts.ctypes.data, None, None, None, None, None, None, 0, err, 256) == -1
    assert b"Failed to open file" in err.value
