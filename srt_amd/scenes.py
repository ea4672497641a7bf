"""Procedural scenes for the path-tracing workloads (deterministic, seeded).

Each generator returns a :class:`SceneData` holding flat world-space arrays in
the layout of ``srt_scene_desc`` (include/srt_hip.h), plus the camera and the
render settings the workload is quoted at.

Workloads
  cornell   36-triangle Cornell box: 5 wall quads on [-1,1]x[0,2]x[-1,1], a short
            (0.6) and a tall (1.2) axis-aligned box with the footprints of the
            reference's example OBJ (example/cornell_box_obj/cornell_box.obj),
            ceiling light quad at y=1.98 with the OBJ's extent and Ke=(17,12,4)
            (intensity |Ke|, colour Ke/|Ke|, as src/importer/assimp.cu:53-55);
            Disney roughness 1, metallic 0, other parameters at the defaults of
            src/device_include/scene/bxdf.h:64-76; camera (0,1,5.5)->(0,1,0.5),
            40 degrees vertical; 256x256, 16 spp, max depth 5.
  spheres   3125 frequency-4 geodesic icospheres (320 triangles each,
            1,000,000 triangles), centres U[-50,50]^3, radii log-uniform in
            [0.2, 3]; per sphere baseColor U[0,1]^3, metallic Bernoulli(0.3),
            roughness U[0.05,1], specular 0.5, clearcoat Bernoulli(0.2); four
            20x20 quad lights at y=70 facing down, Le=(8,8,8); dim constant
            environment; camera (0,0,160)->origin, 40 degrees; 1920x1080, 64 spp,
            depth 8.
  tube      skinned tube: 1001 rings x 100 segments (200,000 triangles), 64-bone
            chain along x, 4 nearest bones per vertex with Dirichlet(1) weights,
            bone b rotates by A sin(2 pi f/60 + phi_b) about a seeded axis; a
            floor quad and one quad light; 1280x720, 16 spp, depth 6.
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

MAT_DTYPE = np.dtype([
    ("type", "<i4"), ("ior", "<f4"), ("metallic", "<f4"), ("subsurface", "<f4"), ("roughness", "<f4"),
    ("specular", "<f4"), ("specular_tint", "<f4"), ("anisotropic", "<f4"), ("sheen", "<f4"),
    ("sheen_tint", "<f4"), ("clearcoat", "<f4"), ("clearcoat_gloss", "<f4"), ("spec_trans", "<f4"),
    ("base_color", "<f4", 3), ("emission_color", "<f4", 3), ("intensity", "<f4"),
])
assert MAT_DTYPE.itemsize == 80

BXDF_DIFFUSE, BXDF_DIFFUSE_TRANSMISSION, BXDF_DIELECTRIC, BXDF_DISNEY = 0, 1, 2, 3


def material(type=BXDF_DISNEY, base_color=(0.8, 0.8, 0.8), emission=(0.0, 0.0, 0.0), **kw):
    """One material record with the reference's BxDF defaults (bxdf.h:64-76).

    ``emission`` is the Ke colour: intensity = |Ke|, emission_color = Ke/|Ke|
    (src/importer/assimp.cu:53-55).
    """
    m = np.zeros((), dtype=MAT_DTYPE)
    m["type"] = type
    defaults = dict(ior=1.5, metallic=0.0, subsurface=0.0, roughness=0.5, specular=0.5, specular_tint=0.0,
                    anisotropic=0.0, sheen=0.0, sheen_tint=0.0, clearcoat=0.0, clearcoat_gloss=0.001, spec_trans=0.0)
    defaults.update(kw)
    for k, v in defaults.items():
        m[k] = v
    m["base_color"] = base_color
    ke = np.asarray(emission, dtype=np.float32)
    inten = np.float32(np.sqrt(np.float32(np.dot(ke, ke))))
    if inten > 0:
        m["emission_color"] = ke / inten
        m["intensity"] = inten
    return m


@dataclasses.dataclass
class SceneData:
    positions: np.ndarray          # (V, 3) float32
    indices: np.ndarray            # (F, 3) uint32
    mesh_first: np.ndarray         # (M + 1,) int32
    mesh_material: np.ndarray      # (M,) int32
    materials: np.ndarray          # (K,) MAT_DTYPE
    normals: np.ndarray | None = None
    tangents: np.ndarray | None = None
    env: tuple = (0.0, 0.0, 0.0)
    camera: dict = dataclasses.field(default_factory=dict)
    render: dict = dataclasses.field(default_factory=dict)
    skin: dict | None = None       # bone_ids (V,4) int32, bone_weights (V,4) float32, frames callable
    name: str = ""
    # instances (None: every mesh is one identity instance in world space):
    # mesh per instance (I,) int32 and 3x4 row-major transforms (I, 3, 4) float32;
    # the meshes are then in object space
    inst_mesh: np.ndarray | None = None
    inst_xf: np.ndarray | None = None

    @property
    def num_triangles(self):
        return int(self.indices.shape[0])


class _Builder:
    def __init__(self):
        self.pos, self.idx, self.nrm = [], [], []
        self.mesh_first, self.mesh_mat = [0], []
        self.nv = 0

    def add_mesh(self, verts, tris, mat, normals=None):
        verts = np.asarray(verts, dtype=np.float32)
        tris = np.asarray(tris, dtype=np.int64) + self.nv
        self.pos.append(verts)
        self.idx.append(tris.astype(np.uint32))
        self.nrm.append(None if normals is None else np.asarray(normals, dtype=np.float32))
        self.nv += verts.shape[0]
        self.mesh_first.append(self.mesh_first[-1] + tris.shape[0])
        self.mesh_mat.append(mat)

    def quad(self, p0, p1, p2, p3, mat, facing):
        """Quad (p0,p1,p2,p3) as two triangles wound so the geometric normal points along `facing`."""
        p = np.asarray([p0, p1, p2, p3], dtype=np.float64)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        if np.dot(n, facing) < 0:
            p = p[::-1]
        fn = np.asarray(facing, dtype=np.float64)
        fn = fn / np.linalg.norm(fn)
        self.add_mesh(p, [[0, 1, 2], [0, 2, 3]], mat, normals=np.tile(fn, (4, 1)))

    def box(self, lo, hi, mat):
        """Axis-aligned box, 12 triangles, outward normals, as one mesh."""
        lo, hi = np.asarray(lo, float), np.asarray(hi, float)
        c = [[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
             [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]]
        faces = [((0, 3, 2, 1), (0, 0, -1)), ((4, 5, 6, 7), (0, 0, 1)), ((0, 4, 7, 3), (-1, 0, 0)),
                 ((1, 2, 6, 5), (1, 0, 0)), ((0, 1, 5, 4), (0, -1, 0)), ((3, 7, 6, 2), (0, 1, 0))]
        verts, tris, norms = [], [], []
        for quad, facing in faces:
            p = np.asarray([c[i] for i in quad], dtype=np.float64)
            n = np.cross(p[1] - p[0], p[2] - p[0])
            if np.dot(n, facing) < 0:
                p = p[::-1]
            b = len(verts)
            verts.extend(p.tolist())
            norms.extend([facing] * 4)
            tris.extend([[b, b + 1, b + 2], [b, b + 2, b + 3]])
        self.add_mesh(verts, tris, mat, normals=norms)

    def finish(self, materials, with_normals, **kw):
        positions = np.concatenate(self.pos).astype(np.float32)
        normals = None
        if with_normals:
            normals = np.concatenate([n for n in self.nrm]).astype(np.float32)
        return SceneData(positions=positions, indices=np.concatenate(self.idx).astype(np.uint32),
                         mesh_first=np.asarray(self.mesh_first, dtype=np.int32),
                         mesh_material=np.asarray(self.mesh_mat, dtype=np.int32),
                         materials=np.asarray(materials, dtype=MAT_DTYPE), normals=normals, **kw)


def cornell(width=256, height=256, spp=16, max_depth=5, rr_depth=4, seed=1, fov=40.0) -> SceneData:
    red = material(base_color=(0.63, 0.065, 0.05), roughness=1.0)
    green = material(base_color=(0.14, 0.45, 0.091), roughness=1.0)
    grey = material(base_color=(0.75, 0.75, 0.75), roughness=1.0)
    light = material(base_color=(0.75, 0.75, 0.75), roughness=1.0, emission=(17.0, 12.0, 4.0))
    mats = [red, green, grey, light]
    b = _Builder()
    # walls face the interior of the box
    b.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), 2, (0, 1, 0))      # floor
    b.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), 2, (0, -1, 0))     # ceiling
    b.quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), 2, (0, 0, 1))    # back
    b.quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1), 0, (1, 0, 0))    # left, red
    b.quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), 1, (-1, 0, 0))       # right, green
    b.box((-0.05, 0.0, 0.00), (0.70, 0.60, 0.75), 2)                           # short box
    b.box((-0.71, 0.0, -0.67), (0.04, 1.20, 0.09), 2)                          # tall box
    b.quad((-0.24, 1.98, -0.22), (0.23, 1.98, -0.22), (0.23, 1.98, 0.16), (-0.24, 1.98, 0.16), 3, (0, -1, 0))
    s = b.finish(mats, with_normals=False, env=(0.0, 0.0, 0.0), name="cornell")
    assert s.num_triangles == 36
    s.camera = dict(position=(0.0, 1.0, 5.5), target=(0.0, 1.0, 0.5), up=(0.0, 1.0, 0.0), fov=fov)
    s.render = dict(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed,
                    use_nee=1, use_mis=1)
    return s


def icosphere_f4():
    """Geodesic icosahedron of frequency 4: 162 unit vertices, 320 outward triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    F = 4
    verts, key = [], {}

    def vid(p):
        p = p / np.linalg.norm(p)
        k = tuple(np.round(p, 9))
        if k not in key:
            key[k] = len(verts)
            verts.append(p)
        return key[k]

    tris = []
    for a, b_, c in f:
        A, B, C = v[a], v[b_], v[c]
        P = {}
        for i in range(F + 1):
            for j in range(F + 1 - i):
                P[(i, j)] = vid(A + (B - A) * (i / F) + (C - A) * (j / F))
        for i in range(F):
            for j in range(F - i):
                tris.append([P[(i, j)], P[(i + 1, j)], P[(i, j + 1)]])
                if i + j < F - 1:
                    tris.append([P[(i + 1, j)], P[(i + 1, j + 1)], P[(i, j + 1)]])
    verts = np.asarray(verts)
    tris = np.asarray(tris, dtype=np.int64)
    # outward winding
    n = np.cross(verts[tris[:, 1]] - verts[tris[:, 0]], verts[tris[:, 2]] - verts[tris[:, 0]])
    flip = np.einsum("ij,ij->i", n, verts[tris].mean(axis=1)) < 0
    tris[flip] = tris[flip][:, ::-1]
    assert verts.shape[0] == 162 and tris.shape[0] == 320
    return verts, tris


def spheres(num_spheres=3125, seed=20261202, width=1920, height=1080, spp=64, max_depth=8, rr_depth=4,
            render_seed=1) -> SceneData:
    rng = np.random.default_rng(seed)
    uv, ut = icosphere_f4()
    centres = rng.uniform(-50.0, 50.0, size=(num_spheres, 3))
    radii = np.exp(rng.uniform(math.log(0.2), math.log(3.0), size=num_spheres))
    base = rng.uniform(0.0, 1.0, size=(num_spheres, 3))
    metallic = (rng.uniform(size=num_spheres) < 0.3).astype(np.float32)
    rough = rng.uniform(0.05, 1.0, size=num_spheres)
    clearcoat = (rng.uniform(size=num_spheres) < 0.2).astype(np.float32)
    mats = np.zeros(num_spheres + 1, dtype=MAT_DTYPE)
    for i in range(num_spheres):
        mats[i] = material(base_color=base[i], metallic=metallic[i], roughness=rough[i], specular=0.5,
                           clearcoat=clearcoat[i])
    mats[num_spheres] = material(base_color=(0.8, 0.8, 0.8), emission=(8.0, 8.0, 8.0))
    V = (centres[:, None, :] + radii[:, None, None] * uv[None, :, :]).reshape(-1, 3)
    T = (ut[None, :, :] + (np.arange(num_spheres) * uv.shape[0])[:, None, None]).reshape(-1, 3)
    mesh_first = list(np.arange(num_spheres + 1) * ut.shape[0])
    mesh_mat = list(range(num_spheres))
    # four quad lights
    lv, lt = [], []
    nv = V.shape[0]
    for k, (cx, cz) in enumerate([(-25.0, -25.0), (25.0, -25.0), (-25.0, 25.0), (25.0, 25.0)]):
        p = np.array([[cx - 10, 70, cz - 10], [cx + 10, 70, cz - 10], [cx + 10, 70, cz + 10], [cx - 10, 70, cz + 10]])
        n = np.cross(p[1] - p[0], p[2] - p[0])
        if n[1] > 0:        # face down
            p = p[::-1]
        base_i = nv + 4 * k
        lv.append(p)
        lt.append([[base_i, base_i + 1, base_i + 2], [base_i, base_i + 2, base_i + 3]])
        mesh_first.append(mesh_first[-1] + 2)
        mesh_mat.append(num_spheres)
    V = np.concatenate([V, np.concatenate(lv)]).astype(np.float32)
    T = np.concatenate([T, np.concatenate(lt)]).astype(np.uint32)
    s = SceneData(positions=V, indices=T, mesh_first=np.asarray(mesh_first, np.int32),
                  mesh_material=np.asarray(mesh_mat, np.int32), materials=mats, env=(0.05, 0.05, 0.07),
                  name="spheres")
    s.camera = dict(position=(0.0, 0.0, 160.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
    s.render = dict(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=render_seed,
                    use_nee=1, use_mis=1)
    return s


def bone_matrices(frame, num_bones=64, length=10.0, amplitude=0.05, seed=7, frames=60):
    """3x4 row-major skinning matrices of a bone chain along x at `frame`.

    Bone b has its joint at x_b = -L/2 + b L/B and rotates by
    amplitude * sin(2 pi frame / frames + phi_b) about a seeded unit axis;
    M_b = M_{b-1} T(p_b) R_b T(-p_b), bind pose identity.
    """
    rng = np.random.default_rng(seed)
    axes = rng.normal(size=(num_bones, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    phases = rng.uniform(0.0, 2.0 * math.pi, size=num_bones)
    out = np.zeros((num_bones, 3, 4), dtype=np.float64)
    M = np.eye(4)
    for b in range(num_bones):
        ang = amplitude * math.sin(2.0 * math.pi * frame / frames + phases[b])
        a = axes[b]
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        R = np.eye(3) + math.sin(ang) * K + (1.0 - math.cos(ang)) * (K @ K)
        p = np.array([-length / 2.0 + b * length / num_bones, 0.0, 0.0])
        L = np.eye(4)
        L[:3, :3] = R
        L[:3, 3] = p - R @ p
        M = M @ L
        out[b] = M[:3, :]
    return out.astype(np.float32)


def tube(rings=1001, segments=100, radius=0.5, length=10.0, num_bones=64, seed=11, width=1280, height=720, spp=16,
         max_depth=6, rr_depth=4, render_seed=1) -> SceneData:
    rng = np.random.default_rng(seed)
    r = np.arange(rings)
    sgm = np.arange(segments)
    x = -length / 2.0 + length * r / (rings - 1)
    th = 2.0 * math.pi * sgm / segments
    X = np.repeat(x, segments)
    C = np.tile(np.cos(th), rings)
    S = np.tile(np.sin(th), rings)
    pos = np.stack([X, radius * C, radius * S], axis=1)
    nrm = np.stack([np.zeros_like(C), C, S], axis=1)
    tan = np.tile(np.array([1.0, 0.0, 0.0]), (rings * segments, 1))
    i0 = (r[:-1, None] * segments + sgm[None, :]).reshape(-1)
    i1 = (r[:-1, None] * segments + (sgm[None, :] + 1) % segments).reshape(-1)
    j0 = i0 + segments
    j1 = i1 + segments
    tris = np.concatenate([np.stack([i0, j0, j1], 1), np.stack([i0, j1, i1], 1)], axis=0)
    # interleave the two triangles of each quad
    tris = tris.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3)
    n = np.cross(pos[tris[:, 1]] - pos[tris[:, 0]], pos[tris[:, 2]] - pos[tris[:, 0]])
    outward = np.einsum("ij,ij->i", n, nrm[tris].mean(axis=1))
    tris[outward < 0] = tris[outward < 0][:, ::-1]
    assert tris.shape[0] == (rings - 1) * segments * 2
    # 4 nearest bones by centre distance along x, Dirichlet(1) weights
    centres = -length / 2.0 + (np.arange(num_bones) + 0.5) * length / num_bones
    d = np.abs(X[:, None] - centres[None, :])
    ids = np.argsort(d, axis=1, kind="stable")[:, :4].astype(np.int32)
    w = rng.dirichlet(np.ones(4), size=X.shape[0]).astype(np.float32)
    tube_mat = material(base_color=(0.8, 0.35, 0.2), roughness=0.4, metallic=0.0, clearcoat=1.0)
    floor_mat = material(base_color=(0.6, 0.6, 0.6), roughness=1.0)
    light_mat = material(base_color=(0.8, 0.8, 0.8), emission=(10.0, 10.0, 10.0))
    b = _Builder()
    b.add_mesh(pos, tris, 0, normals=nrm)
    b.quad((-12, -3, -8), (12, -3, -8), (12, -3, 8), (-12, -3, 8), 1, (0, 1, 0))
    b.quad((-3, 8, -3), (3, 8, -3), (3, 8, 3), (-3, 8, 3), 2, (0, -1, 0))
    s = b.finish([tube_mat, floor_mat, light_mat], with_normals=True, env=(0.2, 0.2, 0.25), name="tube")
    s.tangents = np.concatenate([tan, np.tile([1.0, 0.0, 0.0], (8, 1))]).astype(np.float32)
    nv = s.positions.shape[0]
    bone_ids = np.full((nv, 4), -1, dtype=np.int32)
    bone_w = np.zeros((nv, 4), dtype=np.float32)
    bone_ids[: X.shape[0]] = ids
    bone_w[: X.shape[0]] = w
    s.skin = dict(bone_ids=bone_ids, bone_weights=bone_w, num_bones=num_bones, frames=60,
                  matrices=lambda f: bone_matrices(f, num_bones=num_bones, length=length))
    s.camera = dict(position=(0.0, 2.0, 14.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)
    s.render = dict(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=render_seed,
                    use_nee=1, use_mis=1)
    return s


WORKLOADS = {"cornell": cornell, "spheres": spheres, "tube": tube}


def save_srtscene(s: SceneData, path: str, frames: int | None = None):
    """Write the binary scene read by the C++ passes (srt_amd/host/scene_io.cpp).

    With ``frames`` and a skinned scene, the bone matrices of frames 0..frames-1
    are stored so the renderer can animate without Python.
    """
    nv = s.positions.shape[0]
    with open(path, "wb") as f:
        f.write(b"SRTSCN1\0")
        f.write(np.asarray([nv, s.num_triangles, len(s.mesh_material), len(s.materials),
                            int(s.normals is not None), int(s.tangents is not None)], "<i4").tobytes())
        f.write(np.asarray(s.env, "<f4").tobytes())
        f.write(np.ascontiguousarray(s.positions, "<f4").tobytes())
        if s.normals is not None:
            f.write(np.ascontiguousarray(s.normals, "<f4").tobytes())
        if s.tangents is not None:
            f.write(np.ascontiguousarray(s.tangents, "<f4").tobytes())
        f.write(np.ascontiguousarray(s.indices, "<u4").tobytes())
        f.write(np.ascontiguousarray(s.mesh_first, "<i4").tobytes())
        f.write(np.ascontiguousarray(s.mesh_material, "<i4").tobytes())
        f.write(np.ascontiguousarray(s.materials, MAT_DTYPE).tobytes())
        if s.skin is not None and frames:
            nb = s.skin["num_bones"]
            f.write(np.asarray([nb, frames], "<i4").tobytes())
            f.write(np.ascontiguousarray(s.skin["bone_ids"], "<i4").tobytes())
            f.write(np.ascontiguousarray(s.skin["bone_weights"], "<f4").tobytes())
            for fr in range(frames):
                f.write(np.ascontiguousarray(s.skin["matrices"](fr), "<f4").tobytes())
        else:
            f.write(np.asarray([0, 0], "<i4").tobytes())


def save_obj(s: SceneData, path: str):
    """Write the scene as OBJ + MTL (one group and material per mesh; Ke carries emission)."""
    base = os.path.splitext(path)[0]
    mtl = base + ".mtl"
    with open(mtl, "w") as f:
        for i, m in enumerate(s.materials):
            ke = np.asarray(m["emission_color"], np.float64) * float(m["intensity"])
            kd = m["base_color"]
            f.write(f"newmtl m{i}\n  Kd {kd[0]:.9g} {kd[1]:.9g} {kd[2]:.9g}\n  Ks 0 0 0\n"
                    f"  Ke {ke[0]:.9g} {ke[1]:.9g} {ke[2]:.9g}\n  Ni {float(m['ior']):.9g}\n\n")
    with open(path, "w") as f:
        f.write(f"mtllib {os.path.basename(mtl)}\n")
        for v in s.positions:
            f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for k in range(len(s.mesh_material)):
            f.write(f"g mesh{k}\nusemtl m{int(s.mesh_material[k])}\n")
            for t in s.indices[s.mesh_first[k]:s.mesh_first[k + 1]]:
                f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def _rotation(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1.0 - math.cos(angle)) * (K @ K)


def affine(R=None, t=(0.0, 0.0, 0.0), scale=1.0):
    """3x4 row-major transform [scale R | t] in float32."""
    M = np.zeros((3, 4), np.float64)
    M[:, :3] = (np.eye(3) if R is None else np.asarray(R, np.float64)) * scale
    M[:, 3] = t
    return M.astype(np.float32)


def identity_instanced(s: SceneData) -> SceneData:
    """The same scene with every mesh as one identity instance (two-level form)."""
    out = dataclasses.replace(s)
    n = s.mesh_material.shape[0]
    out.inst_mesh = np.arange(n, dtype=np.int32)
    out.inst_xf = np.tile(np.eye(4, dtype=np.float32)[:3], (n, 1, 1))
    return out


def cornell_instanced(fov=40.0, box_angles=(0.3, -0.25)) -> SceneData:
    """Cornell box whose two boxes are instances of one unit cube mesh (scaled,
    rotated about y, translated) and whose walls and light are identity
    instances of their own meshes: the reference's instance form
    (one mesh, several transformed instances)."""
    red = material(base_color=(0.63, 0.065, 0.05), roughness=1.0)
    green = material(base_color=(0.14, 0.45, 0.091), roughness=1.0)
    grey = material(base_color=(0.75, 0.75, 0.75), roughness=1.0)
    light = material(base_color=(0.75, 0.75, 0.75), roughness=1.0, emission=(17.0, 12.0, 4.0))
    b = _Builder()
    b.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), 2, (0, 1, 0))
    b.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), 2, (0, -1, 0))
    b.quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), 2, (0, 0, 1))
    b.quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1), 0, (1, 0, 0))
    b.quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), 1, (-1, 0, 0))
    b.quad((-0.24, 1.98, -0.22), (0.23, 1.98, -0.22), (0.23, 1.98, 0.16), (-0.24, 1.98, 0.16), 3, (0, -1, 0))
    b.box((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5), 2)           # mesh 6: unit cube on the floor plane
    s = b.finish([red, green, grey, light], with_normals=False, env=(0.0, 0.0, 0.0), name="cornell_instanced")
    xf = [affine() for _ in range(6)]
    xf.append(affine(_rotation((0, 1, 0), box_angles[0]) @ np.diag([0.6, 0.6, 0.6]), (0.32, 0.0, 0.37)))
    xf.append(affine(_rotation((0, 1, 0), box_angles[1]) @ np.diag([0.6, 1.2, 0.6]), (-0.33, 0.0, -0.29)))
    s.inst_mesh = np.array([0, 1, 2, 3, 4, 5, 6, 6], np.int32)
    s.inst_xf = np.stack(xf)
    s.camera = dict(position=(0.0, 1.0, 5.5), target=(0.0, 1.0, 0.5), up=(0.0, 1.0, 0.0), fov=fov)
    s.render = dict(width=256, height=256, spp=16, max_depth=5, rr_depth=4, seed=1, use_nee=1, use_mis=1)
    return s


def spheres_instanced(num_spheres=3125, num_meshes=16, seed=20261202) -> SceneData:
    """The 1M-triangle sphere field in instance form: `num_meshes` unit
    frequency-4 icospheres (one Disney material each) instanced num_spheres
    times with a random rotation, the log-uniform radius as uniform scale and
    the uniform centre as translation; the four quad lights are identity
    instances.  Same sizes and camera as spheres()."""
    rng = np.random.default_rng(seed)
    uv, ut = icosphere_f4()
    b = _Builder()
    mats = []
    for k in range(num_meshes):
        mats.append(material(base_color=rng.uniform(0.0, 1.0, 3), metallic=float(rng.uniform() < 0.3),
                             roughness=float(rng.uniform(0.05, 1.0)), specular=0.5,
                             clearcoat=float(rng.uniform() < 0.2)))
        b.add_mesh(uv, ut, k, normals=uv)
    mats.append(material(base_color=(0.8, 0.8, 0.8), emission=(8.0, 8.0, 8.0)))
    for (cx, cz) in [(-25.0, -25.0), (25.0, -25.0), (-25.0, 25.0), (25.0, 25.0)]:
        b.quad((cx - 10, 70, cz - 10), (cx + 10, 70, cz - 10), (cx + 10, 70, cz + 10), (cx - 10, 70, cz + 10),
               num_meshes, (0, -1, 0))
    s = b.finish(mats, with_normals=True, env=(0.05, 0.05, 0.07), name="spheres_instanced")
    centres = rng.uniform(-50.0, 50.0, size=(num_spheres, 3))
    radii = np.exp(rng.uniform(math.log(0.2), math.log(3.0), size=num_spheres))
    axes = rng.normal(size=(num_spheres, 3))
    angles = rng.uniform(0.0, 2.0 * math.pi, size=num_spheres)
    which = rng.integers(0, num_meshes, size=num_spheres)
    xf = [affine(_rotation(axes[i], angles[i]), centres[i], radii[i]) for i in range(num_spheres)]
    xf += [affine() for _ in range(4)]
    s.inst_mesh = np.concatenate([which, num_meshes + np.arange(4)]).astype(np.int32)
    s.inst_xf = np.stack(xf)
    s.camera = dict(position=(0.0, 0.0, 160.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=40.0)
    s.render = dict(width=1920, height=1080, spp=64, max_depth=8, rr_depth=4, seed=1, use_nee=1, use_mis=1)
    return s
