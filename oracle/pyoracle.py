"""ctypes bindings of the CPU oracle (test infrastructure).

Loaded only by the tests, the smoke check and the CPU-baseline leg of
bench.py, always as the checker or the reported baseline, never as a path of
the product.  The library is oracle/build/libsrt_oracle.so (built by
oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libsrt_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sigs = {
            "orc_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_int, P]),
            "orc_seed_hash": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "orc_path_floats": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P]),
            "orc_mathf": (None, [C.c_int, C.c_int, P, P, P]),
            "orc_bxdf": (None, [C.c_int, P, C.c_int, C.c_int, C.c_int, P, P, P, P, P]),
            "orc_scene_new": (P, [C.c_int, P, P, P, C.c_int, P, C.c_int, P, P, C.c_int, P, P]),
            "orc_scene_free": (None, [P]),
            "orc_scene_set_vertices": (None, [P, P, P, P]),
            "orc_scene_lights": (C.c_int, [P, P, P, P, P, P]),
            "orc_light_sample": (None, [P, C.c_int, P, P]),
            "orc_bvh_new": (P, [P, C.c_int, C.c_int]),
            "orc_bvh_free": (None, [P]),
            "orc_bvh_info": (None, [P, P]),
            "orc_bvh_copy": (None, [P, P, P, P, P]),
            "orc_bvh_karras": (None, [P, P, P]),
            "orc_bvh_refit": (None, [P, P]),
            "orc_trace": (None, [P, P, C.c_int, C.c_int, P, P, C.c_float, P, P, P, P, P, P, P]),
            "orc_camera_basis": (None, [P, P, P, C.c_float, C.c_int, C.c_int, P]),
            "orc_primary_rays": (None, [P, P, P, C.c_float, C.c_int, C.c_int, C.c_uint32, C.c_uint32, P, P]),
            "orc_render_simple": (None, [P, P, P, P, P, C.c_float, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                         P, P]),
            "orc_denoise": (None, [C.c_int, C.c_int, P]),
            "orc_camera_pixel": (None, [P, P, P, C.c_float, C.c_int, C.c_int, C.c_int, P, P]),
            "orc_render_light": (None, [P, P, P, P, P, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_uint32, P]),
            "orc_render": (C.c_double, [P, P, P, P, P, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, P, P]),
            "orc_last_counts": (None, [P]),
            "orc_tonemap": (None, [C.c_int, C.c_float, C.c_int, C.c_int, P]),
            "orc_accumulate": (None, [C.c_int, C.c_int, P, P]),
            "orc_skin": (None, [C.c_int, P, P, P, P, P, C.c_int, P, P, P, P]),
            "orc_scene_set_instances": (C.c_int, [P, C.c_int, P, P]),
            "orc_scene_instance_inverses": (None, [P, P]),
            "orc_scene_light_instances": (None, [P, P]),
            "orc_trace_inst": (None, [P, P, C.c_int, C.c_int, P, P, C.c_float, P, P, P, P, P, P]),
        }
        for k, (r, a) in sigs.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f3(v):
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32(initstate, initseq, n, _p(out))
    return out


def seed_hash(pixel, seed):
    return int(lib().orc_seed_hash(pixel, seed))


def path_floats(pixel, seed, sample, n):
    out = np.zeros(n, np.float32)
    lib().orc_path_floats(pixel, seed, sample, n, _p(out))
    return out


MATH = {"sin": 0, "cos": 1, "log": 2, "exp": 3, "pow": 4}


def mathf(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    yy = np.ascontiguousarray(np.zeros_like(x) if y is None else np.broadcast_to(y, x.shape), np.float32)
    out = np.zeros_like(x)
    lib().orc_mathf(MATH[fn], x.size, _p(x), _p(yy), _p(out))
    return out


def bxdf(material, op, wi=None, wo=None, color=(0.5, 0.5, 0.5), rnd=None, fp64=True, inner=False):
    """op: 'eval' -> (n,3) f; 'pdf' -> (n,) pdf; 'sample' -> dict(f, wi, pdf, delta). Local-frame directions."""
    m = np.ascontiguousarray(material).copy()
    opi = {"eval": 0, "pdf": 1, "sample": 2}[op]
    wo = np.ascontiguousarray(wo, np.float64).reshape(-1, 3)
    n = wo.shape[0]
    wi_ = np.ascontiguousarray(np.zeros((n, 3)) if wi is None else wi, np.float64).reshape(-1, 3)
    rn = np.ascontiguousarray(np.zeros((n, 2)) if rnd is None else rnd, np.float64).reshape(-1, 2)
    col = np.ascontiguousarray(color, np.float64)
    out = np.zeros((n, 8), np.float64)
    lib().orc_bxdf(1 if fp64 else 0, _p(m), opi, 1 if inner else 0, n, _p(wi_), _p(wo), _p(col), _p(rn), _p(out))
    if op == "eval":
        return out[:, :3]
    if op == "pdf":
        return out[:, 0]
    return dict(f=out[:, :3], wi=out[:, 3:6], pdf=out[:, 6], delta=out[:, 7] > 0)


class Scene:
    """Oracle copy of a SceneData (srt_amd.scenes)."""

    def __init__(self, s):
        mf = np.asarray(s.mesh_first)
        if np.any(np.diff(mf) <= 0):     # as srt_scene_upload: every mesh holds triangles
            raise ValueError("a mesh has no triangles")
        self.data = s
        self._arrs = dict(
            pos=np.ascontiguousarray(s.positions, np.float32), idx=np.ascontiguousarray(s.indices, np.uint32),
            mf=np.ascontiguousarray(s.mesh_first, np.int32), mm=np.ascontiguousarray(s.mesh_material, np.int32),
            mats=np.ascontiguousarray(s.materials),
            nrm=None if s.normals is None else np.ascontiguousarray(s.normals, np.float32),
            tan=None if s.tangents is None else np.ascontiguousarray(s.tangents, np.float32),
            env=np.ascontiguousarray(s.env, np.float32))
        a = self._arrs
        self.h = lib().orc_scene_new(a["pos"].shape[0], _p(a["pos"]), _p(a["nrm"]), _p(a["tan"]), a["idx"].shape[0],
                                     _p(a["idx"]), a["mm"].shape[0], _p(a["mf"]), _p(a["mm"]), a["mats"].shape[0],
                                     _p(a["mats"]), _p(a["env"]))
        self.nv = a["pos"].shape[0]
        self.nf = a["idx"].shape[0]
        self.nmesh = a["mm"].shape[0]
        self.ninst = 0
        if getattr(s, "inst_mesh", None) is not None:
            self.set_instances(s.inst_mesh, s.inst_xf)

    def set_instances(self, inst_mesh, inst_xf):
        """Meshes become object space, placed by instances (mesh index, 3x4 transform)."""
        im = np.ascontiguousarray(inst_mesh, np.int32)
        xf = np.ascontiguousarray(inst_xf, np.float32).reshape(-1, 12)
        self._arrs["im"], self._arrs["xf"] = im, xf
        if lib().orc_scene_set_instances(self.h, im.shape[0], _p(im), _p(xf)) != 0:
            raise ValueError("bad instance mesh index or singular transform")
        self.ninst = im.shape[0]

    def instance_inverses(self):
        out = np.zeros((self.ninst, 3, 4), np.float32)
        lib().orc_scene_instance_inverses(self.h, _p(out))
        return out

    def __del__(self):
        try:
            lib().orc_scene_free(self.h)
        except Exception:
            pass

    def set_vertices(self, pos, nrm=None, tan=None):
        p = np.ascontiguousarray(pos, np.float32)
        n = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
        t = None if tan is None else np.ascontiguousarray(tan, np.float32)
        lib().orc_scene_set_vertices(self.h, _p(p), _p(n), _p(t))

    def lights(self):
        n = lib().orc_scene_lights(self.h, None, None, None, None, None)
        cdf = np.zeros(max(n, 1), np.float32)
        mesh = np.zeros(max(n, 1), np.int32)
        acdf = np.zeros(self.nf, np.float32)
        marea = np.zeros(self.nmesh, np.float32)
        ws = np.zeros(1, np.float32)
        lib().orc_scene_lights(self.h, _p(cdf), _p(mesh), _p(acdf), _p(marea), _p(ws))
        return cdf[:n], mesh[:n], acdf, marea, float(ws[0])

    def light_sample(self, rnd):
        rnd = np.ascontiguousarray(rnd, np.float32).reshape(-1, 2)
        out = np.zeros((rnd.shape[0], 8), np.float32)
        lib().orc_light_sample(self.h, rnd.shape[0], _p(rnd), _p(out))
        return out


class Bvh:
    def __init__(self, scene: Scene, max_leaf=2, cluster_size=0):
        self.scene = scene
        self.h = lib().orc_bvh_new(scene.h, max_leaf, cluster_size)

    def __del__(self):
        try:
            lib().orc_bvh_free(self.h)
        except Exception:
            pass

    def info(self):
        i = np.zeros(6, np.int32)
        lib().orc_bvh_info(self.h, _p(i))
        return dict(num_nodes=int(i[0]), root=int(i[1]), max_depth=int(i[2]), clusters=int(i[3]), top_nodes=int(i[4]),
                    cluster_size=int(i[5]))

    def arrays(self):
        inf = self.info()
        n = self.scene.nf
        nodes = np.zeros((inf["num_nodes"], 16), np.uint32)
        order = np.zeros(n, np.uint32)
        tris = np.zeros((n, 12), np.float32)
        codes = np.zeros(n, np.uint64)
        lib().orc_bvh_copy(self.h, _p(nodes), _p(order), _p(tris), _p(codes))
        return nodes, order, tris, codes

    def karras(self):
        n = self.scene.nf
        l = np.zeros(max(n - 1, 1), np.int32)
        r = np.zeros(max(n - 1, 1), np.int32)
        lib().orc_bvh_karras(self.h, _p(l), _p(r))
        return l[: n - 1], r[: n - 1]

    def refit(self):
        lib().orc_bvh_refit(self.h, self.scene.h)


def trace(scene: Scene, origins, dirs, tmin=1e-3, tmax=None, mode="closest", bvh: Bvh | None = None, counters=False):
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    tm = np.full(n, 1e16, np.float32) if tmax is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
    m = {"closest": 0, "any": 1}[mode] + (0 if bvh is not None else 2)
    t = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    prim = np.zeros(n, np.int32)
    vis = np.zeros(n, np.int32)
    tst = np.zeros(n, np.int32)
    lib().orc_trace(scene.h, None if bvh is None else bvh.h, m, n, _p(o), _p(d), C.c_float(tmin), _p(tm), _p(t), _p(u),
                    _p(v), _p(prim), _p(vis), _p(tst))
    if mode == "any":
        return prim
    if counters:
        return t, u, v, prim, vis, tst
    return t, u, v, prim


def trace_inst(scene: Scene, origins, dirs, tmin=1e-3, tmax=None, mode="closest", bvh: Bvh | None = None):
    """Instanced scene: (t, u, v, prim, inst) for closest hits, or 0/1 per ray for any hits."""
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    tm = np.full(n, 1e16, np.float32) if tmax is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
    m = {"closest": 0, "any": 1}[mode] + (0 if bvh is not None else 2)
    t = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    prim = np.zeros(n, np.int32)
    inst = np.zeros(n, np.int32)
    lib().orc_trace_inst(scene.h, None if bvh is None else bvh.h, m, n, _p(o), _p(d), C.c_float(tmin), _p(tm), _p(t),
                         _p(u), _p(v), _p(prim), _p(inst))
    if mode == "any":
        return prim
    return t, u, v, prim, inst


def camera_basis(cam, W, H):
    out = np.zeros(14, np.float32)
    lib().orc_camera_basis(_p(_f3(cam["position"])), _p(_f3(cam["target"])), _p(_f3(cam["up"])),
                           C.c_float(cam["fov"]), W, H, _p(out))
    return out


def primary_rays(cam, W, H, seed, sample=0):
    o = np.zeros((W * H, 3), np.float32)
    d = np.zeros((W * H, 3), np.float32)
    lib().orc_primary_rays(_p(_f3(cam["position"])), _p(_f3(cam["target"])), _p(_f3(cam["up"])),
                           C.c_float(cam["fov"]), W, H, seed, sample, _p(o), _p(d))
    return o, d


def render(scene: Scene, cam, W, H, spp, max_depth=16, rr_depth=4, use_nee=1, use_mis=1, seed=0, fp64=False,
           threads=0, bvh: Bvh | None = None, row_begin=0, row_step=1, sample_begin=0, sample_end=0):
    """Host execution of the reference integrator.  Returns (image (H*W,4), seconds)."""
    if fp64:
        out = np.zeros((W * H, 4), np.float64)
        fo, do = None, out
    else:
        out = np.zeros((W * H, 4), np.float32)
        fo, do = out, None
    secs = lib().orc_render(scene.h, None if bvh is None else bvh.h, _p(_f3(cam["position"])),
                            _p(_f3(cam["target"])), _p(_f3(cam["up"])), C.c_float(cam["fov"]), W, H, spp, max_depth,
                            rr_depth, use_nee, use_mis, seed, 1 if fp64 else 0, threads, row_begin, row_step,
                            sample_begin, sample_end, _p(fo), _p(do))
    return out, secs


SIMPLE = {"Depth": 0, "Normal": 1, "BaseColor": 2, "Ambient": 3, "FaceOrientation": 4}


def render_simple(scene: Scene, cam, W, H, type="Depth", spp=16, seed=0, bvh: Bvh | None = None):
    """Simple NPR pass (float).  Returns (image (H*W,4), (min_depth, max_depth))."""
    out = np.zeros((W * H, 4), np.float32)
    rng2 = np.zeros(2, np.float32)
    t = SIMPLE[type] if isinstance(type, str) else int(type)
    lib().orc_render_simple(scene.h, None if bvh is None else bvh.h, _p(_f3(cam["position"])), _p(_f3(cam["target"])),
                            _p(_f3(cam["up"])), C.c_float(cam["fov"]), W, H, spp, seed, t, _p(out), _p(rng2))
    return out, (float(rng2[0]), float(rng2[1]))


def denoise(rgba, W, H):
    """The a-trous denoiser (float), on a copy."""
    a = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4).copy()
    lib().orc_denoise(W, H, _p(a))
    return a


def camera_pixel(cam, W, H, dirs):
    """Pixel index of each direction pointing away from the image (-1: outside)."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros(d.shape[0], np.int32)
    lib().orc_camera_pixel(_p(_f3(cam["position"])), _p(_f3(cam["target"])), _p(_f3(cam["up"])),
                           C.c_float(cam["fov"]), W, H, d.shape[0], _p(d), _p(out))
    return out


def render_light(scene: Scene, cam, W, H, spp=1, max_depth=16, rr_depth=4, seed=0, bvh: Bvh | None = None):
    """LightTracer pass (float), splats in (launch index, sample, depth) order."""
    out = np.zeros((W * H, 4), np.float32)
    lib().orc_render_light(scene.h, None if bvh is None else bvh.h, _p(_f3(cam["position"])), _p(_f3(cam["target"])),
                           _p(_f3(cam["up"])), C.c_float(cam["fov"]), W, H, spp, max_depth, rr_depth, seed, _p(out))
    return out


def last_counts():
    """(closest-hit traces, shadow rays issued) of the last render call."""
    out = np.zeros(2, np.int64)
    lib().orc_last_counts(_p(out))
    return int(out[0]), int(out[1])


TONEMAP = {"None": 0, "Clamp": 1, "Reinhard": 2, "Uncharted2": 3, "ACES": 4}


def tonemap(rgba, type="None", exposure=0.0, use_gamma=True):
    a = np.ascontiguousarray(rgba, np.float32).copy()
    lib().orc_tonemap(TONEMAP[type] if isinstance(type, str) else type, C.c_float(exposure), 1 if use_gamma else 0,
                      a.shape[0], _p(a))
    return a


def accumulate(acc, rgba, frame_count):
    a = np.ascontiguousarray(acc, np.float32).copy()
    p = np.ascontiguousarray(rgba, np.float32).copy()
    lib().orc_accumulate(a.shape[0], frame_count, _p(a), _p(p))
    return a, p


def skin(orig_pos, orig_nrm, orig_tan, bone_ids, bone_w, bones):
    nv = orig_pos.shape[0]
    op = np.ascontiguousarray(orig_pos, np.float32)
    on = None if orig_nrm is None else np.ascontiguousarray(orig_nrm, np.float32)
    ot = None if orig_tan is None else np.ascontiguousarray(orig_tan, np.float32)
    bi = np.ascontiguousarray(bone_ids, np.int32)
    bw = np.ascontiguousarray(bone_w, np.float32)
    b = np.ascontiguousarray(bones, np.float32)
    p = np.zeros((nv, 3), np.float32)
    n = np.zeros((nv, 3), np.float32) if on is not None else None
    t = np.zeros((nv, 3), np.float32) if ot is not None else None
    lib().orc_skin(nv, _p(op), _p(on), _p(ot), _p(bi), _p(bw), b.shape[0], _p(b), _p(p), _p(n), _p(t))
    return p, n, t
