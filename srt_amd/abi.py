"""ctypes bindings of the C ABI declared in include/srt_hip.h.

The library is srt_amd/libsrt_hip.so, built in-tree by srt_amd/build.py.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .scenes import MAT_DTYPE, SceneData

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrt_hip.so")

SRT_OK = 0
ERRORS = {-1: "SRT_E_INVALID", -2: "SRT_E_NO_DEVICE", -3: "SRT_E_HIP", -4: "SRT_E_OOM", -5: "SRT_E_NOT_FOUND",
          -6: "SRT_E_LIMIT", -7: "SRT_E_STATE"}

# every symbol include/srt_hip.h declares (checked by the CPU tests)
EXPORTED = [
    "srt_create", "srt_destroy", "srt_last_error", "srt_set_option", "srt_hbm_used",
    "srt_scene_upload", "srt_scene_release", "srt_scene_lights",
    "srt_scene_build_bvh", "srt_scene_bvh_info", "srt_scene_bvh_download",
    "srt_scene_set_skin", "srt_scene_skin_refit", "srt_scene_download_vertices",
    "srt_trace_rays", "srt_trace_rays_noalloc", "srt_trace_rays_inst", "srt_scene_set_instances",
    "srt_scene_update_instances", "srt_render", "srt_render_simple", "srt_render_light", "srt_film_download", "srt_tonemap", "srt_accumulate",
    "srt_denoise",
    "srt_stats", "srt_stats_reset", "srt_last_counts", "srt_last_visits",
    "srt_device_alloc", "srt_device_free", "srt_memcpy",
    "srt_mlp_num_params", "srt_mlp_create", "srt_mlp_select", "srt_mlp_set_params", "srt_mlp_get_params", "srt_mlp_set_hparams",
    "srt_mlp_set_bounds", "srt_mlp_infer", "srt_mlp_train_step", "srt_mlp_save", "srt_mlp_load",
    "srt_mlp_train_frame",
    "srt_serve_start", "srt_serve_stop", "srt_submit", "srt_poll", "srt_wait", "srt_serve_stats",
]

TONEMAP = {"None": 0, "Clamp": 1, "Reinhard": 2, "Uncharted2": 3, "ACES": 4}
SIMPLE = {"Depth": 0, "Normal": 1, "BaseColor": 2, "Ambient": 3, "FaceOrientation": 4}


class SrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_int32), ("positions", C.c_void_p), ("normals", C.c_void_p), ("tangents", C.c_void_p),
        ("num_triangles", C.c_int32), ("indices", C.c_void_p), ("num_meshes", C.c_int32),
        ("mesh_first_triangle", C.c_void_p), ("mesh_material", C.c_void_p), ("num_materials", C.c_int32),
        ("materials", C.c_void_p), ("env_color", C.c_float * 3),
    ]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov_deg", C.c_float)]


class PtParams(C.Structure):
    _fields_ = [("samples_per_pixel", C.c_int32), ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
                ("use_nee", C.c_int32), ("use_mis", C.c_int32), ("seed", C.c_uint32), ("mlp_mode", C.c_int32),
                ("reserved", C.c_int32)]


class BvhOpts(C.Structure):
    _fields_ = [("max_leaf", C.c_int32), ("cluster_size", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_uint64), ("camera", Camera), ("params", PtParams), ("width", C.c_int32),
                ("height", C.c_int32), ("rgba_out", C.c_void_p)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("reserved", C.c_int32), ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("done_ms", C.c_double), ("paths", C.c_int64), ("chunks", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    lib = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "srt_create": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(P)]),
        "srt_destroy": (None, [P]),
        "srt_last_error": (C.c_char_p, [P]),
        "srt_set_option": (C.c_int, [P, C.c_char_p, C.c_int64]),
        "srt_hbm_used": (C.c_size_t, [P]),
        "srt_scene_upload": (C.c_int, [P, C.POINTER(SceneDesc), C.POINTER(C.c_uint64)]),
        "srt_scene_release": (C.c_int, [P, C.c_uint64]),
        "srt_scene_lights": (C.c_int, [P, C.c_uint64, P, P, P, P, P]),
        "srt_scene_build_bvh": (C.c_int, [P, C.c_uint64, C.POINTER(BvhOpts)]),
        "srt_scene_bvh_info": (C.c_int, [P, C.c_uint64, P]),
        "srt_scene_bvh_download": (C.c_int, [P, C.c_uint64, P, P, P]),
        "srt_scene_set_skin": (C.c_int, [P, C.c_uint64, P, P]),
        "srt_scene_skin_refit": (C.c_int, [P, C.c_uint64, P, C.c_int32]),
        "srt_scene_download_vertices": (C.c_int, [P, C.c_uint64, P, P, P]),
        "srt_trace_rays": (C.c_int, [P, C.c_uint64, C.c_int32, P, P, C.c_float, P, C.c_int32, P, P, P, P]),
        "srt_trace_rays_noalloc": (C.c_int, [P, C.c_uint64, C.c_int32, P, P, C.c_float, P, C.c_int32, P, P, P, P]),
        "srt_trace_rays_inst": (C.c_int, [P, C.c_uint64, C.c_int32, P, P, C.c_float, P, C.c_int32, P, P, P, P, P]),
        "srt_scene_set_instances": (C.c_int, [P, C.c_uint64, C.c_int32, P, P]),
        "srt_scene_update_instances": (C.c_int, [P, C.c_uint64, P]),
        "srt_render": (C.c_int, [P, C.c_uint64, C.POINTER(Camera), C.POINTER(PtParams), C.c_int32, C.c_int32, P,
                                 C.c_int32]),
        "srt_render_light": (C.c_int, [P, C.c_uint64, C.POINTER(Camera), C.POINTER(PtParams), C.c_int32, C.c_int32, P,
                                       C.c_int32]),
        "srt_render_simple": (C.c_int, [P, C.c_uint64, C.POINTER(Camera), C.c_int32, C.c_int32, C.c_uint32, C.c_int32,
                                        C.c_int32, P, C.c_int32]),
        "srt_film_download": (C.c_int, [P, P, C.c_int64]),
        "srt_tonemap": (C.c_int, [P, P, C.c_int64, C.c_int32, C.c_float, C.c_int32, C.c_int32]),
        "srt_accumulate": (C.c_int, [P, P, P, C.c_int64, C.c_int32, C.c_int32]),
        "srt_denoise": (C.c_int, [P, P, C.c_int32, C.c_int32, C.c_int32]),
        "srt_stats": (C.c_int, [P, P, C.c_int32, P, P, P, C.c_int32]),
        "srt_stats_reset": (None, [P]),
        "srt_last_counts": (C.c_int, [P, P, P]),
        "srt_last_visits": (C.c_int, [P, P]),
        "srt_device_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
        "srt_device_free": (C.c_int, [P, P]),
        "srt_memcpy": (C.c_int, [P, P, P, C.c_size_t, C.c_int32]),
        "srt_mlp_num_params": (C.c_int, []),
        "srt_mlp_create": (C.c_int, [P, C.c_uint32, C.c_float]),
        "srt_mlp_select": (C.c_int, [P, C.c_uint64]),
        "srt_mlp_set_params": (C.c_int, [P, P, C.c_int32]),
        "srt_mlp_get_params": (C.c_int, [P, P, C.c_int32]),
        "srt_mlp_set_hparams": (C.c_int, [P, C.c_float, C.c_float, C.c_float, C.c_float]),
        "srt_mlp_set_bounds": (C.c_int, [P, P, P]),
        "srt_mlp_infer": (C.c_int, [P, C.c_int32, P, P]),
        "srt_mlp_train_step": (C.c_int, [P, C.c_int32, P, P, C.c_int32, P, P]),
        "srt_mlp_save": (C.c_int, [P, C.c_char_p]),
        "srt_mlp_load": (C.c_int, [P, C.c_char_p]),
        "srt_mlp_train_frame": (C.c_int, [P, C.c_uint64, C.POINTER(Camera), C.POINTER(PtParams), C.c_int32,
                                          C.c_int32, C.c_int64, C.c_int32, P, P]),
        "srt_serve_start": (C.c_int, [P, C.c_int64]),
        "srt_serve_stop": (C.c_int, [P, C.c_int32]),
        "srt_submit": (C.c_int, [P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
        "srt_poll": (C.c_int, [P, C.c_uint64, C.POINTER(Result)]),
        "srt_wait": (C.c_int, [P, C.c_uint64, C.c_int32, C.POINTER(Result)]),
        "srt_serve_stats": (C.c_int, [P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """Owns one srt_ctx (device memory, stream, resident scenes)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        r = self.lib.srt_create(device, hbm_budget, C.byref(h))
        if r != SRT_OK:
            raise SrtError(r, "srt_create failed")
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            self.lib.srt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, r):
        if r < 0:
            raise SrtError(r, self.lib.srt_last_error(self.h).decode())
        return r

    def set_option(self, key, value):
        self.check(self.lib.srt_set_option(self.h, key.encode(), int(value)))

    def hbm_used(self):
        return int(self.lib.srt_hbm_used(self.h))

    # ---- scenes ----
    def upload(self, s: SceneData) -> int:
        arrs = dict(
            pos=np.ascontiguousarray(s.positions, np.float32),
            idx=np.ascontiguousarray(s.indices, np.uint32),
            mf=np.ascontiguousarray(s.mesh_first, np.int32),
            mm=np.ascontiguousarray(s.mesh_material, np.int32),
            mats=np.ascontiguousarray(s.materials, MAT_DTYPE),
            nrm=None if s.normals is None else np.ascontiguousarray(s.normals, np.float32),
            tan=None if s.tangents is None else np.ascontiguousarray(s.tangents, np.float32),
        )
        d = SceneDesc()
        d.num_vertices = arrs["pos"].shape[0]
        d.positions = _ptr(arrs["pos"])
        d.normals = _ptr(arrs["nrm"])
        d.tangents = _ptr(arrs["tan"])
        d.num_triangles = arrs["idx"].shape[0]
        d.indices = _ptr(arrs["idx"])
        d.num_meshes = arrs["mm"].shape[0]
        d.mesh_first_triangle = _ptr(arrs["mf"])
        d.mesh_material = _ptr(arrs["mm"])
        d.num_materials = arrs["mats"].shape[0]
        d.materials = _ptr(arrs["mats"])
        for i in range(3):
            d.env_color[i] = s.env[i]
        out = C.c_uint64()
        self.check(self.lib.srt_scene_upload(self.h, C.byref(d), C.byref(out)))
        if getattr(s, "inst_mesh", None) is not None:
            try:
                self.set_instances(out.value, s.inst_mesh, s.inst_xf)
            except SrtError:
                self.lib.srt_scene_release(self.h, out.value)
                raise
        return out.value

    def set_instances(self, scene, inst_mesh, inst_xf):
        """srt_scene_set_instances: meshes become object space, placed by (mesh, 3x4 transform) instances."""
        im = np.ascontiguousarray(inst_mesh, np.int32)
        xf = np.ascontiguousarray(inst_xf, np.float32).reshape(-1, 12)
        self.check(self.lib.srt_scene_set_instances(self.h, scene, im.shape[0], _ptr(im), _ptr(xf)))

    def update_instances(self, scene, inst_xf):
        """srt_scene_update_instances: new transforms for every instance, then the top-level refit."""
        xf = np.ascontiguousarray(inst_xf, np.float32).reshape(-1, 12)
        self.check(self.lib.srt_scene_update_instances(self.h, scene, _ptr(xf)))

    def release(self, scene):
        self.check(self.lib.srt_scene_release(self.h, scene))

    def lights(self, scene, nf, nmesh):
        n = self.check(self.lib.srt_scene_lights(self.h, scene, None, None, None, None, None))
        cdf = np.zeros(max(n, 1), np.float32)
        mesh = np.zeros(max(n, 1), np.int32)
        acdf = np.zeros(nf, np.float32)
        marea = np.zeros(nmesh, np.float32)
        ws = C.c_float()
        self.check(self.lib.srt_scene_lights(self.h, scene, _ptr(cdf), _ptr(mesh), _ptr(acdf), _ptr(marea), C.byref(ws)))
        return cdf[:n], mesh[:n], acdf, marea, ws.value

    def build_bvh(self, scene, max_leaf=0, cluster_size=0):
        o = BvhOpts(max_leaf, cluster_size)
        self.check(self.lib.srt_scene_build_bvh(self.h, scene, C.byref(o)))

    def bvh_info(self, scene):
        info = np.zeros(8, np.int32)
        self.check(self.lib.srt_scene_bvh_info(self.h, scene, _ptr(info)))
        return dict(num_nodes=int(info[0]), root=int(info[1]), max_depth=int(info[2]), clusters=int(info[3]),
                    top_nodes=int(info[4]), cluster_size=int(info[5]), triangles=int(info[6]), max_leaf=int(info[7]))

    def bvh_download(self, scene):
        info = self.bvh_info(scene)
        nodes = np.zeros((info["num_nodes"], 16), np.uint32)
        order = np.zeros(info["triangles"], np.uint32)
        tris = np.zeros((info["triangles"], 12), np.float32)
        self.check(self.lib.srt_scene_bvh_download(self.h, scene, _ptr(nodes), _ptr(order), _ptr(tris)))
        return nodes, order, tris, info

    def set_skin(self, scene, bone_ids, bone_weights):
        bi = np.ascontiguousarray(bone_ids, np.int32)
        bw = np.ascontiguousarray(bone_weights, np.float32)
        self.check(self.lib.srt_scene_set_skin(self.h, scene, _ptr(bi), _ptr(bw)))

    def skin_refit(self, scene, bones_3x4):
        b = np.ascontiguousarray(bones_3x4, np.float32)
        self.check(self.lib.srt_scene_skin_refit(self.h, scene, _ptr(b), b.shape[0]))

    def download_vertices(self, scene, nv, normals=False, tangents=False):
        p = np.zeros((nv, 3), np.float32)
        n = np.zeros((nv, 3), np.float32) if normals else None
        t = np.zeros((nv, 3), np.float32) if tangents else None
        self.check(self.lib.srt_scene_download_vertices(self.h, scene, _ptr(p), _ptr(n), _ptr(t)))
        return p, n, t

    # ---- rays ----
    def trace(self, scene, origins, dirs, tmin=1e-3, tmax=None, any_hit=False):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = o.shape[0]
        tm = np.full(n, 1e16, np.float32) if tmax is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        self.check(self.lib.srt_trace_rays(self.h, scene, n, _ptr(o), _ptr(d), C.c_float(tmin), _ptr(tm),
                                           1 if any_hit else 0, _ptr(t), _ptr(u), _ptr(v), _ptr(prim)))
        return t, u, v, prim

    def trace_inst(self, scene, origins, dirs, tmin=1e-3, tmax=None, any_hit=False):
        """srt_trace_rays_inst: (t, u, v, prim, instance) for closest hits, or 0/1 per ray for any hits."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = o.shape[0]
        tm = np.full(n, 1e16, np.float32) if tmax is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        inst = np.zeros(n, np.int32)
        self.check(self.lib.srt_trace_rays_inst(self.h, scene, n, _ptr(o), _ptr(d), C.c_float(tmin), _ptr(tm),
                                                1 if any_hit else 0, _ptr(t), _ptr(u), _ptr(v), _ptr(prim), _ptr(inst)))
        if any_hit:
            return prim
        return t, u, v, prim, inst

    def trace_noalloc(self, scene, n, origins_dev, dirs_dev, tmax_dev, t_dev, u_dev, v_dev, prim_dev, tmin=1e-3,
                      any_hit=False):
        """srt_trace_rays_noalloc on device arrays (addresses as ints) of n rays."""
        self.check(self.lib.srt_trace_rays_noalloc(self.h, scene, n, C.c_void_p(origins_dev), C.c_void_p(dirs_dev),
                                                   C.c_float(tmin), C.c_void_p(tmax_dev), 1 if any_hit else 0,
                                                   C.c_void_p(t_dev), C.c_void_p(u_dev), C.c_void_p(v_dev),
                                                   C.c_void_p(prim_dev)))

    # ---- render ----
    @staticmethod
    def camera(cam: dict) -> Camera:
        c = Camera()
        for i in range(3):
            c.position[i] = cam["position"][i]
            c.target[i] = cam["target"][i]
            c.up[i] = cam["up"][i]
        c.fov_deg = cam["fov"]
        return c

    def render(self, scene, cam: dict, width, height, spp, max_depth=16, rr_depth=4, use_nee=1, use_mis=1, seed=0,
               out=None, device_ptr=None, mlp_mode=0):
        """Render and return a (H*W, 4) float32 host array, or write to `device_ptr` (device memory)."""
        p = PtParams(spp, max_depth, rr_depth, use_nee, use_mis, seed, mlp_mode, 0)
        c = self.camera(cam)
        if device_ptr is not None:
            self.check(self.lib.srt_render(self.h, scene, C.byref(c), C.byref(p), width, height,
                                           C.c_void_p(device_ptr), 1))
            return None
        if out is None:
            out = np.zeros((width * height, 4), np.float32)
        self.check(self.lib.srt_render(self.h, scene, C.byref(c), C.byref(p), width, height, _ptr(out), 0))
        return out

    def render_resident(self, scene, cam, width, height, spp, max_depth=16, rr_depth=4, use_nee=1, use_mis=1, seed=0,
                        mlp_mode=0):
        """Render into the context's own device film (no host copy)."""
        p = PtParams(spp, max_depth, rr_depth, use_nee, use_mis, seed, mlp_mode, 0)
        c = self.camera(cam)
        self.check(self.lib.srt_render(self.h, scene, C.byref(c), C.byref(p), width, height, None, 0))

    def render_light(self, scene, cam: dict, width, height, spp=1, max_depth=16, rr_depth=4, seed=0, device_ptr=None):
        """LightTracer pass: a (H*W, 4) float32 host array, or written to `device_ptr`."""
        p = PtParams(spp, max_depth, rr_depth, 0, 0, seed, 0, 0)
        c = self.camera(cam)
        if device_ptr is not None:
            self.check(self.lib.srt_render_light(self.h, scene, C.byref(c), C.byref(p), width, height,
                                                 C.c_void_p(device_ptr), 1))
            return None
        out = np.zeros((width * height, 4), np.float32)
        self.check(self.lib.srt_render_light(self.h, scene, C.byref(c), C.byref(p), width, height, _ptr(out), 0))
        return out

    def render_simple(self, scene, cam: dict, width, height, type="Depth", spp=16, seed=0, device_ptr=None):
        """Simple NPR pass: a (H*W, 4) float32 host array, or written to `device_ptr`."""
        t = SIMPLE[type] if isinstance(type, str) else int(type)
        c = self.camera(cam)
        if device_ptr is not None:
            self.check(self.lib.srt_render_simple(self.h, scene, C.byref(c), t, spp, seed, width, height,
                                                  C.c_void_p(device_ptr), 1))
            return None
        out = np.zeros((width * height, 4), np.float32)
        self.check(self.lib.srt_render_simple(self.h, scene, C.byref(c), t, spp, seed, width, height, _ptr(out), 0))
        return out

    def film(self, npix):
        out = np.zeros((npix, 4), np.float32)
        self.check(self.lib.srt_film_download(self.h, _ptr(out), npix))
        return out

    def tonemap(self, rgba, type="None", exposure=0.0, use_gamma=True):
        a = np.ascontiguousarray(rgba, np.float32).copy()
        t = TONEMAP[type] if isinstance(type, str) else int(type)
        self.check(self.lib.srt_tonemap(self.h, _ptr(a), a.shape[0], t, C.c_float(exposure), 1 if use_gamma else 0, 0))
        return a

    # ---- device buffers of the context ----
    def device_alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.srt_device_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def device_free(self, ptr):
        self.check(self.lib.srt_device_free(self.h, C.c_void_p(ptr)))

    def upload_bytes(self, ptr, array):
        a = np.ascontiguousarray(array)
        self.check(self.lib.srt_memcpy(self.h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 0))

    def download_bytes(self, array, ptr):
        assert array.flags["C_CONTIGUOUS"]
        self.check(self.lib.srt_memcpy(self.h, array.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), array.nbytes, 1))

    def denoise_device(self, ptr, width, height):
        """Denoise pass in place on a (H*W) float4 image in device memory."""
        self.check(self.lib.srt_denoise(self.h, C.c_void_p(ptr), width, height, 1))

    def denoise(self, rgba, width, height):
        a = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4).copy()
        self.check(self.lib.srt_denoise(self.h, _ptr(a), width, height, 0))
        return a

    def accumulate(self, acc, rgba, frame_count):
        a = np.ascontiguousarray(acc, np.float32).copy()
        p = np.ascontiguousarray(rgba, np.float32).copy()
        self.check(self.lib.srt_accumulate(self.h, _ptr(a), _ptr(p), a.shape[0], frame_count, 0))
        return a, p

    # ---- radiance MLP ----
    def mlp_create(self, seed=0, lr=1e-3):
        self.check(self.lib.srt_mlp_create(self.h, seed, C.c_float(lr)))

    def mlp_select(self, scene=0):
        """srt_mlp_* calls act on scene's own MLP from now on (0: the context's own)."""
        self.check(self.lib.srt_mlp_select(self.h, scene))

    def mlp_num_params(self):
        return int(self.lib.srt_mlp_num_params())

    def mlp_params(self):
        out = np.zeros(self.mlp_num_params(), np.float32)
        self.check(self.lib.srt_mlp_get_params(self.h, _ptr(out), out.size))
        return out

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.check(self.lib.srt_mlp_set_params(self.h, _ptr(p), p.size))

    def mlp_set_hparams(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.check(self.lib.srt_mlp_set_hparams(self.h, C.c_float(lr), C.c_float(beta1), C.c_float(beta2),
                                                C.c_float(eps)))

    def mlp_set_bounds(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        self.check(self.lib.srt_mlp_set_bounds(self.h, _ptr(lo), _ptr(hi)))

    def mlp_infer(self, in9):
        a = np.ascontiguousarray(in9, np.float32).reshape(-1, 9)
        out = np.zeros((a.shape[0], 3), np.float32)
        self.check(self.lib.srt_mlp_infer(self.h, a.shape[0], _ptr(a), _ptr(out)))
        return out

    def mlp_train_step(self, in9, target, apply=True, want_grads=False):
        a = np.ascontiguousarray(in9, np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = C.c_float()
        g = np.zeros(self.mlp_num_params(), np.float32) if want_grads else None
        self.check(self.lib.srt_mlp_train_step(self.h, a.shape[0], _ptr(a), _ptr(t), 1 if apply else 0,
                                               C.byref(loss), _ptr(g)))
        return (loss.value, g) if want_grads else loss.value

    def mlp_train_frame(self, scene, cam, width, height, spp, max_depth=5, rr_depth=4, seed=0, max_samples=1 << 18,
                        steps=16):
        p = PtParams(spp, max_depth, rr_depth, 1, 1, seed, 0, 0)
        c = self.camera(cam)
        loss = C.c_float()
        n = C.c_int64()
        self.check(self.lib.srt_mlp_train_frame(self.h, scene, C.byref(c), C.byref(p), width, height, max_samples,
                                                steps, C.byref(loss), C.byref(n)))
        return loss.value, n.value

    def mlp_save(self, path):
        self.check(self.lib.srt_mlp_save(self.h, path.encode()))

    def mlp_load(self, path):
        self.check(self.lib.srt_mlp_load(self.h, path.encode()))

    # ---- statistics ----
    def stats(self):
        mx = 64
        names = C.create_string_buffer(4096)
        ms = np.zeros(mx, np.float64)
        la = np.zeros(mx, np.int64)
        it = np.zeros(mx, np.int64)
        k = self.check(self.lib.srt_stats(self.h, names, 4096, _ptr(ms), _ptr(la), _ptr(it), mx))
        raw = names.raw.split(b"\0")
        return {raw[i].decode(): dict(ms=float(ms[i]), launches=int(la[i]), items=int(it[i])) for i in range(k)}

    def stats_reset(self):
        self.lib.srt_stats_reset(self.h)

    def last_visits(self):
        out = np.zeros(4, np.int64)
        self.check(self.lib.srt_last_visits(self.h, _ptr(out)))
        return dict(closest_nodes=int(out[0]), closest_tris=int(out[1]), shadow_nodes=int(out[2]),
                    shadow_tris=int(out[3]))

    def last_counts(self):
        a = C.c_int64()
        b = C.c_int64()
        self.check(self.lib.srt_last_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- serving ----
    def serve_start(self, path_pool=0):
        self.check(self.lib.srt_serve_start(self.h, path_pool))
        self._serve_out = {}

    def serve_stop(self, drain=True):
        self.check(self.lib.srt_serve_stop(self.h, 1 if drain else 0))

    def submit(self, scene, cam: dict, width, height, spp, max_depth=16, rr_depth=4, use_nee=1, use_mis=1, seed=0,
               mlp_mode=0, want_image=True):
        """Queue a render request; returns (ticket, output array filled when the request is done).
        want_image=False: the image is not copied back (load generators); the array is None."""
        out = np.zeros((width * height, 4), np.float32) if want_image else None
        r = Request()
        r.scene = scene
        r.camera = self.camera(cam)
        r.params = PtParams(spp, max_depth, rr_depth, use_nee, use_mis, seed, mlp_mode, 0)
        r.width = width
        r.height = height
        r.rgba_out = out.ctypes.data if out is not None else None
        t = C.c_uint64()
        self.check(self.lib.srt_submit(self.h, C.byref(r), C.byref(t)))
        if out is not None:
            self._serve_out[t.value] = out      # keeps the buffer alive while the dispatcher writes it
        return t.value, out

    def forget_tickets(self):
        """Drop the binding's references to finished requests' buffers (after srt_serve_stop)."""
        self._serve_out = {}

    def poll(self, ticket):
        res = Result()
        self.check(self.lib.srt_poll(self.h, ticket, C.byref(res)))
        return res.as_dict()

    def wait(self, ticket, timeout_ms=-1):
        res = Result()
        self.check(self.lib.srt_wait(self.h, ticket, timeout_ms, C.byref(res)))
        return res.as_dict()

    def serve_stats(self):
        a, b = C.c_int64(), C.c_int64()
        q, x = C.c_int32(), C.c_int32()
        self.check(self.lib.srt_serve_stats(self.h, C.byref(a), C.byref(b), C.byref(q), C.byref(x)))
        return dict(served_paths=a.value, chunks=b.value, queued=q.value, active=x.value)
