"""ctypes binding of libswr (include/swr.h), the MI355X sorted-wavefront path tracer.

This module is the Python-side mirror of the reference's host interface
(src/core/Pathtracer.h:26-69): ``Pathtracer(filename)`` with ``process()``, ``image()``,
``clear()``, ``active()`` and ``terminate()``, plus the lower-level stage primitives the C ABI
exports (sort, keys, intersect, occluded, raygen, render). It loads ``lib/libswr.so`` from
this package directory and raises if the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libswr.so")

SWR_OK = 0
SWR_E_INVALID = -1
SWR_E_HIP = -2
SWR_E_BUDGET = -3
SWR_E_CANCELLED = -4
SWR_E_IO = -5
SWR_E_PARSE = -6
SWR_E_BUSY = -7
SORT_ONESWEEP, SORT_REDUCE_SCAN = 0, 1

MASK_ALL = 0xFFFFFFFF
MASK_SHADOW = 0x0FFFFFFF
MASK_LIGHT = 0xF0000000
MISS = 0xFFFFFFFF

CAMERA_PINHOLE, CAMERA_THINLENS = 0, 1
SAMPLER_STRATIFIED, SAMPLER_INDEPENDENT, SAMPLER_GRID = 0, 1, 2
FILTER_TENT, FILTER_BOX = 0, 1
SHADER_LAMBERT, SHADER_NULL, SHADER_GLOSSY = 0, 1, 2


class CameraDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("translation", C.c_float * 3), ("rotation", C.c_float * 3),
                ("focal_length", C.c_float), ("focal_distance", C.c_float), ("aperture", C.c_float)]


class RenderDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("base", C.c_int32), ("spp", C.c_int32),
                ("sampler", C.c_int32), ("filter", C.c_int32), ("min_depth", C.c_int32),
                ("max_depth", C.c_int32), ("threshold", C.c_float), ("shading_size", C.c_int32),
                ("bin_exponent", C.c_int32), ("iteration", C.c_uint32), ("seed", C.c_uint64),
                ("camera", CameraDesc)]


class ShaderDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("colour", C.c_float * 3), ("reflectance", C.c_float),
                ("roughness", C.c_float)]


class QuadLightDesc(C.Structure):
    _fields_ = [("translation", C.c_float * 3), ("rotation", C.c_float * 3), ("scale", C.c_float * 2),
                ("intensity", C.c_float)]


class TextureDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("colour", C.c_float * 3), ("image", C.c_int32), ("upper", C.c_int32),
                ("lower", C.c_int32), ("mask", C.c_int32)]


TEXTURE_CONSTANT, TEXTURE_IMAGE, TEXTURE_LAYERED = 0, 1, 2


def texture_desc(t: dict, cls=TextureDesc):
    """A texture node dict {type, colour, image, upper, lower, mask} as the C record."""
    return cls(t["type"], (C.c_float * 3)(*t.get("colour", (0.0, 0.0, 0.0))), t.get("image", -1),
               t.get("upper", -1), t.get("lower", -1), t.get("mask", -1))


class Tuning(C.Structure):
    _fields_ = [("sort_keys_per_thread", C.c_int32), ("ray_key_bits", C.c_int32), ("hit_key_bits", C.c_int32),
                ("trace_stack_entries", C.c_int32), ("sort_algorithm", C.c_int32),
                ("shadow_overlap", C.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("state_bytes", C.c_uint64), ("wave_rays", C.c_uint32), ("max_requests", C.c_uint32),
                ("waves_per_check", C.c_uint32), ("iteration_window", C.c_uint32), ("backlog_waves", C.c_uint32)]


class RequestStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_uint32), ("queue_ms", C.c_double),
                ("latency_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerTotals(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("refused", C.c_uint64),
                ("camera_rays", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("waves", C.c_uint64), ("busy_s", C.c_double), ("state_bytes", C.c_uint64),
                ("queue_bytes", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("waves", C.c_uint32), ("max_wave", C.c_uint32),
                ("ms_total", C.c_double), ("ms_raygen", C.c_double), ("ms_sort_rays", C.c_double),
                ("ms_trace", C.c_double), ("ms_sort_hits", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_filter", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# Every entry point of include/swr.h with its ctypes signature.
_P = C.c_void_p
_SIG = {
    "swr_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "swr_destroy": (None, [_P]),
    "swr_last_error": (C.c_char_p, [_P]),
    "swr_set_budget": (C.c_int, [_P, C.c_uint64]),
    "swr_bytes_in_use": (C.c_uint64, [_P]),
    "swr_cancel": (C.c_int, [_P]),
    "swr_reset_cancel": (C.c_int, [_P]),
    "swr_synchronize": (C.c_int, [_P]),
    "swr_last_device_ms": (C.c_double, [_P]),
    "swr_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "swr_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "swr_autotune": (C.c_int, [_P, _P, C.POINTER(RenderDesc), C.POINTER(Tuning)]),
    "swr_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "swr_device_free": (C.c_int, [_P, _P]),
    "swr_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "swr_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "swr_sort_pairs_u32": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, C.c_int]),
    "swr_sort_pairs_u32_noalloc": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, C.c_int]),
    "swr_ray_keys": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_float), _P]),
    "swr_hit_keys": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "swr_mesh_face_keys": (C.c_int, [_P, _P, _P, C.c_size_t, _P]),
    "swr_gather_u32": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, _P]),
    "swr_scene_create": (C.c_int, [_P, C.POINTER(_P)]),
    "swr_scene_destroy": (None, [_P]),
    "swr_scene_add_shader": (C.c_int, [_P, C.POINTER(ShaderDesc)]),
    "swr_scene_add_mesh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_int32]),
    "swr_scene_add_quad_light": (C.c_int, [_P, C.POINTER(QuadLightDesc)]),
    "swr_scene_add_image": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "swr_scene_add_texture": (C.c_int, [_P, C.POINTER(TextureDesc)]),
    "swr_scene_set_shader_texture": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "swr_scene_set_environment": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "swr_scene_commit": (C.c_int, [_P]),
    "swr_scene_load_yaml": (C.c_int, [_P, C.c_char_p, C.POINTER(_P), C.POINTER(RenderDesc)]),
    "swr_scene_num_triangles": (C.c_uint32, [_P]),
    "swr_scene_num_geometries": (C.c_uint32, [_P]),
    "swr_scene_bvh_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "swr_scene_bounds": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "swr_intersect": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_uint32, _P]),
    "swr_occluded": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_uint32, _P]),
    "swr_scene_prim_to_geom": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "swr_scene_set_stack_ring": (C.c_int, [_P, C.c_int32]),
    "swr_scene_stack_ring": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "swr_raygen": (C.c_int, [_P, C.POINTER(RenderDesc), C.c_uint32, C.c_size_t, _P, _P]),
    "swr_render": (C.c_int, [_P, _P, C.POINTER(RenderDesc), _P, _P, C.POINTER(Stats)]),
    "swr_write_exr": (C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32]),
    "swr_batch_create": (C.c_int, [_P, C.POINTER(_P)]),
    "swr_batch_destroy": (None, [_P]),
    "swr_batch_add": (C.c_int, [_P, _P, C.POINTER(RenderDesc), C.POINTER(C.c_int32)]),
    "swr_batch_run": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(Stats)]),
    "swr_batch_result": (C.c_int, [_P, C.c_int32, _P, _P]),
    "swr_batch_autotune": (C.c_int, [_P, C.c_uint32, C.POINTER(Tuning)]),
    "swr_server_create": (C.c_int, [_P, C.POINTER(_P), C.c_int32, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "swr_server_submit": (C.c_int, [_P, C.c_int32, C.POINTER(RenderDesc), C.c_int32, _P, C.POINTER(C.c_uint64)]),
    "swr_server_wait": (C.c_int, [_P, C.c_uint64, C.POINTER(RequestStats)]),
    "swr_server_stats": (C.c_int, [_P, C.POINTER(ServerTotals)]),
    "swr_server_destroy": (None, [_P]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libswr.so and bind every entry point; raises if the library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libswr.so not found at {path}: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by missing_symbols(); calling it raises AttributeError
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIG)


def missing_symbols(path: str = LIB_PATH):
    """Entry points of include/swr.h that the built library does not export."""
    lib = C.CDLL(path)
    return [name for name in _SIG if getattr(lib, name, None) is None]


class SwrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"swr error {code}: {msg}")
        self.code = code


class Context:
    """One device context (the reference's Pathtracer owns the equivalent state)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _P()
        rc = self.lib.swr_create(device, C.byref(h))
        if rc != SWR_OK:
            raise SwrError(rc, "swr_create failed (no HIP device?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.swr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc: int):
        if rc != SWR_OK:
            raise SwrError(rc, self.lib.swr_last_error(self.h).decode())

    # -- memory ---------------------------------------------------------------------------
    def alloc(self, nbytes: int) -> "DeviceArray":
        p = _P()
        self.check(self.lib.swr_device_alloc(self.h, nbytes, C.byref(p)))
        return DeviceArray(self, p.value, nbytes)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = self.alloc(max(a.nbytes, 1))
        if a.nbytes:
            self.check(self.lib.swr_memcpy_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes))
        d.dtype, d.shape = a.dtype, a.shape
        return d

    def set_budget(self, nbytes: int):
        self.check(self.lib.swr_set_budget(self.h, nbytes))

    def bytes_in_use(self) -> int:
        return self.lib.swr_bytes_in_use(self.h)

    def last_device_ms(self) -> float:
        return self.lib.swr_last_device_ms(self.h)

    def tuning(self) -> dict:
        t = Tuning()
        self.check(self.lib.swr_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, sort_keys_per_thread=16, ray_key_bits=32, hit_key_bits=0, trace_stack_entries=0,
                   sort_algorithm=0, shadow_overlap=1):
        t = Tuning(sort_keys_per_thread, ray_key_bits, hit_key_bits, trace_stack_entries, sort_algorithm,
                   shadow_overlap)
        self.check(self.lib.swr_set_tuning(self.h, C.byref(t)))

    def autotune(self, scene: "Scene", desc: "RenderDesc") -> dict:
        t = Tuning()
        self.check(self.lib.swr_autotune(self.h, scene.h, C.byref(desc), C.byref(t)))
        return t.as_dict()

    def synchronize(self):
        self.check(self.lib.swr_synchronize(self.h))

    def cancel(self):
        self.check(self.lib.swr_cancel(self.h))

    def reset_cancel(self):
        self.check(self.lib.swr_reset_cancel(self.h))

    # -- primitives -----------------------------------------------------------------------
    def sort_pairs(self, keys: "DeviceArray", vals: "DeviceArray", n: int, key_bits: int = 32, iota: bool = True):
        self.check(self.lib.swr_sort_pairs_u32(self.h, keys.ptr, vals.ptr, n, key_bits, int(iota)))

    def sort_pairs_noalloc(self, keys: "DeviceArray", vals: "DeviceArray", n: int, key_bits: int = 32,
                           iota: bool = True):
        """swr_sort_pairs_u32_noalloc: the same sort with the library's static scratch (n <= 2^20)."""
        self.check(self.lib.swr_sort_pairs_u32_noalloc(self.h, keys.ptr, vals.ptr, n, key_bits, int(iota)))

    def ray_keys(self, rays: "DeviceArray", n: int, lo, scale, keys: "DeviceArray"):
        lo_a = (C.c_float * 3)(*lo)
        sc_a = (C.c_float * 3)(*scale)
        self.check(self.lib.swr_ray_keys(self.h, rays.ptr, n, lo_a, sc_a, keys.ptr))

    def hit_keys(self, hits: "DeviceArray", n: int, keys: "DeviceArray"):
        self.check(self.lib.swr_hit_keys(self.h, hits.ptr, n, keys.ptr))

    def mesh_face_keys(self, mesh: "DeviceArray", face: "DeviceArray", n: int, keys: "DeviceArray"):
        self.check(self.lib.swr_mesh_face_keys(self.h, mesh.ptr, face.ptr, n, keys.ptr))

    def gather(self, src: "DeviceArray", perm: "DeviceArray", n: int, rec_words: int, out: "DeviceArray"):
        self.check(self.lib.swr_gather_u32(self.h, src.ptr, perm.ptr, n, rec_words, out.ptr))


class DeviceArray:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: Context, ptr: int, nbytes: int):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes
        self.dtype, self.shape = np.dtype(np.uint8), (nbytes,)

    def download(self, dtype=None, shape=None) -> np.ndarray:
        dtype = np.dtype(dtype or self.dtype)
        shape = shape if shape is not None else self.shape
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            self.ctx.check(self.ctx.lib.swr_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def upload_from(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            self.ctx.check(self.ctx.lib.swr_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.check(self.ctx.lib.swr_device_free(self.ctx.h, self.ptr))
        self.ptr = 0


def render_desc(r: dict, cls=RenderDesc):
    """RenderDesc (or a layout-identical ctypes class) from a render dict of swr.scenes."""
    cam = r["camera"]
    cd = cls._fields_[-1][1](cam["type"], (C.c_float * 3)(*cam["translation"]), (C.c_float * 3)(*cam["rotation"]),
                             cam["focal_length"], cam.get("focal_distance", 1000.0), cam.get("aperture", 5.6))
    return cls(r["width"], r["height"], r["base"], r.get("spp", 0), r.get("sampler", 0), r.get("filter", 0),
               r["min_depth"], r["max_depth"], r["threshold"], r.get("shading_size", 4096),
               r.get("bin_exponent", 20), r.get("iteration", 0), r.get("seed", 0), cd)


class Scene:
    """A device-resident scene: swr_scene_* (Pathtracer::construct's scene half)."""

    def __init__(self, ctx: Context, desc: dict | None = None, handle=None):
        self.ctx, self.lib = ctx, ctx.lib
        if handle is not None:
            self.h = handle
            return
        h = _P()
        ctx.check(self.lib.swr_scene_create(ctx.h, C.byref(h)))
        self.h = h
        if desc is not None:
            for sh in desc["shaders"]:
                self.add_shader(sh)
            for im in desc.get("images", []):
                self.add_image(im)
            for t in desc.get("textures", []):
                self.add_texture(t)
            for i, sh in enumerate(desc["shaders"]):
                if sh.get("texture", -1) >= 0:
                    self.set_shader_texture(i, sh["texture"])
            for m in desc["meshes"]:
                self.add_mesh(m["pos4"], m["idx"], m.get("uv"), m["shader"])
            for l in desc["lights"]:
                self.add_quad_light(l)
            self.set_environment(desc.get("environment", (0.0, 0.0, 0.0)))
            self.commit()

    def _rc(self, rc):
        if rc < 0:
            raise SwrError(rc, self.lib.swr_last_error(self.ctx.h).decode())
        return rc

    def add_shader(self, sh: dict) -> int:
        d = ShaderDesc(sh["type"], (C.c_float * 3)(*sh["colour"]), sh.get("reflectance", 1.0),
                       sh.get("roughness", 0.0))
        return self._rc(self.lib.swr_scene_add_shader(self.h, C.byref(d)))

    def add_image(self, rgb) -> int:
        """An (H, W, 3) float image; row 0 is t = 0."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        if rgb.ndim != 3 or rgb.shape[2] != 3:
            raise ValueError("image must be (height, width, 3)")
        return self._rc(self.lib.swr_scene_add_image(self.h, rgb.shape[1], rgb.shape[0], rgb.ctypes.data))

    def add_texture(self, t: dict) -> int:
        d = texture_desc(t)
        return self._rc(self.lib.swr_scene_add_texture(self.h, C.byref(d)))

    def set_shader_texture(self, shader: int, texture: int):
        self._rc(self.lib.swr_scene_set_shader_texture(self.h, shader, texture))

    def add_mesh(self, pos4, idx, uv, shader: int) -> int:
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
        uvp = None
        if uv is not None:
            uv = np.ascontiguousarray(uv, dtype=np.float32)
            uvp = uv.ctypes.data
        return self._rc(self.lib.swr_scene_add_mesh(self.h, pos4.ctypes.data, pos4.shape[0], idx.ctypes.data,
                                                    idx.shape[0], uvp, shader))

    def add_quad_light(self, l: dict) -> int:
        d = QuadLightDesc((C.c_float * 3)(*l["translation"]), (C.c_float * 3)(*l["rotation"]),
                          (C.c_float * 2)(*l["scale"]), l["intensity"])
        return self._rc(self.lib.swr_scene_add_quad_light(self.h, C.byref(d)))

    def set_environment(self, rgb):
        self._rc(self.lib.swr_scene_set_environment(self.h, (C.c_float * 3)(*rgb)))

    def commit(self):
        self._rc(self.lib.swr_scene_commit(self.h))

    def stack_ring(self) -> int:
        r = C.c_int32()
        self.ctx.check(self.lib.swr_scene_stack_ring(self.h, C.byref(r)))
        return r.value

    def set_stack_ring(self, ring: int):
        self.ctx.check(self.lib.swr_scene_set_stack_ring(self.h, ring))

    def close(self):
        if getattr(self, "h", None):
            self.lib.swr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_triangles(self) -> int:
        return self.lib.swr_scene_num_triangles(self.h)

    @property
    def num_geometries(self) -> int:
        return self.lib.swr_scene_num_geometries(self.h)

    def bounds(self):
        b = (C.c_float * 6)()
        self._rc(self.lib.swr_scene_bounds(self.h, b))
        return np.array(b[:3]), np.array(b[3:])

    def bvh_info(self):
        nodes, depth, valid = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._rc(self.lib.swr_scene_bvh_info(self.h, C.byref(nodes), C.byref(depth), C.byref(valid)))
        return {"nodes": nodes.value, "max_depth": depth.value, "valid": bool(valid.value)}

    def prim_to_geom(self, prim: int):
        g, p = C.c_uint32(), C.c_uint32()
        self._rc(self.lib.swr_scene_prim_to_geom(self.h, prim, C.byref(g), C.byref(p)))
        return g.value, p.value

    def intersect(self, rays: np.ndarray, mask: int = MASK_ALL) -> np.ndarray:
        """Closest hits of host rays (n x 8) -> (n x 4) uint32 words (t, u, v bits, prim)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        d_r = self.ctx.upload(rays)
        d_h = self.ctx.alloc(max(16 * n, 16))
        try:
            self.ctx.check(self.lib.swr_intersect(self.ctx.h, self.h, d_r.ptr, n, mask, d_h.ptr))
            return d_h.download(np.uint32, (n, 4))
        finally:
            d_r.free()
            d_h.free()

    def intersect_dev(self, rays: "DeviceArray", n: int, hits: "DeviceArray", mask: int = MASK_ALL):
        """Closest hits of n device rays into device hits (n x 4 words); no host copies."""
        self.ctx.check(self.lib.swr_intersect(self.ctx.h, self.h, rays.ptr, n, mask, hits.ptr))

    def occluded_dev(self, rays: "DeviceArray", n: int, flags: "DeviceArray", mask: int = MASK_SHADOW):
        self.ctx.check(self.lib.swr_occluded(self.ctx.h, self.h, rays.ptr, n, mask, flags.ptr))

    def occluded(self, rays: np.ndarray, mask: int = MASK_SHADOW) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        d_r = self.ctx.upload(rays)
        d_f = self.ctx.alloc(max(4 * n, 4))
        try:
            self.ctx.check(self.lib.swr_occluded(self.ctx.h, self.h, d_r.ptr, n, mask, d_f.ptr))
            return d_f.download(np.uint32, (n,))
        finally:
            d_r.free()
            d_f.free()

    def render(self, desc: RenderDesc, samples: bool = False, pixels: bool = True):
        """One iteration (Pathtracer::process): filtered pixels (H, W, 3) and optionally the
        per-sample (x, y, r, g, b) records; returns (pixels or None, samples or None, stats dict).
        pixels=False leaves the image in device memory (no copy to the host)."""
        spp = desc.spp if desc.spp > 0 else desc.base * desc.base
        pix = np.zeros((desc.height, desc.width, 3), np.float32) if pixels else None
        smp = np.zeros((desc.width * desc.height * spp, 5), np.float32) if samples else None
        st = Stats()
        self.ctx.check(self.lib.swr_render(self.ctx.h, self.h, C.byref(desc),
                                           pix.ctypes.data if pix is not None else None,
                                           smp.ctypes.data if smp is not None else None, C.byref(st)))
        return pix, smp, st.as_dict()


class Batch:
    """Many render jobs traced together in one persistent wavefront loop (swr_batch_*)."""

    def __init__(self, ctx: Context):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        ctx.check(self.lib.swr_batch_create(ctx.h, C.byref(h)))
        self.h = h
        self.descs = []

    def add(self, scene: "Scene", desc: RenderDesc) -> int:
        job = C.c_int32()
        self.ctx.check(self.lib.swr_batch_add(self.h, scene.h, C.byref(desc), C.byref(job)))
        self.descs.append(RenderDesc.from_buffer_copy(desc))
        return job.value

    def run(self, wave_rays: int = 0, paced: bool = False) -> dict:
        st = Stats()
        self.ctx.check(self.lib.swr_batch_run(self.h, wave_rays, 1 if paced else 0, C.byref(st)))
        return st.as_dict()

    def result(self, job: int, samples: bool = False):
        if not 0 <= job < len(self.descs):  # let the library refuse it
            self.ctx.check(self.lib.swr_batch_result(self.h, job, None, None))
        d = self.descs[job]
        spp = d.spp if d.spp > 0 else d.base * d.base
        pix = np.zeros((d.height, d.width, 3), np.float32)
        smp = np.zeros((d.width * d.height * spp, 5), np.float32) if samples else None
        self.ctx.check(self.lib.swr_batch_result(self.h, job, pix.ctypes.data,
                                                 smp.ctypes.data if smp is not None else None))
        return pix, smp

    def autotune(self, wave_rays: int = 0) -> dict:
        t = Tuning()
        self.ctx.check(self.lib.swr_batch_autotune(self.h, wave_rays, C.byref(t)))
        return t.as_dict()

    def close(self):
        if self.h:
            self.lib.swr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over resident scenes (swr_server_*): submit progressive requests from any
    thread, wait for each; results are the mean of the request's filtered iterations."""

    def __init__(self, ctx: Context, scenes, state_bytes: int = 0, wave_rays: int = 0, max_requests: int = 0,
                 waves_per_check: int = 0, iteration_window: int = 0, backlog_waves: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        arr = (_P * len(scenes))(*[sc.h for sc in scenes])
        cfg = ServerConfig(state_bytes, wave_rays, max_requests, waves_per_check, iteration_window, backlog_waves)
        h = _P()
        ctx.check(self.lib.swr_server_create(ctx.h, arr, len(scenes), C.byref(cfg), C.byref(h)))
        self.h = h
        self._out = {}

    def submit(self, scene: int, desc: RenderDesc, iterations: int = 1, keep: bool = True) -> int:
        pix = np.zeros((desc.height, desc.width, 3), np.float32) if keep else None
        rid = C.c_uint64()
        self.ctx.check(self.lib.swr_server_submit(self.h, scene, C.byref(desc), iterations,
                                                  pix.ctypes.data if pix is not None else None, C.byref(rid)))
        self._out[rid.value] = pix
        return rid.value

    def wait(self, request: int):
        st = RequestStats()
        rc = self.lib.swr_server_wait(self.h, request, C.byref(st))
        pix = self._out.pop(request, None)
        if rc != SWR_OK:
            raise SwrError(rc, self.lib.swr_last_error(self.ctx.h).decode())
        return pix, st.as_dict()

    def totals(self) -> dict:
        t = ServerTotals()
        self.ctx.check(self.lib.swr_server_stats(self.h, C.byref(t)))
        return t.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.swr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def raygen(ctx: Context, desc: RenderDesc, sid0: int, n: int):
    d_r = ctx.alloc(max(32 * n, 32))
    d_xy = ctx.alloc(max(8 * n, 8))
    try:
        ctx.check(ctx.lib.swr_raygen(ctx.h, C.byref(desc), sid0, n, d_r.ptr, d_xy.ptr))
        return d_r.download(np.float32, (n, 8)), d_xy.download(np.float32, (n, 2))
    finally:
        d_r.free()
        d_xy.free()


def load_yaml(ctx: Context, path: str):
    """swr_scene_load_yaml: a committed Scene plus its RenderDesc from a reference scene file."""
    h = _P()
    desc = RenderDesc()
    ctx.check(ctx.lib.swr_scene_load_yaml(ctx.h, path.encode(), C.byref(h), C.byref(desc)))
    return Scene(ctx, handle=h), desc


def write_exr(path: str, rgb: np.ndarray):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    rc = load_library().swr_write_exr(path.encode(), rgb.ctypes.data, w, h)
    if rc != SWR_OK:
        raise SwrError(rc, "write_exr failed")


class Pathtracer:
    """The reference's host facade (Pathtracer.h:26-69) over the C ABI: load a scene file,
    process() one iteration at a time into a progressive framebuffer, image(), clear(),
    active(), terminate()."""

    def __init__(self, filename: str, device: int = 0):
        self.ctx = Context(device)
        self.scene, self.desc = load_yaml(self.ctx, filename)
        self.pixels = np.zeros((self.desc.height, self.desc.width, 3), np.float32)
        self.iteration = 0
        self._terminate = False

    def image(self):
        return self.pixels, self.desc.width, self.desc.height

    def clear(self):
        self.pixels[:] = 0.0
        self.iteration = 0

    def active(self) -> bool:
        return not self._terminate

    def terminate(self):
        self._terminate = True
        self.ctx.cancel()

    def process(self) -> int:
        self.desc.iteration = self.iteration
        try:
            pix, _, _ = self.scene.render(self.desc)
        except SwrError as e:
            if e.code == SWR_E_CANCELLED:
                return 0
            raise
        self.pixels += pix
        self.iteration += 1
        return self.iteration
