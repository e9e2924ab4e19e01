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


class Stats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("contributions", C.c_uint64), ("waves", C.c_uint32), ("max_wave", C.c_uint32),
                ("ms_total", C.c_double), ("ms_raygen", C.c_double), ("ms_sort_rays", C.c_double),
                ("ms_trace", C.c_double), ("ms_sort_hits", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_accumulate", C.c_double), ("ms_filter", C.c_double)]

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
    "swr_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "swr_device_free": (C.c_int, [_P, _P]),
    "swr_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "swr_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "swr_sort_pairs_u32": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, C.c_int]),
    "swr_ray_keys": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_float), _P]),
    "swr_hit_keys": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "swr_mesh_face_keys": (C.c_int, [_P, _P, _P, C.c_size_t, _P]),
    "swr_gather_u32": (C.c_int, [_P, _P, _P, C.c_size_t, C.c_int, _P]),
    "swr_scene_create": (C.c_int, [_P, C.POINTER(_P)]),
    "swr_scene_destroy": (None, [_P]),
    "swr_scene_add_shader": (C.c_int, [_P, C.POINTER(ShaderDesc)]),
    "swr_scene_add_mesh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_int32]),
    "swr_scene_add_quad_light": (C.c_int, [_P, C.POINTER(QuadLightDesc)]),
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
    "swr_raygen": (C.c_int, [_P, C.POINTER(RenderDesc), C.c_uint32, C.c_size_t, _P, _P]),
    "swr_render": (C.c_int, [_P, _P, C.POINTER(RenderDesc), _P, _P, C.POINTER(Stats)]),
    "swr_write_exr": (C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32]),
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

    def synchronize(self):
        self.check(self.lib.swr_synchronize(self.h))

    def cancel(self):
        self.check(self.lib.swr_cancel(self.h))

    def reset_cancel(self):
        self.check(self.lib.swr_reset_cancel(self.h))

    # -- primitives -----------------------------------------------------------------------
    def sort_pairs(self, keys: "DeviceArray", vals: "DeviceArray", n: int, key_bits: int = 32, iota: bool = True):
        self.check(self.lib.swr_sort_pairs_u32(self.h, keys.ptr, vals.ptr, n, key_bits, int(iota)))

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
