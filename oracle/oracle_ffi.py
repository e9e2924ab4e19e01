"""ctypes binding of the CPU oracle (liboracle.so). Test infrastructure only: loaded by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


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


class Stats(C.Structure):
    _fields_ = [("camera_rays", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("seconds", C.c_double)]


_P = C.c_void_p
_SIG = {
    "oracle_philox4x32_10": (None, [_P, _P, _P]),
    "oracle_u01": (C.c_float, [C.c_uint32]),
    "oracle_ray_keys": (None, [_P, C.c_size_t, _P, _P, _P]),
    "oracle_morton29": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "oracle_hit_keys": (None, [_P, C.c_size_t, _P]),
    "oracle_stable_sort_perm": (None, [_P, C.c_size_t, _P]),
    "oracle_scene_create": (_P, []),
    "oracle_scene_destroy": (None, [_P]),
    "oracle_scene_add_shader": (C.c_int, [_P, C.POINTER(ShaderDesc)]),
    "oracle_scene_add_mesh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_int32]),
    "oracle_scene_add_quad_light": (C.c_int, [_P, C.POINTER(QuadLightDesc)]),
    "oracle_scene_set_environment": (None, [_P, _P]),
    "oracle_scene_add_image": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "oracle_scene_add_texture": (C.c_int, [_P, C.POINTER(TextureDesc)]),
    "oracle_scene_set_shader_texture": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "oracle_texture_lookup": (None, [_P, C.c_int32, C.c_int32, C.c_float, C.c_float, _P]),
    "oracle_texture_colour": (C.c_int, [_P, C.c_int32, C.c_uint32, C.c_float, C.c_float, _P]),
    "oracle_scene_commit": (C.c_int, [_P]),
    "oracle_scene_num_triangles": (C.c_uint32, [_P]),
    "oracle_scene_triangles": (C.c_int, [_P, _P]),
    "oracle_intersect": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, _P]),
    "oracle_intersect_brute": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, _P]),
    "oracle_occluded": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, _P]),
    "oracle_raygen": (C.c_int, [C.POINTER(RenderDesc), C.c_uint32, C.c_size_t, _P, _P]),
    "oracle_render": (C.c_int, [_P, C.POINTER(RenderDesc), _P, _P, C.c_uint32, C.c_uint32, C.c_int,
                                C.POINTER(Stats)]),
    "oracle_filter": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "oracle_ggx_d_n": (None, [C.c_float, C.c_size_t, _P, _P]),
    "oracle_ggx_g1_n": (None, [C.c_float, C.c_size_t, _P, _P]),
    "oracle_bsdf_eval_n": (None, [C.POINTER(ShaderDesc), C.c_size_t, _P, _P, _P, _P, _P, _P]),
    "oracle_bsdf_sample_n": (None, [C.POINTER(ShaderDesc), C.c_size_t, _P, _P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(path: str) -> C.CDLL:
    """Binds every oracle symbol of the library at path (the oracle, or a copy built with a
    planted mistake by a test)."""
    L = C.CDLL(path)
    for name, (res, args) in _SIG.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def morton29(x, y, z):
    return lib().oracle_morton29(x, y, z)


def ray_keys(rays, lo, scale):
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    keys = np.zeros(n, dtype=np.uint32)
    lo = np.asarray(lo, dtype=np.float32)
    scale = np.asarray(scale, dtype=np.float32)
    lib().oracle_ray_keys(_ptr(rays), n, _ptr(lo), _ptr(scale), _ptr(keys))
    return keys


def hit_keys(prim):
    prim = np.ascontiguousarray(prim, dtype=np.uint32)
    keys = np.zeros(prim.shape[0], dtype=np.uint32)
    lib().oracle_hit_keys(_ptr(prim), prim.shape[0], _ptr(keys))
    return keys


def stable_sort_perm(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.zeros(keys.shape[0], dtype=np.uint32)
    lib().oracle_stable_sort_perm(_ptr(keys), keys.shape[0], _ptr(perm))
    return perm


class Scene:
    """Oracle scene built from the same description the product scene receives."""

    def __init__(self, desc):
        L = lib()
        self.h = L.oracle_scene_create()
        for sh in desc["shaders"]:
            d = ShaderDesc(sh["type"], (C.c_float * 3)(*sh["colour"]), sh.get("reflectance", 1.0),
                           sh.get("roughness", 0.0))
            assert L.oracle_scene_add_shader(self.h, C.byref(d)) >= 0
        for im in desc.get("images", []):
            im = np.ascontiguousarray(im, dtype=np.float32)
            assert L.oracle_scene_add_image(self.h, im.shape[1], im.shape[0], _ptr(im)) >= 0
        for t in desc.get("textures", []):
            d = TextureDesc(t["type"], (C.c_float * 3)(*t.get("colour", (0.0, 0.0, 0.0))), t.get("image", -1),
                            t.get("upper", -1), t.get("lower", -1), t.get("mask", -1))
            assert L.oracle_scene_add_texture(self.h, C.byref(d)) >= 0
        for i, sh in enumerate(desc["shaders"]):
            if sh.get("texture", -1) >= 0:
                assert L.oracle_scene_set_shader_texture(self.h, i, sh["texture"]) == 0
        for m in desc["meshes"]:
            pos = np.ascontiguousarray(m["pos4"], dtype=np.float32)
            idx = np.ascontiguousarray(m["idx"], dtype=np.uint32)
            uv = m.get("uv")
            uv = np.ascontiguousarray(uv, dtype=np.float32) if uv is not None else None
            assert L.oracle_scene_add_mesh(self.h, _ptr(pos), pos.shape[0], _ptr(idx), idx.shape[0], _ptr(uv),
                                           m["shader"]) >= 0
        for l in desc["lights"]:
            d = QuadLightDesc((C.c_float * 3)(*l["translation"]), (C.c_float * 3)(*l["rotation"]),
                              (C.c_float * 2)(*l["scale"]), l["intensity"])
            assert L.oracle_scene_add_quad_light(self.h, C.byref(d)) >= 0
        env = np.asarray(desc.get("environment", (0.0, 0.0, 0.0)), dtype=np.float32)
        L.oracle_scene_set_environment(self.h, _ptr(env))
        assert L.oracle_scene_commit(self.h) == 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    @property
    def num_triangles(self):
        return lib().oracle_scene_num_triangles(self.h)

    def triangles(self):
        out = np.zeros((self.num_triangles, 3, 3), dtype=np.float32)
        lib().oracle_scene_triangles(self.h, _ptr(out))
        return out

    def texture_colour(self, shader: int, prim: int, u: float, v: float):
        out = np.zeros(3, dtype=np.float32)
        assert lib().oracle_texture_colour(self.h, shader, prim, u, v, _ptr(out)) == 0
        return out

    def _query(self, fn, rays, mask, words):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], words) if words > 1 else rays.shape[0], dtype=np.uint32)
        assert fn(self.h, _ptr(rays), rays.shape[0], mask, _ptr(out)) == 0
        return out

    def intersect(self, rays, mask=0xFFFFFFFF):
        return self._query(lib().oracle_intersect, rays, mask, 4)

    def intersect_brute(self, rays, mask=0xFFFFFFFF):
        return self._query(lib().oracle_intersect_brute, rays, mask, 4)

    def occluded(self, rays, mask=0x0FFFFFFF):
        return self._query(lib().oracle_occluded, rays, mask, 1)

    def render(self, rdesc: RenderDesc, samples=False, sid_begin=0, sid_end=0, threads=0, pixels=True):
        """samples: False, True (a new array) or a float32 array of W*H*spp x 5 to fill."""
        spp = rdesc.spp if rdesc.spp > 0 else rdesc.base * rdesc.base
        pix = np.zeros((rdesc.height, rdesc.width, 3), dtype=np.float32) if pixels else None
        if isinstance(samples, np.ndarray):
            assert samples.dtype == np.float32 and samples.flags.c_contiguous
            assert samples.shape == (rdesc.width * rdesc.height * spp, 5)
            smp = samples
        else:
            smp = np.zeros((rdesc.width * rdesc.height * spp, 5), dtype=np.float32) if samples else None
        st = Stats()
        assert lib().oracle_render(self.h, C.byref(rdesc), _ptr(pix), _ptr(smp), sid_begin, sid_end, threads,
                                   C.byref(st)) == 0
        return pix, smp, st


def texture_lookup(rgb, s: float, t: float):
    """StandardTexture's lookup of an (H, W, 3) image at (s, t)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    lib().oracle_texture_lookup(_ptr(rgb), rgb.shape[1], rgb.shape[0], s, t, _ptr(out))
    return out


def raygen(rdesc: RenderDesc, sid0: int, n: int):
    rays = np.zeros((n, 8), dtype=np.float32)
    xy = np.zeros((n, 2), dtype=np.float32)
    assert lib().oracle_raygen(C.byref(rdesc), sid0, n, _ptr(rays), _ptr(xy)) == 0
    return rays, xy


def filter_image(filt, width, height, spp, samples):
    samples = np.ascontiguousarray(samples, dtype=np.float32)
    pix = np.zeros((height, width, 3), dtype=np.float32)
    lib().oracle_filter(filt, width, height, spp, _ptr(samples), _ptr(pix))
    return pix


class Bsdf:
    """One shader's lobe from the oracle (or from a library given by L), vectorised over
    direction pairs: eval(in, out, n) and sample(u, out, n) as in ShaderInterface."""

    def __init__(self, type_, colour=(1.0, 1.0, 1.0), reflectance=1.0, roughness=0.0, L=None):
        self.L = L if L is not None else lib()
        self.d = ShaderDesc(type_, (C.c_float * 3)(*colour), reflectance, roughness)
        self.alpha = roughness

    @staticmethod
    def _v(a, n):
        a = np.asarray(a, dtype=np.float32)
        return np.ascontiguousarray(np.broadcast_to(a, (n, 3)) if a.ndim == 1 else a)

    def eval(self, wi, wo, nrm):
        n = max(np.asarray(wi).reshape(-1, 3).shape[0], np.asarray(wo).reshape(-1, 3).shape[0])
        wi, wo, nrm = self._v(wi, n), self._v(wo, n), self._v(nrm, n)
        f = np.zeros((n, 3), np.float32)
        c = np.zeros(n, np.float32)
        p = np.zeros(n, np.float32)
        self.L.oracle_bsdf_eval_n(C.byref(self.d), n, _ptr(wi), _ptr(wo), _ptr(nrm), _ptr(f), _ptr(c), _ptr(p))
        return f, c, p

    def sample(self, u, wo, nrm):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        n = u.shape[0]
        wo, nrm = self._v(wo, n), self._v(nrm, n)
        wi = np.zeros((n, 3), np.float32)
        f = np.zeros((n, 3), np.float32)
        c = np.zeros(n, np.float32)
        p = np.zeros(n, np.float32)
        ok = np.zeros(n, np.uint8)
        self.L.oracle_bsdf_sample_n(C.byref(self.d), n, _ptr(u), _ptr(wo), _ptr(nrm), _ptr(wi), _ptr(f), _ptr(c),
                                    _ptr(p), _ptr(ok))
        return wi, f, c, p, ok.astype(bool)

    def ggx_d(self, cos_h):
        x = np.ascontiguousarray(cos_h, dtype=np.float32).ravel()
        out = np.zeros_like(x)
        self.L.oracle_ggx_d_n(self.alpha, x.size, _ptr(x), _ptr(out))
        return out.reshape(np.shape(cos_h))

    def ggx_g1(self, cos_v):
        x = np.ascontiguousarray(cos_v, dtype=np.float32).ravel()
        out = np.zeros_like(x)
        self.L.oracle_ggx_g1_n(self.alpha, x.size, _ptr(x), _ptr(out))
        return out.reshape(np.shape(cos_v))
