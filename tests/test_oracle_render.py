"""Pins of the oracle's render path against closed forms (CPU only).

The reference ships no images or fixtures, so the integrator, camera and filter of the CPU
restatement are pinned by what the mathematics fixes: a white-furnace-style convex object under
a constant environment, the irradiance of a rectangular emitter, Russian-roulette
unbiasedness, pixel-centre camera rays and filter identities."""
import math

import numpy as np
import pytest

import oracle_ffi as orc
import swr
from swr import scenes


def _desc(width, height, base, cam, **kw):
    r = {"width": width, "height": height, "base": base, "spp": kw.pop("spp", 0), "sampler": kw.pop("sampler", 0),
         "filter": kw.pop("filter", 0), "min_depth": kw.pop("min_depth", 2), "max_depth": kw.pop("max_depth", 5),
         "threshold": kw.pop("threshold", 0.01), "shading_size": 4096, "bin_exponent": 20,
         "iteration": kw.pop("iteration", 0), "seed": kw.pop("seed", 1), "camera": cam}
    assert not kw
    return swr.render_desc(r, orc.RenderDesc)


def _sphere_scene(albedo, reflectance=1.0, env=(1.0, 1.0, 1.0)):
    pos, idx, uv = scenes.icosphere(3, 0.5, [0.0, 0.0, 0.0])
    return {"shaders": [{"type": 0, "colour": [albedo] * 3, "reflectance": reflectance}],
            "meshes": [{"pos4": pos, "idx": idx, "uv": None, "shader": 0}], "lights": [], "environment": env}


def test_convex_object_under_constant_environment():
    """A convex Lambert object under uniform radiance E with no lights: every camera sample that
    hits it carries exactly albedo * E (one bounce, then the environment), misses carry E."""
    rho = 0.6
    oc = orc.Scene(_sphere_scene(rho))
    cam = scenes.camera_for([0.0, 0.0, 3.0], 90.0, 0.0, focal=120.0)
    d = _desc(24, 24, 2, cam, max_depth=3)
    _, smp, st = oc.render(d, samples=True)
    L = smp[:, 2:5]
    hit = np.abs(L[:, 0] - 1.0) > 1e-3
    assert 0.2 < hit.mean() < 0.9
    np.testing.assert_allclose(L[hit], rho, rtol=2e-6)
    np.testing.assert_allclose(L[~hit], 1.0, rtol=0)
    # a planted mistake in the continuation weight (e.g. a dropped cos/pdf) would give rho / pi


def test_russian_roulette_is_unbiased():
    """Same scene with reflectance 0.5 and threshold 2 (q = 0.25, min depth 0): per-sample values
    are 0 or rho * r / q, their mean stays rho * r."""
    rho, r = 0.6, 0.5
    oc = orc.Scene(_sphere_scene(rho, reflectance=r))
    cam = scenes.camera_for([0.0, 0.0, 1.2], 90.0, 0.0, focal=400.0)  # every sample hits the sphere
    d = _desc(32, 32, 4, cam, max_depth=3, min_depth=0, threshold=2.0)
    _, smp, _ = oc.render(d, samples=True)
    L = smp[:, 2]
    q = 0.25
    assert np.all((L == 0.0) | np.isclose(L, rho * r / q, rtol=1e-5))
    n = L.size
    se = math.sqrt(q * (1 - q) / n) * rho * r / q
    assert abs(L.mean() - rho * r) < 4 * se


def _rect_form_factor(a, b, h):
    """View factor from a differential area to a parallel a x b rectangle centred above it at
    height h (four corner rectangles of (a/2, b/2))."""
    X, Y = (a / 2) / h, (b / 2) / h
    f = (X / math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
         + Y / math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))) / (2 * math.pi)
    return 4 * f


def test_direct_light_matches_rectangle_irradiance():
    """Lambert floor under a downward quad light: outgoing radiance rho * I * F (MIS of light and
    BSDF sampling, nothing else in the scene)."""
    rho, I, a, b, h = 0.5, 3.0, 0.8, 0.6, 1.0
    pos, idx, _ = scenes.quad_mesh([(-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50)])
    desc = {"shaders": [{"type": 0, "colour": [rho] * 3, "reflectance": 1.0}],
            "meshes": [{"pos4": pos, "idx": idx, "uv": None, "shader": 0}],
            # the quad spans local (y, z) scaled by (a, b); Rz(-90) maps local y to world x
            "lights": [{"translation": [0.0, h, 0.0], "rotation": [0.0, 0.0, -90.0], "scale": [a, b],
                        "intensity": I}], "environment": (0.0, 0.0, 0.0)}
    oc = orc.Scene(desc)
    # a camera looking straight down at the origin with a tiny field of view
    cam = scenes.camera_for([0.0, 0.5, 0.0], 0.0, -90.0, focal=1e5)
    d = _desc(1, 1, 64, cam, max_depth=1, min_depth=0, sampler=1)
    _, smp, _ = oc.render(d, samples=True)
    L = smp[:, 2].astype(np.float64)
    want = rho * I * _rect_form_factor(a, b, h)
    se = L.std() / math.sqrt(L.size)
    assert abs(L.mean() - want) < 4 * se + 1e-4, (L.mean(), want, se)
    # dropping the BSDF-sampled half of MIS (max depth 0) must be told apart
    d0 = _desc(1, 1, 64, cam, max_depth=0, min_depth=0, sampler=1)
    _, smp0, _ = oc.render(d0, samples=True)
    assert smp0[:, 2].mean() < want - 10 * se


def _camera_frame(rot_deg):
    def R(axis, deg):
        a = math.radians(deg)
        c, s = math.cos(a), math.sin(a)
        m = np.eye(3)
        i, j = (axis + 1) % 3, (axis + 2) % 3
        m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
        return m
    return R(1, rot_deg[1]) @ R(0, rot_deg[0]) @ R(2, rot_deg[2])


def test_grid_sampler_pixel_centre_rays_closed_form():
    W, H, base = 7, 5, 2
    cam = {"type": 0, "translation": [0.3, -0.2, 1.5], "rotation": [10.0, 35.0, -20.0], "focal_length": 40.0,
           "focal_distance": 1000.0, "aperture": 5.6}
    d = _desc(W, H, base, cam, sampler=2)
    n = W * H * base * base
    rays, xy = orc.raygen(d, 0, n)
    Rm = _camera_frame(cam["rotation"])
    T = np.array(cam["translation"])
    nrm = Rm @ np.array([1.0, 0.0, 0.0])
    nodal = T + nrm * cam["focal_length"]
    for sid in range(n):
        x, rem = divmod(sid, H * base * base)
        y, k = divmod(rem, base * base)
        ix, iy = divmod(k, base)
        sx = x + (ix + 0.5) / base
        sy = y + (iy + 0.5) / base
        fx = (sx * 2 - W) / W * 36
        fy = (sy * 2 - H) / W * 36
        film = T + Rm @ np.array([0.0, fy, fx])
        dv = nodal - film
        dv /= np.linalg.norm(dv)
        np.testing.assert_allclose(xy[sid], [sx, sy], rtol=1e-6)
        np.testing.assert_allclose(rays[sid, 0:3], nodal, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(rays[sid, 4:7], dv, rtol=1e-5, atol=1e-5)
        assert rays[sid, 3] == np.float32(0.001) and rays[sid, 7] == np.float32(1e5)


def test_thin_lens_rays_meet_at_the_focal_point():
    cam = {"type": 1, "translation": [0.0, 0.0, 0.0], "rotation": [0.0, 0.0, 0.0], "focal_length": 50.0,
           "focal_distance": 400.0, "aperture": 2.0}
    W, H = 3, 3
    d = _desc(W, H, 4, cam, sampler=2)
    rays, _ = orc.raygen(d, 0, W * H * 16)
    # Grid sampler: the 16 samples of a pixel differ only in the lens sample, but share the film
    # point; with one pixel-sample per id, check each ray passes through its own focal point.
    nodal = np.array([50.0, 0.0, 0.0])
    for sid in range(W * H * 16):
        x, rem = divmod(sid, H * 16)
        y, k = divmod(rem, 16)
        ix, iy = divmod(k, 4)
        sx, sy = x + (ix + 0.5) / 4, y + (iy + 0.5) / 4
        film = np.array([0.0, (sy * 2 - H) / W * 36, (sx * 2 - W) / W * 36])
        focal = film + (nodal - film) * (400.0 / 50.0)
        o, dv = rays[sid, 0:3].astype(np.float64), rays[sid, 4:7].astype(np.float64)
        t = np.dot(focal - o, dv)
        assert np.linalg.norm(o + t * dv - focal) < 1e-3 * np.linalg.norm(focal)
        # the aperture sample lies in the disk of radius (f / N) / 2 around the nodal point
        assert np.linalg.norm(o[1:] - nodal[1:]) <= 0.5 * 50.0 / 2.0 + 1e-4
        assert abs(o[0] - 50.0) < 1e-4


def test_filters_closed_forms():
    W, H, spp = 6, 4, 4
    rng = np.random.default_rng(0)
    smp = np.zeros((W * H * spp, 5), np.float32)
    for sid in range(W * H * spp):
        x, rem = divmod(sid, H * spp)
        y, k = divmod(rem, spp)
        smp[sid, 0] = x + rng.random()
        smp[sid, 1] = y + rng.random()
    # constant radiance -> constant image, for both filters
    smp[:, 2:5] = [0.25, 0.5, 2.0]
    for f in (0, 1):
        pix = orc.filter_image(f, W, H, spp, smp)
        np.testing.assert_allclose(pix, np.broadcast_to([0.25, 0.5, 2.0], pix.shape), rtol=1e-6)
    # box: the mean of each pixel's own samples
    smp[:, 2] = rng.random(W * H * spp)
    pix = orc.filter_image(1, W, H, spp, smp)
    want = smp[:, 2].reshape(W, H, spp).mean(axis=2).T
    np.testing.assert_allclose(pix[:, :, 0], want, rtol=1e-6)
    # tent with one sample at every pixel centre: neighbours get weight 0 -> identity
    cen = np.zeros((W * H, 5), np.float32)
    for sid in range(W * H):
        x, y = divmod(sid, H)
        cen[sid, 0:2] = [x + 0.5, y + 0.5]
        cen[sid, 2] = rng.random()
    pix = orc.filter_image(0, W, H, 1, cen)
    np.testing.assert_allclose(pix[:, :, 0], cen[:, 2].reshape(W, H).T, rtol=1e-6)


def test_light_geometry_order_and_masks():
    """Lights come after all objects (Pathtracer.cpp:166-272) and are invisible to shadow rays."""
    desc = scenes.box_room(sphere_level=1)
    oc = orc.Scene(desc)
    tris = oc.triangles()
    nobj = sum(m["idx"].shape[0] for m in desc["meshes"])
    assert tris.shape[0] == nobj + 2
    # the last two triangles are the light quad, at the light's height
    np.testing.assert_allclose(tris[nobj:, :, 1], 1.98, atol=1e-6)
    ray = np.array([[0.0, 1.0, 0.0, 0.001, 0.0, 1.0, 0.0, 1e5]], np.float32)
    assert oc.intersect(ray)[0, 3] >= nobj           # camera mask sees the light
    assert oc.intersect(ray, 0x0FFFFFFF)[0, 3] < nobj  # shadow mask passes through to the ceiling
