"""The oracle's texture restatement (StandardTexture, LayeredTexture, ConstantTexture and the
texture coordinate of PolygonObject::texture), pinned by what the mathematics fixes:

* the lookup against the cubic B-spline written as a kernel sum, sum_ij c[j mod H, i mod W]
  beta3(x - i) beta3(y - j) with x = s W - 1/2, in float64 (another formulation of the same
  filter: a dropped weight, a wrong half-texel offset or a clamp instead of the wrap fails it);
* properties that tell the B-spline from other filters: a linear ramp is reproduced exactly
  and x^2 comes back as x^2 + 1/3 (bilinear and Catmull-Rom give other values there);
* periodic wrap: rolling the image by k texels equals shifting the coordinate by k / size;
* which axis is which: s runs along a row, t down the rows from the first;
* layered masks of 1 and 0 select upper and lower; texture coordinates at the triangle's
  corners are the corners' uv (u weights vertex 1, v vertex 2, as the oracle's intersection,
  itself pinned against a double-precision solve elsewhere).

The reference looks images up through OpenImageIO, which is absent here: its filter is restated
from OpenImageIO's documented behaviour for the reference's call (periodic wrap, zero
derivatives, the default smart-bicubic mode, which is bicubic at the finest level). That
restatement is parity unpinned against OpenImageIO itself; no reference output covers it."""
import math

import numpy as np
import pytest

import oracle_ffi as orc


def beta3(x):
    x = abs(x)
    if x < 1.0:
        return 2.0 / 3.0 - x * x + 0.5 * x ** 3
    if x < 2.0:
        return (2.0 - x) ** 3 / 6.0
    return 0.0


def kernel_lookup(img, s, t):
    """Cubic B-spline by its kernel, in float64, periodic in both axes."""
    h, w, _ = img.shape
    s = s - math.floor(s)
    t = t - math.floor(t)
    x, y = s * w - 0.5, t * h - 0.5
    acc = np.zeros(3)
    for j in range(math.floor(y) - 2, math.floor(y) + 4):
        by = beta3(y - j)
        if by == 0.0:
            continue
        for i in range(math.floor(x) - 2, math.floor(x) + 4):
            bx = beta3(x - i)
            if bx:
                acc += img[j % h, i % w].astype(np.float64) * bx * by
    return acc


@pytest.mark.parametrize("w,h", [(16, 16), (7, 3), (1, 5), (2, 2), (40, 24)])
def test_lookup_matches_bspline_kernel_sum(w, h):
    rng = np.random.default_rng(w * 100 + h)
    img = rng.uniform(0.0, 1.0, (h, w, 3)).astype(np.float32)
    pts = list(rng.uniform(-2.5, 3.5, (200, 2)))
    # seams, texel centres and the reduction of tiny negative coordinates to 1.0
    pts += [(0.0, 0.0), (1.0, 1.0), (-1e-9, 0.5), (0.5 / w, 0.5 / h), (1.0 - 0.5 / w, 0.25), (0.999999, -0.999999)]
    for s, t in pts:
        got = orc.texture_lookup(img, float(s), float(t))
        want = kernel_lookup(img, float(np.float32(s)), float(np.float32(t)))
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-6, err_msg=f"s={s} t={t}")


def test_linear_ramp_is_reproduced_and_quadratic_gains_one_third():
    w = 32
    x_idx = np.arange(w, dtype=np.float32)
    ramp = np.zeros((1, w, 3), np.float32)
    ramp[0, :, 0] = 0.25 * x_idx + 1.0
    ramp[0, :, 1] = x_idx * x_idx
    ramp[0, :, 2] = 3.0
    for x in np.linspace(1.5, w - 2.5, 37):  # footprint inside the image: no wrap
        s = (x + 0.5) / w
        got = orc.texture_lookup(ramp, float(s), 0.3)
        xs = float(np.float32(s)) * w - 0.5
        assert got[0] == pytest.approx(0.25 * xs + 1.0, rel=1e-5)
        assert got[1] == pytest.approx(xs * xs + 1.0 / 3.0, rel=1e-5, abs=1e-4)
        assert got[2] == pytest.approx(3.0, rel=1e-6)


def test_periodic_wrap_commutes_with_rolling_the_image():
    rng = np.random.default_rng(7)
    img = rng.uniform(0.0, 1.0, (6, 9, 3)).astype(np.float32)
    for k in (1, 4, 8):
        rolled = np.roll(img, k, axis=1)  # rolled[:, i] = img[:, i - k]
        for s in (0.0, 0.02, 0.5, 0.97):
            a = orc.texture_lookup(img, s, 0.4)
            b = orc.texture_lookup(rolled, s + k / 9.0, 0.4)
            np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
        rolled = np.roll(img, k % 6, axis=0)
        a = orc.texture_lookup(img, 0.3, 0.01)
        b = orc.texture_lookup(rolled, 0.3, 0.01 + (k % 6) / 6.0)
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
    # whole periods
    for s, t in ((0.3, 0.2), (-0.45, 0.9)):
        np.testing.assert_allclose(orc.texture_lookup(img, s, t), orc.texture_lookup(img, s + 2.0, t - 3.0),
                                   rtol=1e-5, atol=1e-6)


def test_s_runs_along_rows_and_t_down_from_the_first_row():
    w, h = 16, 8
    img = np.zeros((h, w, 3), np.float32)
    img[:, :, 0] = np.arange(w)[None, :]  # varies with the column only
    img[:, :, 1] = np.arange(h)[:, None]  # varies with the row only
    a = orc.texture_lookup(img, (5 + 0.5) / w, (3 + 0.5) / h)
    assert a[0] == pytest.approx(5.0, abs=1e-5)
    assert a[1] == pytest.approx(3.0, abs=1e-5)


def test_non_finite_coordinates_read_the_origin():
    rng = np.random.default_rng(3)
    img = rng.uniform(0.0, 1.0, (4, 4, 3)).astype(np.float32)
    o = orc.texture_lookup(img, 0.0, 0.0)
    for bad in (float("nan"), float("inf"), -float("inf")):
        np.testing.assert_array_equal(orc.texture_lookup(img, bad, bad), o)


def _tri_scene(textures, shader_texture, images, uv):
    pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    return {
        "shaders": [{"type": 0, "colour": [0.1, 0.2, 0.3], "reflectance": 1.0, "texture": shader_texture},
                    {"type": 0, "colour": [0.4, 0.5, 0.6], "reflectance": 1.0}],
        "meshes": [{"pos4": pos, "idx": np.array([[0, 1, 2]], np.uint32), "uv": uv, "shader": 0},
                   {"pos4": pos + np.float32([0, 0, 1, 0]), "idx": np.array([[0, 1, 2]], np.uint32), "uv": uv,
                    "shader": 1}],
        "lights": [], "images": images, "textures": textures,
    }


def test_texture_coordinate_at_corners_is_the_corner_uv():
    w = 64
    ramp = np.zeros((w, w, 3), np.float32)
    ramp[:, :, 0] = np.arange(w)[None, :]
    ramp[:, :, 1] = np.arange(w)[:, None]
    uv = np.array([[0.25, 0.3], [0.6, 0.4], [0.35, 0.7]], np.float32)
    oc = orc.Scene(_tri_scene([{"type": 1, "image": 0}], 0, [ramp], uv))
    for (u, v), corner in (((0.0, 0.0), 0), ((1.0, 0.0), 1), ((0.0, 1.0), 2)):
        got = oc.texture_colour(0, 0, u, v)
        want = uv[corner] * w - 0.5  # the ramp reproduces x = s W - 1/2 away from the seam
        np.testing.assert_allclose(got[:2], want, rtol=1e-5, atol=1e-4)
    got = oc.texture_colour(0, 0, 1.0 / 3.0, 1.0 / 3.0)
    np.testing.assert_allclose(got[:2], uv.mean(axis=0) * w - 0.5, rtol=1e-5, atol=1e-4)
    # the untextured shader keeps its constant colour
    np.testing.assert_array_equal(oc.texture_colour(1, 1, 0.2, 0.2), np.float32([0.4, 0.5, 0.6]))


def test_layered_masks_select_upper_and_lower():
    rng = np.random.default_rng(11)
    img = rng.uniform(0.0, 1.0, (5, 5, 3)).astype(np.float32)
    uv = np.array([[0.1, 0.1], [0.9, 0.2], [0.3, 0.8]], np.float32)
    textures = [{"type": 1, "image": 0}, {"type": 0, "colour": [0.7, 0.1, 0.4]},
                {"type": 0, "colour": [1.0, 1.0, 1.0]}, {"type": 0, "colour": [0.0, 0.0, 0.0]},
                {"type": 0, "colour": [0.25, 0.5, 1.0]},
                {"type": 2, "upper": 0, "lower": 1, "mask": 2},
                {"type": 2, "upper": 0, "lower": 1, "mask": 3},
                {"type": 2, "upper": 0, "lower": 1, "mask": 4}]
    plain = orc.Scene(_tri_scene(textures, 0, [img], uv))
    for node, pick in ((5, "upper"), (6, "lower"), (7, "mix")):
        oc = orc.Scene(_tri_scene(textures, node, [img], uv))
        for u, v in ((0.2, 0.3), (0.6, 0.1)):
            up = plain.texture_colour(0, 0, u, v)
            lo = np.float32([0.7, 0.1, 0.4])
            got = oc.texture_colour(0, 0, u, v)
            if pick == "upper":
                np.testing.assert_array_equal(got, up)
            elif pick == "lower":
                np.testing.assert_array_equal(got, lo)
            else:
                m = np.float64([0.25, 0.5, 1.0])
                np.testing.assert_allclose(got, up * m + lo * (1 - m), rtol=1e-6, atol=1e-7)


def test_texture_descriptions_are_validated():
    lib = orc.lib()
    h = lib.oracle_scene_create()
    try:
        d = orc.TextureDesc(1, (orc.C.c_float * 3)(0, 0, 0), 0, -1, -1, -1)
        assert lib.oracle_scene_add_texture(h, orc.C.byref(d)) == -1  # no image 0 yet
        d = orc.TextureDesc(2, (orc.C.c_float * 3)(0, 0, 0), -1, 0, 0, 0)
        assert lib.oracle_scene_add_texture(h, orc.C.byref(d)) == -1  # layers must name earlier nodes
        d = orc.TextureDesc(0, (orc.C.c_float * 3)(1, 1, 1), -1, -1, -1, -1)
        assert lib.oracle_scene_add_texture(h, orc.C.byref(d)) == 0
        assert lib.oracle_scene_set_shader_texture(h, 0, 0) == -1  # no shader 0 yet
    finally:
        lib.oracle_scene_destroy(h)
