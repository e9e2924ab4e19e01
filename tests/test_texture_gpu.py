"""GPU parity of textured shading (StandardTexture, LayeredTexture and ConstantTexture on Lambert
shaders; the texture coordinate of PolygonObject::texture).

The device evaluates each shader's texture graph as a post-order stack program in the shading
kernel; the oracle evaluates the graph recursively. Both draw the same Philox streams, so
per-sample radiance agrees stream for stream up to the rare path that diverges by a last-ulp
transcendental, as for untextured scenes. A batch that mixes a textured and an untextured scene
(the multi-job shading kernel) must equal the solo renders bit for bit."""
import numpy as np
import pytest

import oracle_ffi as orc
import swr
from swr import scenes

pytestmark = pytest.mark.gpu


def _pair(desc, **over):
    r = dict(desc["render"])
    r.update(over)
    return swr.render_desc(r), swr.render_desc(r, orc.RenderDesc)


def _compare_samples(g, o, tol_frac):
    np.testing.assert_array_equal(g[:, 0:2], o[:, 0:2])
    d = np.abs(g[:, 2:5] - o[:, 2:5])
    close = np.all(d <= 1e-4 * (1.0 + np.abs(o[:, 2:5])), axis=1)
    assert close.mean() >= 1.0 - tol_frac, f"{(~close).sum()} of {len(close)} samples differ"
    return close


@pytest.mark.parametrize("filt", [0, 1])
def test_textured_room_matches_oracle_stream_for_stream(ctx, filt):
    desc = scenes.textured_room(width=64, height=48, base=3, sphere_level=3)
    rd, od = _pair(desc, filter=filt)
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    try:
        pg, sg, st = sc.render(rd, samples=True)
        po, so, ost = oc.render(od, samples=True)
        _compare_samples(sg, so, 0.01)
        assert abs(st["closest_rays"] - ost.closest_rays) <= 0.02 * ost.closest_rays
        rel = np.abs(pg - po) / (np.abs(po) + 1e-3)
        assert np.median(rel) < 1e-5
        assert np.mean(rel < 1e-3) > 0.95
    finally:
        sc.close()


def test_textures_change_the_image(ctx):
    """The same room without its textures renders differently where textured surfaces are seen,
    and identically in the first samples' film positions (same camera rays)."""
    desc = scenes.textured_room(width=32, height=32, base=2, sphere_level=2)
    plain = dict(desc)
    plain["shaders"] = [{k: v for k, v in sh.items() if k != "texture"} for sh in desc["shaders"]]
    rd, _ = _pair(desc)
    a, b = swr.Scene(ctx, desc), swr.Scene(ctx, plain)
    try:
        pa, sa, _ = a.render(rd, samples=True)
        pb, sb, _ = b.render(rd, samples=True)
        np.testing.assert_array_equal(sa[:, 0:2], sb[:, 0:2])
        assert np.mean(np.abs(pa - pb) > 1e-3) > 0.2
    finally:
        a.close()
        b.close()


def test_batch_of_textured_and_untextured_scenes_equals_solo(ctx):
    tex = scenes.textured_room(width=40, height=30, base=2, sphere_level=2)
    room = scenes.box_room(seed=9, width=24, height=20, base=2, sphere_level=2)
    scs = [swr.Scene(ctx, tex), swr.Scene(ctx, room)]
    jobs = [(0, _pair(tex)[0]), (1, _pair(room)[0]), (0, _pair(tex, seed=17, iteration=2)[0])]
    try:
        solo = [scs[s].render(rd, samples=True) for s, rd in jobs]
        b = swr.Batch(ctx)
        try:
            for s, rd in jobs:
                b.add(scs[s], rd)
            b.run(2048)
            for j, (pix, smp, _) in enumerate(solo):
                p, sm = b.result(j, samples=True)
                np.testing.assert_array_equal(sm, smp)
                np.testing.assert_array_equal(p, pix)
        finally:
            b.close()
    finally:
        for sc in scs:
            sc.close()


def test_texture_errors(ctx):
    sc = swr.Scene(ctx)
    try:
        sc.add_shader({"type": 0, "colour": [0.5, 0.5, 0.5]})
        with pytest.raises(swr.SwrError):
            sc.add_texture({"type": 1, "image": 0})  # no image yet
        with pytest.raises(swr.SwrError):
            sc.add_image(np.full((2, 2, 3), np.nan, np.float32))
        assert sc.add_image(np.ones((2, 3, 3), np.float32)) == 0
        assert sc.add_texture({"type": 1, "image": 0}) == 0
        with pytest.raises(swr.SwrError):
            sc.add_texture({"type": 2, "upper": 0, "lower": 0, "mask": 1})  # mask names itself
        with pytest.raises(swr.SwrError):
            sc.set_shader_texture(0, 5)
        with pytest.raises(swr.SwrError):
            sc.set_shader_texture(1, 0)
        # a mask chain nested deeper than the evaluation stack fails at commit
        node = 0
        for _ in range(6):
            node = sc.add_texture({"type": 2, "upper": 0, "lower": 0, "mask": node})
        sc.set_shader_texture(0, node)
        pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
        sc.add_mesh(pos, np.array([[0, 1, 2]], np.uint32), None, 0)
        with pytest.raises(swr.SwrError, match="nests"):
            sc.commit()
    finally:
        sc.close()
