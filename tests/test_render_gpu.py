"""GPU parity of the sorted wavefront render (Pathtracer::process replacement).

Both sides draw every random number from Philox keyed by (seed, sample, iteration, depth,
slot), so the GPU and the oracle trace the same paths: per-sample radiance agrees except for
the rare path whose direction differs by a last-ulp transcendental and so hits elsewhere.
The converged-image test (independent streams, 3 sigma of the per-pixel Monte Carlo
standard error) is checked as well, and the result must not depend on how rays are grouped
into waves (queue capacity), which is the batched-equals-solo property."""
import math

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
    """Per-sample radiance: near-identical for all but a small fraction of diverged paths."""
    np.testing.assert_array_equal(g[:, 0:2], o[:, 0:2])  # film positions are bit-identical
    d = np.abs(g[:, 2:5] - o[:, 2:5])
    close = np.all(d <= 1e-4 * (1.0 + np.abs(o[:, 2:5])), axis=1)
    assert close.mean() >= 1.0 - tol_frac, f"{(~close).sum()} of {len(close)} samples differ"
    return close


def test_raygen_matches_oracle(ctx):
    for cam_type in (0, 1):
        desc = scenes.box_room(width=40, height=30, base=3)
        desc["render"]["camera"]["type"] = cam_type
        desc["render"]["camera"]["aperture"] = 8.0
        desc["render"]["camera"]["focal_distance"] = 200.0
        for sampler in (0, 1, 2):
            rd, od = _pair(desc, sampler=sampler)
            n = 40 * 30 * 9
            rg, xg = swr.raygen(ctx, rd, 0, n)
            ro, xo = orc.raygen(od, 0, n)
            np.testing.assert_array_equal(xg, xo)
            np.testing.assert_allclose(rg, ro, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("filt", [0, 1])
def test_box_room_render_matches_oracle_stream_for_stream(ctx, filt):
    desc = scenes.box_room(width=64, height=48, base=3, sphere_level=3)
    rd, od = _pair(desc, filter=filt)
    sc = swr.Scene(ctx, desc)
    oc = orc.Scene(desc)
    pg, sg, st = sc.render(rd, samples=True)
    po, so, ost = oc.render(od, samples=True)
    close = _compare_samples(sg, so, 0.01)
    assert st["camera_rays"] == 64 * 48 * 9
    # ray counts agree up to the diverged paths
    assert abs(st["closest_rays"] - ost.closest_rays) <= 0.02 * ost.closest_rays
    # pixels: the same filter terms, so pixels without a diverged sample agree tightly
    rel = np.abs(pg - po) / (np.abs(po) + 1e-3)
    assert np.median(rel) < 1e-5
    assert np.mean(rel < 1e-3) > 0.95


@pytest.mark.parametrize("filt", [0, 1])
@pytest.mark.parametrize("w,h,base", [(67, 130, 2), (5, 3, 3), (1, 1, 1)])
def test_filter_of_gpu_samples_matches_oracle_filter(ctx, filt, w, h, base):
    """The resolve, applied to the GPU's own samples, against the oracle's resolve of the same
    samples: the same terms summed in another order (sample index outermost), so within a few
    float roundings. Sizes that leave partial tiles and one-pixel images cover the halo edges."""
    desc = scenes.box_room(width=w, height=h, base=base, sphere_level=2)
    rd, _ = _pair(desc, filter=filt)
    sc = swr.Scene(ctx, desc)
    pg, sg, _ = sc.render(rd, samples=True)
    po = orc.filter_image(filt, w, h, base * base, sg)
    np.testing.assert_allclose(pg, po, rtol=1e-5, atol=1e-7)


def test_render_is_deterministic_and_independent_of_wave_size(ctx):
    desc = scenes.box_room(width=48, height=48, base=2, sphere_level=3)
    sc = swr.Scene(ctx, desc)
    outs = []
    for exp in (10, 12, 20):
        rd, _ = _pair(desc, bin_exponent=exp)
        pix, smp, st = sc.render(rd, samples=True)
        outs.append((pix, smp, st))
    for pix, smp, st in outs[1:]:
        np.testing.assert_array_equal(pix, outs[0][0])
        np.testing.assert_array_equal(smp, outs[0][1])
    assert outs[0][2]["waves"] > outs[2][2]["waves"]


def test_glossy_terrain_matches_oracle(ctx):
    desc = scenes.terrain(grid=64, groups=16, width=48, height=27, spp=4, max_depth=8)
    rd, od = _pair(desc)
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    _, sg, _ = sc.render(rd, samples=True)
    _, so, _ = oc.render(od, samples=True)
    _compare_samples(sg, so, 0.02)


def test_images_agree_within_monte_carlo_error(ctx):
    """Independent streams (different seeds): per-pixel means over K = 16 iterations agree within
    3 sigma of the combined standard error for >= 99% of pixel-channels, none beyond 6 sigma."""
    desc = scenes.box_room(width=24, height=24, base=4, sphere_level=3)
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    K = 16
    g, o = [], []
    for it in range(K):
        rd, _ = _pair(desc, iteration=it, seed=1000)
        _, od = _pair(desc, iteration=it, seed=2000)
        g.append(sc.render(rd)[0])
        o.append(oc.render(od)[0])
    g, o = np.array(g, np.float64), np.array(o, np.float64)
    se = np.sqrt(g.var(axis=0, ddof=1) / K + o.var(axis=0, ddof=1) / K) + 1e-12
    z = np.abs(g.mean(axis=0) - o.mean(axis=0)) / se
    assert np.mean(z <= 3.0) >= 0.99
    assert np.max(z) < 6.0


@pytest.mark.parametrize("which", ["box_room", "terrain"])
def test_converged_images_4096spp(ctx, which):
    """Converged images: the GPU and the CPU oracle each at 4096 spp (16 iterations of 256 spp),
    independent seeds; per-pixel means agree within 3 sigma of
    the combined standard error (batch means over the 16 iterations) for >= 99% of pixel-channels,
    none beyond 6 sigma."""
    if which == "box_room":
        desc = scenes.box_room(width=20, height=16, base=16, sphere_level=3)
    else:
        desc = scenes.terrain(grid=64, groups=16, width=20, height=12, spp=256, max_depth=8)
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    K = 16
    g, o = [], []
    for it in range(K):
        rd, _ = _pair(desc, iteration=it, seed=3000, base=16, spp=0, sampler=0)
        _, od = _pair(desc, iteration=it, seed=4000, base=16, spp=0, sampler=0)
        g.append(sc.render(rd)[0])
        o.append(oc.render(od, threads=8)[0])
    g, o = np.array(g, np.float64), np.array(o, np.float64)
    se = np.sqrt(g.var(axis=0, ddof=1) / K + o.var(axis=0, ddof=1) / K) + 1e-12
    z = np.abs(g.mean(axis=0) - o.mean(axis=0)) / se
    assert np.mean(z <= 3.0) >= 0.99, np.sort(z.ravel())[-10:]
    assert np.max(z) < 6.0


def test_render_without_host_pixels(ctx):
    """pixels=None leaves the image on the device: same work (ray counts) as a render that copies
    it, and a render with a copy right after still gives the oracle's image."""
    desc = scenes.box_room(width=48, height=32, base=3, sphere_level=2)
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    rd, od = _pair(desc, iteration=2)
    pix0, _, st0 = sc.render(rd, pixels=False)
    assert pix0 is None and st0["camera_rays"] == 48 * 32 * 9
    pix, smp, st = sc.render(rd, samples=True)
    for k in ("camera_rays", "closest_rays", "shadow_rays", "waves"):
        assert st0[k] == st[k], k
    _, osmp, _ = oc.render(od, samples=True)
    _compare_samples(smp, osmp, 0.02)
    assert np.isfinite(pix).all()


def test_render_errors_and_cancel(ctx):
    desc = scenes.box_room(width=16, height=16, base=2, sphere_level=1)
    sc = swr.Scene(ctx, desc)
    rd, _ = _pair(desc, base=3, spp=8, sampler=0)  # stratified needs spp = base^2
    with pytest.raises(swr.SwrError):
        sc.render(rd)
    # spp = base^2 must fit 32 bits (base 65536 would wrap to 0 in 32-bit arithmetic), and so
    # must the samples of one iteration: refused before anything is allocated or launched
    for over in (dict(base=65536), dict(base=46341), dict(width=4096, height=4096, base=300),
                 dict(base=0, spp=2 ** 31 - 1, sampler=1)):
        rd, _ = _pair(desc, **over)
        with pytest.raises(swr.SwrError) as e:
            sc.render(rd)
        assert e.value.code == swr.SWR_E_INVALID, over
        b = swr.Batch(ctx)
        with pytest.raises(swr.SwrError) as e:
            b.add(sc, rd)
        assert e.value.code == swr.SWR_E_INVALID, over
        b.close()
    rd, _ = _pair(desc)
    ctx.cancel()
    with pytest.raises(swr.SwrError) as e:
        sc.render(rd)
    assert e.value.code == swr.SWR_E_CANCELLED
    ctx.reset_cancel()
    pix, _, _ = sc.render(rd)
    assert np.isfinite(pix).all()


def test_budget_refuses_large_render(ctx):
    c2 = swr.Context(0)
    desc = scenes.box_room(width=16, height=16, base=2, sphere_level=1)
    sc = swr.Scene(c2, desc)
    c2.set_budget(c2.bytes_in_use() + 4096)
    rd, _ = _pair(desc)
    with pytest.raises(swr.SwrError) as e:
        sc.render(rd)
    assert e.value.code == swr.SWR_E_BUDGET
    c2.set_budget(0)
    sc.render(rd)
    sc.close()
    c2.close()


def test_tuning_changes_speed_not_pixels(ctx):
    """Sort pass structures, tile sizes, key widths and traversal stack rings only reorder rays and
    hits or move stack entries between LDS and device memory: every tuning gives the same image,
    bit for bit."""
    desc = scenes.box_room(width=40, height=40, base=2, sphere_level=3)
    sc = swr.Scene(ctx, desc)
    rd, _ = _pair(desc)
    before = ctx.tuning()
    ref = sc.render(rd)[0]
    try:
        for kpt, rb, hb, ring, alg, ov in [(8, 32, 0, 8, 0, 1), (12, 24, 0, 16, 1, 0), (20, 16, 8, 64, 1, 1),
                                           (24, 3, 4, 32, 0, 0), (16, 32, 0, 16, 2, 1), (16, 32, 0, 16, 2, 0)]:
            # ov: the shadow pass beside the next wave's traversal (second stream) or in line
            ctx.set_tuning(kpt, rb, hb, ring, alg, ov)
            np.testing.assert_array_equal(sc.render(rd)[0], ref)
        # With the bucket sort selected, a scene committed now sorts its triangles' 30-bit Morton
        # codes through it (the count is known at launch): the same BVH, the same image.
        sc2 = swr.Scene(ctx, desc)
        np.testing.assert_array_equal(sc2.render(rd)[0], ref)
        sc2.close()
        with pytest.raises(swr.SwrError):
            ctx.set_tuning(16, 32, 0, 16, 3)
        with pytest.raises(swr.SwrError):
            ctx.set_tuning(16, 32, 0, 16, 1, 2)
        with pytest.raises(swr.SwrError):
            ctx.set_tuning(10, 32, 0)
        with pytest.raises(swr.SwrError):
            ctx.set_tuning(16, 32, 0, 12)
        t = ctx.autotune(sc, rd)
        assert t["sort_keys_per_thread"] in (8, 12, 16, 20, 24)
        assert t["trace_stack_entries"] in (8, 16, 32, 64)
        np.testing.assert_array_equal(sc.render(rd)[0], ref)
    finally:
        ctx.set_tuning(**before)
