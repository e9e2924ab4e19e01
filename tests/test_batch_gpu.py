"""GPU tests of the batched driver: many render jobs in one persistent wavefront loop.

Every random number a path draws is keyed by (job seed, sample id, iteration, depth, slot), and
each sample's radiance is accumulated in place in bounce order, so a job's image cannot depend
on which other jobs share its waves or on the wave size: the batched image of every job must
equal, bit for bit, the plain render of that job alone (swr_render)."""
import numpy as np
import pytest

import swr
from swr import scenes

pytestmark = pytest.mark.gpu


def _desc(scene_desc, **over):
    r = dict(scene_desc["render"])
    r.update(over)
    return swr.render_desc(r)


def _mixed_jobs():
    room = scenes.box_room(width=48, height=40, base=2, sphere_level=3)
    room2 = scenes.box_room(seed=7, width=30, height=30, base=3, sphere_level=2)
    soup = scenes.terrain(grid=32, groups=4, width=40, height=24, spp=5, max_depth=4)
    jobs = [
        (0, _desc(room)),
        (1, _desc(room2, filter=1)),
        (2, _desc(soup)),  # independent sampler (spp 5 is not a square)
        (0, _desc(room, width=17, height=9, base=1, seed=11)),
        (2, _desc(soup, seed=99, max_depth=6, sampler=1, spp=3, base=1)),
        (1, _desc(room2, camera=scenes.camera_for([0.3, 1.1, 3.0], 95.0, -5.0), seed=5, iteration=3)),
        (0, _desc(room, sampler=2, base=3, width=20, height=20)),  # grid sampler
    ]
    return [room, room2, soup], jobs


@pytest.fixture(scope="module")
def mixed(ctx):
    descs, jobs = _mixed_jobs()
    scs = [swr.Scene(ctx, d) for d in descs]
    solo = [scs[s].render(rd, samples=True) for s, rd in jobs]
    yield scs, jobs, solo
    for sc in scs:
        sc.close()


@pytest.mark.parametrize("wave_rays,paced", [(0, False), (4096, False), (777, False), (4096, True)])
def test_batch_equals_solo_renders(ctx, mixed, wave_rays, paced):
    scs, jobs, solo = mixed
    b = swr.Batch(ctx)
    try:
        for s, rd in jobs:
            b.add(scs[s], rd)
        st = b.run(wave_rays, paced=paced)
        assert st["camera_rays"] == sum(rd.width * rd.height * (rd.spp or rd.base * rd.base) for _, rd in jobs)
        assert st["closest_rays"] == sum(x[2]["closest_rays"] for x in solo)
        assert st["shadow_rays"] == sum(x[2]["shadow_rays"] for x in solo)
        if wave_rays:
            assert st["max_wave"] <= wave_rays
        for j, (pix_solo, smp_solo, _) in enumerate(solo):
            pix, smp = b.result(j, samples=True)
            np.testing.assert_array_equal(smp, smp_solo)
            np.testing.assert_array_equal(pix, pix_solo)
    finally:
        b.close()


def test_batch_of_tiny_jobs_on_many_scenes_equals_solo(ctx):
    """Scene segments shorter than a traversal tile (128 rays): tiles hold two, three or more
    scenes (and scenes drop out of later waves as their jobs finish), which exercises the split
    of a multi-scene wave into scene-uniform tiles and tiles at scene boundaries."""
    descs = [scenes.box_room(seed=100 + k, width=8, height=8, base=1, sphere_level=1) for k in range(4)]
    scs = [swr.Scene(ctx, d) for d in descs]
    jobs = []
    for k in range(12):
        s = k % 4
        jobs.append((s, _desc(descs[s], width=3 + k % 5, height=2 + k % 3, base=1, spp=1 + k % 4, sampler=1,
                              seed=500 + k)))
    try:
        solo = [scs[s].render(rd, samples=True) for s, rd in jobs]
        for wave_rays in (0, 64, 300):
            b = swr.Batch(ctx)
            try:
                for s, rd in jobs:
                    b.add(scs[s], rd)
                b.run(wave_rays)
                for j, (pix, smp, _) in enumerate(solo):
                    p, sm = b.result(j, samples=True)
                    np.testing.assert_array_equal(sm, smp)
                    np.testing.assert_array_equal(p, pix)
            finally:
                b.close()
    finally:
        for sc in scs:
            sc.close()


def test_batch_autotune_keeps_images(ctx, mixed):
    scs, jobs, solo = mixed
    before = ctx.tuning()
    b = swr.Batch(ctx)
    try:
        for s, rd in jobs:
            b.add(scs[s], rd)
        t = b.autotune(2048)
        assert t["trace_stack_entries"] == 0  # rings were chosen per scene
        b.run(2048)
        for j, (pix_solo, _, _) in enumerate(solo):
            np.testing.assert_array_equal(b.result(j)[0], pix_solo)
    finally:
        b.close()
        ctx.set_tuning(**before)


def test_batch_errors(ctx, mixed):
    scs, jobs, _ = mixed
    b = swr.Batch(ctx)
    try:
        with pytest.raises(swr.SwrError):
            b.run()  # no jobs
        b.add(scs[0], jobs[0][1])
        with pytest.raises(swr.SwrError):
            b.result(0)  # not run yet
        bad = _desc(scenes.box_room(), width=0)
        with pytest.raises(swr.SwrError):
            b.add(scs[0], bad)
        other = swr.Context(0)
        try:
            sc_other = swr.Scene(other, scenes.box_room(width=8, height=8, base=1, sphere_level=1))
            with pytest.raises(swr.SwrError):
                b.add(sc_other, jobs[0][1])  # scene of another context
            sc_other.close()
        finally:
            other.close()
        b.run()
        with pytest.raises(swr.SwrError):
            b.result(5)
        np.testing.assert_array_equal(b.result(0)[0], mixed[2][0][0])
    finally:
        b.close()


def test_batch_cancel(ctx, mixed):
    scs, jobs, _ = mixed
    b = swr.Batch(ctx)
    try:
        for s, rd in jobs:
            b.add(scs[s], rd)
        ctx.cancel()
        with pytest.raises(swr.SwrError) as e:
            b.run(1024)
        assert e.value.code == swr.SWR_E_CANCELLED
    finally:
        ctx.reset_cancel()
        b.close()
    # the context still renders after a cancelled batch
    pix, _, _ = scs[0].render(jobs[0][1])
    np.testing.assert_array_equal(pix, mixed[2][0][0])
