"""GPU tests of the serving loop: progressive requests admitted while the loop runs.

A request's iterations are filtered and added to its framebuffer in iteration order, and its
paths draw random numbers keyed by (seed, sample, iteration, depth, slot) only, so its result is
the same bit for bit alone, in a mix with other requests, and as the mean of plain renders of its
iterations (the reference's pixels += per process(), shown as pixels / iteration)."""
import threading

import numpy as np
import pytest

import swr
from swr import scenes

pytestmark = pytest.mark.gpu


def _desc(scene_desc, **over):
    r = dict(scene_desc["render"])
    r.update(over)
    return swr.render_desc(r)


@pytest.fixture(scope="module")
def resident(ctx):
    descs = [scenes.box_room(width=40, height=32, base=2, sphere_level=3),
             scenes.terrain(grid=32, groups=4, width=36, height=20, spp=4, max_depth=4),
             scenes.sphere_cluster(3, 4, 2, width=24, height=24, base=1)]
    scs = [swr.Scene(ctx, d) for d in descs]
    yield descs, scs
    for sc in scs:
        sc.close()


def _mixed_requests(descs):
    reqs = []
    rng = np.random.default_rng(5)
    for i in range(24):
        s = int(rng.integers(0, 3))
        w, h = int(rng.integers(8, 48)), int(rng.integers(8, 48))
        spp = int(rng.integers(1, 10))
        reqs.append((s, _desc(descs[s], width=w, height=h, spp=spp, base=1, sampler=1, seed=100 + i,
                              filter=int(rng.integers(0, 2))), int(rng.integers(1, 4))))
    return reqs


def test_request_equals_mean_of_plain_iterations(ctx, resident):
    descs, scs = resident
    rd = _desc(descs[0], iteration=7)
    srv = swr.Server(ctx, scs, state_bytes=64 << 20, max_requests=8)
    try:
        pix, st = srv.wait(srv.submit(0, rd, iterations=3))
        assert st["status"] == 0 and st["iterations"] == 3 and st["latency_ms"] > 0
    finally:
        srv.close()
    acc = np.zeros_like(pix)
    for i in range(3):
        rd.iteration = 7 + i
        acc += scs[0].render(rd)[0]
    np.testing.assert_array_equal(pix, acc * np.float32(1.0 / 3.0))


def test_alone_equals_mixed(ctx, resident):
    descs, scs = resident
    reqs = _mixed_requests(descs)
    target = (1, _desc(descs[1], seed=4242), 3)
    srv = swr.Server(ctx, scs, state_bytes=32 << 20, max_requests=6, wave_rays=4096)
    try:
        alone, _ = srv.wait(srv.submit(target[0], target[1], target[2]))
        # the same request in a crowd: admitted behind, between and alongside others
        ids = [srv.submit(s, rd, it) for s, rd, it in reqs[:10]]
        tid = srv.submit(*target)
        ids += [srv.submit(s, rd, it) for s, rd, it in reqs[10:]]
        mixed, st = srv.wait(tid)
        others = [srv.wait(i) for i in ids]
        tot = srv.totals()
    finally:
        srv.close()
    np.testing.assert_array_equal(mixed, alone)
    assert all(o[1]["status"] == 0 for o in others)
    assert tot["completed"] == 2 + len(reqs) and tot["closest_rays"] > 0 and tot["shadow_rays"] > 0
    # every other request also equals its lone rendering (plain renders of its iterations)
    for (s, rd, it), (pix, _) in list(zip(reqs, others))[::5]:
        acc = np.zeros_like(pix)
        base_it = rd.iteration
        for i in range(it):
            rd.iteration = base_it + i
            acc += scs[s].render(rd)[0]
        np.testing.assert_array_equal(pix, acc * np.float32(1.0 / it))


def test_submit_from_threads_and_refusals(ctx, resident):
    descs, scs = resident
    reqs = _mixed_requests(descs)[:12]
    srv = swr.Server(ctx, scs, state_bytes=8 << 20, max_requests=3)
    results = {}
    try:
        with pytest.raises(swr.SwrError) as e:
            srv.submit(0, _desc(descs[0], width=2048, height=2048, base=4))  # can never fit
        assert e.value.code == swr.SWR_E_BUDGET
        with pytest.raises(swr.SwrError):
            srv.submit(7, _desc(descs[0]))  # no such scene
        with pytest.raises(swr.SwrError):
            srv.submit(0, _desc(descs[0]), iterations=0)
        with pytest.raises(swr.SwrError):
            srv.wait(123456)

        def worker(k):
            for j in range(k, len(reqs), 3):
                s, rd, it = reqs[j]
                results[j] = srv.wait(srv.submit(s, rd, it))

        ts = [threading.Thread(target=worker, args=(k,)) for k in range(3)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        tot = srv.totals()
    finally:
        srv.close()
    assert sorted(results) == list(range(len(reqs)))
    assert all(r[1]["status"] == 0 for r in results.values())
    assert tot["refused"] == 1 and tot["completed"] == len(reqs)


def test_destroy_with_work_in_flight(ctx, resident):
    descs, scs = resident
    srv = swr.Server(ctx, scs, state_bytes=16 << 20, max_requests=4)
    for s, rd, it in _mixed_requests(descs):
        srv.submit(s, rd, it, keep=False)
    srv.close()  # stops the thread; unfinished requests are dropped
    # the context is usable again afterwards
    pix, _, _ = scs[0].render(_desc(descs[0]))
    assert np.isfinite(pix).all()
