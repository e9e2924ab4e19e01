"""Out of core: renders whose per-sample state exceeds the device memory budget are traced in
column bands, one band's state at a time (the reference spills ray bins to disk instead,
src/core/DirectionalBins.cpp:37-115, reloaded in Pathtracer.cpp:294-313). A path is a pure
function of its sample id and every pixel is filtered from the same samples in the same order, so
the banded image is bit-identical to the unbudgeted one; bands hold their columns plus the tent
filter's neighbour columns (traced twice, once per neighbouring band)."""
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
def room():
    return scenes.box_room(width=203, height=97, base=3, sphere_level=3)


@pytest.mark.parametrize("filt,fraction", [(0, 5), (1, 7), (0, 40)])
def test_plain_render_under_a_budget_equals_unbudgeted(room, filt, fraction):
    """A fresh context: the unbudgeted render, then a budget that leaves the per-sample arenas
    1/fraction of the frame's state; pixels and every sample record are identical."""
    ctx = swr.Context(0)
    try:
        sc = swr.Scene(ctx, room)
        rd = _desc(room, filter=filt, iteration=3)
        pix, smp, st = sc.render(rd, samples=True)
        total = rd.width * rd.height * 9
        held = ctx.bytes_in_use()  # includes the frame's radiance and film-position arenas, 24 B/sample
        budget = held - 24 * total + (24 + 20) * total // fraction  # (+ 20 B/sample staging for records)
        ctx.set_budget(budget)
        bpix, bsmp, bst = sc.render(rd, samples=True)
        assert ctx.bytes_in_use() <= budget
        np.testing.assert_array_equal(bpix, pix)
        np.testing.assert_array_equal(bsmp, smp)
        assert bst["camera_rays"] > st["camera_rays"] if filt == 0 else bst["camera_rays"] == st["camera_rays"]
        # pixels only (no staging for sample records): still banded, still identical
        np.testing.assert_array_equal(sc.render(rd)[0], pix)
        ctx.set_budget(0)
        sc.close()
    finally:
        ctx.close()


def test_plain_render_refuses_a_budget_without_room_for_one_column(room):
    ctx = swr.Context(0)
    try:
        sc = swr.Scene(ctx, room)
        rd = _desc(room)
        sc.render(rd)
        ctx.set_budget(ctx.bytes_in_use() - 24 * rd.width * rd.height * 9 + 24 * rd.height * 9)
        with pytest.raises(swr.SwrError) as e:
            sc.render(rd)
        assert e.value.code == swr.SWR_E_BUDGET
        ctx.set_budget(0)
        sc.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("filt", [0, 1])
def test_served_request_larger_than_the_state_budget(ctx, room, filt):
    """A server whose per-sample arena holds a fraction of one iteration: the request runs as
    column bands and equals the same request on a server with room to spare, and the mean of
    plain renders of its iterations."""
    sc = swr.Scene(ctx, room)
    rd = _desc(room, filter=filt, iteration=11)
    iters = 3
    big = swr.Server(ctx, [sc], state_bytes=256 << 20, max_requests=8)
    try:
        ref, st = big.wait(big.submit(0, rd, iters))
        assert st["status"] == 0
    finally:
        big.close()
    state = 2 << 20  # arena: (2 MiB * 3/4) / 24 B = 65536 samples, a third of an iteration's 177,219
    small = swr.Server(ctx, [sc], state_bytes=state, max_requests=8)
    try:
        ids = [small.submit(0, rd, iters), small.submit(0, _desc(room, width=31, height=17, seed=5), 2)]
        out = [small.wait(i) for i in ids]
        tot = small.totals()
    finally:
        small.close()
    assert all(o[1]["status"] == 0 for o in out)
    assert tot["state_bytes"] <= state
    np.testing.assert_array_equal(out[0][0], ref)
    acc = None
    for i in range(iters):
        rd.iteration = 11 + i
        p = sc.render(rd)[0]
        acc = p if acc is None else acc + p
    np.testing.assert_array_equal(ref, acc * np.float32(1.0 / iters))
    sc.close()


def test_batch_under_a_budget_runs_in_groups(room):
    """A batch whose jobs' state exceeds the budget runs as groups of consecutive jobs, one loop
    per group: every image is the same as in the unbudgeted run, and so are the ray counts."""
    ctx = swr.Context(0)
    try:
        sc = swr.Scene(ctx, room)
        rds = [_desc(room, width=40 + 17 * j, height=30 + 9 * j, base=1 + j % 3, seed=900 + j, filter=j % 2)
               for j in range(7)]
        b = swr.Batch(ctx)
        for rd in rds:
            b.add(sc, rd)
        st = b.run(4096)
        ref = [b.result(j)[0] for j in range(len(rds))]
        state = sum(24 * rd.width * rd.height * rd.base * rd.base + 12 * rd.width * rd.height for rd in rds)
        held = ctx.bytes_in_use()  # includes the arenas of the run above (the whole state)
        biggest = max(24 * rd.width * rd.height * rd.base * rd.base + 12 * rd.width * rd.height for rd in rds)
        budget = held - state + max(state // 3, biggest + 8192)  # room for about a third of the jobs
        ctx.set_budget(budget)
        st2 = b.run(4096)
        assert ctx.bytes_in_use() <= budget
        for j in range(len(rds)):
            np.testing.assert_array_equal(b.result(j)[0], ref[j], err_msg=f"job {j}")
        assert (st2["camera_rays"], st2["closest_rays"], st2["shadow_rays"]) == \
               (st["camera_rays"], st["closest_rays"], st["shadow_rays"])
        with pytest.raises(swr.SwrError) as e:
            b.result(0, samples=True)  # per-sample records of a grouped run are not kept
        assert e.value.code == swr.SWR_E_INVALID
        # a job that cannot fit alone is refused
        ctx.set_budget(held - state + biggest // 2)
        with pytest.raises(swr.SwrError) as e:
            b.run(4096)
        assert e.value.code == swr.SWR_E_BUDGET
        ctx.set_budget(0)
        assert len(b.run(4096)) > 0  # and runs whole again without a budget
        b.close()
        sc.close()
    finally:
        ctx.close()


def test_batch_and_scene_outlive_their_context(room):
    """Destroying a context with a batch and a scene still alive detaches them: later calls fail
    cleanly, and destroying the handles afterwards is safe."""
    ctx = swr.Context(0)
    sc = swr.Scene(ctx, room)
    b = swr.Batch(ctx)
    b.add(sc, _desc(room, width=16, height=12, base=1))
    b.run()
    ctx.close()
    lib = swr.load_library()
    assert lib.swr_batch_run(b.h, 0, 0, None) == swr.SWR_E_INVALID
    b.close()
    sc.close()
