"""GPU parity of the LBVH/BVH4 build and traversal (rtcIntersect / rtcOccluded replacements).

Closest hits are compared with the oracle's float BVH path (same triangle arithmetic, so t, u,
v and the triangle id are bit-identical) and with its FP64 brute force (id equal except at
documented epsilon ties, t within 1e-5 relative)."""
import numpy as np
import pytest

import oracle_ffi as orc
import swr
from swr import scenes

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def random_rays(rng, n, lo=-1.2, hi=1.2, tmin=0.0, tmax=1e5):
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = tmin
    rays[:, 7] = tmax
    return rays


def check_against_oracle(gpu, ora_bvh, ora_brute=None, allow_tie_frac=1e-3):
    n = gpu.shape[0]
    # bit-identical to the oracle's float path
    same = np.all(gpu == ora_bvh, axis=1) | ((gpu[:, 3] == MISS) & (ora_bvh[:, 3] == MISS))
    assert same.mean() >= 1 - allow_tie_frac, f"{(~same).sum()} of {n} hits differ from the oracle"
    if ora_brute is not None:
        gpu = gpu[:ora_brute.shape[0]]  # the brute force runs on a prefix of the rays
        hit_g = gpu[:, 3] != MISS
        hit_b = ora_brute[:, 3] != MISS
        agree = (gpu[:, 3] == ora_brute[:, 3])
        assert agree.mean() >= 1 - allow_tie_frac
        both = hit_g & hit_b & agree
        tg = gpu[both, 0].view(np.float32).astype(np.float64)
        tb = ora_brute[both, 0].view(np.float32).astype(np.float64)
        # float32 rounding of t follows the magnitudes of the ray origin and the vertices (about 1
        # in these scenes), not t itself: a hit at t = 0.007 carries the error of one at t = 1
        assert np.all(np.abs(tg - tb) <= 1e-5 * np.maximum(np.abs(tb), 1.0))


@pytest.mark.parametrize("ntris,seed", [(1, 0), (3, 1), (5, 2), (100, 3), (3000, 4)])
def test_intersect_soup_vs_oracle(ctx, ntris, seed):
    desc = scenes.random_soup(seed, ntris)
    sc = swr.Scene(ctx, desc)
    info = sc.bvh_info()
    assert info["valid"], info
    oc = orc.Scene(desc)
    rng = np.random.default_rng(seed)
    rays = random_rays(rng, 20000)
    g = sc.intersect(rays)
    check_against_oracle(g, oc.intersect(rays), oc.intersect_brute(rays[:4000]) if ntris <= 3000 else None)


def _soup_desc(tris):
    pos = np.zeros((tris.shape[0] * 3, 4), np.float32)
    pos[:, :3] = tris.reshape(-1, 3)
    idx = np.arange(tris.shape[0] * 3, dtype=np.uint32).reshape(-1, 3)
    return {"shaders": [{"type": 0, "colour": [0.5, 0.5, 0.5]}],
            "meshes": [{"pos4": pos, "idx": idx, "uv": None, "shader": 0}], "lights": [],
            "environment": (0.0, 0.0, 0.0)}


@pytest.mark.parametrize("case", ["small", "coincident", "clusters", "stacked"])
def test_restructured_tree_edge_cases(ctx, case):
    """The build restructures treelets of 7 leaves under every node of at least 3 triangles and
    renumbers the leaves afterwards: scenes at and just above the leaf size (2), treelets whose
    leaves all have the same box (every topology costs the same), many triangles sharing few
    centroids (Morton ties broken by index), and parallel layers (thin boxes, zero-area faces)
    must still give a valid tree whose hits equal the oracle's."""
    rng = np.random.default_rng(sum(map(ord, case)))
    if case == "small":
        sizes = [2, 3, 4, 6, 7, 8, 9, 15, 16, 17, 64]
        descs = [scenes.random_soup(int(rng.integers(1 << 30)), n) for n in sizes]
    elif case == "coincident":
        one = rng.uniform(-0.5, 0.5, (1, 3, 3)).astype(np.float32)
        descs = [_soup_desc(np.repeat(one, n, axis=0)) for n in (3, 7, 8, 9, 33, 200)]
    elif case == "clusters":
        centres = rng.uniform(-1, 1, (5, 1, 3))
        pick = rng.integers(0, 5, 3000)
        offs = rng.uniform(-0.05, 0.05, (3000, 3, 3))
        offs -= offs.mean(axis=1, keepdims=True)  # every triangle's centroid is its cluster's centre
        descs = [_soup_desc((centres[pick] + offs).astype(np.float32))]
    else:
        xy = rng.uniform(-1, 1, (2000, 3, 2))
        z = np.repeat(rng.integers(0, 4, (2000, 1, 1)) * 0.25, 3, axis=1)  # four z = const layers
        descs = [_soup_desc(np.concatenate([xy, z], axis=2).astype(np.float32))]
    ties = case in ("coincident", "stacked")  # equal-t hits on overlapping coplanar triangles
    for desc in descs:
        sc = swr.Scene(ctx, desc)
        info = sc.bvh_info()
        assert info["valid"], (case, info)
        if case == "coincident":  # every split costs the same: even splits keep the tree shallow
            n = desc["meshes"][0]["idx"].shape[0]
            assert info["max_depth"] <= 2 * int(np.ceil(np.log2(n))), (n, info)
        oc = orc.Scene(desc)
        rays = random_rays(rng, 8000)
        g = sc.intersect(rays)
        brute = oc.intersect_brute(rays[:2000])
        check_against_oracle(g, oc.intersect(rays), None if ties else brute)
        if ties:  # FP64 and float32 may pick different triangles at one distance: compare t only
            gb = g[:2000]
            assert np.array_equal(gb[:, 3] == MISS, brute[:, 3] == MISS)
            hit = gb[:, 3] != MISS
            tg = gb[hit, 0].view(np.float32).astype(np.float64)
            tb = brute[hit, 0].view(np.float32).astype(np.float64)
            assert np.all(np.abs(tg - tb) <= 1e-5 * np.maximum(np.abs(tb), 1.0))
        occ = sc.occluded(rays, swr.MASK_SHADOW)
        assert np.array_equal(occ, oc.occluded(rays, swr.MASK_SHADOW))


def test_intersect_box_room_and_masks(ctx):
    desc = scenes.box_room(sphere_level=3)
    sc = swr.Scene(ctx, desc)
    assert sc.bvh_info()["valid"]
    assert sc.num_geometries == len(desc["meshes"]) + len(desc["lights"])
    oc = orc.Scene(desc)
    rng = np.random.default_rng(9)
    rays = random_rays(rng, 50000, lo=-0.9, hi=0.9)
    rays[:, 1] += 1.0
    for mask in (swr.MASK_ALL, swr.MASK_SHADOW):
        g = sc.intersect(rays, mask)
        o = oc.intersect(rays, mask)
        check_against_oracle(g, o, oc.intersect_brute(rays[:3000], mask))
        if mask == swr.MASK_SHADOW:
            # lights (last geometries) are invisible to shadow-masked rays
            geoms = {sc.prim_to_geom(int(p))[0] for p in g[:, 3] if p != MISS}
            assert max(geoms) < len(desc["meshes"])
    # occlusion agrees with the oracle, and with the closest hit where the segment is short
    occ = sc.occluded(rays, swr.MASK_SHADOW)
    assert np.array_equal(occ, oc.occluded(rays, swr.MASK_SHADOW))
    short = rays.copy()
    short[:, 7] = 0.3
    assert np.array_equal(sc.occluded(short, swr.MASK_SHADOW), oc.occluded(short, swr.MASK_SHADOW))


def test_camera_rays_hit_what_the_oracle_hits(ctx):
    desc = scenes.box_room(width=64, height=64, base=2, sphere_level=4)
    rd = swr.render_desc(desc["render"])
    ord_ = swr.render_desc(desc["render"], orc.RenderDesc)
    rays_g, xy_g = swr.raygen(ctx, rd, 0, 64 * 64 * 4)
    rays_o, xy_o = orc.raygen(ord_, 0, 64 * 64 * 4)
    np.testing.assert_allclose(rays_g, rays_o, rtol=2e-6, atol=2e-6)
    np.testing.assert_array_equal(xy_g, xy_o)
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    check_against_oracle(sc.intersect(rays_o), oc.intersect(rays_o))


def test_edge_cases(ctx):
    # axis-parallel rays, rays starting on a triangle, degenerate triangles, empty scene
    pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [2, 2, 2, 0], [2, 2, 2, 0], [2, 2, 2, 0]], np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    desc = {"shaders": [{"type": 0, "colour": [1, 1, 1]}], "lights": [], "environment": (0, 0, 0),
            "meshes": [{"pos4": pos, "idx": idx, "uv": None, "shader": 0}]}
    sc, oc = swr.Scene(ctx, desc), orc.Scene(desc)
    rays = np.array([
        [0.2, 0.2, 1.0, 0.0, 0.0, 0.0, -1.0, 1e5],   # straight down the z axis onto the triangle
        [0.2, 0.2, -1.0, 0.0, 0.0, 0.0, 1.0, 1e5],   # from behind (no culling)
        [0.2, 0.2, 0.0, 0.001, 0.0, 0.0, 1.0, 1e5],  # starting on the triangle: tmin excludes it
        [5.0, 0.2, 0.0, 0.0, -1.0, 0.0, 0.0, 1e5],   # in the triangle's plane
        [0.2, 0.2, 1.0, 0.0, 0.0, 0.0, -1.0, 0.5],   # tmax short of the triangle
        [2.0, 2.0, 5.0, 0.0, 0.0, 0.0, -1.0, 1e5],   # through the degenerate triangle
    ], np.float32)
    g = sc.intersect(rays)
    o = oc.intersect(rays)
    np.testing.assert_array_equal(g[:, 3], o[:, 3])
    assert list(g[:, 3]) == [0, 0, MISS, MISS, MISS, MISS]
    np.testing.assert_array_equal(g[:2], o[:2])
    empty = swr.Scene(ctx, {"shaders": [], "meshes": [], "lights": [], "environment": (0, 0, 0)})
    h = empty.intersect(rays)
    assert np.all(h[:, 3] == MISS)
    assert np.array_equal(h[:, 0].view(np.float32), rays[:, 7])


def test_bvh_invalid_inputs(ctx):
    sc = swr.Scene(ctx)
    with pytest.raises(swr.SwrError):
        sc.add_mesh(np.zeros((3, 4), np.float32), np.array([[0, 1, 3]], np.uint32), None, 0)
    sc.add_shader({"type": 0, "colour": [1, 1, 1]})
    sc.add_mesh(np.zeros((3, 4), np.float32), np.array([[0, 1, 2]], np.uint32), None, 5)
    with pytest.raises(swr.SwrError):
        sc.commit()  # shader 5 does not exist


def test_terrain_full_size(ctx):
    """2,097,152-triangle terrain: BVH invariants and a 50k-ray sample against the oracle."""
    desc = scenes.terrain()
    sc = swr.Scene(ctx, desc)
    assert sc.num_triangles == 2 * 1024 * 1024 + 2
    info = sc.bvh_info()
    assert info["valid"], info
    oc = orc.Scene(desc)
    rng = np.random.default_rng(3)
    rays = random_rays(rng, 50000, lo=-1.0, hi=1.0)
    rays[:, 1] = rng.uniform(0.05, 0.5, 50000)
    check_against_oracle(sc.intersect(rays), oc.intersect(rays))
    assert np.array_equal(sc.occluded(rays), oc.occluded(rays))


@pytest.mark.parametrize("ring", [8, 16, 32, 64])
def test_stack_ring_sizes_agree(ctx, ring):
    """The traversal keeps the top of its stack in an LDS ring and spills the rest to device
    memory. A 30000-triangle soup needs more entries than the small rings hold, so rings of 8 and
    16 spill on many rays; every ring size must give the oracle's hits bit for bit."""
    desc = scenes.random_soup(11, 30000)
    sc = swr.Scene(ctx, desc)
    info = sc.bvh_info()
    assert 3 * info["max_depth"] + 4 > 16, info  # the small rings do spill
    oc = orc.Scene(desc)
    rays = random_rays(np.random.default_rng(11), 20000)
    before = ctx.tuning()
    try:
        ctx.set_tuning(16, 32, 0, 64)
        g64 = sc.intersect(rays)
        ctx.set_tuning(16, 32, 0, ring)
        g = sc.intersect(rays)
        assert np.array_equal(sc.occluded(rays), oc.occluded(rays))
    finally:
        ctx.set_tuning(**before)
    assert np.array_equal(g, g64)
    check_against_oracle(g, oc.intersect(rays))
