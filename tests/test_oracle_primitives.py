"""Pins of the oracle's primitives: Philox known answers, Morton interleave, stable sort."""
import json
import os

import numpy as np
import pytest

import oracle_ffi as orc


# Philox-4x32-10 known-answer vectors (Random123 kat_vectors: counter, key -> output), kept as a
# fixture under tests/golden.
def _load_kat():
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "philox4x32_10_kat.json")) as f:
        doc = json.load(f)
    h = lambda xs: tuple(int(x, 16) for x in xs)
    return [(h(v["ctr"]), h(v["key"]), h(v["out"])) for v in doc["vectors"]]


KAT = _load_kat()


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(x) for x in orc.philox(ctr, key)) == want


def test_u01_range_and_resolution():
    L = orc.lib()
    assert L.oracle_u01(0) == 0.0
    assert L.oracle_u01(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -24)
    assert L.oracle_u01(0x100) == np.float32(2.0 ** -24)


def _morton_brute(x, y, z):
    # 30-bit interleave of (x, y, 2z), x in the top bit of each triple, shifted down by one.
    z2 = z << 1
    code = 0
    for i in range(10):
        code |= ((x >> i) & 1) << (3 * i + 2)
        code |= ((y >> i) & 1) << (3 * i + 1)
        code |= ((z2 >> i) & 1) << (3 * i)
    return code >> 1


def test_morton29_matches_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, y, z = int(rng.integers(1024)), int(rng.integers(1024)), int(rng.integers(512))
        assert orc.morton29(x, y, z) == _morton_brute(x, y, z)
    assert orc.morton29(1023, 1023, 511) == (1 << 29) - 1
    assert orc.morton29(0, 0, 0) == 0
    # the highest bit belongs to x, then y, then z
    assert orc.morton29(512, 0, 0) == 1 << 28
    assert orc.morton29(0, 512, 0) == 1 << 27
    assert orc.morton29(0, 0, 256) == 1 << 26
    # the lowest bits are x0 y0 with z0 above them (z carries one bit fewer)
    assert (orc.morton29(0, 1, 0), orc.morton29(1, 0, 0), orc.morton29(0, 0, 1)) == (1, 2, 4)


def test_ray_keys_octant_and_quantisation():
    lo = np.zeros(3, np.float32)
    scale = np.array([1024.0, 1024.0, 512.0], np.float32)
    rays = np.zeros((8, 8), np.float32)
    for o in range(8):
        rays[o, 4:7] = [(-1.0 if o & 1 else 1.0), (-1.0 if o & 2 else 1.0), (-1.0 if o & 4 else 1.0)]
    rays[:, 0:3] = 0.5
    keys = orc.ray_keys(rays, lo, scale)
    assert [int(k) >> 29 for k in keys] == list(range(8))
    m = orc.morton29(512, 512, 256)
    assert all((int(k) & ((1 << 29) - 1)) == m for k in keys)
    # clamping outside the box
    rays2 = rays[:2].copy()
    rays2[0, 0:3] = [-5.0, -5.0, -5.0]
    rays2[1, 0:3] = [5.0, 5.0, 5.0]
    k2 = orc.ray_keys(rays2, lo, scale)
    assert int(k2[0]) & ((1 << 29) - 1) == 0
    assert int(k2[1]) & ((1 << 29) - 1) == (1 << 29) - 1


def test_hit_keys_reproduce_comparehit_order():
    # CompareHit: (geomID, primID) as signed ints, misses (-1) first.
    rng = np.random.default_rng(2)
    first = np.array([0, 10, 25, 40], np.uint32)   # triangle offsets of 4 meshes (40 triangles)
    geom = rng.integers(-1, 4, 500)
    prim_local = np.array([rng.integers(0, 10 if g < 2 else 15) if g >= 0 else -1 for g in geom])
    prim_global = np.where(geom >= 0, first[np.maximum(geom, 0)] + np.maximum(prim_local, 0), 0xFFFFFFFF)
    keys = orc.hit_keys(prim_global.astype(np.uint32))
    order_keys = np.argsort(keys, kind="stable")
    order_ref = sorted(range(500), key=lambda i: (geom[i], prim_local[i]))
    assert [(geom[i], prim_local[i]) for i in order_keys] == [(geom[i], prim_local[i]) for i in order_ref]


def test_stable_sort_perm_is_stable_and_sorted():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 16, 10000).astype(np.uint32)
    perm = orc.stable_sort_perm(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    # planted mistake: a non-stable order must be told apart
    bad = perm.copy()
    i = int(np.nonzero(keys[perm[:-1]] == keys[perm[1:]])[0][0])
    bad[i], bad[i + 1] = bad[i + 1], bad[i]
    assert not np.array_equal(bad, np.argsort(keys, kind="stable"))
