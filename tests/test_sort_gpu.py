"""GPU parity of the onesweep radix sort: bit-exact stable permutations vs std::stable_sort."""
import numpy as np
import pytest

import oracle_ffi as orc

pytestmark = pytest.mark.gpu


def _sort(ctx, keys, bits=32):
    n = keys.shape[0]
    dk = ctx.upload(keys.astype(np.uint32))
    dv = ctx.alloc(max(4 * n, 4))
    ctx.sort_pairs(dk, dv, n, bits, iota=True)
    out_k = dk.download(np.uint32, (n,))
    out_v = dv.download(np.uint32, (n,))
    dk.free()
    dv.free()
    return out_k, out_v


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 12289, 100000, 1 << 20])
def test_sort_random_u32(ctx, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    k, v = _sort(ctx, keys)
    perm = orc.stable_sort_perm(keys)
    assert np.array_equal(v, perm)
    assert np.array_equal(k, keys[perm])


@pytest.mark.parametrize("distinct", [1, 2, 3, 17, 256, 4097])
def test_sort_many_ties_is_stable(ctx, distinct):
    rng = np.random.default_rng(distinct)
    n = 300000
    keys = (rng.integers(0, distinct, n) * 0x01010101).astype(np.uint32)
    k, v = _sort(ctx, keys)
    assert np.array_equal(v, orc.stable_sort_perm(keys))


@pytest.mark.parametrize("bits", [1, 7, 8, 9, 21, 28, 29])
def test_sort_partial_key_bits(ctx, bits):
    rng = np.random.default_rng(bits)
    n = 200000
    keys = rng.integers(0, 1 << bits, n).astype(np.uint32)
    k, v = _sort(ctx, keys, bits)
    assert np.array_equal(v, orc.stable_sort_perm(keys))


def test_sort_presorted_and_reversed(ctx):
    n = 150000
    asc = np.arange(n, dtype=np.uint32) * 7
    k, v = _sort(ctx, asc)
    assert np.array_equal(v, np.arange(n))
    desc = asc[::-1].copy()
    k, v = _sort(ctx, desc)
    assert np.array_equal(v, orc.stable_sort_perm(desc))


def test_sort_full_size_properties(ctx):
    """16M ray keys (octant + Morton, as the sort/traversal bench draws them): sorted, a
    permutation, and stable, checked by properties that do not need the CPU sort."""
    n = 1 << 24
    rng = np.random.default_rng(7)
    o = rng.random((n, 3), dtype=np.float32)
    z = 1.0 - 2.0 * rng.random(n, dtype=np.float32)
    phi = 2.0 * np.pi * rng.random(n, dtype=np.float32)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4] = r * np.cos(phi)
    rays[:, 5] = r * np.sin(phi)
    rays[:, 6] = z
    lo = np.zeros(3, np.float32)
    scale = np.array([1024.0, 1024.0, 512.0], np.float32)
    d_rays = ctx.upload(rays)
    d_keys = ctx.alloc(4 * n)
    ctx.ray_keys(d_rays, n, lo, scale, d_keys)
    keys = d_keys.download(np.uint32, (n,))
    d_rays.free()
    # keys parity on a sample against the oracle
    sel = rng.integers(0, n, 100000)
    assert np.array_equal(keys[sel], orc.ray_keys(rays[sel], lo, scale))
    dv = ctx.alloc(4 * n)
    ctx.sort_pairs(d_keys, dv, n, 32, iota=True)
    k = d_keys.download(np.uint32, (n,))
    v = dv.download(np.uint32, (n,))
    d_keys.free()
    dv.free()
    assert np.all(k[1:] >= k[:-1])
    assert np.array_equal(np.bincount(v, minlength=n), np.ones(n, np.int64))
    assert np.array_equal(keys[v], k)
    ties = k[1:] == k[:-1]
    assert np.all(v[1:][ties] > v[:-1][ties])


def test_mesh_face_keys_and_hit_keys(ctx):
    rng = np.random.default_rng(11)
    n = 100000
    mesh = rng.integers(0, 64, n).astype(np.uint32)
    face = rng.integers(0, 1 << 21, n).astype(np.uint32)
    dm, df, dk = ctx.upload(mesh), ctx.upload(face), ctx.alloc(4 * n)
    ctx.mesh_face_keys(dm, df, n, dk)
    assert np.array_equal(dk.download(np.uint32, (n,)), ((mesh + 1) << 21) | face)
    hits = np.zeros((n, 4), np.uint32)
    hits[:, 3] = np.where(rng.random(n) < 0.1, 0xFFFFFFFF, rng.integers(0, 1 << 21, n)).astype(np.uint32)
    dh = ctx.upload(hits)
    ctx.hit_keys(dh, n, dk)
    assert np.array_equal(dk.download(np.uint32, (n,)), orc.hit_keys(hits[:, 3]))


@pytest.mark.parametrize("words", [1, 2, 3, 4, 8, 12])
def test_gather(ctx, words):
    rng = np.random.default_rng(words)
    n = 50000
    src = rng.integers(0, 2**32, (n, words), dtype=np.uint64).astype(np.uint32)
    perm = rng.permutation(n).astype(np.uint32)
    ds, dp, do = ctx.upload(src), ctx.upload(perm), ctx.alloc(src.nbytes)
    ctx.gather(ds, dp, n, words, do)
    assert np.array_equal(do.download(np.uint32, (n, words)), src[perm])


@pytest.mark.parametrize("alg", [0, 1])
@pytest.mark.parametrize("kpt", [8, 12, 20, 24])
def test_sort_every_tile_size(ctx, kpt, alg):
    """Both pass structures (onesweep, reduce-then-scan) at every tile size, on sizes that leave
    partial tiles and on skewed keys."""
    before = ctx.tuning()
    try:
        ctx.set_tuning(kpt, 32, 0, 0, alg)
        rng = np.random.default_rng(kpt)
        for n in (1, 5000, 300001):
            keys = rng.integers(0, 1 << 12, n).astype(np.uint32)
            k, v = _sort(ctx, keys, 12)
            assert np.array_equal(v, orc.stable_sort_perm(keys))
        keys = (rng.integers(0, 3, 200000) << 24 | rng.integers(0, 1 << 24, 200000)).astype(np.uint32)
        k, v = _sort(ctx, keys, 32)
        assert np.array_equal(v, orc.stable_sort_perm(keys))
    finally:
        ctx.set_tuning(**before)


@pytest.mark.parametrize("offset_words", [1, 2, 3])
def test_sort_keys_not_16_byte_aligned(ctx, offset_words):
    """Keys and values that start 4, 8 or 12 bytes into an allocation (the histogram then reads
    keys one at a time instead of 16 bytes at a time)."""
    import swr
    n = 50001
    rng = np.random.default_rng(offset_words)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    buf_k = ctx.alloc(4 * (n + 4))
    buf_v = ctx.alloc(4 * (n + 4))
    dk = swr.DeviceArray(ctx, buf_k.ptr + 4 * offset_words, 4 * n)
    dv = swr.DeviceArray(ctx, buf_v.ptr + 4 * offset_words, 4 * n)
    dk.upload_from(keys)
    ctx.sort_pairs(dk, dv, n, 32, iota=True)
    v = dv.download(np.uint32, (n,))
    k = dk.download(np.uint32, (n,))
    dk.ptr = dv.ptr = 0  # views: the allocations are freed below
    buf_k.free()
    buf_v.free()
    perm = orc.stable_sort_perm(keys)
    assert np.array_equal(v, perm)
    assert np.array_equal(k, keys[perm])
