"""GPU parity of the radix sorts (onesweep, reduce-then-scan, bucket sort): bit-exact stable permutations vs std::stable_sort."""
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


class _BucketSort:
    """The context's sort switched to the bucket sort (partition of the key range into 2048
    buckets, each sorted in LDS; sorts of more than 16 key bits) for the duration of a test."""

    def __init__(self, ctx):
        self.ctx = ctx

    def __enter__(self):
        self.before = self.ctx.tuning()
        self.ctx.set_tuning(**dict(self.before, sort_algorithm=2))
        return self

    def __exit__(self, *exc):
        self.ctx.set_tuning(**self.before)


@pytest.mark.parametrize("n", [1, 2, 100, 4097, 16383, 16384, 16385, 300001, 1 << 20])
def test_bucket_sort_random_u32(ctx, n):
    rng = np.random.default_rng(100 + n)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    with _BucketSort(ctx):
        k, v = _sort(ctx, keys)
    perm = orc.stable_sort_perm(keys)
    assert np.array_equal(v, perm)
    assert np.array_equal(k, keys[perm])


@pytest.mark.parametrize("bits", [17, 21, 24, 28, 29, 31])
def test_bucket_sort_key_widths(ctx, bits):
    rng = np.random.default_rng(200 + bits)
    n = 400000
    keys = rng.integers(0, 1 << bits, n).astype(np.uint32)
    with _BucketSort(ctx):
        k, v = _sort(ctx, keys, bits)
    assert np.array_equal(v, orc.stable_sort_perm(keys))


@pytest.mark.parametrize("case", ["equal", "two", "few", "narrow", "skewed", "bimodal", "hit_keys", "descending"])
def test_bucket_sort_distributions(ctx, case):
    """Ties, narrow ranges and clustered keys: buckets of one key value, buckets beyond the LDS
    (sorted through device memory by their workgroup, with none, odd and even pass counts),
    empty buckets, and the hit-key layout ((mesh + 1) << 21 | face)."""
    rng = np.random.default_rng(sum(map(ord, case)))
    n = 300000
    if case == "equal":
        keys = np.full(n, 0xDEADBEEF, np.uint32)
    elif case == "two":
        keys = (rng.integers(0, 2, n) * 0x01010101).astype(np.uint32)
    elif case == "few":
        keys = (rng.integers(0, 17, n) * 0x00F0F0F1).astype(np.uint32)
    elif case == "narrow":
        keys = (0x12340000 + rng.integers(0, 3000, n)).astype(np.uint32)
    elif case == "skewed":
        # 90% of the keys within 2^20 of one value: one bucket far beyond the LDS, three passes
        keys = np.where(rng.random(n) < 0.9, 0x40000000 + rng.integers(0, 1 << 20, n),
                        rng.integers(0, 2**32, n, dtype=np.uint64)).astype(np.uint32)
    elif case == "bimodal":
        # two dense clusters at either end of the range: most buckets empty, two oversized with 16-bit spans
        keys = np.where(rng.random(n) < 0.5, rng.integers(0, 1 << 16, n),
                        0xFFFF0000 + rng.integers(0, 1 << 16, n)).astype(np.uint32)
    elif case == "hit_keys":
        keys = (((rng.integers(0, 64, n) + 1) << 21) | rng.integers(0, 1 << 21, n)).astype(np.uint32)
    else:
        keys = (np.arange(n, dtype=np.uint64)[::-1] * 9973 % 2**32).astype(np.uint32)
    with _BucketSort(ctx):
        k, v = _sort(ctx, keys)
    perm = orc.stable_sort_perm(keys)
    assert np.array_equal(v, perm)
    assert np.array_equal(k, keys[perm])


def test_bucket_sort_carries_values(ctx):
    """Explicit values (not the iota permutation) travel with their keys."""
    rng = np.random.default_rng(31)
    n = 250000
    keys = rng.integers(0, 1 << 30, n).astype(np.uint32)
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    dk, dv = ctx.upload(keys), ctx.upload(vals)
    with _BucketSort(ctx):
        ctx.sort_pairs(dk, dv, n, 30, iota=False)
    perm = orc.stable_sort_perm(keys)
    assert np.array_equal(dk.download(np.uint32, (n,)), keys[perm])
    assert np.array_equal(dv.download(np.uint32, (n,)), vals[perm])
    dk.free()
    dv.free()


@pytest.mark.parametrize("offset_words", [1, 3])
def test_bucket_sort_keys_not_16_byte_aligned(ctx, offset_words):
    import swr
    n = 70001
    rng = np.random.default_rng(40 + offset_words)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    buf_k = ctx.alloc(4 * (n + 4))
    buf_v = ctx.alloc(4 * (n + 4))
    dk = swr.DeviceArray(ctx, buf_k.ptr + 4 * offset_words, 4 * n)
    dv = swr.DeviceArray(ctx, buf_v.ptr + 4 * offset_words, 4 * n)
    dk.upload_from(keys)
    with _BucketSort(ctx):
        ctx.sort_pairs(dk, dv, n, 32, iota=True)
    v = dv.download(np.uint32, (n,))
    dk.ptr = dv.ptr = 0  # views: the allocations are freed below
    buf_k.free()
    buf_v.free()
    assert np.array_equal(v, orc.stable_sort_perm(keys))


def test_bucket_sort_full_size(ctx):
    """2^24 octant + Morton ray keys and 2^24 hit keys (the sort/traversal bench's keys) through
    the bucket sort, bit for bit against the oracle's stable sort of the same keys."""
    from swr import scenes
    n = 1 << 24
    rays = scenes.sort_bench_rays(n)
    d_rays = ctx.upload(rays)
    d_keys = ctx.alloc(4 * n)
    ctx.ray_keys(d_rays, n, (0.0, 0.0, 0.0), (1024.0, 1024.0, 512.0), d_keys)
    d_rays.free()
    keys = d_keys.download(np.uint32, (n,))
    rng = np.random.default_rng(0x5EED0007)
    hkeys = (((rng.integers(0, 64, n) + 1) << 21) | rng.integers(0, 1 << 21, n)).astype(np.uint32)
    dv = ctx.alloc(4 * n)
    with _BucketSort(ctx):
        for kk, bits in ((keys, 32), (hkeys, 28)):
            d_keys.upload_from(kk)
            ctx.sort_pairs(d_keys, dv, n, bits, iota=True)
            v = dv.download(np.uint32, (n,))
            assert np.array_equal(v, orc.stable_sort_perm(kk))
            assert np.array_equal(d_keys.download(np.uint32, (n,)), kk[v])
    d_keys.free()
    dv.free()
