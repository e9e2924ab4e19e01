"""The sort without allocation (swr_sort_pairs_u32_noalloc) against the allocating sort
(swr_sort_pairs_u32), through the C ABI: identical keys and values on two calls in a row (the
second smaller, so stale scratch contents from the first would show), on keys with ties and a
key width below 32 bits, with caller values and with the iota permutation; the device memory the
context accounts for does not change; sizes above the maximum are refused."""
import numpy as np
import pytest

import swr

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 20


def _sort(ctx, fn, keys, vals, bits, iota):
    dk, dv = ctx.upload(keys), ctx.upload(vals)
    fn(dk, dv, len(keys), bits, iota)
    out = dk.download(np.uint32, (len(keys),)), dv.download(np.uint32, (len(keys),))
    dk.free()
    dv.free()
    return out


@pytest.mark.parametrize("bits,iota", [(32, True), (20, False), (8, True)])
def test_noalloc_equals_allocating_sort_twice(ctx, bits, iota):
    rng = np.random.default_rng(bits)
    for n in (MAX_PAIRS, 300001):  # the maximum, then a smaller call on the same scratch
        keys = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)  # below 2^key_bits
        keys[::7] = keys[0]  # ties
        vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        ref = _sort(ctx, ctx.sort_pairs, keys, vals, bits, iota)
        used = ctx.bytes_in_use()
        got = _sort(ctx, ctx.sort_pairs_noalloc, keys, vals, bits, iota)
        assert ctx.bytes_in_use() == used
        np.testing.assert_array_equal(got[0], ref[0], err_msg=f"keys, n={n}")
        np.testing.assert_array_equal(got[1], ref[1], err_msg=f"values, n={n}")
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(got[1], order.astype(np.uint32) if iota else vals[order])


def test_noalloc_refuses_more_than_its_maximum(ctx):
    dk, dv = ctx.alloc(4 * (MAX_PAIRS + 1)), ctx.alloc(4 * (MAX_PAIRS + 1))
    with pytest.raises(swr.SwrError) as e:
        ctx.sort_pairs_noalloc(dk, dv, MAX_PAIRS + 1, 32)
    assert e.value.code == swr.SWR_E_INVALID and "maximum" in str(e.value)
    ctx.sort_pairs_noalloc(dk, dv, 0, 32)  # empty: nothing to do
    dk.free()
    dv.free()
