"""The front end's device radix sort (fmm2d_c_translation_amd/csrc/rsort.h)
against numpy's stable argsort of the same key bits: the sorted keys and the
permutation must match exactly (stability included), for 32- and 64-bit
keys, pairs and keys only, sizes around the 2048-key tile and bit ranges that
end inside a byte.  It orders the points for the tree (pts_tree2d.c sorts
points into boxes) and the refinement, closure and work lists."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev_sort(eng, keys, begin, end, with_vals=True):
    keys = np.ascontiguousarray(keys)
    n = keys.size
    out = np.zeros(max(n, 1), keys.dtype)
    vals = np.zeros(max(n, 1), np.int32) if with_vals else None
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = eng.fmm2dhip_test_sort(P(keys) if n else None, n, keys.itemsize, begin, end, P(out),
                                P(vals) if with_vals else None)
    return rc, out[:n], None if vals is None else vals[:n]


def check(eng, keys, begin, end, with_vals=True):
    rc, out, vals = dev_sort(eng, keys, begin, end, with_vals)
    assert rc == 0, rc
    one = keys.dtype.type(1)
    width = end - begin
    mask = keys.dtype.type(np.iinfo(keys.dtype).max) if width == 8 * keys.itemsize else \
        keys.dtype.type((int(one) << width) - 1)
    digit = (keys >> keys.dtype.type(begin)) & mask
    order = np.argsort(digit, kind="stable")
    np.testing.assert_array_equal(out, keys[order])
    if with_vals:
        np.testing.assert_array_equal(vals, order.astype(np.int32))


@pytest.mark.parametrize("n", [1, 2, 255, 2047, 2048, 2049, 6000, 100_003])
@pytest.mark.parametrize("dtype,begin,end", [(np.uint64, 0, 59), (np.uint64, 20, 59),
                                             (np.uint32, 0, 24), (np.uint32, 0, 15),
                                             (np.uint64, 0, 64), (np.uint32, 3, 4)])
def test_sort_matches_stable_argsort(eng, n, dtype, begin, end):
    rng = np.random.default_rng(n * 131 + begin * 7 + end)
    bits = 8 * np.dtype(dtype).itemsize
    hi = rng.integers(0, 2**31, n, dtype=np.uint64)
    lo = rng.integers(0, 2**33, n, dtype=np.uint64)
    keys = ((hi << np.uint64(33)) | lo)
    if bits == 32:
        keys = keys & np.uint64(0xFFFFFFFF)
    keys = keys.astype(dtype)
    check(eng, keys, begin, end)


def test_sort_many_duplicates_keys_only(eng):
    """Few distinct digits (every tile shares each digit) and a keys-only sort."""
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 4, 300_000).astype(np.uint64) << np.uint64(40)
    keys |= rng.integers(0, 2**20, keys.size).astype(np.uint64)
    check(eng, keys, 40, 48)
    check(eng, keys, 0, 48, with_vals=False)


def test_sort_large_front_end_shape(eng):
    """Two million 59-bit keys (the point sort of a million sources and a
    million targets, the target flag at bit 58), sorted on bits [20, 59)."""
    rng = np.random.default_rng(11)
    n = 2_000_000
    keys = rng.integers(0, 2**58, n, dtype=np.uint64)
    keys[n // 2:] |= np.uint64(1) << np.uint64(58)
    check(eng, keys, 20, 59)


def test_sort_presorted_and_reversed(eng):
    keys = np.arange(50_000, dtype=np.uint32)
    check(eng, keys, 0, 16)
    check(eng, keys[::-1].copy(), 0, 16)


def test_sort_bad_arguments(eng):
    keys = np.arange(10, dtype=np.uint32)
    assert dev_sort(eng, keys, 0, 33)[0] == 102   # past the key width
    assert dev_sort(eng, keys, 5, 5)[0] == 102    # empty bit range
    assert dev_sort(eng, keys.astype(np.uint16), 0, 8)[0] == 102
    rc, out, vals = dev_sort(eng, np.zeros(0, np.uint64), 0, 8)
    assert rc == 0 and out.size == 0
