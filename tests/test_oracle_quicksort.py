"""The oracle's quicksort (include/core/operators/quicksort.h:212-244), the
reference's default sort protocol, pinned by the reference's own test
(tests/test_sorting.cpp:56-96: a shuffled 0..n-1 with one -1, sorted ASC
and DESC, the permutation applied to a copy) and by the plaintext stable
sort in signed order (the index padding makes it stable) with its ranks,
for 32-bit keys (padded to 64 bits) and 64-bit keys (padded to 128), at
sizes that leave one-, two- and many-element segments; every reserved
permutation is consumed."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", [16, 1000, 4099])
def test_reference_quicksort(keys, n):
    v = np.arange(n, dtype=np.int64)
    v[n // 2] = -1
    c = O.Ctx(keys)
    nums = c.secret_share(v.astype(np.uint32), 4)
    c.shuffle(nums)
    shuffled = c.open(nums)
    assert not np.array_equal(shuffled.view(np.int32), v.astype(np.int32))
    b1, b2, b1r, b2r = (c.secret_share(shuffled, 4) for _ in range(4))
    perm = c.quicksort(b1)
    c.apply_elementwise(b2, perm, x_arith=False, perm_arith=True)
    rperm = c.quicksort(b1r, desc=True)
    c.apply_elementwise(b2r, rperm, x_arith=False, perm_arith=True)
    want = np.sort(v).astype(np.int32)
    np.testing.assert_array_equal(c.open(b1).view(np.int32), want)
    np.testing.assert_array_equal(c.open(b2).view(np.int32), want)
    np.testing.assert_array_equal(c.open(b1r).view(np.int32), want[::-1])
    np.testing.assert_array_equal(c.open(b2r).view(np.int32), want[::-1])
    assert c.queue_size() == 0


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 300, 5000])
def test_quicksort_stable_and_ranks(keys, n, desc, width):
    rng = np.random.default_rng(n)
    ut, st = (np.uint32, np.int32) if width == 4 else (np.uint64, np.int64)
    x = rng.integers(0, np.iinfo(ut).max, n, dtype=ut, endpoint=True)
    if n > 4:
        x.view(st)[0] = np.iinfo(st).min  # INT_MIN
        x.view(st)[1] = np.iinfo(st).max  # INT_MAX
        x[2:n // 3] = x[n // 3]  # a long run of equal keys: order decided by the index
    c = O.Ctx(keys)
    v = c.secret_share(x, width)
    p = c.quicksort(v, desc=desc)
    sx = x.view(st)
    order = np.argsort(~sx if desc else sx, kind="stable")
    np.testing.assert_array_equal(c.open(v).view(st), sx[order])
    ranks = np.empty(n, np.int64)
    ranks[order] = np.arange(n)
    np.testing.assert_array_equal(c.open(p, arith=True).view(np.int32), ranks)
    assert c.queue_size() == 0


def test_reference_quicksort_64_bit(keys):
    # tests/test_sorting.cpp:56-96 on int64_t
    n = 1000
    v = np.arange(n, dtype=np.int64) * (1 << 33) - (1 << 40)
    v[n // 2] = -1
    c = O.Ctx(keys)
    nums = c.secret_share(v.view(np.uint64), 8)
    c.shuffle(nums)
    shuffled = c.open(nums)
    b1, b2 = (c.secret_share(shuffled, 8) for _ in range(2))
    perm = c.quicksort(b1)
    c.apply_elementwise(b2, perm, x_arith=False, perm_arith=True)
    want = np.sort(v)
    np.testing.assert_array_equal(c.open(b1).view(np.int64), want)
    np.testing.assert_array_equal(c.open(b2).view(np.int64), want)
    assert c.queue_size() == 0
