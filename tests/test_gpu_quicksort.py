"""GPU quicksort (orqg_quicksort, include/core/operators/quicksort.h:212-244)
against the oracle: shares of the sorted keys and of the sorting
permutation, PRG nonces and the message log, for 32- and 64-bit keys,
across sizes that exercise one-element, two-element and long segments, both
orders; the reference's own test flow (tests/test_sorting.cpp:56-96); and
2^20-row sorts against the plaintext stable sort and its ranks."""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_nonces, _same_shares

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()
    return orq_amd


def _ranks(order):
    r = np.empty(order.size, dtype=np.int64)
    r[order] = np.arange(order.size)
    return r


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 255, 1000, 4097, 70001])
def test_quicksort_matches_oracle(O, G, keys, n, desc, width):
    rng = np.random.default_rng(n + 17)
    ut, st = (np.uint32, np.int32) if width == 4 else (np.uint64, np.int64)
    x = rng.integers(0, np.iinfo(ut).max, n, dtype=ut, endpoint=True)
    if n > 8:
        x.view(st)[0], x.view(st)[1] = np.iinfo(st).min, np.iinfo(st).max
        x[2:n // 4] = x[n // 4]
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ov, gv = oc.secret_share(x, width), gc.secret_share(x, width)
    op, gp = oc.quicksort(ov, desc=desc), gc.quicksort(gv, desc=desc)
    _same_shares(ov, gv)
    _same_shares(op, gp)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    assert gc.queue_size() == 0
    sx = x.view(st)
    order = np.argsort(~sx if desc else sx, kind="stable")
    np.testing.assert_array_equal(gc.open(gv).view(st), sx[order])
    np.testing.assert_array_equal(gc.open(gp).view(np.int32), _ranks(order))


def test_reference_quicksort_flow(O, G, keys):
    # tests/test_sorting.cpp:56-96
    n = 2000
    v = np.arange(n, dtype=np.int64)
    v[n // 2] = -1
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    on, gn = oc.secret_share(v.astype(np.uint32), 4), gc.secret_share(v.astype(np.uint32), 4)
    oc.shuffle(on)
    gc.shuffle(gn)
    _same_shares(on, gn)
    shuffled = gc.open(gn)
    ob = [oc.secret_share(shuffled, 4) for _ in range(4)]
    gb = [gc.secret_share(shuffled, 4) for _ in range(4)]
    op, gp = oc.quicksort(ob[0]), gc.quicksort(gb[0])
    oc.apply_elementwise(ob[1], op, x_arith=False, perm_arith=True)
    gc.apply_elementwise(gb[1], gp)
    opr, gpr = oc.quicksort(ob[2], desc=True), gc.quicksort(gb[2], desc=True)
    oc.apply_elementwise(ob[3], opr, x_arith=False, perm_arith=True)
    gc.apply_elementwise(gb[3], gpr)
    for o, g in zip(ob, gb):
        _same_shares(o, g)
    want = np.sort(v).astype(np.int32)
    for i, w in enumerate([want, want, want[::-1], want[::-1]]):
        np.testing.assert_array_equal(gc.open(gb[i]).view(np.int32), w)


@pytest.mark.parametrize("width", [4, 8])
def test_quicksort_2p20_plaintext(G, keys, width):
    n = 1 << 20
    rng = np.random.default_rng(3)
    st, ut = (np.int32, np.uint32) if width == 4 else (np.int64, np.uint64)
    x = rng.integers(-2**20, 2**20, n).astype(st)
    if width == 8:
        x = x * (1 << 30) + 7  # keys past 32 bits
    gc = G.Ctx(keys)
    gv = gc.secret_share(x.view(ut), width)
    gp = gc.quicksort(gv)
    order = np.argsort(x, kind="stable")
    np.testing.assert_array_equal(gc.open(gv).view(st), x[order])
    np.testing.assert_array_equal(gc.open(gp).view(np.int32), _ranks(order))
    assert gc.queue_size() == 0
