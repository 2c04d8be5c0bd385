"""More than one worker thread (the reference's RunTime with T threads, each
with its own keys): a T-thread GPU context against the T-thread oracle on the
same keys, share for share, every thread's nonces and the message log, for
the operations that support it; the others refuse."""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_shares

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


def _keys(t):
    from orq_amd import keys_from_seed

    return keys_from_seed(0x7E5D0000 + t)


def _pair(O, G, T, batch_size=-12):
    extra = [_keys(t) for t in range(1, T)]
    return (O.Ctx(_keys(0), batch_size, threads=T, thread_keys=extra),
            G.Ctx(_keys(0), batch_size, threads=T, thread_keys=extra))


def _same_nonces(oc, gc, T):
    for t in range(T):
        for p in range(3):
            assert oc.nonce(p, 0, thread=t) == gc.nonce(p, thread=t), (t, "local", p)
            assert oc.nonce(p, 2, thread=t) == gc.nonce(3 + p, thread=t), (t, "pair", p)
        for g in range(3):
            member = 0 if g < 2 else 1
            assert oc.nonce(member, 3 + g, thread=t) == gc.nonce(6 + g, thread=t), (t, "group", g)


@pytest.mark.parametrize("T,n", [(2, 5000), (3, 5000), (4, 300), (3, 70001)])
@pytest.mark.parametrize("width", [4, 8])
def test_primitives_with_threads(O, G, T, n, width):
    rng = np.random.default_rng(T * n + width)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**(8 * width), n, dtype=np.uint64).astype(dt)
    y = rng.integers(0, 2**(8 * width), n, dtype=np.uint64).astype(dt)
    oc, gc = _pair(O, G, T)
    ox, gx = oc.secret_share(x, width), gc.secret_share(x, width)
    oy, gy = oc.secret_share(y, width), gc.secret_share(y, width)
    _same_shares(ox, gx)
    _same_shares(oc.and_b(ox, oy), gc.and_b(gx, gy))
    _same_shares(oc.b2a_bit(ox), gc.b2a_bit(gx))
    oa, ga = oc.secret_share(x, width, arith=True), gc.secret_share(x, width, arith=True)
    ob, gb = oc.secret_share(y, width, arith=True), gc.secret_share(y, width, arith=True)
    _same_shares(oc.mul_a(oa, ob), gc.mul_a(ga, gb))
    for g in range(3):
        oc.reshare(ox, g)
        gc.reshare(gx, g)
    _same_shares(ox, gx)
    _same_nonces(oc, gc, T)
    _same_log(oc, gc)
    np.testing.assert_array_equal(gc.open(gx), x)


def test_reserve_deals_permutations_to_threads(O, G):
    n, count, T = 300, 7, 3
    oc, gc = _pair(O, G, T)
    oc.reserve(n, count)
    gc.reserve(n, count)
    for k in range(count):
        for g in range(3):
            members = [p for p in range(3) if p != 2 - g]
            np.testing.assert_array_equal(gc.peek_perm(k, g, n), oc.peek_perm(k, members[0], g, n),
                                          err_msg=f"perm {k} group {g}")
    _same_nonces(oc, gc, T)


@pytest.mark.parametrize("T", [2, 3])
def test_shuffle_with_threads(O, G, T):
    n = 4099
    x = np.arange(n, dtype=np.uint64)
    oc, gc = _pair(O, G, T)
    ov, gv = oc.secret_share(x, 8), gc.secret_share(x, 8)
    oc.shuffle(ov)
    gc.shuffle(gv)
    _same_shares(ov, gv)
    _same_nonces(oc, gc, T)
    _same_log(oc, gc)
    assert sorted(gc.open(gv).tolist()) == list(range(n))


@pytest.mark.parametrize("T,n", [(2, 3000), (3, 5000), (4, 300)])
@pytest.mark.parametrize("width,desc", [(4, False), (8, True)])
def test_radix_sort_with_threads(O, G, T, n, width, desc):
    rng = np.random.default_rng(n + T)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**(8 * width), n, dtype=np.uint64).astype(dt)
    oc, gc = _pair(O, G, T)
    ov, gv = oc.secret_share(x, width), gc.secret_share(x, width)
    operm = oc.radix_sort(ov, desc=desc)
    gperm = gc.radix_sort(gv, desc=desc)
    _same_shares(ov, gv)
    _same_shares(operm, gperm)
    _same_nonces(oc, gc, T)
    _same_log(oc, gc)
    sdt = np.int32 if width == 4 else np.int64
    signed = x.view(sdt)
    order = np.argsort(~signed if desc else signed, kind="stable")
    np.testing.assert_array_equal(gc.open(gv).view(sdt), signed[order])


def test_radix_sort_with_threads_2p20(G):
    """2^20 rows on 3 threads against the plaintext stable sort and ranks."""
    n, T = 1 << 20, 3
    gc = G.Ctx(_keys(0), threads=T, thread_keys=[_keys(t) for t in range(1, T)])
    x = np.random.default_rng(5).integers(-2**63, 2**63 - 1, n, dtype=np.int64).view(np.uint64)
    v = gc.secret_share(x, 8)
    perm = gc.radix_sort(v)
    order = np.argsort(x.view(np.int64), kind="stable")
    np.testing.assert_array_equal(gc.open(v).view(np.int64), x.view(np.int64)[order])
    ranks = np.empty(n, dtype=np.int64)
    ranks[order] = np.arange(n)
    np.testing.assert_array_equal(gc.open(perm).view(np.int32), ranks)


def test_table_sort_with_threads(O, G):
    from oracle_backend import OracleCtx
    from orq_amd.table import EncodedTable

    T = 2
    extra = [_keys(t) for t in range(1, T)]
    rng = np.random.default_rng(9)
    n = 2000
    data = [rng.integers(0, 7, n), rng.integers(0, 1000, n), rng.integers(0, 1000, n)]
    schema = ["[K]", "[P]", "V"]
    oc = OracleCtx(_keys(0), threads=T, thread_keys=extra)
    gc = G.Ctx(_keys(0), threads=T, thread_keys=extra)
    ot = EncodedTable.secret_share(oc, data, schema, 8)
    gt = EncodedTable.secret_share(gc, data, schema, 8)
    ot.sort(["[K]", ("[P]", "DESC")], protocol="RADIXSORT")
    gt.sort(["[K]", ("[P]", "DESC")], protocol="RADIXSORT")
    for name in ot.column_names():
        _same_shares(ot.schema[name], gt.schema[name])
    got = gt.open()
    order = np.lexsort((-data[1], data[0]))
    np.testing.assert_array_equal(got["[K]"], data[0][order])


@pytest.mark.parametrize("T,n", [(2, 5000), (3, 3001), (4, 300), (3, 70001)])
@pytest.mark.parametrize("width", [4, 8])
def test_circuits_with_threads(O, G, T, n, width):
    """The comparisons (element-wise calls split by element ranges, the packed
    sign-fix call by word ranges), full b2a, the ripple-carry and
    parallel-prefix adders, rca_compare and a2b, each as the T-thread oracle
    gives them."""
    rng = np.random.default_rng(T * n + 3 * width)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**(8 * width), n, dtype=np.uint64).astype(dt)
    y = rng.integers(0, 2**(8 * width), n, dtype=np.uint64).astype(dt)
    y[::7] = x[::7]
    oc, gc = _pair(O, G, T)
    ox, gx = oc.secret_share(x, width), gc.secret_share(x, width)
    oy, gy = oc.secret_share(y, width), gc.secret_share(y, width)
    for op in ["==", "!=", "<", ">", "<=", ">="]:
        _same_shares(oc.cmp(op, ox, oy), gc.cmp(op, gx, gy))
    _same_shares(oc.b2a(ox), gc.b2a(gx))
    for op in ["add", "sub", "ppa_add", "ppa_sub", "rca_lt"]:
        _same_shares(oc.adder(op, ox, oy), gc.adder(op, gx, gy))
    oa, ga = oc.secret_share(x, width, arith=True), gc.secret_share(x, width, arith=True)
    _same_shares(oc.adder("a2b", oa), gc.adder("a2b", ga))
    _same_nonces(oc, gc, T)
    _same_log(oc, gc)
    sdt = np.int32 if width == 4 else np.int64
    got = gc.open(gc.cmp("<", gx, gy))
    np.testing.assert_array_equal(got, (x.view(sdt) < y.view(sdt)).astype(dt))
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(gc.open(gc.adder("add", gx, gy)), x + y)


def test_other_operations_refuse_threads(G):
    gc = G.Ctx(_keys(0), threads=2, thread_keys=[_keys(1)])
    x = gc.secret_share(np.arange(600, dtype=np.uint64), 8)
    with pytest.raises(G.OrqgError, match="worker thread"):
        gc.quicksort(x)
    with pytest.raises(G.OrqgError, match="worker thread"):
        gc.bitonic_sort([x])
