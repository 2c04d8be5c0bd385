"""GPU parity of the ElementwisePermutation operations, compose, shuffles and
table_sort: shares, nonces and message logs of liborqg against the oracle on
the same keys, and the plaintext results of the reference's table tests."""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_nonces, _same_shares
from test_oracle_table import _open, _split, _table

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


@pytest.mark.parametrize("op", ["invert", "reverse", "negate", "b2a", "compose"])
@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_perm_ops(O, G, keys, op, n):
    rng = np.random.default_rng(n)
    p = rng.permutation(n).astype(np.uint32)
    q = rng.permutation(n).astype(np.uint32)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    arith = op != "b2a"
    OP, GP = oc.secret_share(p, 4, arith=arith), gc.secret_share(p, 4, arith=arith)
    if op == "compose":
        OQ, GQ = oc.secret_share(q, 4, arith=True), gc.secret_share(q, 4, arith=True)
        oc.perm_compose(OP, OQ)
        gc.perm_compose(GP, GQ)
        _same_shares(OQ, GQ)
        _same_shares(OP, GP)
        want = q[p]
        out = GQ
    else:
        getattr(oc, "perm_" + op)(OP)
        getattr(gc, "perm_" + op)(GP)
        _same_shares(OP, GP)
        out = GP
        pl = p.astype(np.int64)
        want = {"invert": np.argsort(p), "reverse": n - 1 - pl, "negate": -1 - pl, "b2a": pl}[op]
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    np.testing.assert_array_equal(gc.open(out), np.asarray(want).astype(np.uint32))


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("arith", [False, True])
def test_shuffle_vector(O, G, keys, width, arith):
    n = 5000
    x = np.arange(n, dtype=np.uint64).astype(np.uint32 if width == 4 else np.uint64)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ox, gx = oc.secret_share(x, width, arith=arith), gc.secret_share(x, width, arith=arith)
    oc.shuffle(ox, arith=arith)
    gc.shuffle(gx)
    _same_shares(ox, gx)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    got = gc.open(gx)
    np.testing.assert_array_equal(np.sort(got), x)


def _gpu_table(G, gc, data, schema, width=4):
    from orq_amd.table import EncodedTable

    return EncodedTable.secret_share(gc, data, schema, width)


@pytest.mark.parametrize("protocol", ["RADIXSORT", "QUICKSORT"])
@pytest.mark.parametrize("desc", [False, True])
def test_reference_table_sort_single(O, G, keys, desc, protocol):
    # tests/test_sorting.cpp:214-252: shuffle, then sort on "[0]"
    n, ncol = 1000, 4
    schema = [f"[{i}]" if i % 2 == 0 else str(i) for i in range(ncol)]
    data = [np.arange(n) for _ in range(ncol)]
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ot = _table(oc, data, schema)
    gt = _gpu_table(G, gc, data, schema)
    oc.table_shuffle(*_split(ot))
    gt.shuffle()
    for name in ot:
        _same_shares(ot[name], gt[name])
    a, b = _split(ot, keys=("[0]",))
    oc.table_sort([ot["[0]"]], ["DESC" if desc else "ASC"], None, a, b, protocol)
    gt.sort([("[0]", "DESC" if desc else "ASC")], protocol=protocol)
    for name in ot:
        _same_shares(ot[name], gt[name])
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    got = gt.open()
    want = np.arange(n)[::-1] if desc else np.arange(n)
    for name in schema:
        np.testing.assert_array_equal(got[name], want, err_msg=name)


@pytest.mark.parametrize("protocol", ["RADIXSORT", "QUICKSORT"])
@pytest.mark.parametrize("tiebreak", [False, True])
def test_reference_table_sort_multi(O, G, keys, tiebreak, protocol):
    # tests/test_sorting.cpp:258-318, spec [0] ASC, [2] DESC, [4] ASC
    n, ncol = 3000, 6
    rng = np.random.default_rng(9)
    schema = [f"[{i}]" if i % 2 == 0 else str(i) for i in range(ncol)]
    data = [rng.integers(0, n, n) % (2 if tiebreak else n) for _ in range(ncol)]
    spec = [("[0]", "ASC"), ("[2]", "DESC"), ("[4]", "ASC")]
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ot = _table(oc, data, schema)
    gt = _gpu_table(G, gc, data, schema)
    ks = [k for k, _ in spec]
    a, b = _split(ot, keys=ks)
    oc.table_sort([ot[k] for k in ks], [o for _, o in spec], None, a, b, protocol)
    gt.sort(spec, protocol=protocol)
    assert gc.queue_size() == 0
    for name in ot:
        _same_shares(ot[name], gt[name])
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    order = np.lexsort((data[4], -np.asarray(data[2]), data[0]))
    got = gt.open()
    for i, name in enumerate(schema):
        np.testing.assert_array_equal(got[name], np.asarray(data[i])[order], err_msg=name)


def test_table_sort_valid_bit_and_int64_payload(O, G, keys):
    """[##V] as a single-bit key (DESC) beside an int64 key and payloads,
    with the DEFAULT protocol (quicksort for the int64 key)."""
    n = 4099
    rng = np.random.default_rng(10)
    key = rng.integers(-50, 50, n)
    pay = rng.integers(0, 2**40, n)
    valid = rng.integers(0, 2, n)
    schema = ["[K]", "P", "[##V]"]
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ocols = {}
    for arr, name in zip([key, pay, valid], schema):
        ocols[name] = oc.secret_share(np.asarray(arr, np.int64).view(np.uint64), 8,
                                      arith="[" not in name)
    from orq_amd.table import EncodedTable

    gt = EncodedTable.secret_share(gc, [key, pay, valid], schema, width=8)
    spec = [("[##V]", "DESC"), ("[K]", "ASC")]
    a, b = _split(ocols, keys=("[##V]", "[K]"))
    oc.table_sort([ocols["[##V]"], ocols["[K]"]], ["DESC", "ASC"], [1, 0], a, b, "QUICKSORT")
    gt.sort(spec)
    for name in schema:
        _same_shares(ocols[name], gt[name])
    _same_nonces(oc, gc)
    order = np.lexsort((key, -valid))
    got = gt.open()
    np.testing.assert_array_equal(got["[K]"], key[order])
    np.testing.assert_array_equal(got["P"], pay[order])
    np.testing.assert_array_equal(got["[##V]"], valid[order])


# ---- bitonic sort ----------------------------------------------------------------
@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [2, 256, 4096])
@pytest.mark.parametrize("shape", ["one_key", "two_keys_data"])
def test_bitonic_sort(O, G, keys, width, n, shape):
    rng = np.random.default_rng(n + width)
    ut = np.uint32 if width == 4 else np.uint64
    k1 = rng.integers(-4, 4, n).astype(np.int64)
    k2 = rng.integers(-(2**30), 2**30, n).astype(np.int64)
    a = rng.integers(0, 2**31, n).astype(np.int64)
    b = rng.integers(0, 2**31, n).astype(np.int64)
    oc, gc = O.Ctx(keys), G.Ctx(keys)

    def share(c, v, arith=False):
        return c.secret_share(v.astype(ut), width, arith=arith)

    if shape == "one_key":
        ov, gv = [share(oc, k2)], [share(gc, k2)]
        oc.bitonic_sort(ov, ["DESC"])
        gc.bitonic_sort(gv, ["DESC"])
        oa = ga = ob = gb = []
    else:
        ov, gv = [share(oc, k1), share(oc, k2)], [share(gc, k1), share(gc, k2)]
        oa, ga = [share(oc, a, True)], [share(gc, a, True)]
        ob, gb = [share(oc, b)], [share(gc, b)]
        oc.bitonic_sort(ov, ["ASC", "DESC"], oa, ob)
        gc.bitonic_sort(gv, ["ASC", "DESC"], ga, gb)
    for x, y in zip(ov + oa + ob, gv + ga + gb):
        _same_shares(x, y)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    sdt = np.int32 if width == 4 else np.int64
    if shape == "one_key":
        np.testing.assert_array_equal(gc.open(gv[0]).view(sdt), np.sort(k2.astype(sdt))[::-1])
    else:
        order = np.lexsort((-k2, k1))
        np.testing.assert_array_equal(gc.open(gv[0]).view(sdt), k1[order].astype(sdt))
        np.testing.assert_array_equal(gc.open(gv[1]).view(sdt), k2[order].astype(sdt))


def test_reference_bitonic_tables_gpu(G, keys):
    # tests/test_bitonic_sort.cpp:141-196 through the EncodedTable mirror
    from orq_amd.table import BITONICSORT, EncodedTable
    from test_oracle_table import BITONIC_COLUMNS, INT_MAX, INT_MIN

    gc = G.Ctx(keys)
    t = EncodedTable.secret_share(gc, BITONIC_COLUMNS, ["[col_1]", "[col_2]", "[col_3]"])
    t.sort([("[col_1]", "DESC"), ("[col_2]", "ASC"), ("[col_3]", "DESC")], protocol=BITONICSORT)
    got = t.open()
    np.testing.assert_array_equal(got["[col_1]"], [INT_MAX, 111, 111, 111, 5, 5, 5, 0])
    np.testing.assert_array_equal(got["[col_2]"], [INT_MIN, -14, -14, -4, 13, 13, 13, 0])
    np.testing.assert_array_equal(got["[col_3]"], [INT_MIN, -11233, INT_MIN, 0, 7, 7, 7, 6423])
    t = EncodedTable.secret_share(gc, BITONIC_COLUMNS, ["[col_1]", "col_2", "[col_3]"])
    t.sort([("[col_1]", "DESC"), ("[col_3]", "ASC")], protocol=BITONICSORT)
    got = t.open()
    np.testing.assert_array_equal(got["[col_3]"], [INT_MIN, INT_MIN, -11233, 0, 7, 7, 7, 6423])
    np.testing.assert_array_equal(got["col_2"], [INT_MIN, -14, -14, -4, 13, 13, 13, 0])


def test_bitonic_table_pads_and_unpads(G, keys):
    """A non-power-of-two table is padded with invalid rows, sorted with the
    valid column as the first key (DESC) and cut back to its size."""
    from orq_amd.table import BITONICSORT, EncodedTable

    n = 1000
    rng = np.random.default_rng(3)
    k = rng.integers(-100, 100, n)
    gc = G.Ctx(keys)
    t = EncodedTable.secret_share(gc, [k, np.arange(n)], ["[K]", "P"])
    t.sort([("[##V]", "DESC"), ("[K]", "ASC")], protocol=BITONICSORT)
    assert t.rows == n
    got = t.open()
    np.testing.assert_array_equal(got["[K]"], np.sort(k))
    np.testing.assert_array_equal(got["[##V]"], np.ones(n))
    np.testing.assert_array_equal(k[got["P"]], got["[K]"])
