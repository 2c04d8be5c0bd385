"""Oracle pins for ElementwisePermutation operations, compose, shuffles and
table_sort:

* closed forms on plaintext permutations: compose(sigma, rho) applied to x
  equals applying sigma and then rho, i.e. (rho o sigma)[i] = rho[sigma[i]]
  (shuffle.h:716-743); invert gives p^-1; reverse gives n-1-p; negate gives
  -p-1; b2a keeps the value (permutation.h:177-280);
* the table tests of the reference (tests/test_sorting.cpp:214-318): a table
  whose every column is the row index, shuffled and sorted on "[0]" ASC/DESC,
  has every column equal to i (n-1-i); a multi-key sort equals the
  sequence of single-column sorts from the last key to the first and a
  stable lexicographic numpy sort; and no reserved permutation is left over;
* a shuffle keeps rows together.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _ctx(keys):
    return O.Ctx(keys)


def test_compose_closed_form(keys):
    c = _ctx(keys)
    rng = np.random.default_rng(0)
    n = 777
    s, r = rng.permutation(n).astype(np.uint32), rng.permutation(n).astype(np.uint32)
    S, R = c.secret_share(s, 4, arith=True), c.secret_share(r, 4, arith=True)
    c.perm_compose(S, R)
    comp = c.open(R, arith=True)
    np.testing.assert_array_equal(comp, r[s])
    # applying comp == applying s then r (local_apply_perm: new[p[i]] = x[i])
    x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    y1 = np.empty_like(x)
    y1[comp] = x
    t = np.empty_like(x)
    t[s] = x
    y2 = np.empty_like(x)
    y2[r] = t
    np.testing.assert_array_equal(y1, y2)


def test_perm_ops_closed_forms(keys):
    c = _ctx(keys)
    rng = np.random.default_rng(1)
    n = 600
    p = rng.permutation(n).astype(np.uint32)
    P = c.secret_share(p, 4, arith=True)
    c.perm_invert(P)
    inv = c.open(P, arith=True)
    np.testing.assert_array_equal(inv[p], np.arange(n))
    P = c.secret_share(p, 4, arith=True)
    c.perm_reverse(P)
    np.testing.assert_array_equal(c.open(P, arith=True), (n - 1 - p.astype(np.int64)).astype(np.uint32))
    P = c.secret_share(p, 4, arith=True)
    c.perm_negate(P)
    np.testing.assert_array_equal(c.open(P, arith=True), (-1 - p.astype(np.int64)).astype(np.uint32))
    P = c.secret_share(p, 4)
    c.perm_b2a(P)
    np.testing.assert_array_equal(c.open(P, arith=True), p)


def test_table_shuffle_keeps_rows(keys):
    c = _ctx(keys)
    rng = np.random.default_rng(2)
    n = 1000
    a = rng.permutation(n).astype(np.uint64)
    b = rng.integers(0, 2**32, n, dtype=np.uint64)
    A, B = c.secret_share(a, 8, arith=True), c.secret_share(b, 8)
    c.table_shuffle([A], [B])
    a2, b2 = c.open(A, arith=True), c.open(B)
    np.testing.assert_array_equal(np.sort(a2), np.sort(a))
    row_of = np.argsort(a)  # a is a permutation of 0..n-1: its value names the row
    np.testing.assert_array_equal(b2, b[row_of[a2.astype(np.int64)]])
    assert (a2 != a).mean() > 0.9
    assert c.queue_size() == 0


def _table(c, data, schema):
    """Columns shared in schema order, plus the public valid column, keyed by name."""
    cols = {}
    for arr, name in zip(data, schema):
        cols[name] = c.secret_share(np.asarray(arr, np.int64).astype(np.uint32), 4,
                                    arith="[" not in name)
    ones = np.ones(len(data[0]), np.uint32)
    v = O.Vec(len(ones), 4)
    z = np.zeros_like(ones)
    v.set_shares([ones, z, z])
    cols["[##V]"] = v
    return cols


def _split(cols, keys=()):
    names = sorted(cols)
    a = [cols[k] for k in names if "[" not in k and k not in keys]
    b = [cols[k] for k in names if "[" in k and k not in keys]
    return a, b


def _open(c, cols):
    return {k: c.open(v, arith="[" not in k).view(np.int32) for k, v in cols.items()}


@pytest.mark.parametrize("protocol", ["RADIXSORT", "QUICKSORT"])
@pytest.mark.parametrize("desc", [False, True])
def test_reference_table_sort_single(keys, desc, protocol):
    # tests/test_sorting.cpp:214-252 with num_columns = 4
    n, ncol = 100, 4
    schema = [f"[{i}]" if i % 2 == 0 else str(i) for i in range(ncol)]
    data = [np.arange(n) for _ in range(ncol)]
    c = _ctx(keys)
    cols = _table(c, data, schema)
    c.table_shuffle(*_split(cols))
    shuffled = _open(c, cols)
    assert not np.array_equal(shuffled["[0]"], np.arange(n))
    a, b = _split(cols, keys=("[0]",))
    c.table_sort([cols["[0]"]], ["DESC" if desc else "ASC"], None, a, b, protocol)
    got = _open(c, cols)
    want = np.arange(n)[::-1] if desc else np.arange(n)
    for name in schema:
        np.testing.assert_array_equal(got[name], want, err_msg=name)
    np.testing.assert_array_equal(got["[##V]"], np.ones(n))
    assert c.queue_size() == 0


@pytest.mark.parametrize("protocol", ["RADIXSORT", "QUICKSORT"])
@pytest.mark.parametrize("tiebreak", [False, True])
def test_reference_table_sort_multi(keys, tiebreak, protocol):
    # tests/test_sorting.cpp:258-318 with num_columns = 6: spec [0] ASC, [2] DESC, [4] ASC
    n, ncol = 300, 6
    rng = np.random.default_rng(5)
    schema = [f"[{i}]" if i % 2 == 0 else str(i) for i in range(ncol)]
    data = [rng.integers(0, n, n) % (2 if tiebreak else n) for _ in range(ncol)]
    spec = [("[0]", "ASC"), ("[2]", "DESC"), ("[4]", "ASC")]
    c1, c2 = _ctx(keys), _ctx(keys)
    t1, t2 = _table(c1, data, schema), _table(c2, data, schema)
    ks = [k for k, _ in spec]
    a, b = _split(t1, keys=ks)
    c1.table_sort([t1[k] for k in ks], [o for _, o in spec], None, a, b, protocol)
    assert c1.queue_size() == 0
    for k, o in reversed(spec):  # one column at a time, last key first
        a, b = _split(t2, keys=(k,))
        c2.table_sort([t2[k]], [o], None, a, b, protocol)
    g1, g2 = _open(c1, t1), _open(c2, t2)
    order = np.lexsort((data[4], -np.asarray(data[2]), data[0]))
    for name in schema:
        np.testing.assert_array_equal(g1[name], g2[name], err_msg=name)
        np.testing.assert_array_equal(g1[name], np.asarray(data[schema.index(name)])[order])


def test_table_sort_on_valid_bit(keys):
    """A single-bit key ("[##V]"-style) sorts on its LSB only."""
    n = 200
    rng = np.random.default_rng(6)
    valid = rng.integers(0, 2, n).astype(np.uint32)
    payload = np.arange(n, dtype=np.uint32)
    c = _ctx(keys)
    V, P = c.secret_share(valid, 4), c.secret_share(payload, 4)
    c.table_sort([V], ["DESC"], [1], [], [P])
    order = np.argsort(-valid.astype(np.int64), kind="stable")
    np.testing.assert_array_equal(c.open(V), valid[order])
    np.testing.assert_array_equal(c.open(P), payload[order])
    assert c.queue_size() == 0


# ---- bitonic sort (operators/sorting.h:266-353) ------------------------------------
INT_MIN, INT_MAX = -(2**31), 2**31 - 1
# tests/test_bitonic_sort.cpp:107-109
BITONIC_COLUMNS = [[111, 111, 0, 111, INT_MAX, 5, 5, 5],
                   [-14, -4, 0, -14, INT_MIN, 13, 13, 13],
                   [INT_MIN, 0, 6423, -11233, INT_MIN, 7, 7, 7]]


def _bshare(c, vals, width=4):
    dt = np.int32 if width == 4 else np.int64
    return c.secret_share(np.asarray(vals, dt).view(np.uint32 if width == 4 else np.uint64), width)


@pytest.mark.parametrize("desc", [False, True])
def test_reference_bitonic_vector(keys, desc):
    # tests/test_bitonic_sort.cpp:48-92
    c = _ctx(keys)
    v = _bshare(c, [0, 5, 4, 7, -1, -2, 8, 1])
    c.bitonic_sort([v], ["DESC" if desc else "ASC"])
    want = [8, 7, 5, 4, 1, 0, -1, -2] if desc else [-2, -1, 0, 1, 4, 5, 7, 8]
    np.testing.assert_array_equal(c.open(v).view(np.int32), want)
    for x in ([0, INT_MAX // 2, -1, 42, 87, 100, -444, -13],
              [0, INT_MAX // 2 + 1, -1, 42, 87, 100, -444, -13]):
        v = _bshare(c, x)
        c.bitonic_sort([v])
        np.testing.assert_array_equal(c.open(v).view(np.int32), sorted(x))


def _bitonic_table(c, columns, schema, spec):
    """EncodedTable::sort with BITONICSORT on a table of n = 8 rows (no padding)."""
    cols = {name: c.secret_share(np.asarray(v, np.int32).view(np.uint32), 4,
                                 arith="[" not in name) for v, name in zip(columns, schema)}
    ones = np.ones(len(columns[0]), np.uint32)
    z = np.zeros_like(ones)
    valid = O.Vec(len(ones), 4)
    valid.set_shares([ones, z, z])
    cols["[##V]"] = valid
    ks = [k for k, _ in spec]
    a, b = _split(cols, keys=ks)
    c.bitonic_sort([cols[k] for k in ks], [o for _, o in spec], a, b)
    return {k: c.open(v, arith="[" not in k).view(np.int32) for k, v in cols.items()}


def test_reference_bitonic_tables(keys):
    # tests/test_bitonic_sort.cpp:141-196
    c = _ctx(keys)
    got = _bitonic_table(c, BITONIC_COLUMNS, ["[col_1]", "[col_2]", "[col_3]"],
                         [("[col_1]", "DESC"), ("[col_2]", "ASC"), ("[col_3]", "DESC")])
    np.testing.assert_array_equal(got["[col_1]"], [INT_MAX, 111, 111, 111, 5, 5, 5, 0])
    np.testing.assert_array_equal(got["[col_2]"], [INT_MIN, -14, -14, -4, 13, 13, 13, 0])
    np.testing.assert_array_equal(got["[col_3]"], [INT_MIN, -11233, INT_MIN, 0, 7, 7, 7, 6423])
    got = _bitonic_table(c, BITONIC_COLUMNS, ["[col_1]", "col_2", "[col_3]"],
                         [("[col_1]", "DESC"), ("[col_3]", "ASC")])
    np.testing.assert_array_equal(got["[col_1]"], [INT_MAX, 111, 111, 111, 5, 5, 5, 0])
    np.testing.assert_array_equal(got["[col_3]"], [INT_MIN, INT_MIN, -11233, 0, 7, 7, 7, 6423])
    np.testing.assert_array_equal(got["col_2"], [INT_MIN, -14, -14, -4, 13, 13, 13, 0])


@pytest.mark.parametrize("n", [2, 64, 1024])
def test_bitonic_random_rows(keys, n):
    rng = np.random.default_rng(n)
    k1 = rng.integers(-4, 4, n)
    k2 = rng.integers(-(2**40), 2**40, n)
    a = rng.integers(0, 2**50, n)
    b = rng.integers(0, 2**50, n)
    c = _ctx(keys)
    K1, K2 = _bshare(c, k1, 8), _bshare(c, k2, 8)
    A = c.secret_share(a.astype(np.uint64), 8, arith=True)
    B = c.secret_share(b.astype(np.uint64), 8)
    c.bitonic_sort([K1, K2], ["ASC", "DESC"], [A], [B])
    g1, g2 = c.open(K1).view(np.int64), c.open(K2).view(np.int64)
    order = np.lexsort((-k2, k1))
    np.testing.assert_array_equal(g1, k1[order])
    np.testing.assert_array_equal(g2, k2[order])
    rows = sorted(zip(k1, k2, a, b))
    got = sorted(zip(g1, g2, c.open(A, arith=True).astype(np.int64), c.open(B).astype(np.int64)))
    assert rows == got
