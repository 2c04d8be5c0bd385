"""The host table layer (orq_amd.table: concatenate, distinct, filter,
aggregate, inner_join) run over the CPU oracle, pinned by the reference's
literal join tests (tests/test_join.cpp:29-117) and by SQLite on the same
plaintext: an inner join of a primary-key table with a foreign-key table
followed by GROUP BY key with SUM, COUNT and MAX."""
import sqlite3

import numpy as np
import pytest

from oracle_backend import OracleCtx
from orq_amd.table import EncodedTable


def _ctx(keys):
    return OracleCtx(keys)


def test_reference_multi_key_join(keys):
    # tests/test_join.cpp:29-68
    c = _ctx(keys)
    t1 = EncodedTable.secret_share(c, [[1, 1, 2, 3, 4], [1, 2, 4, 6, 7], [0] * 5],
                                   ["[K1]", "[K2]", "[D]"])
    t2 = EncodedTable.secret_share(c, [[1, 1, 1, 3, 4, 4, 1, 1, 4, 5],
                                       [2, 1, 1, 6, 7, 7, 1, 2, 7, 8],
                                       [0, 1, 1, 2, 3, 5, 8, 3, 1, 4], [0] * 10],
                                   ["[K1]", "[K2]", "[D2]", "C"])
    tj = t1.inner_join(t2, ["[K1]", "[K2]"], [("C", "C", "count")])
    tj.sort(["[K1]", "[K2]"])
    tj.delete_columns(["[D2]"])
    t = tj.open_with_schema()
    np.testing.assert_array_equal(t["C"], [3, 2, 1, 3])
    np.testing.assert_array_equal(t["[K1]"], [1, 1, 3, 4])
    np.testing.assert_array_equal(t["[K2]"], [1, 2, 6, 7])


def test_reference_join_with_valid(keys):
    # tests/test_join.cpp:70-117
    c = _ctx(keys)
    p = EncodedTable.secret_share(c, [[1, 2, 3, 4], [100, 200, 300, 400], [1, 0, 1, 0]],
                                  ["[K]", "[D2]", "[_VALID]"])
    p.filter("[_VALID]")
    f = EncodedTable.secret_share(c, [[1, 1, 1, 1, 2, 3, 3, 3, 3, 5, 5],
                                      [10, 20, 30, 32, 35, 40, 50, 60, 65, 70, 80],
                                      [1, 1, 0, 1, 1, 1, 0, 0, 1, 1, 0]],
                                  ["[K]", "[Data]", "[_VALID]"])
    f.filter("[_VALID]")
    j = p.inner_join(f, ["[K]"], [("[D2]", "[D2]", "copy")])
    j.delete_columns(["[_VALID]"])
    t = j.open_with_schema()
    assert sorted(t["[Data]"].tolist()) == [10, 20, 32, 40, 65]
    d2 = dict(zip(t["[Data]"].tolist(), t["[D2]"].tolist()))
    assert d2 == {10: 100, 20: 100, 32: 100, 40: 300, 65: 300}


def relational_inputs(nr, nl, key_domain, seed):
    rng = np.random.default_rng(seed)
    rkey = rng.permutation(key_domain)[:nr]
    attr = rng.integers(0, 1000, nr)
    lkey = rng.integers(0, key_domain, nl)
    value = rng.integers(0, 10**6, nl)
    return rkey, attr, lkey, value


def sqlite_join_groupby(rkey, attr, lkey, value):
    db = sqlite3.connect(":memory:")
    db.execute("CREATE TABLE R (key INTEGER PRIMARY KEY, attr INTEGER)")
    db.execute("CREATE TABLE L (key INTEGER, value INTEGER)")
    db.executemany("INSERT INTO R VALUES (?, ?)", zip(rkey.tolist(), attr.tolist()))
    db.executemany("INSERT INTO L VALUES (?, ?)", zip(lkey.tolist(), value.tolist()))
    rows = db.execute("SELECT L.key, R.attr, SUM(L.value), COUNT(*), MAX(L.value) "
                      "FROM R JOIN L ON R.key = L.key GROUP BY L.key ORDER BY L.key").fetchall()
    return np.array(rows, dtype=np.int64).reshape(-1, 5)


def run_pipeline(c, rkey, attr, lkey, value, width=4):
    """R JOIN L ON key (copying R.attr down), then GROUP BY key with
    SUM(value), COUNT(*), MAX(value)."""
    r = EncodedTable.secret_share(c, [rkey, attr], ["[key]", "[attr]"], width, name="R")
    l_ = EncodedTable.secret_share(c, [lkey, value, value], ["[key]", "value", "[vmax]"], width,
                                   name="L")
    j = r.inner_join(l_, ["[key]"], [("[attr]", "[attr]", "copy")])
    j.add_columns(["cnt"])
    j.aggregate(["[key]"], [("value", "value", "sum"), ("value", "cnt", "count"),
                            ("[vmax]", "[vmax]", "max")])
    return j


def pipeline_result(j):
    t = j.open_with_schema()
    order = np.argsort(t["[key]"], kind="stable")
    return np.stack([t["[key]"][order], t["[attr]"][order], t["value"][order], t["cnt"][order],
                     t["[vmax]"][order]], axis=1).astype(np.int64)


@pytest.mark.parametrize("nr,nl,dom", [(8, 32, 12), (16, 64, 16), (32, 100, 64)])
def test_join_groupby_matches_sqlite(keys, nr, nl, dom):
    rkey, attr, lkey, value = relational_inputs(nr, nl, dom, nr + nl)
    c = _ctx(keys)
    j = run_pipeline(c, rkey, attr, lkey, value)
    np.testing.assert_array_equal(pipeline_result(j), sqlite_join_groupby(rkey, attr, lkey, value))


def test_table_distinct(keys):
    rng = np.random.default_rng(3)
    k = rng.integers(0, 6, 40)
    v = rng.integers(0, 2, 40)
    c = _ctx(keys)
    t = EncodedTable.secret_share(c, [k, v], ["[K]", "[V2]"])
    t.distinct(["[K]", "[V2]"])
    got = t.open_with_schema()
    pairs = sorted(set(zip(k.tolist(), v.tolist())))
    assert sorted(zip(got["[K]"].tolist(), got["[V2]"].tolist())) == pairs
