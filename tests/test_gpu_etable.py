"""The library's EncodedTable (orqg_et_*): the relational methods as control
flow inside liborqg.  Each test runs the same steps through the host table
layer (orq_amd.table, itself pinned to the oracle and to SQLite) and through
orqg_et_* on twin contexts, and compares every column share for share, the
nonces and the message log; the join pipeline is also checked against
SQLite."""
import numpy as np
import pytest

from orq_amd.table import EncodedTable
from test_gpu_parity import _same_log
from test_oracle_join import pipeline_result, relational_inputs, sqlite_join_groupby

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()
    return orq_amd


def _same(pt, et):
    assert sorted(pt.schema) == et.column_names()
    assert pt.rows == et.rows
    for name in pt.column_names():
        for k, (x, y) in enumerate(zip(pt.schema[name].shares(), et[name].shares())):
            np.testing.assert_array_equal(x, y, err_msg=f"column {name} share {k}")


def _same_state(pc, ec):
    assert [pc.nonce(r) for r in range(10)] == [ec.nonce(r) for r in range(10)]
    _same_log(pc, ec)


def _twins(G, keys, data, schema, width):
    from orq_amd.etable import ETable

    pc, ec = G.Ctx(keys), G.Ctx(keys)
    pt = EncodedTable.secret_share(pc, data, schema, width)
    et = ETable.from_encoded(EncodedTable.secret_share(ec, data, schema, width))
    return pc, ec, pt, et


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("nr,nl,dom", [(16, 64, 16), (256, 2048, 400)])
def test_et_join_groupby_pipeline(G, keys, width, nr, nl, dom):
    """R JOIN L ON key copying R.attr, then GROUP BY key SUM/COUNT/MAX."""
    from orq_amd.etable import ETable
    from test_oracle_join import run_pipeline

    rkey, attr, lkey, value = relational_inputs(nr, nl, dom, nr + nl + width)
    pc, ec = G.Ctx(keys), G.Ctx(keys)
    pt = run_pipeline(pc, rkey, attr, lkey, value, width)
    r = ETable.from_encoded(EncodedTable.secret_share(ec, [rkey, attr], ["[key]", "[attr]"], width))
    l_ = ETable.from_encoded(EncodedTable.secret_share(ec, [lkey, value, value],
                                                       ["[key]", "value", "[vmax]"], width))
    j = r.inner_join(l_, ["[key]"], [("[attr]", "[attr]", "copy")])
    j.add_columns(["cnt"])
    j.aggregate(["[key]"], [("value", "value", "sum"), ("value", "cnt", "count"),
                            ("[vmax]", "[vmax]", "max")])
    _same(pt, j)
    _same_state(pc, ec)
    np.testing.assert_array_equal(pipeline_result(j), sqlite_join_groupby(rkey, attr, lkey, value))


@pytest.mark.parametrize("protocol", ["DEFAULT", "RADIXSORT", "QUICKSORT", "BITONICSORT"])
def test_et_sort_protocols(G, keys, protocol):
    rng = np.random.default_rng(4)
    n = 256  # a power of two: bitonic sort pads nothing
    data = [rng.integers(0, 5, n), rng.integers(-50, 50, n), rng.integers(0, 1000, n)]
    schema = ["[A]", "[B]", "C"]
    pc, ec, pt, et = _twins(G, keys, data, schema, 4)
    spec = ["[A]", ("[B]", "DESC")]
    pt.sort(spec, protocol=protocol)
    et.sort(spec, protocol=protocol)
    _same(pt, et)
    _same_state(pc, ec)
    got = et.open()
    order = np.lexsort((-data[1], data[0]))
    np.testing.assert_array_equal(got["[A]"], data[0][order])
    np.testing.assert_array_equal(got["[B]"], data[1][order])


@pytest.mark.parametrize("width", [4, 8])
def test_et_distinct(G, keys, width):
    rng = np.random.default_rng(3)
    k, v = rng.integers(0, 6, 300), rng.integers(0, 2, 300)
    pc, ec, pt, et = _twins(G, keys, [k, v], ["[K]", "[V2]"], width)
    pt.distinct(["[K]", "[V2]"])
    et.distinct(["[K]", "[V2]"])
    _same(pt, et)
    _same_state(pc, ec)
    got = et.open_with_schema()
    assert sorted(zip(got["[K]"].tolist(), got["[V2]"].tolist())) == \
        sorted(set(zip(k.tolist(), v.tolist())))


@pytest.mark.parametrize("reverse,do_sort,mark_valid", [
    (False, True, True), (True, True, True), (False, True, False), (True, False, True)])
def test_et_aggregate_options(G, keys, reverse, do_sort, mark_valid):
    rng = np.random.default_rng(5)
    n = 200
    data = [rng.integers(0, 9, n), rng.integers(0, 100, n), rng.integers(0, 100, n)]
    schema = ["[K]", "v", "[m]"]
    pc, ec, pt, et = _twins(G, keys, data, schema, 4)
    for t in (pt, et):
        t.add_columns(["cnt", "[mn]"])
    spec = [("v", "v", "sum"), ("v", "cnt", "count"), ("[m]", "[m]", "max"), ("[m]", "[mn]", "min")]
    pt.aggregate(["[K]"], spec, reverse=reverse, do_sort=do_sort, mark_valid=mark_valid)
    et.aggregate(["[K]"], spec, reverse=reverse, do_sort=do_sort, mark_valid=mark_valid)
    _same(pt, et)
    _same_state(pc, ec)


def test_et_concatenate_and_filter(G, keys):
    from orq_amd.etable import ETable

    pc, ec = G.Ctx(keys), G.Ctx(keys)
    a_p = EncodedTable.secret_share(pc, [[1, 2, 3], [4, 5, 6]], ["[X]", "Y"])
    b_p = EncodedTable.secret_share(pc, [[7, 8], [9, 10]], ["[X]", "[Z]"])
    a_e = ETable.from_encoded(EncodedTable.secret_share(ec, [[1, 2, 3], [4, 5, 6]], ["[X]", "Y"]))
    b_e = ETable.from_encoded(EncodedTable.secret_share(ec, [[7, 8], [9, 10]], ["[X]", "[Z]"]))
    tp, te = a_p.concatenate(b_p), a_e.concatenate(b_e)
    _same(tp, te)
    tp.filter("[##TID]")
    te.filter("[##TID]")
    _same(tp, te)
    _same_state(pc, ec)
    got = te.open_with_schema()
    np.testing.assert_array_equal(got["[X]"], [7, 8])
    np.testing.assert_array_equal(got["[Z]"], [9, 10])


def test_et_rejects_bad_calls(G, keys):
    from orq_amd.etable import ETable

    c = G.Ctx(keys)
    t = ETable(c, 4, 8, ["[K]", "v"])
    with pytest.raises(G.OrqgError, match="not boolean-shared"):
        t.sort(["v"])
    with pytest.raises(G.OrqgError, match="no column"):
        t.filter("[nope]")
    with pytest.raises(G.OrqgError, match="encodings differ"):
        t.aggregate(["[K]"], [("v", "[K]", "copy")])
