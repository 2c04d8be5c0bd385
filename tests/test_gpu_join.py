"""GPU parity of the relational pipeline: the same host table code
(concatenate, sort, distinct, aggregate, inner_join) over liborqg and over the
oracle gives the same shares in every column, the same nonces and message
log, and SQLite's result."""
import numpy as np
import pytest

from oracle_backend import OracleCtx
from orq_amd.table import EncodedTable
from test_gpu_parity import _same_log, _same_shares
from test_oracle_join import (pipeline_result, relational_inputs, run_pipeline,
                              sqlite_join_groupby)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()
    return orq_amd


def _same_tables(ot, gt):
    assert sorted(ot.schema) == sorted(gt.schema)
    assert ot.rows == gt.rows
    for name in ot.schema:
        _same_shares(ot.schema[name], gt.schema[name])


def _same_nonce_state(oc, gc):
    for p in range(3):
        assert oc.nonce(p, 0) == gc.nonce(p)
        assert oc.nonce(p, 2) == gc.nonce(3 + p)
    for g in range(3):  # groups {0,1}, {0,2}, {1,2}: read the nonce of a member
        assert oc.nonce(0 if g < 2 else 1, 3 + g) == gc.nonce(6 + g)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("nr,nl,dom", [(16, 64, 16), (256, 2048, 400)])
def test_join_groupby_pipeline(G, keys, width, nr, nl, dom):
    rkey, attr, lkey, value = relational_inputs(nr, nl, dom, nr + nl + width)
    oc, gc = OracleCtx(keys), G.Ctx(keys)
    ot = run_pipeline(oc, rkey, attr, lkey, value, width)
    gt = run_pipeline(gc, rkey, attr, lkey, value, width)
    _same_tables(ot, gt)
    _same_nonce_state(oc, gc)
    _same_log(oc, gc)
    np.testing.assert_array_equal(pipeline_result(gt), sqlite_join_groupby(rkey, attr, lkey, value))


def test_reference_multi_key_join_gpu(G, keys):
    # tests/test_join.cpp:29-68
    gc = G.Ctx(keys)
    t1 = EncodedTable.secret_share(gc, [[1, 1, 2, 3, 4], [1, 2, 4, 6, 7], [0] * 5],
                                   ["[K1]", "[K2]", "[D]"])
    t2 = EncodedTable.secret_share(gc, [[1, 1, 1, 3, 4, 4, 1, 1, 4, 5],
                                        [2, 1, 1, 6, 7, 7, 1, 2, 7, 8],
                                        [0, 1, 1, 2, 3, 5, 8, 3, 1, 4], [0] * 10],
                                   ["[K1]", "[K2]", "[D2]", "C"])
    tj = t1.inner_join(t2, ["[K1]", "[K2]"], [("C", "C", "count")])
    tj.sort(["[K1]", "[K2]"])
    t = tj.open_with_schema()
    np.testing.assert_array_equal(t["C"], [3, 2, 1, 3])


def test_table_distinct_gpu(G, keys):
    rng = np.random.default_rng(3)
    k = rng.integers(0, 6, 300)
    v = rng.integers(0, 2, 300)
    oc, gc = OracleCtx(keys), G.Ctx(keys)
    ot = EncodedTable.secret_share(oc, [k, v], ["[K]", "[V2]"])
    gt = EncodedTable.secret_share(gc, [k, v], ["[K]", "[V2]"])
    ot.distinct(["[K]", "[V2]"])
    gt.distinct(["[K]", "[V2]"])
    _same_tables(ot, gt)
    got = gt.open_with_schema()
    pairs = sorted(set(zip(k.tolist(), v.tolist())))
    assert sorted(zip(got["[K]"].tolist(), got["[V2]"].tolist())) == pairs
