"""GPU parity of the aggregation network and distinct: shares, nonces and
message log of liborqg against the oracle, and the plaintext network."""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_nonces, _same_shares
from test_oracle_agg import A_FNS, B_FNS, UT, plain_network

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


def _run(O, G, keys, kcols, specs, reverse, sel, width):
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    res = []
    for c, is_gpu in ((oc, False), (gc, True)):
        K = [c.secret_share(np.asarray(k).astype(UT[width]), width) for k in kcols]
        S = c.secret_share(np.asarray(sel).astype(UT[width]), width) if sel is not None else None
        bs, as_, outs = [], [], []
        for v, fn, arith in specs:
            inp = c.secret_share(np.asarray(v).astype(UT[width]), width, arith=arith)
            out = c.vec(len(v), width) if is_gpu else O.Vec(len(v), width)
            (as_ if arith else bs).append((inp, out, fn))
            outs.append(out)
        c.aggregate(K, bs, as_, reverse, S)
        res.append(outs)
    for o, g in zip(*res):
        _same_shares(o, g)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    return gc, res[1]


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [2, 64, 4096])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("with_sel", [False, True])
def test_aggregate(O, G, keys, width, n, reverse, with_sel):
    rng = np.random.default_rng(n * 4 + width + int(reverse) + 2 * int(with_sel))
    k1 = np.sort(rng.integers(0, max(n // 8, 2), n))
    k2 = rng.integers(0, 2, n)
    sel = rng.integers(0, 2, n) if with_sel else None
    specs = [(rng.integers(-500, 500, n), fn, False) for fn in B_FNS]
    specs += [(rng.integers(0, 1000, n), fn, True) for fn in A_FNS]
    gc, outs = _run(O, G, keys, [k1, k2], specs, reverse, sel, width)
    want = plain_network([k1, k2], specs, reverse, sel, width)
    for (v, fn, arith), o, w in zip(specs, outs, want):
        np.testing.assert_array_equal(gc.open(o), w, err_msg=fn)


def test_aggregate_without_keys(O, G, keys):
    n = 1024
    rng = np.random.default_rng(5)
    specs = [(rng.integers(-(2**31), 2**31, n), "max", False), (rng.integers(0, 9, n), "sum", True)]
    gc, outs = _run(O, G, keys, [], specs, True, None, 4)
    want = plain_network([], specs, True, None, 4)
    for o, w in zip(outs, want):
        np.testing.assert_array_equal(gc.open(o), w)


def test_aggregate_large(O, G, keys):
    """2^16 rows, int64, two keys, sum / count / max: a multi-batch schedule."""
    n = 1 << 16
    rng = np.random.default_rng(6)
    k1 = np.sort(rng.integers(0, 1000, n))
    v = rng.integers(0, 10**6, n)
    specs = [(v, "max", False), (v, "sum", True), (v, "count", True)]
    gc, outs = _run(O, G, keys, [k1], specs, False, None, 8)
    want = plain_network([k1], specs, False, None, 8)
    for o, w in zip(outs, want):
        np.testing.assert_array_equal(gc.open(o), w)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 2, 7, 5000])
def test_distinct(O, G, keys, width, n):
    rng = np.random.default_rng(n + width)
    k1 = np.sort(rng.integers(0, 5, n))
    k2 = rng.integers(0, 2, n)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    outs = []
    for c, is_gpu in ((oc, False), (gc, True)):
        K = [c.secret_share(np.asarray(k).astype(UT[width]), width) for k in (k1, k2)]
        r = c.vec(n, width) if is_gpu else O.Vec(n, width)
        c.distinct(K, r)
        outs.append(r)
    _same_shares(outs[0], outs[1])
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    want = np.ones(n, bool)
    want[1:] = (k1[1:] != k1[:-1]) | (k2[1:] != k2[:-1])
    np.testing.assert_array_equal(gc.open(outs[1]), want.astype(UT[width]))
