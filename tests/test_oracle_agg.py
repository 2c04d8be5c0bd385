"""Oracle pins for the aggregation network (operators/aggregation.h:27-346)
and distinct (operators/distinct.h:19-44):

* opened results against a plaintext restatement of the network in numpy
  (level by level, every level reading the values before the level), for
  every aggregation function, both directions, with and without the join
  selection column;
* closed forms: forward leaves sum / count / max / min of each contiguous
  group in its first row, reverse in its last row
  (tests/test_operators.cpp:59-200); with no keys a reverse max is a running
  maximum (test_operators.cpp:440-466);
* the literal multi-key distinct of test_operators.cpp:239-257;
* the shares of one aggregation against the same sequence of runtime calls
  written out over the oracle's primitives (call order, operand order and
  which group bits each function uses).
"""
import numpy as np
import pytest

from oracle import oracle as O

UT = {4: np.uint32, 8: np.uint64}
ST = {4: np.int32, 8: np.int64}


def plain_network(keys, specs, reverse=False, sel=None, width=4):
    """specs: list of (values, fn, arith); returns the new columns."""
    n = len(specs[0][0]) if specs else len(keys[0])
    ut, st = UT[width], ST[width]
    outs = []
    for v, fn, arith in specs:
        if fn == "count":
            v = np.asarray(sel, ut) if sel is not None else np.ones(n, ut)
        outs.append(np.asarray(v).astype(ut).copy())
    log = n.bit_length() - 1
    for lv in range(1, log + 1):
        d = n >> lv if not reverse else n >> (log - lv + 1)
        m = n - d
        g = np.ones(m, bool)
        for k in keys:
            k = np.asarray(k)
            g &= k[:m] == k[d:]
        jg = g.copy()
        any_noncopy = any(fn != "copy" for _, fn, _ in specs)
        if sel is not None and any_noncopy:
            s = np.asarray(sel)
            g &= s[d:] == s[:m]
        for i, (_, fn, arith) in enumerate(specs):
            o = outs[i]
            acc, inp = (o[d:], o[:m]) if reverse else (o[:m], o[d:])
            gg = jg if fn == "copy" else g
            if fn in ("sum", "count"):
                new = acc + gg.astype(ut) * inp
            elif fn == "copy":
                new = np.where(gg, inp, acc)
            elif fn == "max":
                new = np.where(gg & (inp.view(st) > acc.view(st)), inp, acc)
            elif fn == "min":
                new = np.where(gg & ~(inp.view(st) > acc.view(st)), inp, acc)
            elif fn == "or":
                new = acc | np.where(gg, inp, 0).astype(ut)
            elif fn == "valid":
                new = np.where(gg, acc & inp, acc)
            new = new.astype(ut)
            if reverse:
                o[d:] = new
            else:
                o[:m] = new
    return outs


B_FNS = ["copy", "max", "min", "or", "valid"]
A_FNS = ["sum", "count", "copy"]


def run_oracle(c, keys, specs, reverse=False, sel=None, width=4):
    K = [c.secret_share(np.asarray(k).astype(UT[width]), width) for k in keys]
    S = c.secret_share(np.asarray(sel).astype(UT[width]), width) if sel is not None else None
    bs, as_, outs = [], [], []
    for v, fn, arith in specs:
        inp = c.secret_share(np.asarray(v).astype(UT[width]), width, arith=arith)
        out = O.Vec(len(v), width)
        (as_ if arith else bs).append((inp, out, fn))
        outs.append((out, arith))
    c.aggregate(K, bs, as_, reverse, S)
    # outputs in spec order
    return [c.open(o, arith=a) for o, a in outs]


def _specs(rng, n, width, fns=None):
    vals = rng.integers(-50, 50, n)
    out = []
    for fn in (fns or B_FNS):
        out.append((vals if fn != "or" else rng.integers(0, 2**20, n), fn, False))
    for fn in (fns or A_FNS):
        if fn in A_FNS:
            out.append((rng.integers(0, 1000, n), fn, True))
    return out


@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("with_sel", [False, True])
@pytest.mark.parametrize("width", [4, 8])
def test_network_matches_plaintext(keys, reverse, with_sel, width):
    n = 64
    rng = np.random.default_rng(int(reverse) * 2 + int(with_sel) + width)
    k1 = np.sort(rng.integers(0, 6, n))
    k2 = rng.integers(0, 2, n)
    sel = rng.integers(0, 2, n) if with_sel else None
    bspecs = [(rng.integers(-50, 50, n), fn, False) for fn in B_FNS]
    aspecs = [(rng.integers(0, 1000, n), fn, True) for fn in A_FNS]
    specs = bspecs + aspecs
    c = O.Ctx(keys)
    got = run_oracle(c, [k1, k2], specs, reverse, sel, width)
    want = plain_network([k1, k2], specs, reverse, sel, width)
    for (v, fn, arith), g, w in zip(specs, got, want):
        np.testing.assert_array_equal(g, w, err_msg=f"{fn} arith={arith}")


@pytest.mark.parametrize("width", [4, 8])
def test_group_closed_forms(keys, width):
    n = 128
    rng = np.random.default_rng(11)
    grp = np.sort(rng.integers(0, 9, n))
    v = rng.integers(-1000, 1000, n)
    specs = [(v, "max", False), (v, "min", False), (v, "sum", True), (v, "count", True)]
    c = O.Ctx(keys)
    fw = run_oracle(c, [grp], specs, False, None, width)
    rv = run_oracle(c, [grp], specs, True, None, width)
    st = ST[width]
    for gid in np.unique(grp):
        rows = np.where(grp == gid)[0]
        vals = v[rows]
        for res, row in ((fw, rows[0]), (rv, rows[-1])):
            assert res[0].view(st)[row] == vals.max()
            assert res[1].view(st)[row] == vals.min()
            assert res[2].view(st)[row] == vals.sum()
            assert res[3][row] == len(rows)


def test_reverse_max_without_keys_is_running_max(keys):
    # tests/test_operators.cpp:440-466
    rng = np.random.default_rng(12)
    x = rng.integers(-(2**31), 2**31, 256)
    c = O.Ctx(keys)
    (y,) = run_oracle(c, [], [(x, "max", False)], reverse=True)
    y = y.view(np.int32)
    assert np.all(np.diff(y) >= 0)
    np.testing.assert_array_equal(y, np.maximum.accumulate(x.astype(np.int32)))


def test_reference_multi_distinct(keys):
    # tests/test_operators.cpp:239-257
    k1 = [0, 0, 0, 1, 1, 1, 2]
    k2 = [0, 0, 1, 1, 1, 0, 2]
    c = O.Ctx(keys)
    K1 = c.secret_share(np.asarray(k1, np.uint32), 4)
    K2 = c.secret_share(np.asarray(k2, np.uint32), 4)
    for ks, want in (([K1], [1, 0, 0, 1, 0, 0, 1]), ([K2], [1, 0, 1, 0, 0, 1, 1]),
                     ([K1, K2], [1, 0, 1, 1, 0, 1, 1])):
        r = O.Vec(7, 4)
        c.distinct(ks, r)
        np.testing.assert_array_equal(c.open(r), want)


# ---- the call sequence written out over the oracle primitives ---------------------
def _mk(shares, width=4):
    v = O.Vec(len(shares[0]), width)
    v.set_shares([np.ascontiguousarray(a, dtype=UT[width]) for a in shares])
    return v


def _cut(v, start, length, width=4):
    return _mk([s[start:start + length] for s in v.shares()], width)


def _ext(sh):  # extend_lsb per share
    return [(0 - (a & np.uint32(1))).astype(np.uint32) for a in sh]


def _mux_b(c, sel, a, b):  # a ^ (extend_lsb(sel) & (b ^ a))
    d = [x ^ y for x, y in zip(b.shares(), a.shares())]
    t = c.and_b(_mk(_ext(sel.shares())), _mk(d)).shares()
    return _mk([x ^ y for x, y in zip(a.shares(), t)])


def test_aggregate_shares_match_call_sequence(keys):
    n, width = 16, 4
    rng = np.random.default_rng(13)
    k = np.sort(rng.integers(0, 4, n)).astype(np.uint32)
    sel = rng.integers(0, 2, n).astype(np.uint32)
    vb = rng.integers(0, 100, n).astype(np.uint32)
    va = rng.integers(0, 100, n).astype(np.uint32)
    c1, c2 = O.Ctx(keys), O.Ctx(keys)
    vecs = []
    for c in (c1, c2):
        vecs.append([c.secret_share(k, 4), c.secret_share(sel, 4), c.secret_share(vb, 4),
                     c.secret_share(va, 4, arith=True)])
    K, SEL, VB, VA = vecs[0]
    OB, OA = O.Vec(n, 4), O.Vec(n, 4)
    c1.aggregate([K], [(VB, OB, "max")], [(VA, OA, "sum")], reverse=False, sel=SEL)

    K, SEL, VB, VA = vecs[1]
    ob, oa = VB, VA  # out = in
    for lv in range(1, 5):
        d = n >> lv
        m = n - d
        g = c2.cmp("==", _cut(K, 0, m), _cut(K, d, m))
        s = [~(x ^ y) for x, y in zip(_cut(SEL, d, m).shares(), _cut(SEL, 0, m).shares())]
        g = c2.and_b(g, _mk(s))
        g = _mk([x & np.uint32(1) for x in g.shares()])
        a, b = _cut(ob, 0, m), _cut(ob, d, m)
        _, bg = c2.compare(b, a)  # b > a
        inner = _mux_b(c2, bg, a, b)
        new = _mux_b(c2, g, a, inner)
        ob = _mk([np.concatenate([x, y[m:]]) for x, y in zip(new.shares(), ob.shares())])
        ga = c2.b2a_bit(g)
        a, b = _cut(oa, 0, m), _cut(oa, d, m)
        t = c2.mul_a(ga, b)
        new = [x + y for x, y in zip(a.shares(), t.shares())]
        oa = _mk([np.concatenate([x, y[m:]]) for x, y in zip(new, oa.shares())])
    for x, y in zip(OB.shares(), ob.shares()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(OA.shares(), oa.shares()):
        np.testing.assert_array_equal(x, y)
    assert c1.nonce(0, 2) == c2.nonce(0, 2)
    assert c1.msglog() == c2.msglog()
