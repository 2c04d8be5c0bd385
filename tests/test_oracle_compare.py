"""The oracle's comparison circuits and full b2a, pinned three ways:

* opened results against plaintext, on the reference's own test vectors
  (tests/test_comparisons.cpp:15-16 and 111-114) and on every pair of a set
  of boundary values (0, +-1, powers of two and their neighbours, INT_MIN,
  INT_MAX);
* the shares themselves against the same circuit written out call by call in
  numpy over the oracle's and_b / b2a_bit / multiply_a primitives, following
  the reference's text (b_shared_vector.h:164-178 and 296-329,
  circuits.h:231-295): a wrong shift kind, level count, call order, packing
  bit order or party role in the fused C++ code changes the shares;
* the PRG nonces consumed, against the closed form of the call schedule.
"""
import numpy as np
import pytest

from oracle import oracle as O

INT_MIN, INT_MAX = -(2**31), 2**31 - 1
# tests/test_comparisons.cpp:15-16
DATA_A = [111, -4, -17, 2345, 999, 0, -28922, 1231241, 0, -23437]
DATA_B = [0, -4, -5, INT_MIN, 999, 70, -243242, INT_MAX, 0, 78]
# tests/test_comparisons.cpp:111-114
CMP_X = [30670, 0, -14730, 25788, -50441, INT_MAX, -19964, INT_MAX, -13582, -64520, 27218, 0,
         34184, INT_MIN, -65282, INT_MIN]
CMP_Y = [-15260, 0, -32847, 30024, -6530, INT_MAX, -1536, 29620, -26019, -64520, INT_MIN,
         -15059, -34111, -7765, 56837, INT_MIN]

UT = {4: np.uint32, 8: np.uint64}
ST = {4: np.int32, 8: np.int64}
OPS = {"==": np.equal, "!=": np.not_equal, "<": np.less, ">": np.greater,
       "<=": np.less_equal, ">=": np.greater_equal}


def boundary_values(width):
    b = 8 * width
    vals = {0, 1, -1, -(2 ** (b - 1)), 2 ** (b - 1) - 1}
    for k in range(b - 1):
        for d in (-1, 0, 1):
            vals.add(2**k + d)
            vals.add(-(2**k) + d)
    vals = [v for v in vals if -(2 ** (b - 1)) <= v < 2 ** (b - 1)]
    return np.array(sorted(vals), dtype=ST[width])


@pytest.mark.parametrize("width", [4, 8])
def test_reference_vectors(keys, width):
    c = O.Ctx(keys)
    a, b = np.array(DATA_A, ST[width]), np.array(DATA_B, ST[width])
    X, Y = c.secret_share(a.view(UT[width]), width), c.secret_share(b.view(UT[width]), width)
    for op, f in OPS.items():
        np.testing.assert_array_equal(c.open(c.cmp(op, X, Y)), f(a, b).astype(UT[width]), op)
    x, y = np.array(CMP_X, ST[width]), np.array(CMP_Y, ST[width])
    X, Y = c.secret_share(x.view(UT[width]), width), c.secret_share(y.view(UT[width]), width)
    eq, gt = c.compare(X, Y)
    np.testing.assert_array_equal(c.open(eq), (x == y).astype(UT[width]))
    np.testing.assert_array_equal(c.open(gt), (x > y).astype(UT[width]))


@pytest.mark.parametrize("width", [4, 8])
def test_all_boundary_pairs(keys, width):
    v = boundary_values(width)
    x, y = np.repeat(v, v.size), np.tile(v, v.size)
    c = O.Ctx(keys)
    X, Y = c.secret_share(x.view(UT[width]), width), c.secret_share(y.view(UT[width]), width)
    eq, gt = c.compare(X, Y)
    np.testing.assert_array_equal(c.open(eq), (x == y).astype(UT[width]))
    np.testing.assert_array_equal(c.open(gt), (x > y).astype(UT[width]))
    for op in ("<", ">="):
        np.testing.assert_array_equal(c.open(c.cmp(op, X, Y)), OPS[op](x, y).astype(UT[width]))


# ---- the circuits written out call by call -----------------------------------------
def _mk(s, width):
    v = O.Vec(len(s[0]), width)
    v.set_shares([np.ascontiguousarray(a, dtype=UT[width]) for a in s])
    return v


def _sar(a, k, width):
    return (a.view(ST[width]) >> k).view(UT[width])


def _bit_same(c, X, Y, width):
    e = [~(a ^ b) for a, b in zip(X.shares(), Y.shares())]
    levels = {4: 5, 8: 6}[width]
    for l in range(levels):
        t = [_sar(a, 1 << l, width) for a in e]
        e = c.and_b(_mk(e, width), _mk(t, width)).shares()
    return e


def _pack(bits, width):
    b = 8 * width
    n = bits.size
    nc = n // b + (n % b > 0)
    pad = np.zeros(nc * b, dtype=np.uint64)
    pad[:n] = bits
    w = (pad.reshape(nc, b) << np.arange(b, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    return w.astype(UT[width])


def _unpack(words, n, width):
    b = 8 * width
    w = words.astype(np.uint64)
    bits = (w[:, None] >> np.arange(b, dtype=np.uint64)) & np.uint64(1)
    return bits.reshape(-1)[:n].astype(UT[width])


def _popcount_parity(a):
    a = a.astype(np.uint64)
    p = np.zeros_like(a)
    for i in range(64):
        p ^= (a >> np.uint64(i)) & np.uint64(1)
    return p


def _compare_by_calls(c, X, Y, width):
    b = 8 * width
    n = X.n
    e = _bit_same(c, X, Y, width)
    g = [_sar(a, 1, width) ^ a for a in e]
    g = c.and_b(_mk(g, width), X).shares()
    g = [_popcount_parity(a).astype(UT[width]) for a in g]
    msb = lambda a: ((a >> (b - 1)) & 1).astype(np.uint64)
    s1 = [_pack(msb(a), width) for a in X.shares()]
    s2 = [_pack(msb(a), width) for a in Y.shares()]
    r = [_pack((a & 1).astype(np.uint64), width) for a in g]
    r = [a ^ d for a, d in zip(r, s2)]
    s2 = [a ^ d for a, d in zip(s2, s1)]
    s2 = [~a for a in s2]
    r = [~a for a in r]
    s2 = c.and_b(_mk(s2, width), _mk(r, width)).shares()
    s2 = [~a for a in s2]
    s1 = [a ^ d for a, d in zip(s1, s2)]
    gt = [_unpack(a, n, width) for a in s1]
    eq = [a & 1 for a in e]
    return eq, gt


def _nonce_sched(n, width, calls=1, batch_size=-12):
    total = 0
    for size in O.batches(n, batch_size):
        total += calls * ((size * width + (1 << 20) - 1) >> 20)
    return total


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 3000])
def test_compare_shares_match_call_sequence(keys, width, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2**63, n, dtype=np.uint64).astype(UT[width])
    y = rng.integers(0, 2**63, n, dtype=np.uint64).astype(UT[width])
    y[::3] = x[::3]
    y[1::5] = x[1::5] ^ UT[width](1 << (8 * width - 1))
    c1, c2 = O.Ctx(keys), O.Ctx(keys)
    X1, Y1 = c1.secret_share(x, width), c1.secret_share(y, width)
    X2, Y2 = c2.secret_share(x, width), c2.secret_share(y, width)
    before = c1.nonce(0, 2)
    eq1, gt1 = c1.compare(X1, Y1)
    eq2, gt2 = _compare_by_calls(c2, X2, Y2, width)
    for k in range(3):
        np.testing.assert_array_equal(eq1.shares()[k], eq2[k])
        np.testing.assert_array_equal(gt1.shares()[k], gt2[k])
    for p in range(3):
        for role in range(3 + 4):
            assert c1.nonce(p, role) == c2.nonce(p, role)
    levels = {4: 5, 8: 6}[width]
    nc = n // (8 * width) + (n % (8 * width) > 0)
    assert c1.nonce(0, 2) - before == (levels + 1) * _nonce_sched(n, width) + _nonce_sched(nc, width)
    assert c1.msglog() == c2.msglog()


@pytest.mark.parametrize("width", [4, 8])
def test_equality_shares_match_call_sequence(keys, width):
    n = 1000
    rng = np.random.default_rng(7)
    x = rng.integers(0, 2**63, n, dtype=np.uint64).astype(UT[width])
    y = x.copy()
    y[::2] ^= UT[width](1) << rng.integers(0, 8 * width, n // 2).astype(UT[width])
    c1, c2 = O.Ctx(keys), O.Ctx(keys)
    X1, Y1 = c1.secret_share(x, width), c1.secret_share(y, width)
    X2, Y2 = c2.secret_share(x, width), c2.secret_share(y, width)
    z = c1.cmp("==", X1, Y1).shares()
    e = _bit_same(c2, X2, Y2, width)
    for k in range(3):
        np.testing.assert_array_equal(z[k], e[k] & 1)
    ne = c1.cmp("!=", X1, Y1).shares()
    e = _bit_same(c2, X2, Y2, width)
    for k in range(3):  # not_b_1 on every share: !(s & 1)
        np.testing.assert_array_equal(ne[k], (e[k] & 1) ^ 1)
    np.testing.assert_array_equal(c1.open(c1.cmp("==", X1, Y1)), (x == y).astype(UT[width]))


@pytest.mark.parametrize("width", [4, 8])
def test_b2a_matches_call_sequence(keys, width):
    n = 700
    rng = np.random.default_rng(3)
    x = rng.integers(0, 2**64, n, dtype=np.uint64).astype(UT[width])
    x[:4] = [0, 1, UT[width](-1 % 2 ** (8 * width)), UT[width](1 << (8 * width - 1))]
    c1, c2 = O.Ctx(keys), O.Ctx(keys)
    X1, X2 = c1.secret_share(x, width), c2.secret_share(x, width)
    y1 = c1.b2a(X1)
    res = [np.zeros(n, UT[width]) for _ in range(3)]
    for i in range(8 * width):
        bit = [(a >> UT[width](i)) & UT[width](1) for a in X2.shares()]
        ba = c2.b2a_bit(_mk(bit, width))
        if i > 0:
            pub = [np.full(n, 1 << i, UT[width]), np.zeros(n, UT[width]), np.zeros(n, UT[width])]
            ba = c2.mul_a(ba, _mk(pub, width))
        res = [r + a for r, a in zip(res, ba.shares())]
    for k in range(3):
        np.testing.assert_array_equal(y1.shares()[k], res[k])
    assert c1.nonce(1, 2) == c2.nonce(1, 2)
    assert c1.msglog() == c2.msglog()
    np.testing.assert_array_equal(c1.open(y1, arith=True), x)
