"""Oracle pins for the boolean adders and a2b (operators/circuits.h:35-216,
a_shared_vector.h:138-144, replicated_3pc.h:259-262):

* plaintext closed forms: x + y and x - y mod 2^bits for both adders; the
  sign of x - y (mod 2^bits) for rca_compare; a2b keeps the value;
* every pair of boundary values (0, +-1, powers of two and neighbours, MIN,
  MAX) for the adders and rca_compare;
* owner-d secret sharing against the raw keystream of the owner's local PRG:
  s_d, s_{d+1} are its first and second getNext calls, s_{d+2} the rest
  (replicated_3pc.h:384-421);
* the ripple carry adder's shares against its call sequence written out over
  the oracle's and_b on bit-packed words.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_compare import boundary_values

UT = {4: np.uint32, 8: np.uint64}
ST = {4: np.int32, 8: np.int64}


def _rand(n, width, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, n, dtype=np.uint64).astype(UT[width])


@pytest.mark.parametrize("width", [4, 8])
def test_adders_closed_forms(keys, width):
    n = 300
    x, y = _rand(n, width, 1), _rand(n, width, 2)
    c = O.Ctx(keys)
    X, Y = c.secret_share(x, width), c.secret_share(y, width)
    with np.errstate(over="ignore"):
        for op, want in (("add", x + y), ("sub", x - y), ("ppa_add", x + y), ("ppa_sub", x - y),
                         ("rca_lt", ((x - y).view(ST[width]) < 0).astype(UT[width]))):
            np.testing.assert_array_equal(c.open(c.adder(op, X, Y)), want, err_msg=op)
    XA = c.secret_share(x, width, arith=True)
    np.testing.assert_array_equal(c.open(c.adder("a2b", XA)), x)


@pytest.mark.parametrize("width", [4, 8])
def test_adders_boundary_pairs(keys, width):
    v = boundary_values(width).view(UT[width])
    x, y = np.repeat(v, v.size), np.tile(v, v.size)
    c = O.Ctx(keys)
    X, Y = c.secret_share(x, width), c.secret_share(y, width)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(c.open(c.adder("add", X, Y)), x + y)
        np.testing.assert_array_equal(c.open(c.adder("ppa_sub", X, Y)), x - y)
        np.testing.assert_array_equal(c.open(c.adder("rca_lt", X, Y)),
                                      ((x - y).view(ST[width]) < 0).astype(UT[width]))


@pytest.mark.parametrize("owner", [0, 1, 2])
def test_secret_share_owner_layout(keys, owner):
    n, width = 100, 8
    x = _rand(n, width, 3)
    c = O.Ctx(keys)
    v = c.secret_share_from(x, width, owner)
    key = keys[32 * owner: 32 * owner + 32]  # the owner's local PRG
    k0, _ = O.keystream(key, 0, n * width)
    k1, _ = O.keystream(key, 1, n * width)
    s = v.shares()
    np.testing.assert_array_equal(s[owner], k0.view(UT[width]))
    np.testing.assert_array_equal(s[(owner + 1) % 3], k1.view(UT[width]))
    np.testing.assert_array_equal(s[0] ^ s[1] ^ s[2], x)
    assert c.nonce(owner, 0) == 2


def _mk(shares, width):
    v = O.Vec(len(shares[0]), width)
    v.set_shares([np.ascontiguousarray(a, dtype=UT[width]) for a in shares])
    return v


def _pack(bits, width):
    b = 8 * width
    n = bits.size
    nc = n // b + (n % b > 0)
    pad = np.zeros(nc * b, dtype=np.uint64)
    pad[:n] = bits
    return (pad.reshape(nc, b) << np.arange(b, dtype=np.uint64)).sum(axis=1, dtype=np.uint64) \
        .astype(UT[width])


def _unpack(words, n, width):
    b = 8 * width
    w = words.astype(np.uint64)
    return ((w[:, None] >> np.arange(b, dtype=np.uint64)) & np.uint64(1)).reshape(-1)[:n]


@pytest.mark.parametrize("width", [4, 8])
def test_rca_shares_match_call_sequence(keys, width):
    n = 70
    x, y = _rand(n, width, 4), _rand(n, width, 5)
    c1, c2 = O.Ctx(keys), O.Ctx(keys)
    X1, Y1 = c1.secret_share(x, width), c1.secret_share(y, width)
    X2, Y2 = c2.secret_share(x, width), c2.secret_share(y, width)
    z1 = c1.adder("add", X1, Y1)
    b = 8 * width
    xs, ys = X2.shares(), Y2.shares()
    axb = [a ^ d for a, d in zip(xs, ys)]
    nc = n // b + (n % b > 0)
    carry = [np.zeros(nc, UT[width]) for _ in range(3)]
    out = [np.zeros(n, np.uint64) for _ in range(3)]
    for i in range(b):
        axb_i = [_pack(((a >> UT[width](i)) & UT[width](1)).astype(np.uint64), width) for a in axb]
        a_i = [_pack(((a >> UT[width](i)) & UT[width](1)).astype(np.uint64), width) for a in xs]
        for k in range(3):
            out[k] |= _unpack(axb_i[k] ^ carry[k], n, width) << np.uint64(i)
        if i == b - 1:
            break
        carry = [cc ^ a for cc, a in zip(carry, a_i)]
        carry = c2.and_b(_mk(carry, width), _mk(axb_i, width)).shares()
        carry = [cc ^ a for cc, a in zip(carry, a_i)]
    for k in range(3):
        np.testing.assert_array_equal(z1.shares()[k], out[k].astype(UT[width]))
    assert c1.nonce(0, 2) == c2.nonce(0, 2)
    assert c1.msglog() == c2.msglog()
