"""GPU parity of the fused comparison circuits and the full b2a: the shares,
nonces and message log of liborqg against the oracle on the same keys, and the
opened results against plaintext (reference vectors from
tests/test_comparisons.cpp)."""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_nonces, _same_shares
from test_oracle_compare import CMP_X, CMP_Y, DATA_A, DATA_B, OPS, ST, UT, boundary_values

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


def _inputs(n, width, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2**64, n, dtype=np.uint64).astype(UT[width])
    y = rng.integers(0, 2**64, n, dtype=np.uint64).astype(UT[width])
    y[::3] = x[::3]  # equal pairs
    y[1::7] = x[1::7] ^ UT[width](1 << (8 * width - 1))  # differ only in the sign bit
    y[2::11] = x[2::11] ^ UT[width](1)  # differ only in the LSB
    return x, y


def _both(O, G, keys, x, y, width):
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    return oc, gc, (oc.secret_share(x, width), oc.secret_share(y, width)), (
        gc.secret_share(x, width), gc.secret_share(y, width))


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 65, 3000, 70001])
@pytest.mark.parametrize("op", list(OPS))
def test_cmp_ops(O, G, keys, width, n, op):
    x, y = _inputs(n, width, n * 7 + width)
    oc, gc, (OX, OY), (GX, GY) = _both(O, G, keys, x, y, width)
    oz, gz = oc.cmp(op, OX, OY), gc.cmp(op, GX, GY)
    _same_shares(oz, gz)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    want = OPS[op](x.view(ST[width]), y.view(ST[width])).astype(UT[width])
    np.testing.assert_array_equal(gc.open(gz), want)


@pytest.mark.parametrize("width", [4, 8])
def test_compare_reference_vectors(O, G, keys, width):
    x, y = np.array(CMP_X, ST[width]), np.array(CMP_Y, ST[width])
    oc, gc, (OX, OY), (GX, GY) = _both(O, G, keys, x.view(UT[width]), y.view(UT[width]), width)
    (oe, og), (ge, gg) = oc.compare(OX, OY), gc.compare(GX, GY)
    _same_shares(oe, ge)
    _same_shares(og, gg)
    _same_nonces(oc, gc)
    np.testing.assert_array_equal(gc.open(ge), (x == y).astype(UT[width]))
    np.testing.assert_array_equal(gc.open(gg), (x > y).astype(UT[width]))
    a, b = np.array(DATA_A, ST[width]), np.array(DATA_B, ST[width])
    GA, GB = gc.secret_share(a.view(UT[width]), width), gc.secret_share(b.view(UT[width]), width)
    for op, f in OPS.items():
        np.testing.assert_array_equal(gc.open(gc.cmp(op, GA, GB)), f(a, b).astype(UT[width]))


@pytest.mark.parametrize("width", [4, 8])
def test_compare_boundary_pairs(O, G, keys, width):
    v = boundary_values(width)
    x, y = np.repeat(v, v.size), np.tile(v, v.size)
    oc, gc, (OX, OY), (GX, GY) = _both(O, G, keys, x.view(UT[width]), y.view(UT[width]), width)
    (oe, og), (ge, gg) = oc.compare(OX, OY), gc.compare(GX, GY)
    _same_shares(oe, ge)
    _same_shares(og, gg)
    np.testing.assert_array_equal(gc.open(gg), (x > y).astype(UT[width]))


def test_compare_multi_chunk_batches(O, G, keys):
    """n large enough that each runtime batch spans several 1 MiB PRG chunks."""
    n = (1 << 21) + 5
    x, y = _inputs(n, 8, 11)
    oc, gc, (OX, OY), (GX, GY) = _both(O, G, keys, x, y, 8)
    (oe, og), (ge, gg) = oc.compare(OX, OY), gc.compare(GX, GY)
    _same_shares(oe, ge)
    _same_shares(og, gg)
    _same_nonces(oc, gc)
    _same_log(oc, gc)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_b2a(O, G, keys, width, n):
    rng = np.random.default_rng(n + width)
    x = rng.integers(0, 2**64, n, dtype=np.uint64).astype(UT[width])
    x[:1] = UT[width](1 << (8 * width - 1))
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ox, gx = oc.secret_share(x, width), gc.secret_share(x, width)
    oy, gy = oc.b2a(ox), gc.b2a(gx)
    _same_shares(oy, gy)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    np.testing.assert_array_equal(gc.open(gy), x)
