"""GPU parity of the boolean adders, rca_compare, a2b and owner-d secret
sharing: shares, nonces and message log against the oracle, plus the
plaintext closed forms."""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_nonces, _same_shares
from test_oracle_adder import ST, UT, _rand

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


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 5000, 70001])
@pytest.mark.parametrize("op", ["add", "sub", "ppa_add", "ppa_sub", "rca_lt"])
def test_adders(O, G, keys, width, n, op):
    x, y = _rand(n, width, n + 1), _rand(n, width, n + 2)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    oz = oc.adder(op, oc.secret_share(x, width), oc.secret_share(y, width))
    gz = gc.adder(op, gc.secret_share(x, width), gc.secret_share(y, width))
    _same_shares(oz, gz)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    with np.errstate(over="ignore"):
        want = {"add": x + y, "sub": x - y, "ppa_add": x + y, "ppa_sub": x - y,
                "rca_lt": ((x - y).view(ST[width]) < 0).astype(UT[width])}[op]
    np.testing.assert_array_equal(gc.open(gz), want)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 1000, 70001])
def test_a2b(O, G, keys, width, n):
    x = _rand(n, width, n + 7)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    oz = oc.adder("a2b", oc.secret_share(x, width, arith=True))
    gz = gc.adder("a2b", gc.secret_share(x, width, arith=True))
    _same_shares(oz, gz)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    np.testing.assert_array_equal(gc.open(gz), x)


@pytest.mark.parametrize("owner", [0, 1, 2])
@pytest.mark.parametrize("arith", [False, True])
def test_secret_share_from(O, G, keys, owner, arith):
    x = _rand(3000, 8, owner)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ov = oc.secret_share_from(x, 8, owner, arith=arith)
    gv = gc.secret_share_from(x, 8, owner, arith=arith)
    _same_shares(ov, gv)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    np.testing.assert_array_equal(gc.open(gv), x)
