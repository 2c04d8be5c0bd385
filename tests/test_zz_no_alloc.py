"""orqg_distinct_noalloc (the distinct whose scratch is a fixed static device
array) against orqg_distinct, through the C ABI: two calls in a row, the
second on a smaller input, each compared share for share, with the PRG
nonces and the message log, against orqg_distinct on a second context with
the same keys; and an input over the array's size is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()
    return orq_amd


def _keys_table(G, ctx, n, width, seed):
    rng = np.random.default_rng(seed)
    k1 = rng.integers(0, 5, n)
    k2 = rng.integers(0, 3, n)
    order = np.lexsort((k2, k1))  # sorted rows: distinct flags the first of each group
    dt = np.uint32 if width == 4 else np.uint64
    return [ctx.secret_share(k[order].astype(dt), width) for k in (k1, k2)], k1[order], k2[order]


def _call(G, fn, ctx, cols, res):
    arr = (ctypes.c_void_p * len(cols))(*[c.h for c in cols])
    ctx._check(getattr(G.lib(), fn)(ctx.h, arr, len(cols), res.h), fn)


@pytest.mark.parametrize("width", [4, 8])
def test_distinct_noalloc_matches_distinct(G, width):
    keys = G.keys_from_seed(0x0D15C7)
    a, b = G.Ctx(keys), G.Ctx(keys)
    for call, n in enumerate([(1 << 20) + 3, 1000]):
        cols_a, k1, k2 = _keys_table(G, a, n, width, call)
        cols_b, _, _ = _keys_table(G, b, n, width, call)
        ra, rb = a.vec(n, width), b.vec(n, width)
        _call(G, "orqg_distinct", a, cols_a, ra)
        _call(G, "orqg_distinct_noalloc", b, cols_b, rb)
        for sa, sb in zip(ra.shares(), rb.shares()):
            np.testing.assert_array_equal(sa, sb, err_msg=f"call {call}: shares differ")
        for role in range(10):
            assert a.nonce(role) == b.nonce(role)
        assert a.msglog() == b.msglog()
        want = np.ones(n, dtype=np.uint64)
        want[1:] = (k1[1:] != k1[:-1]) | (k2[1:] != k2[:-1])
        np.testing.assert_array_equal(b.open(rb).astype(np.uint64), want)
        a.open(ra)  # keeps the two message logs in step for the next call


def test_distinct_noalloc_refuses_oversize(G):
    c = G.Ctx(G.keys_from_seed(1))
    n = (1 << 21) + 1
    k = c.vec(n, 8)
    r = c.vec(n, 8)
    arr = (ctypes.c_void_p * 1)(k.h)
    rc = G.lib().orqg_distinct_noalloc(c.h, arr, 1, r.h)
    assert rc == -1
    assert b"48" in G.lib().orqg_last_error(c.h) or b"scratch" in G.lib().orqg_last_error(c.h)
