"""The batched radix sort (orqg_radix_sort_batch): many requests, each on its
own context, sorted with shared segmented launches.  A batch must give every
request exactly what orqg_radix_sort on its context gives alone -- shares of
the sorted vector and of the permutation, every PRG nonce, the permutation
queue and the message log -- whatever the launch shapes the tuner picks; one
small batch is also checked against the oracle directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()  # fails loudly when the HIP library is missing
    return orq_amd


def _keys(i):
    from orq_amd import keys_from_seed

    return keys_from_seed(0x0BA7C400 + i)


def _plain(rng, n, width):
    if width == 4:
        k = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    else:
        k = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64).view(np.uint64)
    if n > 8:
        k[3] = k[7]  # a duplicate
    return k


def _nonces(c):
    return [c.nonce(r) for r in range(10)]


def _run(G, reqs, seed=0):
    """reqs: (n, width, bits or None, desc).  Runs them batched and one by
    one on twin contexts; compares everything and returns the batch's
    contexts."""
    rng = np.random.default_rng(seed)
    plains = [_plain(rng, n, w) for n, w, _, _ in reqs]
    bc = [G.Ctx(_keys(i)) for i in range(len(reqs))]
    sc = [G.Ctx(_keys(i)) for i in range(len(reqs))]
    bv = [c.secret_share(p, w) for c, p, (_, w, _, _) in zip(bc, plains, reqs)]
    sv = [c.secret_share(p, w) for c, p, (_, w, _, _) in zip(sc, plains, reqs)]
    bits = [b if b is not None else 8 * w for _, w, b, _ in reqs]
    desc = [d for _, _, _, d in reqs]
    bperm = G.radix_sort_batch(bv, bits=bits, desc=desc)
    for q, (c, v) in enumerate(zip(sc, sv)):
        sperm = c.radix_sort(v, bits=bits[q], desc=desc[q])
        n, w = reqs[q][0], reqs[q][1]
        for k, (x, y) in enumerate(zip(bv[q].shares(), v.shares())):
            np.testing.assert_array_equal(x, y, err_msg=f"request {q} ({n} rows): v share {k}")
        for k, (x, y) in enumerate(zip(bperm[q].shares(), sperm.shares())):
            np.testing.assert_array_equal(x, y, err_msg=f"request {q} ({n} rows): perm share {k}")
        assert _nonces(bc[q]) == _nonces(c), f"request {q}: nonces"
        assert bc[q].msglog() == c.msglog(), f"request {q}: message log"
        assert bc[q].queue_size() == c.queue_size()
        # and the plaintext result: stable sort on the low `bits` bits
        if n:
            sdt = np.int32 if w == 4 else np.int64
            got = bc[q].open(bv[q]).view(sdt)
            if bits[q] == 8 * w:
                key = plains[q].view(sdt).astype(np.int64) if w == 4 else plains[q].view(sdt)
                order = np.argsort(~key if desc[q] else key, kind="stable")
            else:
                low = (plains[q].astype(np.uint64) & np.uint64((1 << bits[q]) - 1)).astype(np.int64)
                order = np.argsort(-low if desc[q] else low, kind="stable")
            np.testing.assert_array_equal(got, plains[q].view(sdt)[order], err_msg=f"request {q}")
    return bc


def test_batch_mixed_sizes_widths_orders(G):
    """Sizes from 1 row to several superblocks, both record layouts, full and
    partial key widths (requests leave the shared launches at different bits),
    ascending and descending, and an empty request."""
    _run(G, [(1, 8, None, False), (3, 4, None, True), (1000, 8, None, False),
             (4097, 8, 20, True), (0, 8, None, False), (777, 4, None, False),
             (5000, 4, 9, False), (12289, 8, None, True)])


def test_batch_with_a_large_request(G):
    """A request above the small-permutation size (its permutations come from
    the windowed generator) beside small ones."""
    _run(G, [(70001, 8, None, False), (2048, 8, None, True), (100, 4, None, False)], seed=3)


def test_batch_many_small_requests_twice(G, monkeypatch):
    """Forty requests of 2^10 rows, twice: the second call runs on the launch
    shapes the first one tuned; then once more with tuning off."""
    reqs = [(1024, 8, None, i % 2 == 1) for i in range(40)]
    _run(G, reqs, seed=5)
    _run(G, reqs, seed=6)
    monkeypatch.setenv("ORQG_TUNE", "0")
    _run(G, reqs[:7], seed=7)


def test_batch_against_oracle(G):
    """Two requests of one batch against the oracle's radix_sort on the same keys."""
    from oracle import oracle as O

    rng = np.random.default_rng(9)
    plains = [_plain(rng, 300, 8), _plain(rng, 1025, 4)]
    widths = [8, 4]
    gc = [G.Ctx(_keys(10 + i)) for i in range(2)]
    oc = [O.Ctx(_keys(10 + i), -12) for i in range(2)]
    gv = [c.secret_share(p, w) for c, p, w in zip(gc, plains, widths)]
    ov = [c.secret_share(p, w) for c, p, w in zip(oc, plains, widths)]
    gperm = G.radix_sort_batch(gv, desc=[False, True])
    for q in range(2):
        operm = oc[q].radix_sort(ov[q], desc=q == 1)
        for k, (x, y) in enumerate(zip(ov[q].shares(), gv[q].shares())):
            np.testing.assert_array_equal(x, y, err_msg=f"request {q}: v share {k}")
        for k, (x, y) in enumerate(zip(operm.shares(), gperm[q].shares())):
            np.testing.assert_array_equal(x, y, err_msg=f"request {q}: perm share {k}")
        assert oc[q].msglog() == gc[q].msglog()


def test_batch_rejects_bad_arguments(G):
    c = G.Ctx(_keys(20))
    v = c.secret_share(np.arange(10, dtype=np.uint64), 8)
    w = c.secret_share(np.arange(10, dtype=np.uint64), 8)
    with pytest.raises(G.OrqgError, match="appears twice"):
        G.radix_sort_batch([v, w])
    with pytest.raises(G.OrqgError, match="bits"):
        G.radix_sort_batch([v], bits=[65])
