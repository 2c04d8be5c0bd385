"""The five workload configurations, and the headline sort, at their stated
sizes on the GPU.

C1  radix sort of 2^16 boolean-shared keys drawn uniform on [0, 2^32)
    (splitmix64 seed 0x0A11CE01, int32 shares, so ordered as signed int32),
    no payload: shares, PRG nonces and message log against the oracle; the
    opened keys against the plaintext stable sort and the opened permutation
    against its ranks (the reference's check, tests/test_sorting.cpp:134-174).
C2  primitives at 2^24 boolean-shared uint64: xor_b and and_b shares against
    the oracle; xor, and, ==, <, full b2a against plaintext closed forms; the
    keystream of 2^24 64-bit words (128 nonces) against the oracle's AES-CTR.
C3  the table {[K] Zipf(1.1) int64 over [0, 2^20), P1 and [P2] uniform
    uint32, [##V]} x 2^24: shuffle, then sort on [K] by radix sort and, on a
    fresh copy, by bitonic sort: keys in order and every (K, P1, P2) row kept.
C4  R (2^20 rows, unique keys) joined with L (2^22 rows, keys uniform on
    [0, 2^20)) copying R's attribute, then GROUP BY key SUM, COUNT, MAX:
    against SQLite on the same plaintext.
C5  a served stream of requests of 2^10 .. 2^22 rows with Zipf keys, four
    slots: every status OK, each result's shares equal to the same request
    served alone, and the plaintext answer right.
M   the headline: radix sort of 2^24 int64 Zipf keys, against the plaintext
    stable sort and its ranks.
"""
import numpy as np
import pytest

from test_gpu_parity import _same_log, _same_nonces, _same_shares

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()  # fails loudly when the HIP library is missing
    return orq_amd


def zipf(n, seed, s=1.1, domain=1 << 20):
    """Zipf(s) rank - 1 over [0, domain) by inversion of the exact CDF,
    uniforms from splitmix64."""
    from orq_amd import splitmix64

    cdf = np.cumsum(np.arange(1, domain + 1, dtype=np.float64) ** -s)
    cdf /= cdf[-1]
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return np.searchsorted(cdf, u, side="right").astype(np.int64)


def _ranks(order):
    r = np.empty(order.size, dtype=np.int64)
    r[order] = np.arange(order.size)
    return r


def test_c1_radix_sort_2p16_uint32(O, G):
    from orq_amd import keys_from_seed, splitmix64

    n = 1 << 16
    keys = keys_from_seed(0x0A11CE01)
    x = (splitmix64(0x0A11CE01, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ov, gv = oc.secret_share(x, 4), gc.secret_share(x, 4)
    operm, gperm = oc.radix_sort(ov), gc.radix_sort(gv)
    _same_shares(ov, gv)
    _same_shares(operm, gperm)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    assert gc.queue_size() == 0
    signed = x.view(np.int32)
    order = np.argsort(signed, kind="stable")
    np.testing.assert_array_equal(gc.open(gv).view(np.int32), signed[order])
    np.testing.assert_array_equal(gc.open(gperm).view(np.int32), _ranks(order))


def test_c2_primitives_2p24_uint64(O, G):
    from orq_amd import keys_from_seed, splitmix64

    n = 1 << 24
    keys = keys_from_seed(0x0A11CE02)
    x = splitmix64(0x0A11CE02, n)
    y = splitmix64(0x0A11CE12, n)
    y[::7] = x[::7]  # equal pairs for == and <
    y[1::11] = x[1::11] ^ np.uint64(1)  # pairs that differ in the last bit only
    oc, gc = O.Ctx(keys), G.Ctx(keys)
    ox, gx = oc.secret_share(x, 8), gc.secret_share(x, 8)
    oy, gy = oc.secret_share(y, 8), gc.secret_share(y, 8)
    _same_shares(ox, gx)
    oz, gx_y = oc.xor_b(ox, oy), gc.xor_b(gx, gy)
    _same_shares(oz, gx_y)
    del oz
    oz, gz = oc.and_b(ox, oy), gc.and_b(gx, gy)
    _same_shares(oz, gz)
    _same_nonces(oc, gc)
    _same_log(oc, gc)  # before any open, which only the GPU side does below
    del ox, oy, oz, oc
    np.testing.assert_array_equal(gc.open(gx_y), x ^ y)
    np.testing.assert_array_equal(gc.open(gz), x & y)
    sx, sy = x.view(np.int64), y.view(np.int64)
    np.testing.assert_array_equal(gc.open(gc.cmp("==", gx, gy)), (x == y).astype(np.uint64))
    np.testing.assert_array_equal(gc.open(gc.cmp("<", gx, gy)), (sx < sy).astype(np.uint64))
    np.testing.assert_array_equal(gc.open(gc.b2a(gx)), x)  # mod 2^64
    # the keystream of 2^24 64-bit words on the pair PRG of party 1
    n0 = gc.nonce(4)
    ks = gc.keystream(4, 8 * n)
    want, nxt = O.keystream(keys[4 * 32:5 * 32], n0, 8 * n)
    np.testing.assert_array_equal(ks, want)
    assert gc.nonce(4) == nxt == n0 + 128


def _rows(k, p1, p2):
    order = np.lexsort((p2, p1, k))
    return np.stack([k[order], p1[order], p2[order]])


@pytest.mark.parametrize("proto", ["RADIXSORT", "BITONICSORT"])
def test_c3_table_shuffle_sort_2p24(G, proto):
    from orq_amd import keys_from_seed, splitmix64
    from orq_amd.table import EncodedTable

    n = 1 << 24
    k = zipf(n, 0x0A11CE03)
    p1 = (splitmix64(0x0A11CE13, n) >> np.uint64(32)).astype(np.int64)
    p2 = (splitmix64(0x0A11CE23, n) >> np.uint64(32)).astype(np.int64)
    gc = G.Ctx(keys_from_seed(0x0A11CE03))
    t = EncodedTable.secret_share(gc, [k, p1, p2], ["[K]", "P1", "[P2]"], width=8)
    t.shuffle()
    t.sort([("[K]", "ASC")], protocol=proto)
    got = t.open()
    np.testing.assert_array_equal(got["[K]"].astype(np.int64), np.sort(k))
    assert (got["[##V]"] == 1).all()
    np.testing.assert_array_equal(
        _rows(got["[K]"].astype(np.int64), got["P1"].astype(np.int64),
              got["[P2]"].astype(np.int64)), _rows(k, p1, p2))


def test_c4_join_groupby_2p22_2p20(G):
    from orq_amd import keys_from_seed, splitmix64
    from test_oracle_join import pipeline_result, run_pipeline, sqlite_join_groupby

    nl, nr = 1 << 22, 1 << 20
    rng = np.random.default_rng(0x0A11CE04)
    rkey = rng.permutation(1 << 20)[:nr].astype(np.int64)
    attr = (splitmix64(0x0A11CE14, nr) >> np.uint64(32)).astype(np.int64)
    lkey = (splitmix64(0x0A11CE24, nl) % np.uint64(1 << 20)).astype(np.int64)
    value = (splitmix64(0x0A11CE34, nl) % np.uint64(10**6)).astype(np.int64)
    gc = G.Ctx(keys_from_seed(0x0A11CE04))
    j = run_pipeline(gc, rkey, attr, lkey, value, width=8)
    np.testing.assert_array_equal(pipeline_result(j), sqlite_join_groupby(rkey, attr, lkey, value))


def test_c5_served_stream_up_to_2p22(G):
    from orq_amd.serve import REQ_GROUPBY_SUM, REQ_SORT, Job, Server
    from test_gpu_serve import _check_plain, _shares

    sizes = [1 << 22, 1 << 10, 3001, 1 << 21, 70001, (1 << 18) + 7, 1 << 20, 40000]

    def make():
        jobs, plain = [], []
        r = np.random.default_rng(0x0A11CE05)
        for i, n in enumerate(sizes):
            k = zipf(n, 0x0A11CE05 + 17 * i)
            v = r.integers(0, 2**32, n)
            kb, _ = _shares(k, r, 8)
            _, va = _shares(v, r, 8)
            op = REQ_SORT if i % 2 == 0 else REQ_GROUPBY_SUM
            jobs.append(Job(i, op, kb, va, G.keys_from_seed(0x5E000000 + i)))
            plain.append((op, k, v))
        return jobs, plain

    mixed, plain = make()
    srv = Server(slots=4)
    for j in mixed:
        srv.submit(j)
    done = srv.drain()
    st = srv.stats()
    srv.close()
    assert len(done) == len(sizes)
    for j in done:
        assert j.completion[0] == 0, (j.id, j.completion, j.error)
    assert st["failed"] == 0 and st["max_running"] > 1
    alone, _ = make()
    srv = Server(slots=1)
    for j in alone:
        srv.submit(j)
        srv.drain()
    srv.close()
    for a, b, (op, k, v) in zip(mixed, alone, plain):
        assert b.completion[0] == 0, (b.id, b.error)
        for x, y in ((a.out_key, b.out_key), (a.out_val, b.out_val), (a.out_valid, b.out_valid)):
            np.testing.assert_array_equal(x, y, err_msg=f"request {a.id} differs from alone")
        _check_plain(a, op, k, v)


def test_headline_radix_sort_2p24_int64(G):
    from orq_amd import keys_from_seed

    n = 1 << 24
    k = zipf(n, 0x0A11CE03)
    gc = G.Ctx(keys_from_seed(0x0A11CE03))
    gv = gc.secret_share(k.view(np.uint64), 8)
    gperm = gc.radix_sort(gv)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(gc.open(gv).view(np.int64), k[order])
    np.testing.assert_array_equal(gc.open(gperm).view(np.int32), _ranks(order))
    assert gc.queue_size() == 0
