"""The serving loop (continuous batching of secret-shared requests): each
result equals the same request run alone and the table-layer flow it stands
for (bit-exact shares), the plaintext answer is right, and admission keeps to
the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UT = {4: np.uint32, 8: np.uint64}


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()
    return orq_amd


def _shares(x, rng, width):
    """A client's boolean and arithmetic sharings of x, [3][n] each."""
    ut = UT[width]
    x = np.asarray(x).astype(ut)
    b = rng.integers(0, 2**64, (2, x.size), dtype=np.uint64).astype(ut)
    bs = np.stack([b[0], b[1], x ^ b[0] ^ b[1]])
    with np.errstate(over="ignore"):
        a = rng.integers(0, 2**64, (2, x.size), dtype=np.uint64).astype(ut)
        as_ = np.stack([a[0], a[1], x - a[0] - a[1]])
    return bs, as_


def _make_jobs(G, sizes, width=8, seed=0):
    from orq_amd.serve import REQ_GROUPBY_SUM, REQ_SORT, Job

    rng = np.random.default_rng(seed)
    jobs, plain = [], []
    for i, n in enumerate(sizes):
        k = rng.integers(0, max(n // 4, 2), n)
        v = rng.integers(0, 10**6, n)
        kb, _ = _shares(k, rng, width)
        _, va = _shares(v, rng, width)
        op = REQ_SORT if i % 2 == 0 else REQ_GROUPBY_SUM
        jobs.append(Job(i, op, kb, va, G.keys_from_seed(1000 + i)))
        plain.append((op, k, v))
    return jobs, plain


def _check_plain(job, op, k, v):
    from orq_amd.serve import REQ_SORT

    sdt = np.int64 if job.out_key.dtype == np.uint64 else np.int32
    key = (job.out_key[0] ^ job.out_key[1] ^ job.out_key[2]).view(sdt)
    val = (job.out_val[0] + job.out_val[1] + job.out_val[2]).view(sdt)
    valid = job.out_valid[0] ^ job.out_valid[1] ^ job.out_valid[2]
    if op == REQ_SORT:
        order = np.argsort(k, kind="stable")
        np.testing.assert_array_equal(key, k[order])
        np.testing.assert_array_equal(val, v[order])
        assert valid.all()
    else:
        keys_u, inv = np.unique(k, return_inverse=True)
        sums = np.bincount(inv, weights=v).astype(np.int64)
        m = valid == 1
        got = sorted(zip(key[m].tolist(), val[m].tolist()))
        assert got == list(zip(keys_u.tolist(), sums.tolist()))


def _table_flow(G, job):
    """The same request through the host table layer on a fresh context."""
    from orq_amd.serve import REQ_SORT
    from orq_amd.table import EncodedTable

    c = G.Ctx(bytes(job.req.keys))
    c.default_sort = "RADIXSORT"  # the server's contexts: DEFAULT_SORT_PROTO=RADIXSORT
    w = job.key_in.dtype.itemsize
    K = c.vec(job.n, w, G.ORQG_BSHARED)
    K.set_shares(list(job.key_in))
    V = c.vec(job.n, w, G.ORQG_ASHARED)
    V.set_shares(list(job.val_in))
    t = EncodedTable(c, "req", {"[K]": K, "V": V}, w)
    if job.op == REQ_SORT:
        t.sort(["[K]"])
    else:
        t.aggregate(["[K]"], [("V", "V", "sum")])
    return [np.stack(x.shares()) for x in (t["[K]"], t["V"], t["[##V]"])]


@pytest.mark.parametrize("width", [4, 8])
def test_server_matches_table_layer(G, width):
    """Every request equals the table layer's EncodedTable::sort / aggregate
    with their default options and the server's DEFAULT protocol (radix
    sort, the reference built with DEFAULT_SORT_PROTO=RADIXSORT)."""
    from orq_amd.serve import Server

    jobs, plain = _make_jobs(G, [1024, 3000, 4096, 777], width=width, seed=1)
    srv = Server(slots=2)
    for j in jobs:
        srv.submit(j)
    done = srv.drain()
    srv.close()
    assert len(done) == len(jobs)
    for j, (op, k, v) in zip(jobs, plain):
        assert j.completion[0] == 0, j.completion
        ref = _table_flow(G, j)
        for got, want in zip((j.out_key, j.out_val, j.out_valid), ref):
            np.testing.assert_array_equal(got, want)
        _check_plain(j, op, k, v)


def test_mix_equals_alone(G):
    from orq_amd.serve import Server

    sizes = [1 << 10, 1 << 14, 3000, 1 << 12, 1 << 16, 5000, 1 << 11, 20000]
    mix, plain = _make_jobs(G, sizes, seed=2)
    alone, _ = _make_jobs(G, sizes, seed=2)
    srv = Server(slots=4)
    for j in mix:
        srv.submit(j)
    srv.drain()
    st = srv.stats()
    srv.close()
    assert st["max_running"] > 1  # requests did overlap
    srv1 = Server(slots=1)
    for j in alone:
        srv1.submit(j)
        srv1.drain()
    srv1.close()
    for a, b, (op, k, v) in zip(mix, alone, plain):
        assert a.completion[0] == 0 and b.completion[0] == 0
        for x, y in ((a.out_key, b.out_key), (a.out_val, b.out_val), (a.out_valid, b.out_valid)):
            np.testing.assert_array_equal(x, y)
        _check_plain(a, op, k, v)


def test_admission_budget(G):
    from orq_amd.serve import Server

    sizes = [1 << 14] * 6
    jobs, _ = _make_jobs(G, sizes, seed=3)
    # the footprint estimate of one 2^14 request is ~ 100 MB: a 250 MB budget
    # admits two at a time
    srv = Server(slots=4, hbm_budget=250 << 20)
    for j in jobs:
        srv.submit(j)
    big, _ = _make_jobs(G, [1 << 20], seed=4)
    big[0].id = 99
    big[0].req.id = 99
    srv.submit(big[0])
    done = srv.drain()
    st = srv.stats()
    srv.close()
    status = {j.id: j.completion[0] for j in done}
    assert status[99] == -3  # ORQG_E_BUDGET: larger than the whole budget
    assert all(status[i] == 0 for i in range(6))
    assert st["max_running"] <= 2
    assert st["peak_reserved"] <= 250 << 20


def test_many_overlapping_small_requests(G):
    """48 small requests over four slots: every slot keeps re-keying its
    context (host-to-device copies) while the others generate permutations
    (graph launches of gen_perm steps).  Every request must complete OK,
    and each result must equal the same request served alone.  (With a
    stream-captured graph and null-stream copies this mix failed requests
    and left a slot's stream stuck capturing.)"""
    from orq_amd.serve import Server

    rng = np.random.default_rng(11)
    sizes = [int(x) for x in rng.integers(1 << 10, 1 << 15, 48)]
    mix, plain = _make_jobs(G, sizes, seed=5)
    alone, _ = _make_jobs(G, sizes, seed=5)
    srv = Server(slots=4)
    for j in mix:
        srv.submit(j)
    done = srv.drain()
    st = srv.stats()
    srv.close()
    bad = [(j.id, j.completion[0], j.error) for j in done if j.completion[0] != 0]
    assert not bad, bad
    assert st["failed"] == 0 and st["max_running"] > 1
    srv1 = Server(slots=1)
    for j in alone[:8]:
        srv1.submit(j)
        srv1.drain()
    srv1.close()
    for a, b, (op, k, v) in zip(mix[:8], alone[:8], plain[:8]):
        assert b.completion[0] == 0, b.error
        for x, y in ((a.out_key, b.out_key), (a.out_val, b.out_val), (a.out_valid, b.out_valid)):
            np.testing.assert_array_equal(x, y)
        _check_plain(a, op, k, v)


@pytest.mark.parametrize("slots", [4, 12])
def test_batched_sorts_across_slots(G, slots):
    """Served with batch_rows set: the small radix sorts of concurrent
    requests meet and run as shared batches (a batch holds several sorts),
    a large request's sorts run on their own, and every result equals the
    same request served alone, unbatched."""
    from orq_amd.serve import Server

    rng = np.random.default_rng(12)
    sizes = [int(x) for x in rng.integers(1 << 10, 1 << 14, 40)] + [1 << 17]
    mix, plain = _make_jobs(G, sizes, seed=6)
    alone, _ = _make_jobs(G, sizes, seed=6)
    srv = Server(slots=slots, batch_rows=1 << 16, batch_window_us=3000)
    for j in mix:
        srv.submit(j)
    done = srv.drain()
    st = srv.stats()
    srv.close()
    bad = [(j.id, j.completion[0], j.error) for j in done if j.completion[0] != 0]
    assert not bad, bad
    assert st["sort_batches"] > 0 and st["sorts_batched"] > st["sort_batches"]
    srv1 = Server(slots=1)
    for j in alone:
        srv1.submit(j)
    srv1.drain()
    srv1.close()
    for a, b, (op, k, v) in zip(mix, alone, plain):
        assert b.completion[0] == 0, b.error
        for x, y in ((a.out_key, b.out_key), (a.out_val, b.out_val), (a.out_valid, b.out_valid)):
            np.testing.assert_array_equal(x, y, err_msg=f"request {a.id} ({a.n} rows)")
        _check_plain(a, op, k, v)
