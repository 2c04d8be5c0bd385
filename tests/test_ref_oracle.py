"""The oracle against the reference's own code.

oracle/_ref/liborqref.so compiles the reference's headers in place (its
AESPRGAlgorithm/CommonPRG, gen_perm and HMShardedPermutationGenerator,
ZeroSharingGenerator, Replicated_3PC and Protocol::reshare, and the runtime's
Task batching) and runs the three parties on three host threads with explicit
keys.  These tests require the oracle to give the same keystream, the same
permutations, and the same shares of every party, call after call (so the
PRG nonces must advance alike), with the same bytes sent per party.

They run where /root/reference exists (the library is built from it); they
are skipped elsewhere.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R

pytestmark = pytest.mark.skipif(not R.available(), reason="the reference is not present here")

KEYS = bytes((i * 37 + 11) & 0xFF for i in range(320))
_DT = {4: np.uint32, 8: np.uint64}


def faithful(v):
    """The oracle vector as [party][idx][n]."""
    return np.stack([np.stack([v.get(p, 0), v.get(p, 1)]) for p in range(3)])


def to_oracle(arr):
    v = O.Vec(arr.shape[2], arr.itemsize)
    for p in range(3):
        for k in range(2):
            v.set(p, k, arr[p, k])
    return v


def sent_per_party(log):
    out = [0, 0, 0]
    for _op, frm, _to, nbytes in log:
        out[frm] += nbytes
    return out


@pytest.mark.parametrize("width", [1, 4, 8])
@pytest.mark.parametrize("count", [1, 3, 1000, 300_001])
def test_keystream(width, count):
    key = bytes(range(7, 39))
    for skip in (0, 5):
        ref = R.keystream(key, skip, width, count)
        ks, _ = O.keystream(key, skip, width * count)
        np.testing.assert_array_equal(ks.view(ref.dtype), ref)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 100, 4099, 70_000, (1 << 20) + 3])
def test_gen_perm(n):
    key = bytes(range(40, 72))
    count = 3 if n < (1 << 20) else 2
    ref = R.gen_perm(key, 9, n, count)
    nonce = 9
    for k in range(count):
        perm, nonce, _ = O.gen_perm(key, nonce, n)
        np.testing.assert_array_equal(perm, ref[k], err_msg=f"permutation {k} of size {n}")


@pytest.mark.parametrize("n", [5, 1000, 66_000])
def test_sharded_permutations(n):
    count = 3
    ref = R.Ctx(KEYS).sharded_perms(n, count)
    oc = O.Ctx(KEYS, gen_everyone=True)
    oc.reserve(n, count)
    for k in range(count):
        for p in range(3):
            for g, members in enumerate([(0, 1), (0, 2), (1, 2), (0, 1, 2)]):
                if p in members:
                    np.testing.assert_array_equal(oc.peek_perm(k, p, g, n), ref[k, p, g],
                                                  err_msg=f"perm {k} party {p} group {g}")
                else:
                    assert (ref[k, p, g] == -1).all()


@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 3000, 65_536, 1 << 22, (1 << 24) + 1])
@pytest.mark.parametrize("batches", [-12, -1, -5, 256, 1000])
def test_batch_schedule(n, batches):
    if batches > 0:
        size = batches
    else:
        size = R.equal_batches(0, n, batches)
    expect = [min(size, n - s) for s in range(0, n, size)] if size > 0 else []
    assert O.batches(n, batches) == expect


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 300, 5000, 70_001])
@pytest.mark.parametrize("batch_size", [-12, 1024])
def test_primitive_sequence(width, n, batch_size):
    """secret sharing from every owner, and_b, mul_a, b2a_bit, reshare in every
    group, open: every party's two shares after each call, and bytes sent."""
    rng = np.random.default_rng(n * width)
    dt = _DT[width]
    x = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    y = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    rc = R.Ctx(KEYS, batch_size)
    oc = O.Ctx(KEYS, batch_size)

    def same(o_vec, r_arr, what):
        np.testing.assert_array_equal(faithful(o_vec), r_arr, err_msg=what)

    # boolean and arithmetic sharing from each data owner
    rx = rc.secret_share(x, width, owner=0, binary=True)
    ox = oc.secret_share_from(x, width, 0)
    same(ox, rx, "secret_share_b owner 0")
    ry = rc.secret_share(y, width, owner=2, binary=True)
    oy = oc.secret_share_from(y, width, 2)
    same(oy, ry, "secret_share_b owner 2")
    rxa = rc.secret_share(x, width, owner=1, binary=False)
    oxa = oc.secret_share_from(x, width, 1, arith=True)
    same(oxa, rxa, "secret_share_a owner 1")
    rya = rc.secret_share(y, width, owner=0, binary=False)
    oya = oc.secret_share_from(y, width, 0, arith=True)
    same(oya, rya, "secret_share_a owner 0")

    rz = rc.and_b(rx, ry)
    same(oc.and_b(ox, oy), rz, "and_b")
    rm = rc.mul_a(rxa, rya)
    same(oc.mul_a(oxa, oya), rm, "multiply_a")
    rb = rc.b2a_bit(rz)
    ob = oc.b2a_bit(to_oracle(rz))
    same(ob, rb, "b2a_bit")

    # a second and_b (the nonces must have advanced alike), then every group's
    # resharing, binary and arithmetic
    rzz = rc.and_b(rx, ry)
    oc_z = oc.and_b(ox, oy)
    same(oc_z, rzz, "and_b (second call)")
    for g in range(3):
        rzz = rc.reshare(rzz, g, binary=True)
        oc.reshare(oc_z, g)
        same(oc_z, rzz, f"binary reshare group {g}")
    rmm = rm
    om = to_oracle(rm)
    for g in range(3):
        rmm = rc.reshare(rmm, g, binary=False)
        oc.reshare(om, g, arith=True)
        same(om, rmm, f"arithmetic reshare group {g}")

    opened = rc.open(rzz, binary=True)
    for p in range(3):
        np.testing.assert_array_equal(opened[p], x & y)
    np.testing.assert_array_equal(oc.open(oc_z), x & y)
    opened = rc.open(rmm, binary=False)
    for p in range(3):
        np.testing.assert_array_equal(opened[p], x * y)
    np.testing.assert_array_equal(oc.open(om, arith=True), x * y)

    assert sent_per_party(oc.msglog()) == [rc.bytes_sent(p) for p in range(3)]


# ---- composite operators: the reference's own runtime, three processes ------

def _sent(oc):
    return sent_per_party(oc.msglog())


@pytest.mark.parametrize("width,bits,desc", [(4, 32, False), (4, 32, True), (4, 20, False),
                                             (8, 64, False), (8, 64, True), (8, 40, True)])
@pytest.mark.parametrize("n", [1, 2, 777, 4096])
def test_radix_sort_vs_reference(width, bits, desc, n):
    rng = np.random.default_rng(n + bits)
    dt = _DT[width]
    x = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    x[: n // 3] = x[n // 2]  # ties
    v, perm, sent = R.radix_sort(KEYS, x, bits=bits, desc=desc)
    oc = O.Ctx(KEYS)
    ov = oc.secret_share(x, width)
    op = oc.radix_sort(ov, bits=bits, desc=desc)
    np.testing.assert_array_equal(faithful(ov), v)
    np.testing.assert_array_equal(faithful(op), perm)
    assert _sent(oc) == sent


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 1000, 5000])
def test_quicksort_vs_reference(desc, n, width):
    rng = np.random.default_rng(n)
    x = rng.integers(0, np.iinfo(_DT[width]).max, n, dtype=_DT[width], endpoint=True)
    x[::7] = x[n // 2]  # ties
    v, perm, sent = R.quicksort(KEYS, x, desc=desc)
    oc = O.Ctx(KEYS)
    ov = oc.secret_share(x, width)
    op = oc.quicksort(ov, desc=desc)
    np.testing.assert_array_equal(faithful(ov), v)
    np.testing.assert_array_equal(faithful(op), perm)
    assert _sent(oc) == sent


def _same_table(ref_table, ot):
    assert sorted(ref_table) == sorted(ot.schema)
    for name, (enc, arr) in ref_table.items():
        assert enc == (1 if name.startswith("[") else 0), name
        np.testing.assert_array_equal(faithful(ot.schema[name]), arr, err_msg=f"column {name}")


def _host():
    from oracle_backend import OracleCtx
    from orq_amd import table

    return OracleCtx, table


@pytest.mark.parametrize("width,nkeys,proto,desc", [
    (4, 1, "QUICKSORT", False), (4, 2, "QUICKSORT", True), (4, 1, "RADIXSORT", False),
    (4, 2, "RADIXSORT", False), (8, 1, "RADIXSORT", True), (8, 1, "QUICKSORT", False),
    (8, 2, "QUICKSORT", True), (4, 1, "BITONICSORT", False),
    (4, 2, "BITONICSORT", True), (8, 2, "BITONICSORT", False)])
def test_table_sort_vs_reference(width, nkeys, proto, desc):
    OracleCtx, T = _host()
    n = 300
    rng = np.random.default_rng(nkeys * 10 + width)
    cols = [rng.integers(-20, 20, n) for _ in range(nkeys)] + [rng.integers(0, 1000, n),
                                                              rng.integers(-9, 1000, n)]
    schema = ["[K1]", "[K2]"][:nkeys] + ["[P]", "A"]
    ref_table, sent = R.table_sort(KEYS, cols, width, nkeys,
                                   {"BITONICSORT": R.BITONICSORT_ID, "QUICKSORT": R.QUICKSORT_ID,
                                    "RADIXSORT": R.RADIXSORT_ID}[proto], desc)
    oc = OracleCtx(KEYS)
    ot = T.EncodedTable.secret_share(oc, cols, schema, width)
    order = "DESC" if desc else T.ASC
    ot.sort([(k, order) for k in schema[:nkeys]], protocol=getattr(T, proto))
    _same_table(ref_table, ot)
    assert _sent(oc.c) == sent


# The reference's DEFAULT sort protocol is quicksort (operators/sorting.h:6-9)
# unless it is built with DEFAULT_SORT_PROTO set; the table layer's DEFAULT is
# its context's default_sort.  Both settings are checked, against the shipped
# build and against the build with radix sort as its default.
_SORT_DEFAULTS = ["QUICKSORT", "RADIXSORT"]


@pytest.mark.parametrize("default_sort", _SORT_DEFAULTS)
@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("nr,nl,dom", [(8, 32, 12), (40, 160, 64)])
def test_join_groupby_vs_reference(width, nr, nl, dom, default_sort):
    OracleCtx, _ = _host()
    from test_oracle_join import relational_inputs, run_pipeline

    rkey, attr, lkey, value = relational_inputs(nr, nl, dom, nr + nl + width)
    ref_table, sent = R.join_groupby(KEYS, rkey, attr, lkey, value, width,
                                     default_sort=default_sort)
    oc = OracleCtx(KEYS, default_sort=default_sort)
    ot = run_pipeline(oc, rkey, attr, lkey, value, width)
    _same_table(ref_table, ot)
    assert _sent(oc.c) == sent


@pytest.mark.parametrize("default_sort", _SORT_DEFAULTS)
@pytest.mark.parametrize("width", [4, 8])
def test_distinct_vs_reference(width, default_sort):
    OracleCtx, T = _host()
    rng = np.random.default_rng(3)
    k, v = rng.integers(0, 6, 200), rng.integers(0, 2, 200)
    ref_table, sent = R.table_distinct(KEYS, k, v, width, default_sort=default_sort)
    oc = OracleCtx(KEYS, default_sort=default_sort)
    ot = T.EncodedTable.secret_share(oc, [k, v], ["[K]", "[V2]"], width)
    ot.distinct(["[K]", "[V2]"])
    _same_table(ref_table, ot)
    assert _sent(oc.c) == sent


@pytest.mark.parametrize("default_sort", _SORT_DEFAULTS)
@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("reverse", [False, True])
def test_aggregate_vs_reference(width, reverse, default_sort):
    OracleCtx, T = _host()
    rng = np.random.default_rng(7 + width)
    n = 150
    cols = [rng.integers(0, 9, n), rng.integers(0, 1000, n), np.zeros(n, np.int64),
            rng.integers(-50, 1000, n), rng.integers(-50, 1000, n)]
    ref_table, sent = R.table_aggregate(KEYS, cols, width, reverse, default_sort=default_sort)
    oc = OracleCtx(KEYS, default_sort=default_sort)
    ot = T.EncodedTable.secret_share(oc, cols, ["[K]", "A", "C", "[B]", "[M]"], width)
    ot.aggregate(["[K]"], [("A", "A", "sum"), ("A", "C", "count"), ("[B]", "[B]", "max"),
                           ("[M]", "[M]", "min")], reverse=reverse)
    _same_table(ref_table, ot)
    assert _sent(oc.c) == sent


@pytest.mark.parametrize("width", [4, 8])
def test_table_shuffle_vs_reference(width):
    OracleCtx, T = _host()
    rng = np.random.default_rng(11)
    cols = [rng.integers(0, 100, 77), rng.integers(0, 100, 77)]
    ref_table, sent = R.table_shuffle(KEYS, cols, width)
    oc = OracleCtx(KEYS)
    ot = T.EncodedTable.secret_share(oc, cols, ["[K]", "A"], width)
    ot.shuffle()
    _same_table(ref_table, ot)
    assert _sent(oc.c) == sent


def test_c1_radix_sort_2p16_vs_reference():
    """The radix sort workload of 2^16 keys uniform on [0, 2^32) (splitmix64
    seed 0x0A11CE01, int32 shares), at its full size."""
    from orq_amd import keys_from_seed, splitmix64

    n = 1 << 16
    keys = keys_from_seed(0x0A11CE01)
    x = (splitmix64(0x0A11CE01, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    v, perm, sent = R.radix_sort(keys, x)
    oc = O.Ctx(keys)
    ov = oc.secret_share(x, 4)
    op = oc.radix_sort(ov)
    np.testing.assert_array_equal(faithful(ov), v)
    np.testing.assert_array_equal(faithful(op), perm)
    assert _sent(oc) == sent


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 64, 3001])
def test_circuits_vs_reference(width, n):
    """==, !=, <, >, <=, >=, full b2a, the ripple-carry + and -, and a2b."""
    dt = _DT[width]
    rng = np.random.default_rng(n * width)
    x = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    y = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    y[::5] = x[::5]
    edge = np.array([0, 1, np.iinfo(dt).max, np.iinfo(dt).max // 2, np.iinfo(dt).max // 2 + 1],
                    dtype=dt)
    x[: min(n, 5)] = edge[: min(n, 5)]
    y[: min(n, 5)] = edge[::-1][: min(n, 5)]
    res, sent = R.circuits(KEYS, x, y)
    oc = O.Ctx(KEYS)
    ox, oy = oc.secret_share(x, width), oc.secret_share(y, width)
    oa = oc.secret_share(x, width, arith=True)
    got = {op: oc.cmp(op, ox, oy) for op in ["==", "!=", "<", ">", "<=", ">="]}
    got["b2a"] = oc.b2a(ox)
    got["+"] = oc.adder("add", ox, oy)
    got["-"] = oc.adder("sub", ox, oy)
    got["a2b"] = oc.adder("a2b", oa)
    for k in R.CIRCUIT_OUTPUTS:
        np.testing.assert_array_equal(faithful(got[k]), res[k], err_msg=k)
    assert _sent(oc) == sent


# ---- the reference's runtime with T worker threads ---------------------------

def _thread_keys(T):
    from orq_amd import keys_from_seed

    return keys_from_seed(0x7E5D0000), [keys_from_seed(0x7E5D0000 + t) for t in range(1, T)]


@pytest.mark.parametrize("T,n,width,batch_size", [(2, 5000, 4, -12), (3, 5000, 8, -12),
                                                  (4, 300, 4, -12), (2, 70_001, 4, -12),
                                                  (3, 3000, 8, 512)])
def test_radix_sort_threads_vs_reference(T, n, width, batch_size):
    """The runtime's thread ranges (runtime.h:198-247), each thread's own
    randomness (setup.h:48-57) and the reserved permutations dealt to threads
    (runtime.h:749-788), through the reference's own RunTime."""
    k0, extra = _thread_keys(T)
    dt = _DT[width]
    x = np.random.default_rng(T * n).integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    v, perm, sent = R.radix_sort(k0, x, batch_size=batch_size, thread_keys=extra)
    oc = O.Ctx(k0, batch_size, threads=T, thread_keys=extra)
    ov = oc.secret_share(x, width)
    op = oc.radix_sort(ov)
    np.testing.assert_array_equal(faithful(ov), v)
    np.testing.assert_array_equal(faithful(op), perm)
    assert _sent(oc) == sent


@pytest.mark.parametrize("T,n,width,batch_size", [(2, 5000, 4, -12), (3, 3001, 8, -12),
                                                  (4, 70_001, 4, -12), (3, 2000, 8, 512)])
def test_circuits_threads_vs_reference(T, n, width, batch_size):
    """The comparison circuits, full b2a, the adders and a2b under the
    reference's RunTime with T worker threads: every runtime call (over n
    elements, or over the packed words of the sign fix and the ripple-carry
    adder) split into the threads' ranges, each with its thread's randomness."""
    k0, extra = _thread_keys(T)
    dt = _DT[width]
    rng = np.random.default_rng(T * n + width)
    x = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    y = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    y[::7] = x[::7]
    res, sent = R.circuits(k0, x, y, batch_size=batch_size, thread_keys=extra)
    oc = O.Ctx(k0, batch_size, threads=T, thread_keys=extra)
    ox, oy = oc.secret_share(x, width), oc.secret_share(y, width)
    oa = oc.secret_share(x, width, arith=True)
    got = {op: oc.cmp(op, ox, oy) for op in ["==", "!=", "<", ">", "<=", ">="]}
    got["b2a"] = oc.b2a(ox)
    got["+"] = oc.adder("add", ox, oy)
    got["-"] = oc.adder("sub", ox, oy)
    got["a2b"] = oc.adder("a2b", oa)
    for k in R.CIRCUIT_OUTPUTS:
        np.testing.assert_array_equal(faithful(got[k]), res[k], err_msg=k)
    assert _sent(oc) == sent


@pytest.mark.parametrize("T", [2, 3])
def test_table_sort_threads_vs_reference(T):
    OracleCtx, Tb = _host()
    k0, extra = _thread_keys(T)
    n = 999
    rng = np.random.default_rng(T)
    cols = [rng.integers(-20, 20, n), rng.integers(-20, 20, n), rng.integers(0, 9, n),
            rng.integers(0, 9, n)]
    ref_table, sent = R.table_sort(k0, cols, 8, 2, R.RADIXSORT_ID, thread_keys=extra)
    oc = OracleCtx(k0, threads=T, thread_keys=extra)
    ot = Tb.EncodedTable.secret_share(oc, cols, ["[K1]", "[K2]", "[P]", "A"], 8)
    ot.sort(["[K1]", "[K2]"], protocol=Tb.RADIXSORT)
    _same_table(ref_table, ot)
    assert _sent(oc.c) == sent
