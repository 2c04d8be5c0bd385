"""The CPU oracle, pinned against known answers, closed forms and invariants.

Each check is chosen so that a plausible restatement mistake fails it:
keystream framing (counter start, nonce per 1 MiB chunk), batch schedule,
zero-sharing sign/orientation, the b2a_bit party roles, round order and
scatter direction of sharded permutations, the rejection rule and refill
accounting of gen_perm, padding/sign handling of radix sort.
"""
import numpy as np
import pytest

from oracle import oracle as O


def test_aes_gcm_known_answers():
    # AES-GCM specification test case 14: K = 0^32, IV = 0^12, P = 0^16 gives
    # C = cea7403d...; GCM's keystream = CTR from counter block IV||BE32(2).
    ks, nxt = O.keystream(bytes(32), 0, 32)
    assert ks.tobytes().hex() == ("cea7403d4d606b6e074ec5d3baf39d18"
                                  "726003ca37a62a74d1a2f58e7506358e")
    assert nxt == 1
    ks, nxt = O.keystream(bytes(range(32)), 5, 32)
    assert ks.tobytes().hex() == ("e66badc133a753d9b48a5d04097e24b2"
                                  "25f5cbbe1d147d782ea91513b5c8400c")
    assert nxt == 6


def test_keystream_one_nonce_per_mebibyte():
    key = bytes(range(1, 33))
    whole, nxt = O.keystream(key, 7, 3 * (1 << 20) + 100)
    assert nxt == 11
    parts = [O.keystream(key, 7 + i, 1 << 20)[0] for i in range(3)]
    parts.append(O.keystream(key, 10, 100)[0])
    np.testing.assert_array_equal(whole, np.concatenate(parts))
    # a chunk is a plain CTR stream: its second block starts 16 bytes in
    a, _ = O.keystream(key, 3, 48)
    b, _ = O.keystream(key, 3, 16)
    np.testing.assert_array_equal(a[:16], b)
    assert not np.array_equal(a[16:32], b)


def test_batch_schedule_matches_reference_sizes():
    # compute_equal_batches with batch_size -12, one worker (task.h:34-44)
    assert O.batches(1 << 16) == [5632] * 11 + [3584]
    assert O.batches(1 << 22) == [349696] * 11 + [347648]
    assert O.batches(1 << 24) == [1398272] * 11 + [1396224]
    assert O.batches(1 << 25) == [2796288] * 11 + [2795264]
    assert O.batches(100) == [256] * 0 + [100]
    assert O.batches(1000, 300) == [300, 300, 300, 100]


@pytest.fixture
def ctx(keys):
    return O.Ctx(keys)


@pytest.mark.parametrize("width", [4, 8])
def test_zero_sharings_cancel(ctx, width):
    # x AND 0 leaves only the zero sharing: its XOR over the three parties
    # must be 0; likewise x * 0 sums to 0 for the arithmetic sharing.
    n = 3000
    dt = np.uint32 if width == 4 else np.uint64
    x = np.arange(n, dtype=dt) * dt(2654435761)
    X = ctx.secret_share(x, width)
    Z = ctx.secret_share(np.zeros(n, dt), width)
    r = ctx.and_b(X, Z)
    s = r.shares()
    assert (s[0] != 0).any()
    np.testing.assert_array_equal(s[0] ^ s[1] ^ s[2], np.zeros(n, dt))
    Xa = ctx.secret_share(x, width, arith=True)
    Za = ctx.secret_share(np.zeros(n, dt), width, arith=True)
    s = ctx.mul_a(Xa, Za).shares()
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(s[0] + s[1] + s[2], np.zeros(n, dt))


@pytest.mark.parametrize("width", [4, 8])
def test_primitives_closed_forms(ctx, width):
    rng = np.random.default_rng(3)
    n = 6001
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**63, n, dtype=np.uint64).astype(dt)
    y = rng.integers(0, 2**63, n, dtype=np.uint64).astype(dt)
    X, Y = ctx.secret_share(x, width), ctx.secret_share(y, width)
    np.testing.assert_array_equal(ctx.open(X), x)
    np.testing.assert_array_equal(ctx.open(ctx.and_b(X, Y)), x & y)
    np.testing.assert_array_equal(ctx.open(ctx.b2a_bit(X), arith=True), x & dt(1))
    Xa, Ya = ctx.secret_share(x, width, True), ctx.secret_share(y, width, True)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(ctx.open(ctx.mul_a(Xa, Ya), arith=True), x * y)
    # replicated consistency: party p's second share is party p+1's first
    for p in range(3):
        np.testing.assert_array_equal(X.get(p, 1), X.get((p + 1) % 3, 0))


def test_b2a_bit_every_input_combination(ctx):
    # all 8 assignments of the three share bits, repeated over many batches
    n = 8 * 512
    bits = np.arange(n) % 8
    s = [((bits >> k) & 1).astype(np.uint32) for k in range(3)]
    v = O.Vec(n, 4)
    v.set_shares(s)
    y = ctx.b2a_bit(v)
    np.testing.assert_array_equal(ctx.open(y, arith=True), (s[0] ^ s[1] ^ s[2]))


def test_message_log_and_b(ctx):
    n = 5000
    x = ctx.secret_share(np.arange(n, dtype=np.uint64), 8)
    ctx.clear_msglog()
    ctx.and_b(x, x)
    log = ctx.msglog()
    assert sorted(log) == sorted([("and_b", p, (p + 2) % 3, n * 8) for p in range(3)])


def test_reshare_changes_shares_not_values(ctx):
    n = 2000
    x = np.arange(n, dtype=np.uint32)
    for arith in (False, True):
        v = ctx.secret_share(x, 4, arith)
        before = v.shares()
        for g in range(3):
            ctx.reshare(v, g, arith)
        after = v.shares()
        assert all(not np.array_equal(a, b) for a, b in zip(before, after))
        np.testing.assert_array_equal(ctx.open(v, arith), x)


def _words(key, nonce, count):
    raw, _ = O.keystream(key, nonce, 4 * count)
    return raw.view("<u4")


def test_gen_perm_n1_closed_form():
    # n = 1: mask = 1, accept the first word with an even value; the permutation
    # is [0]; 4 nonces per refill of 2^20 words, one refill up front.
    key = bytes(range(2, 34))
    perm, nonce, draws = O.gen_perm(key, 9, 1)
    w = _words(key, 9, 64)
    first_even = int(np.argmax((w & 1) == 0))
    assert draws == first_even + 1
    assert list(perm) == [0]
    assert nonce == 9 + 4


def test_gen_perm_n2_closed_form():
    # n = 2: mask = 3; step 0 accepts v <= 1 and swaps (0, v); step 1 accepts v == 0.
    key = bytes(range(3, 35))
    perm, nonce, draws = O.gen_perm(key, 0, 2)
    w = (_words(key, 0, 256) & 3).tolist()
    i0 = next(i for i, v in enumerate(w) if v <= 1)
    i1 = next(i for i, v in enumerate(w) if i > i0 and v == 0)
    assert draws == i1 + 1
    assert list(perm) == ([0, 1] if w[i0] == 0 else [1, 0])
    assert nonce == 4


def test_gen_perm_refill_accounting():
    # n = 2^16: ~2^17 * H_n draws, several refills; nonces advance by
    # 4 * (1 + draws // 2^20) and the result is a bijection.
    key = bytes(range(4, 36))
    n = 1 << 16
    perm, nonce, draws = O.gen_perm(key, 100, n)
    assert draws > (1 << 20)
    assert nonce == 100 + 4 * (1 + draws // (1 << 20))
    assert np.array_equal(np.sort(perm), np.arange(n))
    # uniformity smoke: position of element 0 spread over the range
    assert 0 < int(np.argmax(perm == 0)) < n


def test_group_members_generate_identical_perms(ctx):
    n = 700
    ctx.reserve(n, 2)
    for k in range(2):
        for g in range(3):
            members = [p for p in range(3) if p != 2 - g]
            a = ctx.peek_perm(k, members[0], g, n)
            b = ctx.peek_perm(k, members[1], g, n)
            np.testing.assert_array_equal(a, b)
        assert not np.array_equal(ctx.peek_perm(k, 0, 0, n), ctx.peek_perm(k, 0, 1, n))


@pytest.mark.parametrize("inverse", [False, True])
def test_sharded_perm_is_composition_of_local_perms(ctx, inverse):
    n = 1000
    x = np.arange(n, dtype=np.uint64) * np.uint64(7919)
    v = ctx.secret_share(x, 8)
    ctx.reserve(n, 1)
    pis = [ctx.peek_perm(0, [p for p in range(3) if p != 2 - g][0], g, n) for g in range(3)]
    ctx.apply_sharded(v, arith=False, inverse=inverse)
    want = x.copy()
    if not inverse:
        for g in (0, 1, 2):  # local_apply_perm: new[pi[i]] = x[i]
            t = np.empty_like(want)
            t[pis[g]] = want
            want = t
    else:
        for g in (2, 1, 0):  # local_apply_inverse_perm: new[i] = x[pi[i]]
            want = want[pis[g]]
    np.testing.assert_array_equal(ctx.open(v), want)


def test_sharded_perm_then_inverse_is_identity(ctx):
    n = 513
    x = np.arange(n, dtype=np.uint32)
    v = ctx.secret_share(x, 4)
    ctx.reserve(n, 1)
    pis = [ctx.peek_perm(0, [p for p in range(3) if p != 2 - g][0], g, n) for g in range(3)]
    ctx.apply_sharded(v)
    y = ctx.open(v)
    assert (y != x).mean() > 0.9
    for g in (2, 1, 0):
        y = y[pis[g]]
    np.testing.assert_array_equal(y, x)


def test_elementwise_perm_scatters(ctx):
    n = 900
    rng = np.random.default_rng(9)
    x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    sigma = rng.permutation(n).astype(np.uint32)
    v = ctx.secret_share(x, 4)
    p = ctx.secret_share(sigma, 4, arith=True)
    ctx.apply_elementwise(v, p)
    want = np.empty_like(x)
    want[sigma] = x
    np.testing.assert_array_equal(ctx.open(v), want)


def _check_sort(ctx, plain, width, desc):
    n = plain.size
    sdt = np.int32 if width == 4 else np.int64
    v = ctx.secret_share(plain, width)
    perm = ctx.radix_sort(v, desc=desc)
    signed = plain.view(sdt)
    order = np.argsort(~signed if desc else signed, kind="stable")
    np.testing.assert_array_equal(ctx.open(v).view(sdt), signed[order])
    ranks = np.empty(n, dtype=np.int64)
    ranks[order] = np.arange(n)
    np.testing.assert_array_equal(ctx.open(perm, arith=True).view(np.int32), ranks)
    assert ctx.queue_size() == 0  # every reserved permutation consumed
    # applying the sorting permutation to a second copy sorts it the same way
    w = ctx.secret_share(plain, width)
    ctx.apply_elementwise(w, perm)
    np.testing.assert_array_equal(ctx.open(w).view(sdt), signed[order])


@pytest.mark.parametrize("desc", [False, True])
def test_radix_sort_reference_case_int32(ctx, desc):
    # test_sorting.cpp:134-174: 0..n-1 with v[n/2] = -1, shuffled
    n = 1000
    v = np.arange(n, dtype=np.int32)
    v[n // 2] = -1
    v = np.random.default_rng(4).permutation(v)
    _check_sort(ctx, v.view(np.uint32), 4, desc)


@pytest.mark.parametrize("desc", [False, True])
def test_radix_sort_int64_edges(ctx, desc):
    rng = np.random.default_rng(6)
    v = rng.integers(-2**63, 2**63 - 1, 500, dtype=np.int64)
    v[:4] = [-2**63, 2**63 - 1, -1, 0]
    v[10:30] = v[5]
    _check_sort(ctx, v.view(np.uint64), 8, desc)


def test_radix_sort_int32_limits(ctx):
    # INT_MIN / INT_MAX literals as in test_comparisons.cpp
    v = np.array([2**31 - 1, -2**31, 0, -1, 1, -2**31, 2**31 - 1, 5], dtype=np.int32)
    _check_sort(ctx, v.view(np.uint32), 4, False)


def test_radix_sort_partial_bits(ctx):
    # sorting on the low 8 bits only: order by (key & 0xff) unsigned, stable
    rng = np.random.default_rng(8)
    k = rng.integers(0, 2**16, 300).astype(np.uint32)
    v = ctx.secret_share(k, 4)
    ctx.radix_sort(v, bits=8)
    order = np.argsort(k & 0xFF, kind="stable")
    np.testing.assert_array_equal(ctx.open(v), k[order])


@pytest.mark.parametrize("width", [4, 8])
def test_local_ops_closed_forms(keys, width):
    """xor_b, add_a, sub_a, not_b, neg_a (protocol.h:561-649, replicated_3pc.h:199):
    every party applies the op to both of its shares, so the opened value is
    the op on the plaintexts (mod 2^bits), and each of the three distinct
    shares is the op on the inputs' shares."""
    dt = np.uint32 if width == 4 else np.uint64
    rng = np.random.default_rng(width)
    x = rng.integers(0, 2**63, 1000, dtype=np.uint64).astype(dt)
    y = rng.integers(0, 2**63, 1000, dtype=np.uint64).astype(dt)
    c = O.Ctx(keys)
    X, Y = c.secret_share(x, width), c.secret_share(y, width)
    A, B = c.secret_share(x, width, arith=True), c.secret_share(y, width, arith=True)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(c.open(c.xor_b(X, Y)), x ^ y)
        np.testing.assert_array_equal(c.open(c.not_b(X)), ~x)
        np.testing.assert_array_equal(c.open(c.add_a(A, B), arith=True), x + y)
        np.testing.assert_array_equal(c.open(c.sub_a(A, B), arith=True), x - y)
        np.testing.assert_array_equal(c.open(c.local("neg_a", A), arith=True), dt(0) - x)
        z = c.sub_a(A, B)
        for p in range(3):
            np.testing.assert_array_equal(z.get(p, 0), A.get(p, 0) - B.get(p, 0))
            np.testing.assert_array_equal(z.get(p, 1), A.get(p, 1) - B.get(p, 1))


@pytest.mark.parametrize("width", [4, 8])
def test_unary_and_prefix_sum_closed_forms(keys, width):
    """Share-local maps (mapping_access_vector.h:376-446,1312-1316,1528-1547)
    open to the same map of the plaintext, and prefix_sum opens to the
    running sum mod 2^bits."""
    dt = np.uint32 if width == 4 else np.uint64
    sdt = np.int32 if width == 4 else np.int64
    bits = 8 * width
    rng = np.random.default_rng(3 + width)
    x = rng.integers(0, 2**63, 777, dtype=np.uint64).astype(dt)
    x[:4] = [0, 1, dt(1) << dt(bits - 1), ~dt(0)]
    c = O.Ctx(keys)
    X = c.secret_share(x, width)
    with np.errstate(over="ignore"):
        for s in (0, 1, 5, bits - 1, bits, bits + 3):
            sh = dt(min(s, bits - 1))
            want_shl = np.zeros_like(x) if s >= bits else (x << dt(s))
            want_lsr = np.zeros_like(x) if s >= bits else (x >> dt(s))
            np.testing.assert_array_equal(c.open(c.unary("<<", X, s)), want_shl)
            np.testing.assert_array_equal(c.open(c.unary("lsr", X, s)), want_lsr)
            np.testing.assert_array_equal(c.open(c.unary(">>", X, s)).view(sdt),
                                          x.view(sdt) >> sdt(sh))
        par = np.array([bin(int(v)).count("1") & 1 for v in x], dtype=dt)
        np.testing.assert_array_equal(c.open(c.unary("parity", X)), par)
        np.testing.assert_array_equal(c.open(c.unary("extend_lsb", X)), dt(0) - (x & dt(1)))
        np.testing.assert_array_equal(c.open(c.unary("ltz", X)), (x.view(sdt) < 0).astype(dt))
        m = 0x00F0F0F0F0F0F0F0 & ((1 << bits) - 1)
        np.testing.assert_array_equal(c.open(c.unary("mask", X, m)), x & dt(m))
        A = c.secret_share(x, width, arith=True)
        c.prefix_sum(A)
        np.testing.assert_array_equal(c.open(A, arith=True), np.cumsum(x, dtype=dt))
