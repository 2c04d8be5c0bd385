"""GPU parity: liborqg (HIP, gfx950) against the CPU oracle on the same keys.

Every test drives the same call sequence through the oracle (per-party
simulation) and through the C ABI of the GPU library, then compares
  * the three distinct shares bit for bit,
  * the PRG nonce of every key afterwards,
  * the message log (op, from, to, bytes),
and the reconstructed values against plaintext closed forms.
"""
import numpy as np
import pytest

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


def _pair(O, G, keys, batch_size=-12):
    return O.Ctx(keys, batch_size), G.Ctx(keys, batch_size)


def _same_shares(ov, gv):
    os_, gs = ov.shares(), gv.shares()
    for k in range(3):
        np.testing.assert_array_equal(os_[k], gs[k], err_msg=f"share {k} differs")
    # party p's second share in the oracle is s_{p+1}
    for p in range(3):
        np.testing.assert_array_equal(ov.get(p, 1), gs[(p + 1) % 3])


def _same_nonces(oc, gc):
    # oracle roles per party: 0 local, 1 prev, 2 next, 3+g group g
    for p in range(3):
        assert oc.nonce(p, 0) == gc.nonce(p), f"local key {p}"
        assert oc.nonce(p, 2) == gc.nonce(3 + p), f"pair key {p}"
        assert oc.nonce(p, 1) == gc.nonce(3 + (p + 2) % 3), f"prev key of party {p}"
    for g in range(3):
        members = [p for p in range(3) if p != 2 - g]
        for p in members:
            assert oc.nonce(p, 3 + g) == gc.nonce(6 + g), f"group {g} seen by party {p}"


def _same_log(oc, gc):
    assert oc.msglog() == gc.msglog()


def test_keystream_kat(G):
    """AES-GCM spec test case 14 (K = 0^32, IV = 0^12): first block of the
    keystream; second vector K = 00..1f, nonce 5 (libsodium == OpenSSL)."""
    keys = bytes(32) + bytes(range(32)) + bytes(32 * 8)
    c = G.Ctx(keys)
    ks = c.keystream(0, 32)
    assert ks.tobytes().hex() == ("cea7403d4d606b6e074ec5d3baf39d18"
                                  "726003ca37a62a74d1a2f58e7506358e")
    assert c.nonce(0) == 1
    for _ in range(5):
        c.keystream(1, 16)
    assert c.nonce(1) == 5
    ks = c.keystream(1, 32)
    assert ks.tobytes().hex() == ("e66badc133a753d9b48a5d04097e24b2"
                                  "25f5cbbe1d147d782ea91513b5c8400c")


@pytest.mark.parametrize("nbytes", [1, 17, (1 << 20) + 5, 3 * (1 << 20)])
def test_keystream_matches_oracle(O, G, keys, nbytes):
    c = G.Ctx(keys)
    got = c.keystream(4, nbytes)
    want, nxt = O.keystream(keys[4 * 32:5 * 32], 0, nbytes)
    np.testing.assert_array_equal(got, want)
    assert c.nonce(4) == nxt


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 257, 5000, 70001])
def test_share_and_mult(O, G, keys, width, n):
    rng = np.random.default_rng(n + width)
    dt = np.uint32 if width == 4 else np.uint64
    hi = 2**32 if width == 4 else 2**63
    x = rng.integers(0, hi, n, dtype=np.uint64).astype(dt)
    y = rng.integers(0, hi, n, dtype=np.uint64).astype(dt)
    oc, gc = _pair(O, G, keys)
    ox, gx = oc.secret_share(x, width), gc.secret_share(x, width)
    oy, gy = oc.secret_share(y, width), gc.secret_share(y, width)
    _same_shares(ox, gx)
    oz, gz = oc.and_b(ox, oy), gc.and_b(gx, gy)
    _same_shares(oz, gz)
    np.testing.assert_array_equal(gc.open(gz), x & y)
    np.testing.assert_array_equal(oc.open(oz), x & y)
    oa, ga = oc.secret_share(x, width, arith=True), gc.secret_share(x, width, arith=True)
    ob, gb = oc.secret_share(y, width, arith=True), gc.secret_share(y, width, arith=True)
    om, gm = oc.mul_a(oa, ob), gc.mul_a(ga, gb)
    _same_shares(om, gm)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(gc.open(gm), (x * y).astype(dt))
        np.testing.assert_array_equal(oc.open(om, arith=True), (x * y).astype(dt))
    obb, gbb = oc.b2a_bit(ox), gc.b2a_bit(gx)
    _same_shares(obb, gbb)
    np.testing.assert_array_equal(gc.open(gbb), (x & 1).astype(dt))
    np.testing.assert_array_equal(oc.open(obb, arith=True), (x & 1).astype(dt))
    _same_nonces(oc, gc)
    _same_log(oc, gc)


@pytest.mark.parametrize("width", [4, 8])
def test_reshare(O, G, keys, width):
    n = 3001
    rng = np.random.default_rng(7)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**31, n).astype(dt)
    oc, gc = _pair(O, G, keys)
    ov, gv = oc.secret_share(x, width), gc.secret_share(x, width)
    before = gv.shares()
    for g in range(3):
        oc.reshare(ov, g)
        gc.reshare(gv, g)
        _same_shares(ov, gv)
    after = gv.shares()
    assert not np.array_equal(before[0], after[0])
    np.testing.assert_array_equal(gc.open(gv), x)
    _same_nonces(oc, gc)


@pytest.mark.parametrize("path", ["small", "windows"])
@pytest.mark.parametrize("n", [1, 2, 3, 100, 1000, 4099])
def test_gen_perm_matches_oracle(O, G, keys, monkeypatch, n, path):
    """Both generators: one workgroup per local permutation (the default at
    these sizes) and the windowed pipeline (forced by a small-path limit of 0)."""
    if path == "windows":
        monkeypatch.setenv("ORQG_GP_SMALL_MAX", "0")
    oc, gc = _pair(O, G, keys)
    count = 3
    oc.reserve(n, count)
    gc.reserve(n, count)
    for k in range(count):
        for g in range(3):
            members = [p for p in range(3) if p != 2 - g]
            want = oc.peek_perm(k, members[0], g, n)
            np.testing.assert_array_equal(oc.peek_perm(k, members[1], g, n), want)
            got = gc.peek_perm(k, g, n)
            np.testing.assert_array_equal(got, want, err_msg=f"perm {k} group {g}")
            assert np.array_equal(np.sort(got), np.arange(n))
    _same_nonces(oc, gc)


@pytest.mark.parametrize("n,count,predict", [
    ((1 << 16) + 1, 3, 4),    # every permutation takes 8 nonces, 4 predicted
    (1 << 17, 4, None),       # about 3.1 refills each: the prediction misses often
    (5000, 3, 12),            # 4 each, 12 predicted
])
def test_gen_perm_small_path_renonce(O, G, keys, monkeypatch, n, count, predict):
    """The small path when the predicted start nonces are wrong for some
    permutations: those are drawn again from the exact ones."""
    monkeypatch.setenv("ORQG_GP_SMALL_MAX", str(1 << 17))
    if predict is not None:
        monkeypatch.setenv("ORQG_GP_SMALL_PREDICT", str(predict))
    oc, gc = _pair(O, G, keys)
    oc.reserve(n, count)
    gc.reserve(n, count)
    for k in range(count):
        for g in range(3):
            members = [p for p in range(3) if p != 2 - g]
            np.testing.assert_array_equal(gc.peek_perm(k, g, n), oc.peek_perm(k, members[0], g, n),
                                          err_msg=f"perm {k} group {g}")
    _same_nonces(oc, gc)


def test_gen_perm_large(O, G, keys):
    """2^17 rows: many windows per permutation, the tail with rare acceptances."""
    n = 1 << 17
    oc, gc = _pair(O, G, keys)
    oc.reserve(n, 1)
    gc.reserve(n, 1)
    for g in range(3):
        members = [p for p in range(3) if p != 2 - g]
        np.testing.assert_array_equal(gc.peek_perm(0, g, n), oc.peek_perm(0, members[0], g, n))
    _same_nonces(oc, gc)


@pytest.mark.parametrize("n", [(1 << 20) + 3, 1 << 22, 1 << 24])
def test_gen_perm_many_tiles(O, G, keys, n):
    """Windows of hundreds of tiles: the dense look-back over long tile chains,
    sparse super-tiles, several permutations in a row (the second starts at
    the first one's refill boundary), a non-power-of-two size, and the sort
    benchmark's size (2^24)."""
    count = 2 if n < (1 << 22) else 1
    oc, gc = _pair(O, G, keys)
    oc.reserve(n, count)
    gc.reserve(n, count)
    for k in range(count):
        for g in range(3):
            members = [p for p in range(3) if p != 2 - g]
            np.testing.assert_array_equal(gc.peek_perm(k, g, n), oc.peek_perm(k, members[0], g, n),
                                          err_msg=f"perm {k} group {g}")
    _same_nonces(oc, gc)


@pytest.mark.parametrize("n,count,l2cap,shift", [
    (50000, 2, 64, 0),          # undecided lists over a tiny cap: windows retried at W/4
    ((1 << 20) + 3, 2, 0, 3),   # k_lo margin / 8: some speculative windows redrawn
    ((1 << 20) + 3, 1, 0, 12),  # no margin: nearly every speculative window redrawn
])
def test_gen_perm_forced_redraws(O, G, keys, monkeypatch, n, count, l2cap, shift):
    """The retry and redraw paths of the window pipeline, forced through the
    test hooks (a smaller undecided-list cap, a smaller k_lo margin), still
    give the oracle's permutations and nonces."""
    if l2cap:
        monkeypatch.setenv("ORQG_GP_L2CAP", str(l2cap))
    if shift:
        monkeypatch.setenv("ORQG_GP_BOUND_SHIFT", str(shift))
    oc, gc = _pair(O, G, keys)
    oc.reserve(n, count)
    gc.reserve(n, count)
    for k in range(count):
        for g in range(3):
            members = [p for p in range(3) if p != 2 - g]
            np.testing.assert_array_equal(gc.peek_perm(k, g, n), oc.peek_perm(k, members[0], g, n),
                                          err_msg=f"perm {k} group {g}")
    _same_nonces(oc, gc)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("inverse", [False, True])
def test_apply_sharded(O, G, keys, width, inverse):
    n = 2049
    rng = np.random.default_rng(11)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**31, n).astype(dt)
    oc, gc = _pair(O, G, keys)
    ov, gv = oc.secret_share(x, width), gc.secret_share(x, width)
    oc.apply_sharded(ov, arith=False, inverse=inverse)
    gc.apply_sharded(gv, inverse=inverse)
    _same_shares(ov, gv)
    _same_nonces(oc, gc)
    _same_log(oc, gc)


def test_apply_elementwise(O, G, keys):
    n = 1500
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2**63, n, dtype=np.uint64)
    sigma = rng.permutation(n).astype(np.uint32)
    oc, gc = _pair(O, G, keys)
    ov, gv = oc.secret_share(x, 8), gc.secret_share(x, 8)
    op, gp = oc.secret_share(sigma, 4, arith=True), gc.secret_share(sigma, 4, arith=True)
    oc.apply_elementwise(ov, op, x_arith=False, perm_arith=True)
    gc.apply_elementwise(gv, gp)
    _same_shares(ov, gv)
    want = np.empty_like(x)
    want[sigma] = x
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    np.testing.assert_array_equal(gc.open(gv), want)


def _sort_case(O, G, keys, keys_plain, width, desc):
    n = keys_plain.size
    oc, gc = _pair(O, G, keys)
    ov, gv = oc.secret_share(keys_plain, width), gc.secret_share(keys_plain, width)
    operm = oc.radix_sort(ov, desc=desc)
    gperm = gc.radix_sort(gv, desc=desc)
    _same_shares(ov, gv)
    _same_shares(operm, gperm)
    _same_nonces(oc, gc)
    _same_log(oc, gc)
    assert gc.queue_size() == 0
    sdt = np.int32 if width == 4 else np.int64
    signed = keys_plain.view(sdt)
    # ASC and DESC are both stable w.r.t. the original index (index padding:
    # sorting.h:431-462); ~x reverses signed order without overflow
    order = np.argsort(~signed if desc else signed, kind="stable")
    got = gc.open(gv).view(sdt)
    np.testing.assert_array_equal(got, signed[order])
    perm = gc.open(gperm).view(np.int32)
    ranks = np.empty(n, dtype=np.int64)
    ranks[order] = np.arange(n)
    np.testing.assert_array_equal(perm, ranks)


@pytest.mark.parametrize("desc", [False, True])
def test_radix_sort_int32(O, G, keys, desc):
    rng = np.random.default_rng(1)
    k = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)
    k[500] = 0xFFFFFFFF  # -1
    k[10] = 0x80000000   # INT_MIN
    k[11] = 0x7FFFFFFF   # INT_MAX
    k[20:30] = k[40]     # duplicates
    _sort_case(O, G, keys, k, 4, desc)


@pytest.mark.parametrize("desc", [False, True])
def test_radix_sort_int64(O, G, keys, desc):
    rng = np.random.default_rng(2)
    k = rng.integers(-2**63, 2**63 - 1, 777, dtype=np.int64).view(np.uint64)
    k[3] = np.uint64(1 << 63)
    k[4] = np.uint64((1 << 63) - 1)
    k[100:140] = k[7]
    _sort_case(O, G, keys, k, 8, desc)


@pytest.mark.parametrize("n", [1, 2, 255, 4097])
def test_radix_sort_sizes(O, G, keys, n):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 50, n).astype(np.uint32)
    _sort_case(O, G, keys, k, 4, False)


def test_radix_sort_many_superblocks(G, keys):
    """2^20 + 5 int64 rows: 257 superblocks of 4096 rows, more than the per-bit
    kernels' one workgroup per CU, so a workgroup takes several (the next
    one's words are prefetched) and the last is ragged.  Checked against the
    plaintext stable sort and its ranks; the oracle is too slow at this size."""
    n = (1 << 20) + 5
    rng = np.random.default_rng(9)
    k = rng.integers(-2**40, 2**40, n, dtype=np.int64)
    k[::97] = k[5]  # duplicates: stability decides their order
    gc = G.Ctx(keys, -12)
    gv = gc.secret_share(k.view(np.uint64), 8)
    gperm = gc.radix_sort(gv)
    order = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(gc.open(gv).view(np.int64), k[order])
    ranks = np.empty(n, dtype=np.int64)
    ranks[order] = np.arange(n)
    np.testing.assert_array_equal(gc.open(gperm).view(np.int32), ranks)
    assert gc.queue_size() == 0


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 4097, 70001])
def test_local_ops(O, G, keys, width, n):
    """xor_b, not_b (boolean) and add_a, sub_a (arithmetic): shares equal the
    oracle's, values equal the closed forms, and no message or randomness is
    used (nonces and log unchanged)."""
    rng = np.random.default_rng(n * width)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**63, n, dtype=np.uint64).astype(dt)
    y = rng.integers(0, 2**63, n, dtype=np.uint64).astype(dt)
    oc, gc = _pair(O, G, keys)
    ox, gx = oc.secret_share(x, width), gc.secret_share(x, width)
    oy, gy = oc.secret_share(y, width), gc.secret_share(y, width)
    oa, ga = oc.secret_share(x, width, arith=True), gc.secret_share(x, width, arith=True)
    ob, gb = oc.secret_share(y, width, arith=True), gc.secret_share(y, width, arith=True)
    log0 = gc.msglog()
    with np.errstate(over="ignore"):
        for name, ov, gv, want in [
            ("xor_b", oc.xor_b(ox, oy), gc.xor_b(gx, gy), x ^ y),
            ("not_b", oc.not_b(ox), gc.not_b(gx), ~x),
            ("add_a", oc.add_a(oa, ob), gc.add_a(ga, gb), x + y),
            ("sub_a", oc.sub_a(oa, ob), gc.sub_a(ga, gb), x - y),
        ]:
            _same_shares(ov, gv)
            np.testing.assert_array_equal(gc.open(gv), want, err_msg=name)
    _same_nonces(oc, gc)
    assert [m for m in gc.msglog() if m[0] != "open"] == [m for m in log0 if m[0] != "open"]


@pytest.mark.parametrize("width", [4, 8])
def test_unary_ops(O, G, keys, width):
    """orqg_vec_unary: shares equal the oracle's for every map and shift
    amount (the oracle's maps are pinned by closed forms); in place too."""
    n = 5003
    rng = np.random.default_rng(17 + width)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**63, n, dtype=np.uint64).astype(dt)
    oc, gc = _pair(O, G, keys)
    ox, gx = oc.secret_share(x, width), gc.secret_share(x, width)
    bits = 8 * width
    cases = [("<<", s) for s in (0, 3, bits - 1, bits + 1)] + \
            [(">>", s) for s in (1, bits - 1, bits + 2)] + [("lsr", 7), ("lsr", bits)] + \
            [("parity", 0), ("extend_lsb", 0), ("ltz", 0), ("mask", 0x5555_5555_5555_5555)]
    for op, arg in cases:
        arg = arg & ((1 << 64) - 1)
        _same_shares(oc.unary(op, ox, arg), gc.unary(op, gx, arg))
    gc.unary("mask", gx, 0xFF, out=gx)  # in place
    _same_shares(oc.unary("mask", ox, 0xFF), gx)
    _same_nonces(oc, gc)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 70001, (1 << 20) + 3])
def test_prefix_sum(O, G, keys, width, n):
    """orqg_prefix_sum (one tile, tile edges, many tiles): shares equal the
    oracle's inclusive scans, the value is the running sum mod 2^bits."""
    rng = np.random.default_rng(n + width)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**63, n, dtype=np.uint64).astype(dt)
    oc, gc = _pair(O, G, keys)
    oa, ga = oc.secret_share(x, width, arith=True), gc.secret_share(x, width, arith=True)
    oc.prefix_sum(oa)
    gc.prefix_sum(ga)
    _same_shares(oa, ga)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(gc.open(ga), np.cumsum(x, dtype=dt))
