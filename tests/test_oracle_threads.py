"""The oracle's multi-threaded runtime schedule: T worker threads, each with
its own PRG objects (setup_random_generation per thread,
include/backend/common/setup.h:48-57).

  * thread ranges: RunTime::getThreadBatchBoundaries
    (include/backend/common/runtime.h:198-247), known answers worked out by
    hand from its arithmetic;
  * a T-thread call is, thread by thread, the single-thread call on the
    thread's range with that thread's keys (its own batch schedule, its own
    nonces);
  * reserve deals the permutations to the threads in contiguous runs
    (runtime.h:749-788), thread t drawing from its own group PRGs.
"""
import numpy as np
import pytest

from oracle import oracle as O
from orq_amd import keys_from_seed


def _keys(t):
    return keys_from_seed(0x7E5D0000 + t)


def _ctx(T, batch_size=-12):
    return O.Ctx(_keys(0), batch_size, threads=T, thread_keys=[_keys(t) for t in range(1, T)])


@pytest.mark.parametrize("n,bs,T,want", [
    # B = 256 (negative batch size); 1000 = 3 whole batches + 232: one each
    (1000, -12, 3, [(0, 256), (256, 512), (512, 1000)]),
    # 5000 = 19 whole + 136; 19 = 3 x 6 + 1: thread 0 takes 7, the others 6,
    # the last also the 136
    (5000, -12, 3, [(0, 1792), (1792, 3328), (3328, 5000)]),
    # no whole batch: everything is the last thread's partial batch
    (100, -12, 4, [(0, 0), (0, 0), (0, 0), (0, 100)]),
    # B = 1000: 2 whole + 500, one each, the last takes the 500
    (2500, 1000, 2, [(0, 1000), (1000, 2500)]),
    (1 << 24, -12, 1, [(0, 1 << 24)]),
])
def test_thread_bounds(n, bs, T, want):
    assert O.thread_bounds(n, bs, T) == want


@pytest.mark.parametrize("T", [2, 3])
@pytest.mark.parametrize("op", ["and_b", "mul_a", "b2a_bit", "reshare", "secret_share"])
def test_thread_slices_are_single_thread_calls(op, T):
    n, w = 5000, 8
    rng = np.random.default_rng(T)
    x = rng.integers(0, 2**63, n, dtype=np.uint64)
    y = rng.integers(0, 2**63, n, dtype=np.uint64)
    c = _ctx(T)
    bounds = O.thread_bounds(n, -12, T)

    def run(ctx, xs, ys):
        arith = op == "mul_a"
        if op == "secret_share":
            return ctx.secret_share(xs, w)
        vx, vy = ctx.secret_share(xs, w, arith), ctx.secret_share(ys, w, arith)
        if op == "and_b":
            return ctx.and_b(vx, vy)
        if op == "mul_a":
            return ctx.mul_a(vx, vy)
        if op == "b2a_bit":
            return ctx.b2a_bit(vx)
        ctx.reshare(vx, 1)
        return vx

    got = run(c, x, y).shares()
    for t, (s, e) in enumerate(bounds):
        if s == e:
            continue
        one = O.Ctx(_keys(t))
        want = run(one, x[s:e], y[s:e]).shares()
        for k in range(3):
            np.testing.assert_array_equal(got[k][s:e], want[k], err_msg=f"thread {t} share {k}")
        for p in range(3):
            for role in range(7):
                assert c.nonce(p, role, thread=t) == one.nonce(p, role), (t, p, role)


def test_reserve_deals_permutations_to_threads():
    n, count, T = 300, 7, 3
    c = _ctx(T)
    c.reserve(n, count)
    # 7 = 3 x 2 + 1: thread 0 makes 3, threads 1 and 2 make 2 each
    owner = [0, 0, 0, 1, 1, 2, 2]
    nonce = {}
    for k in range(count):
        t = owner[k]
        for g in range(3):
            key = _keys(t)[32 * (6 + g):32 * (7 + g)]
            want, nxt, _ = O.gen_perm(key, nonce.get((t, g), 0), n)
            nonce[(t, g)] = nxt
            members = [p for p in range(3) if p != 2 - g]
            for p in members:
                np.testing.assert_array_equal(c.peek_perm(k, p, g, n), want,
                                              err_msg=f"perm {k} group {g} party {p}")


@pytest.mark.parametrize("T", [2, 3])
@pytest.mark.parametrize("width,desc", [(4, False), (8, True)])
def test_radix_sort_with_threads(T, width, desc):
    n = 3000
    rng = np.random.default_rng(7 + T)
    dt = np.uint32 if width == 4 else np.uint64
    x = rng.integers(0, 2**(8 * width), n, dtype=np.uint64).astype(dt)
    c = _ctx(T)
    v = c.secret_share(x, width)
    perm = c.radix_sort(v, desc=desc)
    sdt = np.int32 if width == 4 else np.int64
    signed = x.view(sdt)
    order = np.argsort(~signed if desc else signed, kind="stable")
    np.testing.assert_array_equal(c.open(v).view(sdt), signed[order])
    ranks = np.empty(n, dtype=np.int64)
    ranks[order] = np.arange(n)
    np.testing.assert_array_equal(c.open(perm, arith=True).view(np.int32), ranks)
    # every thread drew: its pair nonces moved
    for t in range(T):
        assert c.nonce(0, 2, thread=t) > 0
