"""ctypes binding of oracle/_ref/liborqref.so: the reference's own randomness,
3-party replicated primitives and batch schedule, compiled from its headers in
place (see ref_driver.cpp and the Makefile's `ref` target).

Test infrastructure only, used to check the oracle.  The library is built
where /root/reference exists; `available()` is False elsewhere.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "liborqref.so")
_PARTY = os.path.join(_HERE, "_ref", "orq_ref_party")
# the same program built with the reference's default sort protocol set to
# radix sort (operators/sorting.h:6-9)
_PARTY_RADIX = os.path.join(_HERE, "_ref", "orq_ref_party_radix")
REF_INCLUDE = "/root/reference/include"
_lib = None

_DT = {4: np.uint32, 8: np.uint64}


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def party_built():
    """The party programs are present (they travel with the tree; no sources
    are needed to run them)."""
    return os.path.exists(_PARTY) and os.path.exists(_PARTY_RADIX)


def available():
    built = lambda: all(os.path.exists(f) for f in (_LIB_PATH, _PARTY, _PARTY_RADIX))
    if built():
        return True
    if os.path.isdir(REF_INCLUDE):
        build()
        return built()
    return False


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("the reference's library is not built (needs /root/reference)")
        L = ctypes.CDLL(_LIB_PATH)
        vp, sz, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int
        L.rf_create.restype = vp
        L.rf_create.argtypes = [ctypes.c_char_p, ctypes.c_long]
        L.rf_destroy.argtypes = [vp]
        L.rf_bytes_sent.restype = u64
        L.rf_bytes_sent.argtypes = [vp, i32]
        for name, args in [
            ("rf_and_b", [vp, i32, sz, vp, vp, vp]),
            ("rf_mul_a", [vp, i32, sz, vp, vp, vp]),
            ("rf_b2a_bit", [vp, i32, sz, vp, vp]),
            ("rf_reshare", [vp, i32, sz, vp, i32, i32]),
            ("rf_secret_share", [vp, i32, sz, vp, i32, i32, vp]),
            ("rf_open", [vp, i32, sz, vp, i32, vp]),
            ("rf_zero_sharing", [vp, i32, sz, i32, vp]),
            ("rf_sharded_perms", [vp, sz, i32, vp]),
        ]:
            f = getattr(L, name)
            f.restype = i32
            f.argtypes = args
        L.rf_keystream.restype = i32
        L.rf_keystream.argtypes = [ctypes.c_char_p, u64, i32, sz, vp]
        L.rf_gen_perm.restype = i32
        L.rf_gen_perm.argtypes = [ctypes.c_char_p, u64, sz, i32, vp]
        L.rf_equal_batches.restype = ctypes.c_long
        L.rf_equal_batches.argtypes = [sz, sz, ctypes.c_long]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"reference {what} failed ({rc})")


def keystream(key: bytes, skip: int, width: int, count: int):
    dt = {1: np.uint8, 4: np.uint32, 8: np.uint64}[width]
    out = np.zeros(count, dtype=dt)
    _check(lib().rf_keystream(key, skip, width, count, _ptr(out)), "keystream")
    return out


def gen_perm(key: bytes, skip: int, n: int, count: int):
    out = np.zeros((count, n), dtype=np.int32)
    _check(lib().rf_gen_perm(key, skip, n, count, _ptr(out)), "gen_perm")
    return out


def equal_batches(start, end, batches):
    return lib().rf_equal_batches(start, end, batches)


class Ctx:
    """The reference's three parties (Replicated_3PC, one worker thread each)
    with explicit keys; replicated vectors are arrays [party][idx][n]."""

    def __init__(self, keys: bytes, batch_size=-12):
        assert len(keys) == 320
        self.h = lib().rf_create(keys, batch_size)
        if not self.h:
            raise RuntimeError("rf_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().rf_destroy(self.h)
            self.h = None

    def bytes_sent(self, party):
        return lib().rf_bytes_sent(self.h, party)

    @staticmethod
    def _buf(n, width):
        return np.zeros((3, 2, n), dtype=_DT[width])

    def secret_share(self, data, width, owner=0, binary=True):
        d = np.ascontiguousarray(data, dtype=_DT[width])
        out = self._buf(d.size, width)
        _check(lib().rf_secret_share(self.h, width, d.size, _ptr(d), owner, int(binary), _ptr(out)),
               "secret_share")
        return out

    def and_b(self, x, y):
        z = np.zeros_like(x)
        _check(lib().rf_and_b(self.h, x.itemsize, x.shape[2], _ptr(x), _ptr(y), _ptr(z)), "and_b")
        return z

    def mul_a(self, x, y):
        z = np.zeros_like(x)
        _check(lib().rf_mul_a(self.h, x.itemsize, x.shape[2], _ptr(x), _ptr(y), _ptr(z)), "mul_a")
        return z

    def b2a_bit(self, x):
        y = np.zeros_like(x)
        _check(lib().rf_b2a_bit(self.h, x.itemsize, x.shape[2], _ptr(x), _ptr(y)), "b2a_bit")
        return y

    def reshare(self, v, group, binary=True):
        v = np.ascontiguousarray(v).copy()
        _check(lib().rf_reshare(self.h, v.itemsize, v.shape[2], _ptr(v), group, int(binary)),
               "reshare")
        return v

    def open(self, v, binary=True):
        out = np.zeros((3, v.shape[2]), dtype=v.dtype)
        _check(lib().rf_open(self.h, v.itemsize, v.shape[2], _ptr(v), int(binary), _ptr(out)), "open")
        return out

    def zero_sharing(self, n, width, binary=True):
        out = np.zeros((3, n), dtype=_DT[width])
        _check(lib().rf_zero_sharing(self.h, width, n, int(binary), _ptr(out)), "zero_sharing")
        return out

    def sharded_perms(self, n, count):
        """out[k, party, group] for groups {01},{02},{12},{012}; -1 where the
        party is not in the group."""
        out = np.full((count, 3, 4, n), -1, dtype=np.int32)
        _check(lib().rf_sharded_perms(self.h, n, count, _ptr(out)), "sharded_perms")
        return out


# ---- the reference's runtime and operators, three party processes ----------

RADIX, QUICKSORT, TABLE_SORT, JOIN_GROUPBY, DISTINCT, AGGREGATE, SHUFFLE, CIRCUITS, RADIX_BENCH = range(1, 10)
# orq::SortingProtocol (operators/sorting.h:15-21)
BITONICSORT_ID, QUICKSORT_ID, RADIXSORT_ID = 0, 1, 2


def run_parties(keys: bytes, scenario: int, params, payload: bytes, batch_size=-12, timeout=900,
                default_sort="QUICKSORT", thread_keys=()):
    """Run `scenario` of orq_ref_party as parties 0, 1, 2 connected by a
    socket pair per pair of parties and worker thread; returns each party's
    output bytes.  default_sort: the reference's DEFAULT sort protocol,
    QUICKSORT (as it ships) or RADIXSORT.  thread_keys: the key blocks of
    worker threads 1, 2, ... (thread 0 has `keys`)."""
    import socket
    import tempfile

    if not party_built():
        raise RuntimeError("the reference's party program is not built (needs /root/reference)")
    threads = 1 + len(thread_keys)
    assert len(keys) == 320 and all(len(k) == 320 for k in thread_keys) and len(params) <= 8
    prm = list(params) + [0] * (8 - len(params))
    hdr = np.array([batch_size, scenario] + prm, dtype=np.int64).tobytes()
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "in")
        with open(inp, "wb") as f:
            f.write(np.int64(threads).tobytes() + keys + b"".join(thread_keys) + hdr + payload)
        pairs = [{(0, 1): socket.socketpair(), (0, 2): socket.socketpair(),
                  (1, 2): socket.socketpair()} for _ in range(threads)]

        def end(t, p, q):  # thread t's end at party p of its connection to party q
            a, b = min(p, q), max(p, q)
            return pairs[t][(a, b)][0 if p == a else 1].fileno()

        exe = {"QUICKSORT": _PARTY, "RADIXSORT": _PARTY_RADIX}[default_sort]
        procs, outs = [], []
        try:
            for p in range(3):
                fds = []
                for t in range(threads):
                    fds += [end(t, p, (p + 1) % 3), end(t, p, (p + 2) % 3)]
                out = os.path.join(d, f"out{p}")
                outs.append(out)
                procs.append(subprocess.Popen([exe, str(p), ",".join(map(str, fds)), inp, out],
                                              pass_fds=fds, stderr=subprocess.PIPE))
        finally:
            for ps in pairs:
                for s in ps.values():
                    s[0].close()
                    s[1].close()
        errs = []
        for p, proc in enumerate(procs):
            try:
                _, err = proc.communicate(timeout=timeout)
            except subprocess.TimeoutExpired:
                for q in procs:
                    q.kill()
                raise
            if proc.returncode != 0:
                errs.append(f"party {p}: rc {proc.returncode}: {err.decode(errors='replace')}")
        if errs:
            raise RuntimeError("; ".join(errs))
        res = []
        for out in outs:
            with open(out, "rb") as f:
                res.append(f.read())
        return res


class Reader:
    """Sequential reader of one party's output."""

    def __init__(self, b):
        self.b, self.off = b, 0

    def arr(self, dtype, n):
        a = np.frombuffer(self.b, dtype=dtype, count=n, offset=self.off)
        self.off += a.nbytes
        return a.copy()

    def pair(self, dtype, n):
        return np.stack([self.arr(dtype, n), self.arr(dtype, n)])

    def u64(self):
        return int(self.arr(np.uint64, 1)[0])

    def done(self):
        return self.off == len(self.b)


def radix_sort(keys, data, bits=None, desc=False, batch_size=-12, thread_keys=()):
    """secret_share_b(data, 0), then radix_sort: ([party][idx][n] shares of
    the sorted vector, of the permutation, bytes sent per party)."""
    d = np.ascontiguousarray(data)
    w = d.itemsize
    bits = 8 * w if bits is None else bits
    outs = run_parties(keys, RADIX, [w, bits, int(desc), 0, 0, 0, 0, d.size], d.tobytes(),
                       batch_size, thread_keys=thread_keys)
    return _sorted_outputs(outs, d.dtype, d.size)


def quicksort(keys, data, desc=False, batch_size=-12):
    """secret_share_b(data, 0), then quicksort (32- or 64-bit keys)."""
    d = np.ascontiguousarray(data)
    outs = run_parties(keys, QUICKSORT, [d.itemsize, 8 * d.itemsize, int(desc), 0, 0, 0, 0, d.size],
                       d.tobytes(), batch_size)
    return _sorted_outputs(outs, d.dtype, d.size)


def _sorted_outputs(outs, dtype, n):
    v, perm, sent = [], [], []
    for b in outs:
        r = Reader(b)
        v.append(r.pair(dtype, n))
        perm.append(r.pair(np.uint32, n))
        sent.append(r.u64())
        assert r.done()
    return np.stack(v), np.stack(perm), sent


def _tables(outs, width):
    """Each party's table: {name: (encoding, [2][n] shares)} -> combined
    {name: (encoding, [party][idx][n])}, and bytes sent per party."""
    dt = _DT[width]
    per, sent = [], []
    for b in outs:
        r = Reader(b)
        cols = {}
        for _ in range(r.u64()):
            name = r.arr(np.uint8, r.u64()).tobytes().decode()
            enc = r.u64()
            n = r.u64()
            cols[name] = (enc, r.pair(dt, n))
        sent.append(r.u64())
        assert r.done()
        per.append(cols)
    assert per[0].keys() == per[1].keys() == per[2].keys()
    table = {k: (per[0][k][0], np.stack([per[p][k][1] for p in range(3)])) for k in per[0]}
    return table, sent


def _payload(cols, width):
    dt = _DT[width]
    return b"".join(np.asarray(c).astype(np.int64).astype(dt).tobytes() for c in cols)


def table_sort(keys, cols, width, nkeys, protocol, desc=False, batch_size=-12, thread_keys=()):
    """{[K1], ([K2],) [P], A} from `cols`, then sort on the key columns."""
    n = len(cols[0])
    outs = run_parties(keys, TABLE_SORT, [width, nkeys, protocol, int(desc), 0, 0, 0, n],
                       _payload(cols, width), batch_size, thread_keys=thread_keys)
    return _tables(outs, width)


def join_groupby(keys, rkey, attr, lkey, value, width, batch_size=-12, default_sort="QUICKSORT"):
    outs = run_parties(keys, JOIN_GROUPBY, [width, len(rkey), len(lkey), 0, 0, 0, 0, 0],
                       _payload([rkey, attr, lkey, value], width), batch_size,
                       default_sort=default_sort)
    return _tables(outs, width)


def table_distinct(keys, k, v, width, batch_size=-12, default_sort="QUICKSORT"):
    outs = run_parties(keys, DISTINCT, [width, 0, 0, 0, 0, 0, 0, len(k)],
                       _payload([k, v], width), batch_size, default_sort=default_sort)
    return _tables(outs, width)


def table_aggregate(keys, cols, width, reverse=False, batch_size=-12, default_sort="QUICKSORT"):
    """{[K], A, C, [B], [M]}, aggregate on [K]: sum A, count into C, max [B], min [M]."""
    outs = run_parties(keys, AGGREGATE, [width, int(reverse), 0, 0, 0, 0, 0, len(cols[0])],
                       _payload(cols, width), batch_size, default_sort=default_sort)
    return _tables(outs, width)


def table_shuffle(keys, cols, width, batch_size=-12):
    outs = run_parties(keys, SHUFFLE, [width, 0, 0, 0, 0, 0, 0, len(cols[0])],
                       _payload(cols, width), batch_size)
    return _tables(outs, width)


CIRCUIT_OUTPUTS = ["==", "!=", "<", ">", "<=", ">=", "b2a", "+", "-", "a2b"]


def circuits(keys, x, y, batch_size=-12, thread_keys=()):
    """x, y boolean-shared and x arithmetic-shared from party 0, then the
    outputs CIRCUIT_OUTPUTS in order: {name: [party][idx][n]}, bytes sent."""
    x, y = np.ascontiguousarray(x), np.ascontiguousarray(y, dtype=x.dtype)
    outs = run_parties(keys, CIRCUITS, [x.itemsize, 0, 0, 0, 0, 0, 0, x.size],
                       x.tobytes() + y.tobytes(), batch_size, thread_keys=thread_keys)
    res, sent = {k: [] for k in CIRCUIT_OUTPUTS}, []
    for b in outs:
        r = Reader(b)
        for k in CIRCUIT_OUTPUTS:
            res[k].append(r.pair(x.dtype, x.size))
        sent.append(r.u64())
        assert r.done()
    return {k: np.stack(v) for k, v in res.items()}, sent


def radix_bench(keys, data, thread_keys=(), batch_size=-12, timeout=900):
    """The reference's radix_sort timed in its three party processes: (the
    sorted shares [party][idx][n], reserve seconds, total seconds), the times
    the slowest party's."""
    d = np.ascontiguousarray(data)
    outs = run_parties(keys, RADIX_BENCH, [d.itemsize, 8 * d.itemsize, 0, 0, 0, 0, 0, d.size],
                       d.tobytes(), batch_size, timeout=timeout, thread_keys=thread_keys)
    v, t_res, t_all = [], 0.0, 0.0
    for b in outs:
        r = Reader(b)
        v.append(r.pair(d.dtype, d.size))
        t_res = max(t_res, float(r.arr(np.float64, 1)[0]))
        t_all = max(t_all, float(r.arr(np.float64, 1)[0]))
        r.u64()
        assert r.done()
    return np.stack(v), t_res, t_all
