"""ctypes binding of the CPU oracle (liborqoracle.so).

Test infrastructure only: the tests, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module; the GPU library never
does.  The shared library is a restatement of ORQ's 3-party replicated
protocol path (see orq_oracle.cpp for the per-function reference citations).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborqoracle.so")
_lib = None

_DT = {4: np.uint32, 8: np.uint64}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, sz, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int
        L.oc_create.restype = vp
        L.oc_create.argtypes = [ctypes.c_char_p, ctypes.c_long, i32]
        L.oc_destroy.argtypes = [vp]
        L.oc_prg_nonce.restype = u64
        L.oc_prg_nonce.argtypes = [vp, i32, i32]
        L.oc_create_threads.restype = vp
        L.oc_create_threads.argtypes = [ctypes.c_char_p, i32, ctypes.c_long, i32]
        L.oc_prg_nonce_thread.restype = u64
        L.oc_prg_nonce_thread.argtypes = [vp, i32, i32, i32]
        L.oc_thread_bounds.restype = None
        L.oc_thread_bounds.argtypes = [sz, ctypes.c_long, i32, vp]
        L.oc_keystream.restype = u64
        L.oc_keystream.argtypes = [ctypes.c_char_p, u64, vp, sz]
        L.oc_gen_perm.argtypes = [ctypes.c_char_p, ctypes.POINTER(u64), vp, sz, ctypes.POINTER(u64)]
        L.oc_vec_new.restype = vp
        L.oc_vec_new.argtypes = [sz, i32]
        L.oc_vec_free.argtypes = [vp]
        L.oc_vec_set.argtypes = [vp, i32, i32, vp]
        L.oc_vec_get.argtypes = [vp, i32, i32, vp]
        for name, args in [
            ("oc_secret_share", [vp, vp, vp, i32]),
            ("oc_open", [vp, vp, i32, vp]),
            ("oc_and_b", [vp, vp, vp, vp]),
            ("oc_mul_a", [vp, vp, vp, vp]),
            ("oc_b2a_bit", [vp, vp, vp]),
            ("oc_reshare", [vp, vp, i32, i32]),
            ("oc_reserve", [vp, sz, sz]),
            ("oc_peek_perm", [vp, sz, i32, i32, vp]),
            ("oc_apply_sharded", [vp, vp, i32, i32]),
            ("oc_apply_elementwise", [vp, vp, i32, vp, i32]),
            ("oc_radix_sort", [vp, vp, i32, i32, vp]),
            ("oc_compare", [vp, vp, vp, vp, vp]),
            ("oc_b2a", [vp, vp, vp]),
            ("oc_cmp", [vp, i32, vp, vp, vp]),
            ("oc_perm_invert", [vp, vp, i32]),
            ("oc_perm_b2a", [vp, vp]),
            ("oc_perm_negate", [vp, vp]),
            ("oc_perm_reverse", [vp, vp]),
            ("oc_perm_compose", [vp, vp, vp, i32]),
            ("oc_shuffle", [vp, vp, i32]),
            ("oc_table_shuffle", [vp, vp, sz, vp, sz]),
            ("oc_table_sort", [vp, vp, sz, vp, vp, vp, sz, vp, sz, i32]),
            ("oc_quicksort", [vp, vp, i32, vp]),
            ("oc_bitonic_sort", [vp, vp, sz, vp, vp, sz, vp, sz]),
            ("oc_aggregate", [vp, vp, sz, vp, vp, vp, sz, vp, vp, vp, sz, i32, vp]),
            ("oc_distinct", [vp, vp, sz, vp]),
            ("oc_local", [vp, i32, vp, vp, vp]),
            ("oc_unary", [vp, i32, vp, u64, vp]),
            ("oc_prefix_sum", [vp, vp]),
            ("oc_vec_resize", [vp, sz]),
            ("oc_adder", [vp, i32, vp, vp, vp]),
            ("oc_secret_share_from", [vp, vp, vp, i32, i32]),
            ("oc_vec_tail", [vp, sz]),
            ("oc_vec_copy_rows", [vp, sz, vp, sz, sz]),
            ("oc_vec_broadcast_row", [vp, sz, vp, sz, sz]),
        ]:
            f = getattr(L, name)
            f.restype = i32
            f.argtypes = args
        L.oc_batches.restype = sz
        L.oc_batches.argtypes = [sz, ctypes.c_long, vp, sz]
        L.oc_queue_size.restype = sz
        L.oc_queue_size.argtypes = [vp]
        L.oc_msglog_count.restype = sz
        L.oc_msglog_count.argtypes = [vp]
        L.oc_msglog_get.argtypes = [vp, sz, ctypes.c_char_p, ctypes.POINTER(i32),
                                    ctypes.POINTER(i32), ctypes.POINTER(u64)]
        L.oc_msglog_clear.argtypes = [vp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def keystream(key: bytes, nonce: int, nbytes: int):
    out = np.zeros(nbytes, dtype=np.uint8)
    nxt = lib().oc_keystream(key, nonce, _ptr(out), nbytes)
    return out, nxt


def gen_perm(key: bytes, nonce: int, n: int):
    out = np.zeros(n, dtype=np.int32)
    nn = ctypes.c_uint64(nonce)
    draws = ctypes.c_uint64(0)
    lib().oc_gen_perm(key, ctypes.byref(nn), _ptr(out), n, ctypes.byref(draws))
    return out, nn.value, draws.value


def batches(n, batch_size=-12):
    k = lib().oc_batches(n, batch_size, None, 0)
    out = np.zeros(max(k, 1), dtype=np.uint64)
    lib().oc_batches(n, batch_size, _ptr(out), out.size)
    return [int(x) for x in out[:k]]


def thread_bounds(n, batch_size, threads):
    """The threads' element ranges of one runtime call over n elements."""
    out = np.zeros(2 * threads, dtype=np.uint64)
    lib().oc_thread_bounds(n, batch_size, threads, _ptr(out))
    return [(int(out[2 * t]), int(out[2 * t + 1])) for t in range(threads)]


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed")


class Vec:
    """A replicated vector: party p holds (s_p, s_{p+1})."""

    def __init__(self, n, width):
        self.n, self.width = n, width
        self.h = lib().oc_vec_new(n, width)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_vec_free(self.h)
            self.h = None

    def set(self, party, idx, arr):
        a = np.ascontiguousarray(arr, dtype=_DT[self.width])
        assert a.size == self.n
        lib().oc_vec_set(self.h, party, idx, _ptr(a))

    def get(self, party, idx):
        a = np.zeros(self.n, dtype=_DT[self.width])
        lib().oc_vec_get(self.h, party, idx, _ptr(a))
        return a

    def shares(self):
        """The three distinct shares (s_0, s_1, s_2) as seen by their first holder."""
        return [self.get(p, 0) for p in range(3)]

    def set_shares(self, s):
        for p in range(3):
            self.set(p, 0, s[p])
            self.set(p, 1, s[(p + 1) % 3])


class Ctx:
    """The three parties of one ORQ run.  threads > 1: that many worker
    threads, thread 0 with `keys` and thread t with thread_keys[t - 1]."""

    def __init__(self, keys: bytes, batch_size=-12, gen_everyone=False, threads=1,
                 thread_keys=()):
        assert len(keys) == 320
        self.threads = threads
        if threads == 1:
            self.h = lib().oc_create(keys, batch_size, int(gen_everyone))
        else:
            assert len(thread_keys) == threads - 1 and all(len(k) == 320 for k in thread_keys)
            self.h = lib().oc_create_threads(keys + b"".join(thread_keys), threads, batch_size,
                                             int(gen_everyone))
            if not self.h:
                raise RuntimeError("oracle oc_create_threads: thread count")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_destroy(self.h)
            self.h = None

    def nonce(self, party, role, thread=0):
        return lib().oc_prg_nonce_thread(self.h, thread, party, role)

    def secret_share(self, data, width, arith=False):
        a = np.ascontiguousarray(data, dtype=_DT[width])
        v = Vec(a.size, width)
        _check(lib().oc_secret_share(self.h, _ptr(a), v.h, int(arith)), "secret_share")
        return v

    def open(self, v, arith=False):
        out = np.zeros(v.n, dtype=_DT[v.width])
        _check(lib().oc_open(self.h, v.h, int(arith), _ptr(out)), "open")
        return out

    LOCAL = {"xor_b": 0, "add_a": 1, "sub_a": 2, "not_b": 3, "neg_a": 4}

    def local(self, op, x, y=None):
        """Message-free share ops: xor_b, add_a, sub_a, not_b, neg_a."""
        z = Vec(x.n, x.width)
        _check(lib().oc_local(self.h, self.LOCAL[op], x.h, (y or x).h, z.h), op)
        return z

    UNARY = {"<<": 0, ">>": 1, "lsr": 2, "parity": 3, "extend_lsb": 4, "ltz": 5, "mask": 6}

    def unary(self, op, x, arg=0):
        """Share-local maps: '<<', '>>' (arithmetic), 'lsr', 'parity',
        'extend_lsb', 'ltz', 'mask' (x & arg)."""
        z = Vec(x.n, x.width)
        _check(lib().oc_unary(self.h, self.UNARY[op], x.h, arg, z.h), op)
        return z

    def prefix_sum(self, v):
        _check(lib().oc_prefix_sum(self.h, v.h), "prefix_sum")

    def xor_b(self, x, y):
        return self.local("xor_b", x, y)

    def add_a(self, x, y):
        return self.local("add_a", x, y)

    def sub_a(self, x, y):
        return self.local("sub_a", x, y)

    def not_b(self, x):
        return self.local("not_b", x)

    def and_b(self, x, y):
        z = Vec(x.n, x.width)
        _check(lib().oc_and_b(self.h, x.h, y.h, z.h), "and_b")
        return z

    def mul_a(self, x, y):
        z = Vec(x.n, x.width)
        _check(lib().oc_mul_a(self.h, x.h, y.h, z.h), "mul_a")
        return z

    def b2a_bit(self, x):
        y = Vec(x.n, x.width)
        _check(lib().oc_b2a_bit(self.h, x.h, y.h), "b2a_bit")
        return y

    def reshare(self, v, group, arith=False):
        _check(lib().oc_reshare(self.h, v.h, group, int(arith)), "reshare")

    def reserve(self, n, count):
        _check(lib().oc_reserve(self.h, n, count), "reserve")

    def queue_size(self):
        return lib().oc_queue_size(self.h)

    def peek_perm(self, index, party, group, n):
        out = np.zeros(n, dtype=np.int32)
        _check(lib().oc_peek_perm(self.h, index, party, group, _ptr(out)), "peek_perm")
        return out

    def apply_sharded(self, x, arith=False, inverse=False):
        _check(lib().oc_apply_sharded(self.h, x.h, int(arith), int(inverse)), "apply_sharded")

    def apply_elementwise(self, x, perm, x_arith=False, perm_arith=True):
        _check(lib().oc_apply_elementwise(self.h, x.h, int(x_arith), perm.h, int(perm_arith)),
               "apply_elementwise")

    def radix_sort(self, v, bits=None, desc=False):
        if bits is None:
            bits = 8 * v.width
        perm = Vec(v.n, 4)
        _check(lib().oc_radix_sort(self.h, v.h, bits, int(desc), perm.h), "radix_sort")
        return perm

    def compare(self, x, y):
        """x._compare(y): (eq, gt) boolean shares, signed order."""
        eq, gt = Vec(x.n, x.width), Vec(x.n, x.width)
        _check(lib().oc_compare(self.h, x.h, y.h, eq.h, gt.h), "compare")
        return eq, gt

    CMP_OPS = {"==": 0, "!=": 1, "<": 2, ">": 3, "<=": 4, ">=": 5}

    def cmp(self, op, x, y):
        """BSharedVector comparison operator `op` ('==', '!=', '<', '>', '<=', '>=')."""
        z = Vec(x.n, x.width)
        _check(lib().oc_cmp(self.h, self.CMP_OPS[op], x.h, y.h, z.h), "cmp")
        return z

    def b2a(self, x):
        y = Vec(x.n, x.width)
        _check(lib().oc_b2a(self.h, x.h, y.h), "b2a")
        return y

    def perm_identity(self, n):
        p = Vec(n, 4)
        ident = np.arange(n, dtype=np.uint32)
        z = np.zeros(n, np.uint32)
        p.set_shares([ident, z, z])  # public_share: s_0 = value
        return p

    def perm_invert(self, p):
        _check(lib().oc_perm_invert(self.h, p.h, 1), "perm_invert")

    def perm_b2a(self, p):
        _check(lib().oc_perm_b2a(self.h, p.h), "perm_b2a")

    def perm_negate(self, p):
        _check(lib().oc_perm_negate(self.h, p.h), "perm_negate")

    def perm_reverse(self, p):
        _check(lib().oc_perm_reverse(self.h, p.h), "perm_reverse")

    def perm_compose(self, sigma, rho, arith=True):
        _check(lib().oc_perm_compose(self.h, sigma.h, rho.h, int(arith)), "perm_compose")

    def shuffle(self, x, arith=False):
        _check(lib().oc_shuffle(self.h, x.h, int(arith)), "shuffle")

    @staticmethod
    def _vec_array(cols):
        arr = (ctypes.c_void_p * max(len(cols), 1))(*[c.h for c in cols])
        return arr, len(cols)

    def table_shuffle(self, a_cols, b_cols):
        a, na = self._vec_array(a_cols)
        b, nb = self._vec_array(b_cols)
        _check(lib().oc_table_shuffle(self.h, a, na, b, nb), "table_shuffle")

    PROTOCOLS = {"RADIXSORT": 0, "QUICKSORT": 1}

    def table_sort(self, keys, orders=None, single_bit=None, a_cols=(), b_cols=(),
                   protocol="RADIXSORT"):
        k, nk = self._vec_array(keys)
        o = (ctypes.c_int * nk)(*[1 if x in (1, "DESC") else 0 for x in (orders or [0] * nk)])
        sb = (ctypes.c_int * nk)(*[int(bool(x)) for x in (single_bit or [0] * nk)])
        a, na = self._vec_array(list(a_cols))
        b, nb = self._vec_array(list(b_cols))
        _check(lib().oc_table_sort(self.h, k, nk, o, sb, a, na, b, nb, self.PROTOCOLS[protocol]),
               "table_sort")

    def quicksort(self, v, desc=False):
        """operators::quicksort of 32-bit B-shared keys; returns the A-shared
        sorting permutation (quicksort.h:212-244)."""
        perm = Vec(v.n, 4)
        _check(lib().oc_quicksort(self.h, v.h, int(desc), perm.h), "quicksort")
        return perm

    def bitonic_sort(self, keys, orders=None, a_cols=(), b_cols=()):
        k, nk = self._vec_array(keys)
        o = (ctypes.c_int * nk)(*[1 if x in (1, "DESC") else 0 for x in (orders or [0] * nk)])
        a, na = self._vec_array(list(a_cols))
        b, nb = self._vec_array(list(b_cols))
        _check(lib().oc_bitonic_sort(self.h, k, nk, o, a, na, b, nb), "bitonic_sort")

    AGG = {"sum": 0, "count": 1, "copy": 2, "max": 3, "min": 4, "or": 5, "valid": 6}

    def aggregate(self, keys, b_specs=(), a_specs=(), reverse=False, sel=None):
        """aggregators::aggregate; specs are (in, out, fn-name) tuples."""
        k, nk = self._vec_array(list(keys))
        bi, _ = self._vec_array([x[0] for x in b_specs])
        bo, nb = self._vec_array([x[1] for x in b_specs])
        bf = (ctypes.c_int * max(nb, 1))(*[self.AGG[x[2]] for x in b_specs])
        ai, _ = self._vec_array([x[0] for x in a_specs])
        ao, na = self._vec_array([x[1] for x in a_specs])
        af = (ctypes.c_int * max(na, 1))(*[self.AGG[x[2]] for x in a_specs])
        _check(lib().oc_aggregate(self.h, k, nk, bi, bo, bf, nb, ai, ao, af, na, int(reverse),
                                  sel.h if sel is not None else None), "aggregate")

    def distinct(self, keys, res):
        k, nk = self._vec_array(list(keys))
        _check(lib().oc_distinct(self.h, k, nk, res.h), "distinct")

    ADDER = {"add": 0, "sub": 1, "ppa_add": 2, "ppa_sub": 3, "rca_lt": 4, "a2b": 5}

    def adder(self, op, x, y=None):
        """ripple_carry_adder / parallel_prefix_adder / rca_compare / a2b."""
        z = Vec(x.n, x.width)
        _check(lib().oc_adder(self.h, self.ADDER[op], x.h, (y or x).h, z.h), op)
        return z

    def secret_share_from(self, data, width, owner, arith=False):
        a = np.ascontiguousarray(data, dtype=_DT[width])
        v = Vec(a.size, width)
        _check(lib().oc_secret_share_from(self.h, _ptr(a), v.h, int(arith), owner),
               "secret_share_from")
        return v

    def msglog(self):
        L = lib()
        out = []
        op = ctypes.create_string_buffer(24)
        f, t, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        for i in range(L.oc_msglog_count(self.h)):
            L.oc_msglog_get(self.h, i, op, ctypes.byref(f), ctypes.byref(t), ctypes.byref(b))
            out.append((op.value.decode(), f.value, t.value, b.value))
        return out

    def clear_msglog(self):
        lib().oc_msglog_clear(self.h)
