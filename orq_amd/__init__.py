"""orq_amd -- MI355X execution backend for ORQ's 3-party replicated oblivious operators.

The compute lives in ``liborqg.so`` (hand-written HIP kernels for gfx950 behind
the C ABI declared in ``include/orqg.h``).  This module is a thin ctypes
binding used by the tests and the benchmark; it never computes anything on the
CPU.  If the library is missing it raises instead of falling back.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborqg.so")
_lib = None

ORQG_BSHARED, ORQG_ASHARED = 1, 0
ASC, DESC = 0, 1
_DT = {4: np.uint32, 8: np.uint64}

ROLE_LOCAL = (0, 1, 2)
ROLE_PAIR = (3, 4, 5)
ROLE_GROUP = (6, 7, 8, 9)


class OrqgError(RuntimeError):
    pass


class _Config(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("keys", (ctypes.c_uint8 * 32) * 10),
        ("batch_size", ctypes.c_long),
        ("hbm_budget", ctypes.c_size_t),
        ("threads", ctypes.c_int),
        ("thread_keys", ctypes.c_char_p),
    ]


class _Msg(ctypes.Structure):
    _fields_ = [("op", ctypes.c_char * 24), ("from_", ctypes.c_int), ("to", ctypes.c_int),
                ("bytes", ctypes.c_uint64)]


EXPORTS = [
    "orqg_ctx_create", "orqg_ctx_destroy", "orqg_last_error", "orqg_sync", "orqg_stream",
    "orqg_prg_nonce", "orqg_prg_nonce_thread", "orqg_dev_alloc", "orqg_dev_free", "orqg_dev_to_host", "orqg_vec_create", "orqg_vec_free", "orqg_vec_size", "orqg_vec_width",
    "orqg_vec_data", "orqg_vec_write", "orqg_vec_read", "orqg_vec_copy", "orqg_prg_keystream",
    "orqg_secret_share", "orqg_open", "orqg_public_share", "orqg_xor_b", "orqg_add_a",
    "orqg_sub_a", "orqg_not_b", "orqg_and_b", "orqg_mul_a", "orqg_b2a_bit", "orqg_reshare",
    "orqg_compare", "orqg_cmp", "orqg_b2a",
    "orqg_perm_identity", "orqg_perm_invert", "orqg_perm_b2a", "orqg_perm_negate",
    "orqg_perm_reverse", "orqg_perm_compose", "orqg_shuffle", "orqg_table_shuffle",
    "orqg_table_sort", "orqg_bitonic_sort", "orqg_vec_resize", "orqg_vec_tail",
    "orqg_aggregate", "orqg_distinct", "orqg_distinct_noalloc", "orqg_vec_unary",
    "orqg_prefix_sum", "orqg_quicksort", "orqg_table_sort_protocol",
    "orqg_vec_read_share_files", "orqg_vec_write_share_files", "orqg_table_read_share_csv",
    "orqg_table_write_share_csv", "orqg_table_read_csv_data", "orqg_public_fill", "orqg_vec_copy_rows",
    "orqg_vec_broadcast_row", "orqg_ctx_rekey", "orqg_ctx_set_default_sort",
    "orqg_ctx_default_sort", "orqg_build_info", "orqg_add_b", "orqg_sub_b", "orqg_ppa", "orqg_rca_compare",
    "orqg_a2b", "orqg_secret_share_from", "orqg_server_create", "orqg_server_submit",
    "orqg_server_poll", "orqg_server_pending", "orqg_server_get_stats", "orqg_server_now_ms",
    "orqg_server_destroy",
    "orqg_reserve_perms", "orqg_perm_queue_size", "orqg_peek_perm", "orqg_apply_sharded_perm",
    "orqg_et_create", "orqg_et_free", "orqg_et_rows", "orqg_et_ncols", "orqg_et_column_name",
    "orqg_et_column", "orqg_et_set_column", "orqg_et_add_columns", "orqg_et_delete_columns",
    "orqg_et_project", "orqg_et_filter", "orqg_et_resize", "orqg_et_tail", "orqg_et_shuffle",
    "orqg_et_sort", "orqg_et_distinct", "orqg_et_aggregate", "orqg_et_concatenate",
    "orqg_et_inner_join",
    "orqg_apply_elementwise_perm", "orqg_radix_sort", "orqg_radix_sort_batch",
    "orqg_msglog_count", "orqg_msglog_get",
    "orqg_msglog_clear", "orqg_timer_enable", "orqg_timer_get", "orqg_timer_reset",
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OrqgError(f"{LIB_PATH} is missing: build it with `python orq_amd/build.py` "
                            "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
        sig = {
            "orqg_ctx_create": (i32, [ctypes.POINTER(_Config), ctypes.POINTER(vp)]),
            "orqg_ctx_destroy": (i32, [vp]),
            "orqg_last_error": (ctypes.c_char_p, [vp]),
            "orqg_sync": (i32, [vp]),
            "orqg_stream": (vp, [vp]),
            "orqg_prg_nonce": (i32, [vp, i32, ctypes.POINTER(u64)]),
            "orqg_prg_nonce_thread": (i32, [vp, i32, i32, ctypes.POINTER(u64)]),
            "orqg_dev_alloc": (i32, [vp, sz, ctypes.POINTER(vp)]),
            "orqg_dev_free": (i32, [vp]),
            "orqg_dev_to_host": (i32, [vp, vp, vp, sz]),
            "orqg_vec_create": (i32, [vp, sz, i32, i32, ctypes.POINTER(vp)]),
            "orqg_vec_free": (i32, [vp]),
            "orqg_vec_size": (sz, [vp]),
            "orqg_vec_width": (i32, [vp]),
            "orqg_vec_data": (vp, [vp]),
            "orqg_vec_write": (i32, [vp, vp, vp]),
            "orqg_vec_read": (i32, [vp, vp, vp]),
            "orqg_vec_copy": (i32, [vp, vp, vp]),
            "orqg_prg_keystream": (i32, [vp, i32, sz, vp]),
            "orqg_secret_share": (i32, [vp, vp, vp]),
            "orqg_open": (i32, [vp, vp, vp]),
            "orqg_public_share": (i32, [vp, vp, vp]),
            "orqg_xor_b": (i32, [vp, vp, vp, vp]),
            "orqg_add_a": (i32, [vp, vp, vp, vp]),
            "orqg_sub_a": (i32, [vp, vp, vp, vp]),
            "orqg_not_b": (i32, [vp, vp, vp]),
            "orqg_and_b": (i32, [vp, vp, vp, vp]),
            "orqg_mul_a": (i32, [vp, vp, vp, vp]),
            "orqg_b2a_bit": (i32, [vp, vp, vp]),
            "orqg_compare": (i32, [vp, vp, vp, vp, vp]),
            "orqg_cmp": (i32, [vp, i32, vp, vp, vp]),
            "orqg_b2a": (i32, [vp, vp, vp]),
            "orqg_perm_identity": (i32, [vp, vp, i32]),
            "orqg_perm_invert": (i32, [vp, vp]),
            "orqg_perm_b2a": (i32, [vp, vp]),
            "orqg_perm_negate": (i32, [vp, vp]),
            "orqg_perm_reverse": (i32, [vp, vp]),
            "orqg_perm_compose": (i32, [vp, vp, vp]),
            "orqg_shuffle": (i32, [vp, vp]),
            "orqg_table_shuffle": (i32, [vp, vp, sz, vp, sz]),
            "orqg_table_sort": (i32, [vp, vp, sz, vp, vp, vp, sz, vp, sz]),
            "orqg_bitonic_sort": (i32, [vp, vp, sz, vp, vp, sz, vp, sz]),
            "orqg_vec_resize": (i32, [vp, vp, sz]),
            "orqg_vec_tail": (i32, [vp, vp, sz]),
            "orqg_aggregate": (i32, [vp, vp, sz, vp, sz, vp, sz, i32, vp]),
            "orqg_distinct": (i32, [vp, vp, sz, vp]),
            "orqg_distinct_noalloc": (i32, [vp, vp, sz, vp]),
            "orqg_vec_unary": (i32, [vp, i32, vp, u64, vp]),
            "orqg_prefix_sum": (i32, [vp, vp]),
            "orqg_quicksort": (i32, [vp, vp, i32, vp]),
            "orqg_vec_read_share_files": (i32, [vp, vp, vp]),
            "orqg_vec_write_share_files": (i32, [vp, vp, vp]),
            "orqg_table_read_share_csv": (i32, [vp, vp, vp, vp, sz]),
            "orqg_table_write_share_csv": (i32, [vp, vp, vp, vp, sz]),
            "orqg_table_read_csv_data": (i32, [vp, ctypes.c_char_p, vp, vp, sz, i32]),
            "orqg_table_sort_protocol": (i32, [vp, vp, sz, vp, vp, vp, sz, vp, sz, i32]),
            "orqg_public_fill": (i32, [vp, vp, u64, i32]),
            "orqg_vec_copy_rows": (i32, [vp, vp, sz, vp, sz, sz]),
            "orqg_vec_broadcast_row": (i32, [vp, vp, sz, vp, sz, sz]),
            "orqg_add_b": (i32, [vp, vp, vp, vp]),
            "orqg_ctx_rekey": (i32, [vp, ctypes.c_char_p]),
            "orqg_ctx_set_default_sort": (i32, [vp, i32]),
            "orqg_ctx_default_sort": (i32, [vp]),
            "orqg_build_info": (ctypes.c_char_p, []),
            "orqg_server_create": (i32, [vp, vp]),
            "orqg_server_submit": (i32, [vp, vp]),
            "orqg_server_poll": (i32, [vp, vp, i32, vp, i32]),
            "orqg_server_pending": (i32, [vp, vp, vp]),
            "orqg_server_get_stats": (i32, [vp, vp]),
            "orqg_server_now_ms": (ctypes.c_double, [vp]),
            "orqg_server_destroy": (i32, [vp]),
            "orqg_sub_b": (i32, [vp, vp, vp, vp]),
            "orqg_ppa": (i32, [vp, vp, vp, i32, vp]),
            "orqg_rca_compare": (i32, [vp, vp, vp, vp]),
            "orqg_a2b": (i32, [vp, vp, vp]),
            "orqg_secret_share_from": (i32, [vp, vp, vp, i32]),
            "orqg_reshare": (i32, [vp, vp, i32]),
            "orqg_reserve_perms": (i32, [vp, sz, sz]),
            "orqg_perm_queue_size": (sz, [vp]),
            "orqg_peek_perm": (i32, [vp, sz, i32, vp]),
            "orqg_apply_sharded_perm": (i32, [vp, vp, i32]),
            "orqg_apply_elementwise_perm": (i32, [vp, vp, vp]),
            "orqg_radix_sort": (i32, [vp, vp, i32, i32, vp]),
            "orqg_et_create": (i32, [vp, i32, sz, vp, sz, ctypes.POINTER(vp)]),
            "orqg_et_free": (i32, [vp]),
            "orqg_et_rows": (sz, [vp]),
            "orqg_et_ncols": (sz, [vp]),
            "orqg_et_column_name": (ctypes.c_char_p, [vp, sz]),
            "orqg_et_column": (vp, [vp, ctypes.c_char_p]),
            "orqg_et_set_column": (i32, [vp, ctypes.c_char_p, vp]),
            "orqg_et_add_columns": (i32, [vp, vp, sz]),
            "orqg_et_delete_columns": (i32, [vp, vp, sz]),
            "orqg_et_project": (i32, [vp, vp, sz]),
            "orqg_et_filter": (i32, [vp, ctypes.c_char_p]),
            "orqg_et_resize": (i32, [vp, sz]),
            "orqg_et_tail": (i32, [vp, sz]),
            "orqg_et_shuffle": (i32, [vp]),
            "orqg_et_sort": (i32, [vp, vp, vp, sz, i32]),
            "orqg_et_distinct": (i32, [vp, vp, sz]),
            "orqg_et_aggregate": (i32, [vp, vp, sz, vp, sz, i32, i32, i32, ctypes.c_char_p]),
            "orqg_et_concatenate": (i32, [vp, vp, ctypes.POINTER(vp)]),
            "orqg_et_inner_join": (i32, [vp, vp, vp, sz, vp, sz, i32, ctypes.POINTER(vp)]),
            "orqg_radix_sort_batch": (i32, [ctypes.POINTER(vp), ctypes.POINTER(vp),
                                            ctypes.POINTER(i32), ctypes.POINTER(i32),
                                            ctypes.POINTER(vp), sz]),
            "orqg_msglog_count": (sz, [vp]),
            "orqg_msglog_get": (i32, [vp, sz, ctypes.POINTER(_Msg)]),
            "orqg_msglog_clear": (i32, [vp]),
            "orqg_timer_enable": (i32, [vp, i32]),
            "orqg_timer_get": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(u64)]),
            "orqg_timer_reset": (i32, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def build_info():
    """The source hash compiled into liborqg.so (orqg_build_info)."""
    return lib().orqg_build_info().decode()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def splitmix64(seed, count):
    """splitmix64 stream (Steele, Lea, Flood 2014) as a uint64 array."""
    M = (1 << 64) - 1
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & M) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def keys_from_seed(seed):
    """The 10 AES-256 keys (local x3, pairwise x3, group x4) from a 64-bit seed."""
    return splitmix64(seed, 40).astype("<u8").tobytes()


class Vec:
    """A device-resident replicated vector: shares s_0, s_1, s_2 stored [n][3]."""

    def __init__(self, ctx, n, width, encoding=ORQG_BSHARED):
        self.ctx, self.n, self.width = ctx, n, width
        self.encoding = encoding
        h = ctypes.c_void_p()
        ctx._check(lib().orqg_vec_create(ctx.h, n, width, encoding, ctypes.byref(h)), "vec_create")
        self.h = h
        ctx._vecs.add(self)

    def free(self):
        if getattr(self, "h", None):
            lib().orqg_vec_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def shares(self):
        out = np.zeros((3, self.n), dtype=_DT[self.width])
        self.ctx._check(lib().orqg_vec_read(self.ctx.h, self.h, _ptr(out)), "vec_read")
        return [out[0], out[1], out[2]]

    def set_shares(self, s):
        a = np.ascontiguousarray(np.stack([np.asarray(x, dtype=_DT[self.width]) for x in s]))
        self.ctx._check(lib().orqg_vec_write(self.ctx.h, self.h, _ptr(a)), "vec_write")


class Ctx:
    """The three parties on one GPU.  threads > 1: that many worker threads of
    the reference's RunTime, thread 0 with `keys` and thread t with
    thread_keys[t - 1] (320 bytes each)."""

    def __init__(self, keys: bytes, batch_size=-12, device=0, hbm_budget=0, threads=1,
                 thread_keys=()):
        cfg = _Config()
        cfg.device = device
        ctypes.memmove(ctypes.addressof(cfg.keys), keys, 320)
        cfg.batch_size = batch_size
        cfg.hbm_budget = hbm_budget
        cfg.threads = threads
        if threads > 1:
            if len(thread_keys) != threads - 1 or any(len(k) != 320 for k in thread_keys):
                raise OrqgError("Ctx: threads - 1 key sets of 320 bytes are needed")
            self._thread_keys = b"".join(thread_keys)
            cfg.thread_keys = self._thread_keys
        h = ctypes.c_void_p()
        rc = lib().orqg_ctx_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise OrqgError(f"orqg_ctx_create failed ({rc})")
        self.h = h
        import weakref

        self._vecs = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            lib().orqg_ctx_destroy(self.h)
            self.h = None
            for v in list(self._vecs):
                v.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().orqg_last_error(self.h)
            raise OrqgError(f"{what}: {msg.decode() if msg else rc} (code {rc})")

    def sync(self):
        self._check(lib().orqg_sync(self.h), "sync")

    SORT_PROTOCOLS = {"RADIXSORT": 0, "QUICKSORT": 1}

    @property
    def default_sort(self):
        """The protocol SortingProtocol::DEFAULT stands for in table sorts (the
        reference's DEFAULT_SORT_PROTO, operators/sorting.h:6-9)."""
        code = lib().orqg_ctx_default_sort(self.h)
        return {v: k for k, v in self.SORT_PROTOCOLS.items()}[code]

    @default_sort.setter
    def default_sort(self, protocol):
        self._check(lib().orqg_ctx_set_default_sort(self.h, self.SORT_PROTOCOLS[protocol]),
                    "ctx_set_default_sort")

    def nonce(self, role, thread=0):
        v = ctypes.c_uint64()
        self._check(lib().orqg_prg_nonce_thread(self.h, thread, role, ctypes.byref(v)),
                    "prg_nonce")
        return v.value

    def vec(self, n, width, encoding=ORQG_BSHARED):
        return Vec(self, n, width, encoding)

    def secret_share(self, data, width, arith=False):
        a = np.ascontiguousarray(data, dtype=_DT[width])
        v = Vec(self, a.size, width, ORQG_ASHARED if arith else ORQG_BSHARED)
        self._check(lib().orqg_secret_share(self.h, _ptr(a), v.h), "secret_share")
        return v

    def open(self, v):
        out = np.zeros(v.n, dtype=_DT[v.width])
        self._check(lib().orqg_open(self.h, v.h, _ptr(out)), "open")
        return out

    def public_share(self, data, width, arith=True):
        a = np.ascontiguousarray(data, dtype=_DT[width])
        v = Vec(self, a.size, width, ORQG_ASHARED if arith else ORQG_BSHARED)
        self._check(lib().orqg_public_share(self.h, _ptr(a), v.h), "public_share")
        return v

    def _binop(self, fn, x, y, enc):
        z = Vec(self, x.n, x.width, enc)
        self._check(getattr(lib(), fn)(self.h, x.h, y.h, z.h), fn)
        return z

    def xor_b(self, x, y):
        return self._binop("orqg_xor_b", x, y, ORQG_BSHARED)

    def and_b(self, x, y):
        return self._binop("orqg_and_b", x, y, ORQG_BSHARED)

    def add_a(self, x, y):
        return self._binop("orqg_add_a", x, y, ORQG_ASHARED)

    def sub_a(self, x, y):
        return self._binop("orqg_sub_a", x, y, ORQG_ASHARED)

    def mul_a(self, x, y):
        return self._binop("orqg_mul_a", x, y, ORQG_ASHARED)

    def not_b(self, x):
        z = Vec(self, x.n, x.width, ORQG_BSHARED)
        self._check(lib().orqg_not_b(self.h, x.h, z.h), "not_b")
        return z

    def b2a_bit(self, x):
        y = Vec(self, x.n, x.width, ORQG_ASHARED)
        self._check(lib().orqg_b2a_bit(self.h, x.h, y.h), "b2a_bit")
        return y

    CMP_OPS = {"==": 0, "!=": 1, "<": 2, ">": 3, "<=": 4, ">=": 5}

    def compare(self, x, y):
        """x._compare(y): (eq, gt) boolean shares of x == y and x > y (signed)."""
        eq = Vec(self, x.n, x.width, ORQG_BSHARED)
        gt = Vec(self, x.n, x.width, ORQG_BSHARED)
        self._check(lib().orqg_compare(self.h, x.h, y.h, eq.h, gt.h), "compare")
        return eq, gt

    def cmp(self, op, x, y):
        """Comparison operator op ('==', '!=', '<', '>', '<=', '>=') on boolean shares."""
        z = Vec(self, x.n, x.width, ORQG_BSHARED)
        self._check(lib().orqg_cmp(self.h, self.CMP_OPS[op], x.h, y.h, z.h), "cmp")
        return z

    def b2a(self, x):
        """Full-width boolean-to-arithmetic conversion (BSharedVector::b2a)."""
        y = Vec(self, x.n, x.width, ORQG_ASHARED)
        self._check(lib().orqg_b2a(self.h, x.h, y.h), "b2a")
        return y

    UNARY = {"<<": 0, ">>": 1, "lsr": 2, "parity": 3, "extend_lsb": 4, "ltz": 5, "mask": 6}

    def unary(self, op, x, arg=0, out=None):
        """Share-local map: '<<', '>>' (arithmetic), 'lsr', 'parity',
        'extend_lsb', 'ltz', 'mask' (x & arg); out may be x."""
        z = out if out is not None else Vec(self, x.n, x.width, x.encoding)
        self._check(lib().orqg_vec_unary(self.h, self.UNARY[op], x.h, arg, z.h), "vec_unary")
        return z

    def prefix_sum(self, v):
        """SharedVector::prefix_sum, in place."""
        self._check(lib().orqg_prefix_sum(self.h, v.h), "prefix_sum")

    def reshare(self, v, group):
        self._check(lib().orqg_reshare(self.h, v.h, group), "reshare")

    def keystream(self, role, nbytes):
        L = lib()
        d = ctypes.c_void_p()
        self._check(L.orqg_dev_alloc(self.h, max(nbytes, 16), ctypes.byref(d)), "dev_alloc")
        try:
            self._check(L.orqg_prg_keystream(self.h, role, nbytes, d), "keystream")
            out = np.zeros(nbytes, dtype=np.uint8)
            self._check(L.orqg_dev_to_host(self.h, _ptr(out), d, nbytes), "dev_to_host")
        finally:
            L.orqg_dev_free(d)
        return out

    def reserve(self, n, count):
        self._check(lib().orqg_reserve_perms(self.h, n, count), "reserve_perms")

    def queue_size(self):
        return lib().orqg_perm_queue_size(self.h)

    def peek_perm(self, index, group, n):
        out = np.zeros(n, dtype=np.int32)
        self._check(lib().orqg_peek_perm(self.h, index, group, _ptr(out)), "peek_perm")
        return out

    def apply_sharded(self, x, inverse=False):
        self._check(lib().orqg_apply_sharded_perm(self.h, x.h, int(inverse)), "apply_sharded_perm")

    def apply_elementwise(self, x, perm):
        self._check(lib().orqg_apply_elementwise_perm(self.h, x.h, perm.h), "apply_elementwise")

    # ---- ElementwisePermutation, shuffle, table sort ----
    def perm_identity(self, n, encoding=ORQG_ASHARED):
        p = Vec(self, n, 4, encoding)
        self._check(lib().orqg_perm_identity(self.h, p.h, encoding), "perm_identity")
        return p

    def perm_invert(self, p):
        self._check(lib().orqg_perm_invert(self.h, p.h), "perm_invert")

    def perm_b2a(self, p):
        self._check(lib().orqg_perm_b2a(self.h, p.h), "perm_b2a")
        p.encoding = ORQG_ASHARED

    def perm_negate(self, p):
        self._check(lib().orqg_perm_negate(self.h, p.h), "perm_negate")

    def perm_reverse(self, p):
        self._check(lib().orqg_perm_reverse(self.h, p.h), "perm_reverse")

    def perm_compose(self, sigma, rho):
        """rho <- compose_permutations(sigma, rho); sigma is consumed."""
        self._check(lib().orqg_perm_compose(self.h, sigma.h, rho.h), "perm_compose")

    def shuffle(self, x):
        self._check(lib().orqg_shuffle(self.h, x.h), "shuffle")

    @staticmethod
    def _vec_array(cols):
        arr = (ctypes.c_void_p * max(len(cols), 1))(*[c.h for c in cols])
        return arr, len(cols)

    def table_shuffle(self, a_cols, b_cols):
        a, na = self._vec_array(a_cols)
        b, nb = self._vec_array(b_cols)
        self._check(lib().orqg_table_shuffle(self.h, a, na, b, nb), "table_shuffle")

    PROTOCOLS = {"RADIXSORT": 0, "QUICKSORT": 1}

    def table_sort(self, keys, orders=None, single_bit=None, a_cols=(), b_cols=(),
                   protocol="RADIXSORT"):
        k, nk = self._vec_array(keys)
        o = (ctypes.c_int * nk)(*[(DESC if x in (DESC, "DESC") else ASC) for x in
                                  (orders or [ASC] * nk)])
        sb = (ctypes.c_int * nk)(*[int(bool(x)) for x in (single_bit or [0] * nk)])
        a, na = self._vec_array(list(a_cols))
        b, nb = self._vec_array(list(b_cols))
        self._check(lib().orqg_table_sort_protocol(self.h, k, nk, o, sb, a, na, b, nb,
                                                   self.PROTOCOLS[protocol]), "table_sort")

    # ---- share files (the reference's CSV formats) ----
    @staticmethod
    def _paths(files):
        assert len(files) == 3, "one file per party"
        bufs = [str(f).encode() for f in files]
        return (ctypes.c_char_p * 3)(*bufs), bufs

    @staticmethod
    def _names(names):
        bufs = [n.encode() for n in names]
        return (ctypes.c_char_p * max(len(bufs), 1))(*bufs), bufs

    def read_share_files(self, v, files):
        """EVector(size, file) for the three parties' files."""
        arr, _keep = self._paths(files)
        self._check(lib().orqg_vec_read_share_files(self.h, v.h, arr), "vec_read_share_files")

    def write_share_files(self, v, files):
        """EVector::output to the three parties' files."""
        arr, _keep = self._paths(files)
        self._check(lib().orqg_vec_write_share_files(self.h, v.h, arr), "vec_write_share_files")

    def table_read_share_csv(self, files, names, cols):
        arr, _k1 = self._paths(files)
        nm, _k2 = self._names(names)
        c, nc = self._vec_array(list(cols))
        self._check(lib().orqg_table_read_share_csv(self.h, arr, nm, c, nc), "table_read_share_csv")

    def table_write_share_csv(self, files, names, cols):
        arr, _k1 = self._paths(files)
        nm, _k2 = self._names(names)
        c, nc = self._vec_array(list(cols))
        self._check(lib().orqg_table_write_share_csv(self.h, arr, nm, c, nc),
                    "table_write_share_csv")

    def table_read_csv_data(self, path, names, cols, input_party=0):
        nm, _k = self._names(names)
        c, nc = self._vec_array(list(cols))
        self._check(lib().orqg_table_read_csv_data(self.h, str(path).encode(), nm, c, nc,
                                                   input_party), "table_read_csv_data")

    def quicksort(self, v, desc=False):
        """operators::quicksort of 32-bit B-shared keys, in place; returns the
        A-shared sorting permutation."""
        perm = Vec(self, v.n, 4, ORQG_ASHARED)
        self._check(lib().orqg_quicksort(self.h, v.h, DESC if desc else ASC, perm.h), "quicksort")
        return perm

    def bitonic_sort(self, keys, orders=None, a_cols=(), b_cols=()):
        k, nk = self._vec_array(keys)
        o = (ctypes.c_int * nk)(*[(DESC if x in (DESC, "DESC") else ASC) for x in
                                  (orders or [ASC] * nk)])
        a, na = self._vec_array(list(a_cols))
        b, nb = self._vec_array(list(b_cols))
        self._check(lib().orqg_bitonic_sort(self.h, k, nk, o, a, na, b, nb), "bitonic_sort")

    def resize(self, v, n):
        self._check(lib().orqg_vec_resize(self.h, v.h, n), "vec_resize")
        v.n = n

    def tail(self, v, n):
        self._check(lib().orqg_vec_tail(self.h, v.h, n), "vec_tail")
        v.n = n

    AGG = {"sum": 0, "count": 1, "copy": 2, "max": 3, "min": 4, "or": 5, "valid": 6}

    class _AggSpec(ctypes.Structure):
        _fields_ = [("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("func", ctypes.c_int)]

    def _specs(self, specs):
        arr = (self._AggSpec * max(len(specs), 1))()
        for i, (inp, out, fn) in enumerate(specs):
            arr[i].inp, arr[i].out, arr[i].func = inp.h, out.h, self.AGG[fn]
        return arr, len(specs)

    def aggregate(self, keys, b_specs=(), a_specs=(), reverse=False, sel=None):
        """aggregators::aggregate; specs are (in, out, function-name) tuples."""
        k, nk = self._vec_array(list(keys))
        b, nb = self._specs(list(b_specs))
        a, na = self._specs(list(a_specs))
        self._check(lib().orqg_aggregate(self.h, k, nk, b, nb, a, na, int(reverse),
                                         sel.h if sel is not None else None), "aggregate")
        for _, out, fn in b_specs:
            out.encoding = ORQG_BSHARED
        for _, out, fn in a_specs:
            out.encoding = ORQG_ASHARED

    def distinct(self, keys, res):
        k, nk = self._vec_array(list(keys))
        self._check(lib().orqg_distinct(self.h, k, nk, res.h), "distinct")

    def public_fill(self, v, value, encoding=ORQG_BSHARED):
        self._check(lib().orqg_public_fill(self.h, v.h, value, encoding), "public_fill")
        v.encoding = encoding

    def adder(self, op, x, y=None):
        """op: add / sub (ripple carry, the default adder), ppa_add / ppa_sub,
        rca_lt (rca_compare), a2b."""
        L = lib()
        if op == "a2b":
            z = Vec(self, x.n, x.width, ORQG_BSHARED)
            self._check(L.orqg_a2b(self.h, x.h, z.h), "a2b")
            return z
        z = Vec(self, x.n, x.width, ORQG_BSHARED)
        if op in ("add", "sub", "rca_lt"):
            fn = {"add": L.orqg_add_b, "sub": L.orqg_sub_b, "rca_lt": L.orqg_rca_compare}[op]
            self._check(fn(self.h, x.h, y.h, z.h), op)
        else:
            self._check(L.orqg_ppa(self.h, x.h, y.h, int(op == "ppa_sub"), z.h), op)
        return z

    def secret_share_from(self, data, width, owner, arith=False):
        a = np.ascontiguousarray(data, dtype=_DT[width])
        v = Vec(self, a.size, width, ORQG_ASHARED if arith else ORQG_BSHARED)
        self._check(lib().orqg_secret_share_from(self.h, _ptr(a), v.h, owner), "secret_share_from")
        return v

    def rekey(self, keys):
        """Fresh keys: nonces back to 0, permutation queue and message log cleared."""
        assert len(keys) == 32 * 10
        self._check(lib().orqg_ctx_rekey(self.h, bytes(keys)), "ctx_rekey")

    def copy_rows(self, src, src_start, dst, dst_start, count):
        self._check(lib().orqg_vec_copy_rows(self.h, src.h, src_start, dst.h, dst_start, count),
                    "vec_copy_rows")

    def broadcast_row(self, src, row, dst, dst_start, count):
        self._check(lib().orqg_vec_broadcast_row(self.h, src.h, row, dst.h, dst_start, count),
                    "vec_broadcast_row")

    def radix_sort(self, v, bits=None, desc=False):
        if bits is None:
            bits = 8 * v.width
        perm = Vec(self, v.n, 4, ORQG_ASHARED)
        self._check(lib().orqg_radix_sort(self.h, v.h, DESC if desc else ASC, bits, perm.h),
                    "radix_sort")
        return perm

    def msglog(self):
        L = lib()
        out = []
        m = _Msg()
        for i in range(L.orqg_msglog_count(self.h)):
            L.orqg_msglog_get(self.h, i, ctypes.byref(m))
            out.append((m.op.decode(), m.from_, m.to, m.bytes))
        return out

    def clear_msglog(self):
        lib().orqg_msglog_clear(self.h)

    def timer_enable(self, on=True):
        lib().orqg_timer_enable(self.h, int(on))

    def timer(self, family):
        ms, cnt = ctypes.c_double(), ctypes.c_uint64()
        lib().orqg_timer_get(self.h, family.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        return ms.value, cnt.value

    def timer_reset(self):
        lib().orqg_timer_reset(self.h)


def radix_sort_batch(vecs, bits=None, desc=None):
    """radix_sort of several requests at once, each vector on its own context
    (orqg_radix_sort_batch): vecs[q] is sorted in place on its context, as
    Ctx.radix_sort would, and the A-shared sorting permutations are returned
    in order.  bits / desc: per request (default: full width, ascending)."""
    m = len(vecs)
    if bits is None:
        bits = [8 * v.width for v in vecs]
    if desc is None:
        desc = [False] * m
    perms = [Vec(v.ctx, v.n, 4, ORQG_ASHARED) for v in vecs]
    vp = ctypes.c_void_p
    ctxs = (vp * m)(*[v.ctx.h for v in vecs])
    hv = (vp * m)(*[v.h for v in vecs])
    hp = (vp * m)(*[p.h for p in perms])
    orders = (ctypes.c_int * m)(*[DESC if d else ASC for d in desc])
    nb = (ctypes.c_int * m)(*bits)
    rc = lib().orqg_radix_sort_batch(ctxs, hv, orders, nb, hp, m)
    if rc != 0:
        vecs[0].ctx._check(rc, "radix_sort_batch")
    return perms
