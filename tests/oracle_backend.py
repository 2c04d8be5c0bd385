"""The CPU oracle behind the interface of orq_amd.Ctx that the host table
layer (orq_amd.table) uses, so the same table code runs over the oracle and
over liborqg and their shares can be compared column by column."""
import ctypes

import numpy as np

from oracle import oracle as O

A, B = 0, 1  # ORQG_ASHARED, ORQG_BSHARED
_UT = {4: np.uint32, 8: np.uint64}


def _tag(v, enc):
    v.encoding = enc
    return v


class OracleCtx:
    def __init__(self, keys, batch_size=-12, threads=1, thread_keys=(), default_sort="QUICKSORT"):
        self.c = O.Ctx(keys, batch_size, threads=threads, thread_keys=thread_keys)
        # the protocol SortingProtocol::DEFAULT stands for (DEFAULT_SORT_PROTO)
        self.default_sort = default_sort

    # vectors
    def vec(self, n, width, encoding=B):
        return _tag(O.Vec(n, width), encoding)

    def secret_share(self, data, width, arith=False):
        return _tag(self.c.secret_share(data, width, arith=arith), A if arith else B)

    def public_share(self, data, width, arith=True):
        a = np.ascontiguousarray(data, dtype=_UT[width])
        v = O.Vec(a.size, width)
        z = np.zeros_like(a)
        v.set_shares([a, z, z])
        return _tag(v, A if arith else B)

    def open(self, v):
        return self.c.open(v, arith=v.encoding == A)

    def resize(self, v, n):
        O.lib().oc_vec_resize(v.h, n)
        v.n = n

    def tail(self, v, n):
        O.lib().oc_vec_tail(v.h, n)
        v.n = n

    def copy_rows(self, src, s0, dst, d0, count):
        O.lib().oc_vec_copy_rows(src.h, s0, dst.h, d0, count)

    def broadcast_row(self, src, row, dst, d0, count):
        O.lib().oc_vec_broadcast_row(src.h, row, dst.h, d0, count)

    # primitives
    def and_b(self, x, y):
        return _tag(self.c.and_b(x, y), B)

    def not_b(self, x):
        z = O.Vec(x.n, x.width)
        z.set_shares([~s for s in x.shares()])
        return _tag(z, B)

    # operators
    def table_shuffle(self, a_cols, b_cols):
        self.c.table_shuffle(a_cols, b_cols)

    def table_sort(self, keys, orders=None, single_bit=None, a_cols=(), b_cols=(),
                   protocol="RADIXSORT"):
        self.c.table_sort(keys, orders, single_bit, a_cols, b_cols, protocol)

    def bitonic_sort(self, keys, orders=None, a_cols=(), b_cols=()):
        self.c.bitonic_sort(keys, orders, a_cols, b_cols)

    def distinct(self, keys, res):
        self.c.distinct(keys, res)
        res.encoding = B

    def aggregate(self, keys, b_specs=(), a_specs=(), reverse=False, sel=None):
        self.c.aggregate(keys, b_specs, a_specs, reverse, sel)
        for _, out, _ in b_specs:
            out.encoding = B
        for _, out, _ in a_specs:
            out.encoding = A

    def queue_size(self):
        return self.c.queue_size()

    def nonce(self, party, role):
        return self.c.nonce(party, role)

    def msglog(self):
        return self.c.msglog()
