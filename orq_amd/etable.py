"""ctypes wrapper of the library's EncodedTable (orqg_et_* in include/orqg.h):
the relational methods -- sort, distinct, aggregate, concatenate, inner join --
run as control flow inside liborqg.  Names, column conventions and options
are those of orq_amd.table.EncodedTable (the reference's EncodedTable)."""
import ctypes

import numpy as np

from . import ASC, DESC, OrqgError, _DT, _ptr, lib

PROTOCOLS = {"RADIXSORT": 0, "QUICKSORT": 1, "BITONICSORT": 2, "DEFAULT": 3}
AGG = {"sum": 0, "count": 1, "copy": 2, "max": 3, "min": 4, "or": 5, "valid": 6}
VALID = "[##V]"


class _EtAgg(ctypes.Structure):
    _fields_ = [("data", ctypes.c_char_p), ("result", ctypes.c_char_p), ("func", ctypes.c_int)]


def _names(names):
    bufs = [n.encode() for n in names]
    return (ctypes.c_char_p * max(len(bufs), 1))(*bufs), len(bufs)


def _specs(spec):
    spec = list(spec)
    arr = (_EtAgg * max(len(spec), 1))()
    for i, (data, result, fn) in enumerate(spec):
        arr[i].data, arr[i].result, arr[i].func = data.encode(), result.encode(), AGG[fn]
    return arr, len(spec)


class Column:
    """A column lent by an ETable: valid until the table changes or is freed."""

    def __init__(self, ctx, h, n, width):
        self.ctx, self.h, self.n, self.width = ctx, h, n, width

    def shares(self):
        out = np.zeros((3, self.n), dtype=_DT[self.width])
        self.ctx._check(lib().orqg_vec_read(self.ctx.h, self.h, _ptr(out)), "vec_read")
        return [out[0], out[1], out[2]]


class ETable:
    def __init__(self, ctx, width, rows=0, names=(), _h=None):
        self.ctx, self.width = ctx, width
        if _h is None:
            h = ctypes.c_void_p()
            arr, n = _names(names)
            self._check(lib().orqg_et_create(ctx.h, width, rows, arr, n, ctypes.byref(h)),
                        "et_create")
            _h = h
        self.h = _h

    @classmethod
    def from_encoded(cls, t):
        """A copy of an orq_amd.table.EncodedTable's columns (valid included)."""
        e = cls(t.ctx, t.width, t.rows)
        for name in t.column_names():
            e._check(lib().orqg_et_set_column(e.h, name.encode(), t.schema[name].h), "set_column")
        return e

    def _check(self, rc, what):
        if rc:
            self.ctx._check(rc, what)

    def free(self):
        if getattr(self, "h", None):
            lib().orqg_et_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def rows(self):
        return lib().orqg_et_rows(self.h)

    def column_names(self):
        return [lib().orqg_et_column_name(self.h, i).decode()
                for i in range(lib().orqg_et_ncols(self.h))]

    def __getitem__(self, name):
        h = lib().orqg_et_column(self.h, name.encode())
        if not h:
            raise KeyError(name)
        return Column(self.ctx, ctypes.c_void_p(h), self.rows, self.width)

    def add_columns(self, names):
        arr, n = _names(names)
        self._check(lib().orqg_et_add_columns(self.h, arr, n), "et_add_columns")

    def delete_columns(self, names):
        arr, n = _names(names)
        self._check(lib().orqg_et_delete_columns(self.h, arr, n), "et_delete_columns")

    def project(self, cols):
        arr, n = _names(cols)
        self._check(lib().orqg_et_project(self.h, arr, n), "et_project")

    def filter(self, col):
        self._check(lib().orqg_et_filter(self.h, col.encode()), "et_filter")

    def resize(self, n):
        self._check(lib().orqg_et_resize(self.h, n), "et_resize")

    def tail(self, n):
        self._check(lib().orqg_et_tail(self.h, n), "et_tail")

    def shuffle(self):
        self._check(lib().orqg_et_shuffle(self.h), "et_shuffle")
        return self

    def sort(self, spec, protocol="DEFAULT"):
        spec = [(s, ASC) if isinstance(s, str) else (s[0], s[1]) for s in spec]
        arr, n = _names([k for k, _ in spec])
        orders = (ctypes.c_int * n)(*[DESC if o in (DESC, "DESC") else ASC for _, o in spec])
        self._check(lib().orqg_et_sort(self.h, arr, orders, n, PROTOCOLS[protocol]), "et_sort")
        return self

    def distinct(self, keys):
        arr, n = _names(keys)
        self._check(lib().orqg_et_distinct(self.h, arr, n), "et_distinct")
        return self

    def aggregate(self, keys, spec, reverse=False, do_sort=True, mark_valid=True, table_id=None):
        arr, n = _names(keys)
        sp, ns = _specs(spec)
        self._check(lib().orqg_et_aggregate(self.h, arr, n, sp, ns, int(reverse), int(do_sort),
                                            int(mark_valid),
                                            table_id.encode() if table_id else None),
                    "et_aggregate")
        return self

    def concatenate(self, other):
        h = ctypes.c_void_p()
        self._check(lib().orqg_et_concatenate(self.h, other.h, ctypes.byref(h)), "et_concatenate")
        return ETable(self.ctx, self.width, _h=h)

    def inner_join(self, right, keys, spec=(), trim_invalid=True):
        arr, n = _names(keys)
        sp, ns = _specs(spec)
        h = ctypes.c_void_p()
        self._check(lib().orqg_et_inner_join(self.h, right.h, arr, n, sp, ns, int(trim_invalid),
                                             ctypes.byref(h)), "et_inner_join")
        return ETable(self.ctx, self.width, _h=h)

    def open(self):
        sdt = np.int64 if self.width == 8 else np.int32
        return {c: self.ctx.open(self[c]).view(sdt) for c in self.column_names()}

    def open_with_schema(self, remove_invalid=True):
        cols = self.open()
        valid = cols.pop(VALID)
        if remove_invalid:
            keep = valid != 0
            cols = {k: v[keep] for k, v in cols.items()}
        return cols
