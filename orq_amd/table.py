"""Host mirror of orq::EncodedTable for the shuffle / sort path
(include/core/containers/tabular/encoded_table.h).

Columns live on the device as liborqg vectors.  As in the reference:
  * the schema is a std::map, so columns are visited in sorted name order
    (encoded_table.h:147);
  * a column whose name contains '[' is boolean-shared, any other is
    arithmetic-shared (EncodedTable::isBShared, encoded_table.h:1672);
  * every table carries the valid column "[##V]", a public sharing of ones
    (configureValid, encoded_table.h:274-285);
  * sort keys named "[##V]" or "[##TID]" are sorted on one bit
    (encoded_table.h:915-935).
"""
import numpy as np

from . import ASC, DESC, ORQG_ASHARED, ORQG_BSHARED

VALID = "[##V]"
JOIN_ID = "[##TID]"
RADIXSORT = "RADIXSORT"
BITONICSORT = "BITONICSORT"


def is_bshared(name):
    return "[" in name


class EncodedTable:
    def __init__(self, ctx, name, columns, width):
        self.ctx, self.name, self.width = ctx, name, width
        self.schema = dict(columns)
        n = {v.n for v in self.schema.values()}
        assert len(n) <= 1, "columns differ in size"
        self.rows = n.pop() if n else 0
        if VALID not in self.schema:
            self.schema[VALID] = self._ones()

    def _ones(self):
        return self.ctx.public_share(np.ones(self.rows, dtype=np.uint64 if self.width == 8
                                             else np.uint32), self.width, arith=False)

    @classmethod
    def secret_share(cls, ctx, data, schema, width=4, name="table"):
        """Share each column from party 0 in the order given (the reference's
        secret_share(table_data, schema) helper), then add the valid column."""
        cols = {}
        for arr, col in zip(data, schema):
            a = np.asarray(arr).astype(np.int64).astype(np.uint64 if width == 8 else np.uint32)
            cols[col] = ctx.secret_share(a, width, arith=not is_bshared(col))
        return cls(ctx, name, cols, width)

    def column_names(self):
        return sorted(self.schema)

    def __getitem__(self, col):
        return self.schema[col]

    def shuffle(self):
        """EncodedTable::shuffle (encoded_table.h:1003-1018)."""
        names = self.column_names()
        a = [self.schema[c] for c in names if not is_bshared(c)]
        b = [self.schema[c] for c in names if is_bshared(c)]
        self.ctx.table_shuffle(a, b)
        return self

    def resize(self, n):
        """EncodedTable::resize (encoded_table.h:1426-1436): new rows are zero shares."""
        for v in self.schema.values():
            self.ctx.resize(v, n)
        self.rows = n

    def tail(self, n):
        """EncodedTable::tail (encoded_table.h:1407-1418): keep the last n rows."""
        for v in self.schema.values():
            self.ctx.tail(v, n)
        self.rows = n

    def pad_power_of_two(self):
        """encoded_table.h:1446-1466 with pad value 0."""
        self.resize(1 << max(self.rows - 1, 0).bit_length())

    def sort(self, spec, to_be_sorted=None, protocol=RADIXSORT):
        """EncodedTable::sort (encoded_table.h:900-998).  `spec` is a list of
        column names (all ascending) or of (name, order) pairs."""
        if protocol not in (RADIXSORT, BITONICSORT):
            raise ValueError(f"sorting protocol {protocol!r} is not provided")
        spec = [(s, ASC) if isinstance(s, str) else (s[0], s[1]) for s in spec]
        keys = [k for k, _ in spec]
        for k in keys:
            if not is_bshared(k):
                raise ValueError(f"sort key {k!r} must be a boolean-shared column")
        orders = [DESC if o in (DESC, "DESC") else ASC for _, o in spec]
        single = [k in (VALID, JOIN_ID) for k in keys]
        tbs = self.column_names() if to_be_sorted is None else list(to_be_sorted)
        data = [c for c in self.column_names() if c not in keys and c in tbs]
        a = [self.schema[c] for c in data if not is_bshared(c)]
        b = [self.schema[c] for c in data if is_bshared(c)]
        if protocol == BITONICSORT:
            original = self.rows
            self.pad_power_of_two()
            self.ctx.bitonic_sort([self.schema[k] for k in keys], orders, a, b)
            if VALID in keys:  # padded rows are invalid: drop them again
                if orders[keys.index(VALID)] == ASC:
                    self.tail(original)
                else:
                    self.resize(original)
            return self
        self.ctx.table_sort([self.schema[k] for k in keys], orders, single, a, b)
        return self

    def open(self):
        """{name: signed plaintext column}: XOR for boolean, sum for arithmetic."""
        sdt = np.int64 if self.width == 8 else np.int32
        return {c: self.ctx.open(v).view(sdt) for c, v in sorted(self.schema.items())}
