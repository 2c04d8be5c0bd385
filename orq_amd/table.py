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
UNIQ = "[##UNIQ]"
VALID_TEMP = "[##VALID_TEMP]"
RESERVED = (JOIN_ID, VALID, UNIQ)  # encoded_table.h:38
NON_AGGREGATIONS = ("copy", "valid")  # AggregationSelector::isAggregation
RADIXSORT = "RADIXSORT"
BITONICSORT = "BITONICSORT"
QUICKSORT = "QUICKSORT"
# SortingProtocol::DEFAULT: the backend context's `default_sort`, the
# reference's compile-time DEFAULT_SORT_PROTO (sorting.h:6-9), QUICKSORT
# unless set.
DEFAULT = "DEFAULT"


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

    @classmethod
    def with_schema(cls, ctx, name, schema, rows, width=4):
        """EncodedTable(name, schema, rows) (encoded_table.h:244-249): zero
        columns, B-shared when the name has '[', and the valid column set to
        ones (configureValid)."""
        cols = {c: ctx.vec(rows, width, ORQG_BSHARED if is_bshared(c) else ORQG_ASHARED)
                for c in schema}
        t = cls(ctx, name, cols, width)
        if VALID in schema:
            ctx.public_fill(t.schema[VALID], 1, ORQG_BSHARED)
        return t

    # ---- share files (encoded_table.h:382-628) ----
    def input_csv_table_data(self, path, input_party=0):
        """inputCSVTableData: the input party's plaintext CSV, then every
        column secret-shared from it in schema order."""
        names = self.column_names()
        self.ctx.table_read_csv_data(path, names, [self.schema[c] for c in names], input_party)

    def input_csv_table_secret_shares(self, files):
        """inputCSVTableSecretShares from the three parties' files."""
        names = self.column_names()
        self.ctx.table_read_share_csv(files, names, [self.schema[c] for c in names])

    def output_csv_table_secret_shares(self, files):
        """outputCSVTableSecretShares to the three parties' files."""
        names = self.column_names()
        self.ctx.table_write_share_csv(files, names, [self.schema[c] for c in names])

    def input_secret_shares(self, col, files):
        """inputSecretShares: one column from the three parties' files."""
        self.ctx.read_share_files(self.schema[col], files)

    def output_secret_shares(self, col, files):
        """outputSecretShares: one column to the three parties' files."""
        self.ctx.write_share_files(self.schema[col], files)

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

    def sort(self, spec, to_be_sorted=None, protocol=DEFAULT):
        """EncodedTable::sort (encoded_table.h:900-998).  `spec` is a list of
        column names (all ascending) or of (name, order) pairs."""
        if protocol == DEFAULT:
            protocol = getattr(self.ctx, "default_sort", QUICKSORT)
        if protocol not in (RADIXSORT, BITONICSORT, QUICKSORT):
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
        self.ctx.table_sort([self.schema[k] for k in keys], orders, single, a, b,
                            protocol=protocol)
        return self

    # ---- columns (encoded_table.h:724-803) ----
    def _zero_col(self, name, rows=None):
        enc = ORQG_BSHARED if is_bshared(name) else ORQG_ASHARED
        return self.ctx.vec(self.rows if rows is None else rows, self.width, enc)

    def add_columns(self, names, rows=None):
        """addColumns: zero-initialised columns, B-shared when the name has '['."""
        for c in names:
            self.schema[c] = self._zero_col(c, rows)

    def delete_columns(self, names):
        for c in names:
            self.schema.pop(c, None)

    def project(self, cols):
        """Keep only `cols` and the reserved columns."""
        keep = set(cols) | set(RESERVED)
        self.delete_columns([c for c in self.column_names() if c not in keep])

    def filter(self, col):
        """filter (encoded_table.h:298-314): valid &= col."""
        v = self.schema[col] if isinstance(col, str) else col
        self.schema[VALID] = self.ctx.and_b(self.schema[VALID], v)

    def head(self, n):
        self.resize(n)

    # ---- distinct (encoded_table.h:1199-1237) ----
    def distinct_into(self, keys, res):
        """Mark rows that differ from the previous row on some key."""
        self.ctx.distinct([self.schema[k] for k in keys], self.schema[res])

    def distinct(self, keys):
        """Sort on valid + keys, keep the first row of each key group."""
        self.sort([VALID] + list(keys))
        self.add_columns([UNIQ])
        self.distinct_into(keys, UNIQ)
        self.filter(UNIQ)
        self.delete_columns([UNIQ])
        return self

    # ---- concatenate (encoded_table.h:1548-1616) ----
    def concatenate(self, other):
        names = [JOIN_ID] + [c for c in self.column_names() if c != JOIN_ID]
        names += [c for c in other.column_names() if c not in self.schema]
        n = self.rows + other.rows
        cols = {c: EncodedTable._zero_static(self.ctx, c, n, self.width) for c in names}
        t = EncodedTable.__new__(EncodedTable)
        t.ctx, t.name, t.width, t.schema, t.rows = self.ctx, self.name + "+" + other.name, \
            self.width, cols, n
        t.schema[VALID] = t._ones()  # configureValid, overwritten below
        for c in self.column_names():
            self.ctx.copy_rows(self.schema[c], 0, t.schema[c], 0, self.rows)
        for c in other.column_names():
            self.ctx.copy_rows(other.schema[c], 0, t.schema[c], self.rows, other.rows)
        # table id: 0 above, a secret sharing of 1 repeated over the other table's rows
        one = self.ctx.secret_share(np.ones(1, np.uint64 if self.width == 8 else np.uint32),
                                    self.width)
        self.ctx.broadcast_row(one, 0, t.schema[JOIN_ID], self.rows, other.rows)
        return t

    @staticmethod
    def _zero_static(ctx, name, rows, width):
        enc = ORQG_BSHARED if is_bshared(name) else ORQG_ASHARED
        return ctx.vec(rows, width, enc)

    # ---- aggregate (encoded_table.h:1059-1188) ----
    def aggregate(self, keys, spec, reverse=False, do_sort=True, mark_valid=True, table_id=None):
        """`spec`: (input column, result column, function) with function one
        of sum, count, copy (A-shared) or copy, max, min, or, valid (B-shared)."""
        keys = list(keys)
        if do_sort:
            keys = [VALID] + keys
            self.sort(keys)
        original = self.rows
        self.pad_power_of_two()
        key_vecs = [self.schema[k] for k in keys]
        b_agg, a_agg, has_any = [], [], False
        for data, result, fn in spec:
            if is_bshared(data) != is_bshared(result):
                raise ValueError(f"aggregation {data} -> {result}: encodings differ")
            if fn not in NON_AGGREGATIONS:
                has_any = True
            (b_agg if is_bshared(data) else a_agg).append(
                (self.schema[data], self.schema[result], fn))
        sel = None
        if mark_valid:
            self.add_columns([UNIQ])
            self.distinct_into(keys, UNIQ)
            if table_id is not None:
                sel = self.schema[table_id]
        self.ctx.aggregate(key_vecs, b_agg, a_agg, reverse, sel)
        if original < self.rows:
            self.resize(original)
        if mark_valid:
            if has_any:
                if reverse:
                    # valid[0 .. n-1) &= uniq[1 .. n): the last row of each group stays
                    n = self.rows
                    if n > 1:
                        v = self._zero_col(VALID, n - 1)
                        u = self._zero_col(UNIQ, n - 1)
                        self.ctx.copy_rows(self.schema[VALID], 0, v, 0, n - 1)
                        self.ctx.copy_rows(self.schema[UNIQ], 1, u, 0, n - 1)
                        z = self.ctx.and_b(v, u)
                        self.ctx.copy_rows(z, 0, self.schema[VALID], 0, n - 1)
                else:
                    self.filter(UNIQ)
            self.delete_columns([UNIQ])
        return self

    # ---- joins (operators/join.h:31-233) ----
    def inner_join(self, right, keys, spec=(), trim_invalid=True):
        """this (unique keys) JOIN right (foreign keys) ON keys, with the
        aggregations in `spec` (e.g. copying this table's columns down)."""
        concat = self.concatenate(right)
        out = set(keys) | set(right.schema)
        for data, result, _ in spec:
            out |= {data, result}
        concat.project(out)
        keys = [VALID] + list(keys)
        concat.sort(keys + [JOIN_ID])
        concat.add_columns([UNIQ, VALID_TEMP])
        spec = list(spec) + [(VALID_TEMP, VALID_TEMP, "valid")]
        concat.distinct_into(keys, UNIQ)
        # inner join: a row is valid if its key group is not a singleton
        concat.schema[VALID_TEMP] = self.ctx.and_b(concat.schema[VALID],
                                                  self.ctx.not_b(concat.schema[UNIQ]))
        concat.delete_columns([UNIQ])
        concat.aggregate(keys, spec, reverse=True, do_sort=False, table_id=JOIN_ID)
        concat.filter(VALID_TEMP)
        concat.delete_columns([VALID_TEMP, JOIN_ID])
        if trim_invalid:
            L, R = self.rows, right.rows
            w = 8 * self.width
            c = len(self.schema) - 1
            if R < L * (w + c - 1) // (c + 2):
                concat.sort([(VALID, "DESC")])
                concat.head(R)
        return concat

    def open_with_schema(self, remove_invalid=True):
        """encoded_table.h:682-717: every column but the valid one, and with
        remove_invalid only the rows whose valid bit is set."""
        cols = self.open()
        valid = cols.pop(VALID)
        if remove_invalid:
            keep = valid != 0
            cols = {k: v[keep] for k, v in cols.items()}
        return cols

    def open(self):
        """{name: signed plaintext column}: XOR for boolean, sum for arithmetic."""
        sdt = np.int64 if self.width == 8 else np.int32
        return {c: self.ctx.open(v).view(sdt) for c, v in sorted(self.schema.items())}
