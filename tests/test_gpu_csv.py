"""Share files in the reference's CSV formats, through the C ABI, on the
reference's own fixtures (tests/golden/, copied from ORQ's tests/data/),
following tests/test_io.cpp:7-113:
  * inputCSVTableData of employees.csv by party 0: the opened table is the
    CSV (zero rows past its 25), and the shares equal the oracle's secret
    sharing of the same columns in schema order;
  * inputCSVTableSecretShares of employees_secrets_3pc_{0,1,2}.csv opens
    to the same table, and outputCSVTableSecretShares writes the three
    files back byte for byte;
  * inputSecretShares / outputSecretShares of one column and the EVector
    file constructor on employees_age_secrets_3pc_*.csv, written back byte
    for byte; a bitonic sort of it opens sorted;
  * a file whose two copies of a share disagree is refused."""
import os
import shutil

import numpy as np
import pytest

from test_gpu_parity import _same_shares

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AGES_25 = [25, 30, 35, 40, 45, 50, 55, 60, 65, 70, 75, 80, 85, 90, 95, 100, 25, 30, 35, 40, 45,
           50, 55, 60, 25]
IDS = list(range(1, 26))
REGIONS = [45, 70] * 12 + [45]
AGES_32 = AGES_25 + [0] * 7
SCHEMA = ["[EMPLOYEE_ID]", "[REGION]", "[AGE]"]


def _same_gpu(a, b):
    for k, (x, y) in enumerate(zip(a.shares(), b.shares())):
        np.testing.assert_array_equal(x, y, err_msg=f"share {k} differs")


def _files(stem):
    return [os.path.join(HERE, f"{stem}_3pc_{p}.csv") for p in range(3)]


@pytest.fixture(scope="module")
def G():
    import orq_amd

    orq_amd.lib()
    return orq_amd


def _table(G, keys):
    from orq_amd.table import EncodedTable

    c = G.Ctx(keys)
    return c, EncodedTable.with_schema(c, "Employee_Profile", SCHEMA, 32, width=4)


def _check_employees(t):
    got = t.open()
    for name, col in (("[AGE]", AGES_25), ("[EMPLOYEE_ID]", IDS), ("[REGION]", REGIONS)):
        np.testing.assert_array_equal(got[name].astype(np.int64), col + [0] * 7, err_msg=name)


def test_input_csv_table_data(G, keys):
    from oracle import oracle as O

    c, t = _table(G, keys)
    t.input_csv_table_data(os.path.join(HERE, "employees.csv"), 0)
    _check_employees(t)
    np.testing.assert_array_equal(t.open()["[##V]"], [1] * 25 + [0] * 7)
    # the same plaintext columns shared by party 0 in schema order on the oracle
    oc = O.Ctx(keys)
    plain = {"[##V]": [1] * 25 + [0] * 7, "[AGE]": AGES_32, "[EMPLOYEE_ID]": IDS + [0] * 7,
             "[REGION]": REGIONS + [0] * 7}
    for name in t.column_names():
        ov = oc.secret_share(np.asarray(plain[name], np.int64).astype(np.uint32), 4)
        _same_shares(ov, t[name])


def test_input_and_output_csv_table_secret_shares(G, keys, tmp_path):
    c, t = _table(G, keys)
    t.input_csv_table_secret_shares(_files("employees_secrets"))
    _check_employees(t)
    out = [str(tmp_path / f"out_{p}.csv") for p in range(3)]
    t.output_csv_table_secret_shares(out)
    for mine, ref in zip(out, _files("employees_secrets")):
        assert open(mine).read() == open(ref).read(), mine
    # and what was written reads back to the same table
    c2, t2 = _table(G, keys)
    t2.input_csv_table_secret_shares(out)
    for name in t.column_names():
        _same_gpu(t[name], t2[name])


def test_column_and_vector_share_files(G, keys, tmp_path):
    c, t = _table(G, keys)
    t.input_secret_shares("[AGE]", _files("employees_age_secrets"))
    np.testing.assert_array_equal(t.open()["[AGE]"], AGES_32)
    out = [str(tmp_path / f"age_{p}.csv") for p in range(3)]
    t.output_secret_shares("[AGE]", out)
    for mine, ref in zip(out, _files("employees_age_secrets")):
        assert open(mine).read() == open(ref).read(), mine
    # the BSharedVector file constructor, then a bitonic sort (test_io.cpp:87-113)
    v = c.vec(32, 4)
    c.read_share_files(v, _files("employees_age_secrets"))
    np.testing.assert_array_equal(c.open(v), AGES_32)
    c.bitonic_sort([v])
    np.testing.assert_array_equal(c.open(v), sorted(AGES_32))
    c.write_share_files(v, out)
    w = c.vec(32, 4)
    c.read_share_files(w, out)
    _same_gpu(v, w)
    ref = c.vec(32, 4)
    c.read_share_files(ref, _files("employees_sorted_age_secrets"))
    np.testing.assert_array_equal(c.open(ref), sorted(AGES_32))


def test_inconsistent_share_files_refused(G, keys, tmp_path):
    files = []
    for p, f in enumerate(_files("employees_age_secrets")):
        dst = tmp_path / f"bad_{p}.csv"
        shutil.copy(f, dst)
        files.append(str(dst))
    rows = open(files[1]).read().splitlines()
    a, b = rows[3].split(",")
    rows[3] = f"{int(a) ^ 1},{b}"  # party 1's copy of s_1 in row 3 no longer matches party 0's
    open(files[1], "w").write("\n".join(rows) + "\n")
    c = G.Ctx(keys)
    v = c.vec(32, 4)
    with pytest.raises(G.OrqgError, match="different copies"):
        c.read_share_files(v, files)
