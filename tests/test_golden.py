"""The reference's own 3PC share fixtures (tests/golden/*.csv, copied from
ORQ's tests/data/) pin the replicated share layout, the XOR reconstruction and
the expected sorted result (tests/test_io.cpp:7-113)."""
import csv
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# test_io.cpp:71-73
AGES_32 = [25, 30, 35, 40, 45, 50, 55, 60, 65, 70, 75, 80, 85, 90, 95, 100, 25, 30, 35, 40, 45,
           50, 55, 60, 25, 0, 0, 0, 0, 0, 0, 0]


def read_party_file(name, p):
    with open(os.path.join(HERE, f"{name}_3pc_{p}.csv")) as f:
        rows = list(csv.reader(f))
    if rows and not rows[0][0].lstrip("-").isdigit():
        header, rows = rows[0], rows[1:]
    else:
        header = None
    a = np.array([[int(x) for x in r] for r in rows], dtype=np.int64).astype(np.int32)
    return header, a.view(np.uint32)


def shares_of(name, col=0):
    """The three distinct shares s_0, s_1, s_2 of column pair `col`."""
    parts = [read_party_file(name, p)[1] for p in range(3)]
    s = [parts[p][:, 2 * col] for p in range(3)]
    for p in range(3):  # party p holds (s_p, s_{p+1})
        np.testing.assert_array_equal(parts[p][:, 2 * col + 1], s[(p + 1) % 3])
    return s


def test_age_shares_reconstruct():
    s = shares_of("employees_age_secrets")
    np.testing.assert_array_equal((s[0] ^ s[1] ^ s[2]).view(np.int32), AGES_32)


def test_sorted_age_shares_reconstruct_sorted():
    s = shares_of("employees_sorted_age_secrets")
    np.testing.assert_array_equal((s[0] ^ s[1] ^ s[2]).view(np.int32), sorted(AGES_32))


def test_table_shares_reconstruct_plaintext():
    header, _ = read_party_file("employees_secrets", 0)
    cols = [h[:-2] for h in header[::2]]
    with open(os.path.join(HERE, "employees.csv")) as f:
        rows = list(csv.reader(f))
    plain = {h: [int(r[i]) for r in rows[1:]] for i, h in enumerate(rows[0])}
    for c, name in enumerate(cols):
        s = shares_of("employees_secrets", c)
        got = (s[0] ^ s[1] ^ s[2]).view(np.int32)
        if name == "[##V]":
            assert list(got[: len(plain["[AGE]"])]) == [1] * len(plain["[AGE]"])
            continue
        want = plain[name] + [0] * (len(got) - len(plain[name]))
        np.testing.assert_array_equal(got, want)


def test_oracle_radix_sort_of_fixture_shares():
    from oracle import oracle as O
    from orq_amd import keys_from_seed

    s = shares_of("employees_age_secrets")
    ctx = O.Ctx(keys_from_seed(5))
    v = O.Vec(len(AGES_32), 4)
    v.set_shares(s)
    ctx.radix_sort(v)
    sorted_ref = shares_of("employees_sorted_age_secrets")
    np.testing.assert_array_equal(ctx.open(v), sorted_ref[0] ^ sorted_ref[1] ^ sorted_ref[2])


@pytest.mark.gpu
def test_gpu_radix_sort_of_fixture_shares():
    import orq_amd
    from oracle import oracle as O

    s = shares_of("employees_age_secrets")
    keys = orq_amd.keys_from_seed(5)
    g, o = orq_amd.Ctx(keys), O.Ctx(keys)
    gv = g.vec(len(AGES_32), 4)
    gv.set_shares(s)
    ov = O.Vec(len(AGES_32), 4)
    ov.set_shares(s)
    gp, op = g.radix_sort(gv), o.radix_sort(ov)
    for a, b in zip(gv.shares() + gp.shares(), ov.shares() + op.shares()):
        np.testing.assert_array_equal(a, b)
    sorted_ref = shares_of("employees_sorted_age_secrets")
    np.testing.assert_array_equal(g.open(gv), sorted_ref[0] ^ sorted_ref[1] ^ sorted_ref[2])
