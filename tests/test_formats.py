"""The reference's on-disk formats (include/matrix.h:209-257 writers,
:396-533 readers) through libgraphhip: byte layouts written out by hand with
struct, hand-written text files with their expected CSR, and round trips.
Host code only."""
import struct

import numpy as np
import pytest

from graphkokkos_amd import SpMtx, write_matrix
from graphkokkos_amd._lib import GraphHipError


def rand_csr(rng, m, n, density=0.2):
    rst, cols, vals = [0], [], []
    for _ in range(m):
        c = np.sort(rng.choice(n, size=rng.binomial(n, density), replace=False))
        cols.append(c)
        vals.append(rng.integers(-1000, 1000, size=len(c)))
        rst.append(rst[-1] + len(c))
    return SpMtx(m, n, np.array(rst, np.int32), np.concatenate(cols).astype(np.int32),
                 np.concatenate(vals).astype(np.int32), "MCIG")


def assert_same(a, b, values=True):
    assert (a.m, a.n, a.nz) == (b.m, b.n, b.nz)
    assert np.array_equal(a.Rst, b.Rst) and np.array_equal(a.Col, b.Col)
    if values:
        assert np.array_equal(np.asarray(a.Val, np.int64), np.asarray(b.Val, np.int64))


def test_bin_layout_is_the_references(tmp_path):
    rng = np.random.default_rng(1)
    a = rand_csr(rng, 7, 7)
    p = tmp_path / "a.bin"
    write_matrix(a, p, "bin")
    raw = p.read_bytes()
    m, n, nz = struct.unpack_from("<3Q", raw, 4)
    assert raw[:4] == b"MCIG" and (m, n, nz) == (a.m, a.n, a.nz)
    off = 4 + 24
    rst = np.frombuffer(raw, np.int32, m + 1, off)
    col = np.frombuffer(raw, np.int32, nz, off + 4 * (m + 1))
    val = np.frombuffer(raw, np.int32, nz, off + 4 * (m + 1) + 4 * nz)
    assert len(raw) == off + 4 * (m + 1) + 8 * nz
    assert np.array_equal(rst, a.Rst) and np.array_equal(col, a.Col) and np.array_equal(val, a.Val)


def test_bin_reader_on_hand_made_bytes(tmp_path):
    # 3x4 matrix: row 0 = {1: 5, 3: -2}, row 1 = {}, row 2 = {0: 7}
    raw = b"MCIG" + struct.pack("<3Q", 3, 4, 3) + struct.pack("<4i", 0, 2, 2, 3) + \
        struct.pack("<3i", 1, 3, 0) + struct.pack("<3i", 5, -2, 7)
    p = tmp_path / "h.bin"
    p.write_bytes(raw)
    a = SpMtx.from_file(p, "bin")
    assert (a.m, a.n, a.nz) == (3, 4, 3)
    assert a.Rst.tolist() == [0, 2, 2, 3] and a.Col.tolist() == [1, 3, 0]
    assert np.asarray(a.Val).tolist() == [5, -2, 7]


@pytest.mark.parametrize("fmt", ["bin", "crs"])
def test_round_trips(tmp_path, fmt):
    rng = np.random.default_rng(2)
    for m in (1, 5, 40):
        a = rand_csr(rng, m, m)
        p = tmp_path / f"r{m}.{fmt}"
        write_matrix(a, p, fmt)
        assert_same(SpMtx.from_file(p, fmt), a)


def test_mtx_writer_round_trip_and_truss_numbers(tmp_path):
    rng = np.random.default_rng(3)
    a = rand_csr(rng, 30, 30)
    keep = a.Col != np.repeat(np.arange(a.m), np.diff(a.Rst))  # the mtx reader drops self loops
    rows = np.repeat(np.arange(a.m), np.diff(a.Rst))[keep]
    rst = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=a.m))]).astype(np.int32)
    b = SpMtx(a.m, a.n, rst, a.Col[keep], np.asarray(a.Val)[keep], "MCIG")
    p = tmp_path / "b.mtx"
    write_matrix(b, p, "mtx")
    assert_same(SpMtx.from_file(p, "mtx"), b)
    # per-entry results (e.g. truss numbers) written as the values
    truss = (np.arange(b.nz) % 7 + 2).astype(np.uint32)
    write_matrix(b, p, "mtx", values=truss.astype(np.int32))
    c = SpMtx.from_file(p, "mtx")
    assert np.array_equal(np.asarray(c.Val, np.int64), truss.astype(np.int64))


def test_crs_text(tmp_path):
    p = tmp_path / "a.crs"
    p.write_text("3 4 MCIG\n2 0 1 2\n10 20 30 40\n0 1 1 4\n")
    a = SpMtx.from_file(p, "crs")
    assert (a.m, a.n, a.nz) == (3, 3, 4)  # n = m (the reference leaves it unset)
    assert a.Rst.tolist() == [0, 1, 1, 4] and a.Col.tolist() == [2, 0, 1, 2]
    assert np.asarray(a.Val).tolist() == [10, 20, 30, 40]


def test_graph_text(tmp_path):
    p = tmp_path / "g.graph"
    p.write_text("4 5\n1 2\n0\n\n0 1\n")
    a = SpMtx.from_file(p, "graph")
    assert (a.m, a.n, a.nz) == (4, 4, 5)
    assert a.Rst.tolist() == [0, 2, 3, 3, 5] and a.Col.tolist() == [1, 2, 0, 0, 1]
    assert np.asarray(a.Val).tolist() == [1] * 5


def test_rmat_text_is_symmetric_sorted_deduplicated(tmp_path):
    p = tmp_path / "e.rmat"
    p.write_text("4 4 6\n0 1\n1 0\n2 3\n3 3\n0 2\n3 3\n")
    a = SpMtx.from_file(p, "rmat")
    assert (a.m, a.n) == (4, 4)
    assert [a.row(i).tolist() for i in range(4)] == [[1, 2], [0], [0, 3], [2, 3]]
    assert np.asarray(a.Val).tolist() == [1] * a.nz


@pytest.mark.parametrize("fmt,text", [
    ("graph", "2 1\n0 1\n"),            # more columns than nz
    ("graph", "2 1\n5\n"),              # column out of range
    ("crs", "2 2 MCIG\n0 1\n1 1\n0 2 1\n"),  # row pointers decrease
    ("rmat", "3 3 1\n0 7\n"),           # pair out of range
    ("rmat", "3 4 1\n0 1\n"),           # not square
])
def test_malformed_files_are_rejected(tmp_path, fmt, text):
    p = tmp_path / f"bad.{fmt}"
    p.write_text(text)
    with pytest.raises(GraphHipError):
        SpMtx.from_file(p, fmt)


def test_short_bin_and_unknown_format(tmp_path):
    p = tmp_path / "s.bin"
    p.write_bytes(b"MCIG" + struct.pack("<3Q", 4, 4, 10) + b"\x00" * 8)
    with pytest.raises(GraphHipError):
        SpMtx.from_file(p, "bin")
    with pytest.raises(GraphHipError):
        SpMtx.from_file(p, "xyz")
