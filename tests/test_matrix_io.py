"""Sparse-matrix files as SparseMatrix::read / read_matlab / print_matlab handle
them (sparse_matrix.C:398-491, 686-745, 1050-1330), on the reference's own
fixtures (tests/matrices/geom_*_extraction_op.*, copied under
tests/golden/matrices): Matlab text, gzipped Matlab text and PETSc binary with
32- and 64-bit indices must describe the same matrices; print_matlab output
reads back exactly; malformed files raise."""
import gzip
import os

import numpy as np
import pytest

import libmesh_amd as L

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "matrices")


def parse_text(path):
    """Independent reading of the fixture text: size comment, 1-based entries."""
    opener = gzip.open if path.endswith(".gz") else open
    m = n = None
    ijv = []
    with opener(path, "rt") as f:
        for line in f:
            t = line.split()
            if line.startswith("% Size"):
                m, n = int(t[3]), int(t[4])
            elif len(t) == 3 and t[0].isdigit():
                ijv.append((int(t[0]) - 1, int(t[1]) - 1, float(t[2])))
    return m, n, ijv


def dense(m, n, i, j, v):
    A = np.zeros((m, n))
    A[i, j] = v
    return A


@pytest.mark.parametrize("g", [1, 2, 4, 8])
def test_matlab_fixture_read(g):
    path = os.path.join(GOLD, f"geom_{g}_extraction_op.m")
    m, n, i, j, v = L.read_matrix(path)
    pm, pn, ijv = parse_text(path)
    assert (m, n) == (pm, pn)
    assert len(v) == len(ijv)
    assert np.array_equal(i, [e[0] for e in ijv]) and np.array_equal(j, [e[1] for e in ijv])
    assert np.array_equal(v, [e[2] for e in ijv])
    if g != 8:  # testReadMatlab: the gzipped file gives the same matrix and norms
        m2, n2, i2, j2, v2 = L.read_matrix(path + ".gz")
        A, B = dense(m, n, i, j, v), dense(m2, n2, i2, j2, v2)
        assert np.array_equal(A, B)
        assert np.abs(A).sum(0).max() == np.abs(B).sum(0).max()  # l1
        assert np.abs(A).sum(1).max() == np.abs(B).sum(1).max()  # linfty


def test_petsc_binary_fixtures_match_matlab():
    """The reference's binary and text files of geom_1 differ in the last bit of
    some entries (1.0000000000000002 in the binary, 1.0 in the text), so they
    agree to one unit in the last place; the two binary widths agree exactly."""
    m, n, i, j, v = L.read_matrix(os.path.join(GOLD, "geom_1_extraction_op.m"))
    A = dense(m, n, i, j, v)
    B = {}
    for ext in ("petsc32", "petsc64"):
        mb, nb, ib, jb, vb = L.read_matrix(os.path.join(GOLD, f"geom_1_extraction_op.{ext}"))
        assert (mb, nb) == (m, n)
        B[ext] = dense(mb, nb, ib, jb, vb)
        assert np.abs(B[ext] - A).max() <= 2.3e-16 * np.abs(A).max()
        assert np.array_equal(B[ext] != 0, A != 0)
    assert np.array_equal(B["petsc32"], B["petsc64"])


@pytest.mark.parametrize("g", [2, 8])
def test_print_matlab_round_trip(g, tmp_path):
    m, n, i, j, v = L.read_matrix(os.path.join(GOLD, f"geom_{g}_extraction_op.m"))
    rp = np.zeros(m + 1, dtype=np.int32)
    np.add.at(rp, i + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    out = tmp_path / "a.m"
    L.write_matrix_matlab(out, m, n, rp, j, v)
    m2, n2, i2, j2, v2 = L.read_matrix(out)
    keep = v != 0.0  # print_matlab skips exact zeros
    assert (m2, n2) == (m, n)
    assert np.array_equal(i2, i[keep]) and np.array_equal(j2, j[keep]) and np.array_equal(v2, v[keep])
    text = out.read_text().splitlines()
    assert text[2] == f"% Size = {m} {n}" and text[5] == "zzz = [" and text[-1] == "Mat_sparse = spconvert(zzz);"


@pytest.mark.parametrize("body,msg", [
    ("1 1  1.0\nzzz = [\n];\n", "premature"),
    ("zzz = [\n2 1  1.0\n1 1  1.0\n];\n", "out-of-order"),
    ("zzz = [\n0 1  1.0\n];\n", "1-based"),
    ("zzz = [\n1 1  1.0\n", "Incomplete"),
])
def test_malformed_matlab_raises(body, msg, tmp_path):
    p = tmp_path / "bad.m"
    p.write_text(body)
    with pytest.raises(L.LmhError, match=msg):
        L.read_matrix(p)


def test_unknown_extension_raises(tmp_path):
    p = tmp_path / "a.h5"
    p.write_bytes(b"\x89HDF")
    with pytest.raises(L.LmhError, match="Unrecognized"):
        L.read_matrix(p)
