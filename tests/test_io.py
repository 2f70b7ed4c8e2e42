"""Graph files (host code, no GPU): the METIS .graph reader and partition
writer of hunyuangraph_io.h:25-205 and the Matrix Market converter of
matrix_to_graph/matrix_to_graph.c.

Pinned by hand-derived CSR of the files under tests/golden/io, by every error
reason the reference's reader gives, by write -> read round trips, and for
Matrix Market by scipy's independent reader (scipy.io.mmread) and writer
(scipy.io.mmwrite)."""
import json
import os
import zlib

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
KEYS = ["xadj", "adjncy", "vwgt", "adjwgt"]


def _expect(got, want):
    for arr, key in zip(got, KEYS):
        assert arr.tolist() == want[key], key


@pytest.mark.parametrize("name", ["metis_manual.graph", "weighted_011.graph", "crlf_isolated.graph",
                                  "general_real.mtx", "symmetric_pattern.mtx"])
def test_fixtures(hgp, name):
    want = json.load(open(os.path.join(GOLD, "expected.json")))[name]
    path = os.path.join(GOLD, name)
    got = hgp.read_mtx(path) if name.endswith(".mtx") else hgp.read_graph(path)
    _expect(got, want)


def _write(tmp_path, text, name="g.graph"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


# Each malformed file the reference's reader rejects, with the reason it gives.
BAD_GRAPHS = [
    ("7\n", "numbers of vertices and edges"),
    ("% only comments\n", "premature end of file"),
    ("0 1\n", "must be positive"),
    ("3 0\n2\n1\n\n", "must be positive"),
    ("3 1 112\n2\n1\n\n", "file format"),
    ("3 1 -1\n2\n1\n\n", "file format"),
    ("3 1 010\n1 2\n-1 1\n1\n", "must be >= 0"),
    ("3 1 010\n1 2\n\n1\n", "has no weight"),
    ("3 1\n4\n1\n\n", "out of bounds"),
    ("3 1\n0\n1\n\n", "out of bounds"),
    ("3 1 001\n2 0\n1 1\n\n", "must be positive"),
    ("3 1 001\n2 -3\n1 1\n\n", "must be positive"),
    ("3 1 001\n2\n1 1\n\n", "premature end of line"),
    ("3 1\n2 3\n1\n1\n", "more edges"),
    ("3 2\n2\n1\n\n", "each edge is counted once"),
    ("3 1\n2\n1\n", "premature end of file while reading vertex 3"),
    ("3 1 0 2\n2\n1\n\n", "one vertex weight"),
]


@pytest.mark.parametrize("text,reason", BAD_GRAPHS)
def test_read_graph_rejects(hgp, tmp_path, text, reason):
    with pytest.raises(hgp.HgpError) as e:
        hgp.read_graph(_write(tmp_path, text))
    assert e.value.code in (-1, -4)   # invalid input, unsupported
    assert reason in str(e.value)


def test_missing_file(hgp, tmp_path):
    with pytest.raises(hgp.HgpError) as e:
        hgp.read_graph(str(tmp_path / "absent.graph"))
    assert e.value.code == -6
    with pytest.raises(hgp.HgpError) as e:
        hgp.read_mtx(str(tmp_path / "absent.mtx"))
    assert e.value.code == -6


def test_reader_quirks(hgp, tmp_path):
    # a digit of fmt other than 1 leaves its field out, as "%03d" digit tests do
    x, a, vw, aw = hgp.read_graph(_write(tmp_path, "3 1 2\n2\n1\n\n"))
    assert a.tolist() == [1, 0] and aw.tolist() == [1, 1]
    # parsing a line stops at the first thing that is not a number
    x, a, vw, aw = hgp.read_graph(_write(tmp_path, "3 2 % header comment\n2 3 # rest ignored\n1x\n1\n"))
    assert x.tolist() == [0, 2, 3, 4] and a.tolist() == [1, 2, 0, 0]
    # lines after the last vertex are not read; blank lines are vertices
    x, a, vw, aw = hgp.read_graph(_write(tmp_path, "4 1\n\n3\n2\n\ntrailing text\n"))
    assert x.tolist() == [0, 0, 1, 2, 2]
    # fmt 1xx: the vertex size leads each line and is read and ignored
    x, a, vw, aw = hgp.read_graph(_write(tmp_path, "3 1 111\n9 2 2 7\n9 3 1 7\n9 1\n"))
    assert a.tolist() == [1, 0] and vw.tolist() == [2, 3, 1] and aw.tolist() == [7, 7]
    # vertex weight 0 is allowed (>= 0)
    x, a, vw, aw = hgp.read_graph(_write(tmp_path, "2 1 10\n0 2\n5 1\n"))
    assert vw.tolist() == [0, 5]


@pytest.mark.parametrize("gen", ["grid", "rgg", "rmat"])
@pytest.mark.parametrize("weights", [0, 1, 10, 11])
def test_graph_round_trip(hgp, tmp_path, gen, weights):
    if gen == "grid":
        g = hgp.gen_grid2d(17, 23)
    elif gen == "rgg":
        g = hgp.gen_rgg3d(3000, hgp.rgg_radius(3000, 10), seed=3, vw=(0, 9), ew=(1, 40))
    else:
        g = hgp.gen_rmat(12, 20000, seed=4)
    x, a, vw, aw = g
    p = str(tmp_path / "rt.graph")
    hgp.write_graph(p, x, a, vw if weights >= 10 else None, aw if weights % 10 else None)
    header = open(p).readline().split()
    assert header[:2] == [str(len(x) - 1), str(len(a) // 2)]
    assert (len(header) == 3) == (weights != 0)
    x2, a2, vw2, aw2 = hgp.read_graph(p)
    np.testing.assert_array_equal(x2, x)
    np.testing.assert_array_equal(a2, a)
    np.testing.assert_array_equal(vw2, vw if weights >= 10 else np.ones_like(vw))
    np.testing.assert_array_equal(aw2, aw if weights % 10 else np.ones_like(aw))


def test_write_partition(hgp, tmp_path):
    part = np.array([3, 0, 2, 2, 1, 0, 3], np.int32)
    path = hgp.write_partition(str(tmp_path / "g.graph"), part, 4)
    assert path == str(tmp_path / "g.graph.part.4")
    assert open(path).read() == "3\n0\n2\n2\n1\n0\n3\n"
    big = (np.arange(300000, dtype=np.int64) * 7919 % 1000).astype(np.int32)
    path = hgp.write_partition(str(tmp_path / "h"), big, 1000)
    assert np.loadtxt(path, dtype=np.int32).tolist() == big.tolist()
    with pytest.raises(hgp.HgpError):
        hgp.write_partition(str(tmp_path / "no_dir" / "g"), part, 4)


def _scipy_graph(path):
    """pattern(A + A^T) without the diagonal, from scipy's own reader."""
    A = sp.coo_matrix(scipy.io.mmread(path))
    n = A.shape[0]
    P = sp.coo_matrix((np.ones(A.nnz), (A.row, A.col)), shape=(n, n)).tocsr()
    C = (P + P.T).tocsr()
    C.setdiag(0)
    C.eliminate_zeros()
    C.sort_indices()
    return C.indptr.astype(np.int64), C.indices.astype(np.int32)


@pytest.mark.parametrize("field", ["real", "integer", "complex", "pattern"])
@pytest.mark.parametrize("symmetry", ["general", "symmetric", "hermitian", "skew-symmetric"])
def test_mtx_against_scipy(hgp, tmp_path, field, symmetry):
    if symmetry == "hermitian" and field != "complex":
        pytest.skip("hermitian needs complex values")
    if symmetry == "skew-symmetric" and field == "pattern":
        pytest.skip("no skew-symmetric pattern matrices")
    rng = np.random.default_rng(zlib.crc32(f"{field}/{symmetry}".encode()))
    n = 200
    M = sp.random(n, n, density=0.03, random_state=rng, format="coo")
    vals = rng.integers(1, 9, M.nnz).astype(np.float64)
    M = sp.coo_matrix((vals, (M.row, M.col)), shape=(n, n))
    if field == "complex":
        M = M.astype(np.complex128) * (1 + 2j)
    if symmetry == "symmetric":
        M = M + M.T
    elif symmetry == "hermitian":
        M = M + M.conj().T
        M = M - sp.diags(M.diagonal().imag * 1j)   # a hermitian diagonal is real
    elif symmetry == "skew-symmetric":
        M = M - M.T
    M = sp.coo_matrix(M)
    M.eliminate_zeros()
    if field == "integer":
        M = sp.coo_matrix((np.rint(M.data.real).astype(np.int64), (M.row, M.col)), shape=(n, n))
    p = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(p, M, field=field, symmetry=symmetry)
    x, a, vw, aw = hgp.read_mtx(p)
    ex, ea = _scipy_graph(p)
    np.testing.assert_array_equal(x, ex)
    np.testing.assert_array_equal(a, ea)
    assert (vw == 1).all() and (aw == 1).all()


def test_mtx_unordered_duplicates(hgp, tmp_path):
    """Entries in any order, repeated, both triangles: still one sorted, unique row each."""
    rng = np.random.default_rng(11)
    n = 500
    r = rng.integers(1, n + 1, 6000)
    c = rng.integers(1, n + 1, 6000)
    body = "".join(f"{i} {j} {rng.random():.3f}\n" for i, j in zip(r, c))
    p = _write(tmp_path, f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(r)}\n" + body, "u.mtx")
    x, a, vw, aw = hgp.read_mtx(p)
    ex, ea = _scipy_graph(p)
    np.testing.assert_array_equal(x, ex)
    np.testing.assert_array_equal(a, ea)
    # symmetric and loop-free, as the partitioner requires
    rows = np.repeat(np.arange(n), np.diff(x))
    assert not (rows == a).any()
    fwd = set(zip(rows.tolist(), a.tolist()))
    assert fwd == set((v, u) for u, v in fwd)


BAD_MTX = [
    ("3 3 1\n1 2\n", "banner"),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "coordinate"),
    ("%%MatrixMarket matrix coordinate real general\n3 4 1\n1 2 1\n", "square"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 2 1\n", "premature end of file"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 4 1\n", "out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n0 1 1\n", "out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n3 3\n", "rows, columns and entries"),
]


@pytest.mark.parametrize("text,reason", BAD_MTX)
def test_read_mtx_rejects(hgp, tmp_path, text, reason):
    with pytest.raises(hgp.HgpError) as e:
        hgp.read_mtx(_write(tmp_path, text, "bad.mtx"))
    assert reason in str(e.value)


def test_mtx_to_graph_file(hgp, tmp_path):
    """The converter's product is a .graph file that the reader takes back whole."""
    x, a, vw, aw = hgp.read_mtx(os.path.join(GOLD, "general_real.mtx"))
    p = str(tmp_path / "general_real.graph")
    hgp.write_graph(p, x, a)
    assert open(p).read() == "5 5\n2 4\n1 3 5\n2 5\n1\n2 3\n"
    got = hgp.read_graph(p)
    for u, v in zip(got, (x, a, vw, aw)):
        np.testing.assert_array_equal(u, v)
