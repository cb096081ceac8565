"""MatrixMarket reader (src/input.jl:26-115 readMM) on the host: the parsed
entries against scipy.io.mmread on the same files, and readMM's errors.
No GPU: sev_mtx_info / sev_mtx_read only parse."""
import ctypes as C
import gzip

import numpy as np
import pytest
import scipy.io

HEADER_INT = "%%MatrixMarket matrix coordinate integer general\n"
HEADER_REAL = "%%MatrixMarket matrix coordinate real general\n"


@pytest.fixture(scope="module")
def lib(sev):
    return sev.load()


def _entries(lib, path):
    rows, cols, entries, is_real = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
    rc = lib.sev_mtx_info(str(path).encode(), C.byref(rows), C.byref(cols), C.byref(entries),
                          C.byref(is_real))
    if rc != 0:
        raise RuntimeError(lib.sev_last_error().decode())
    E = entries.value
    I = np.empty(E, np.int64)
    J = np.empty(E, np.int64)
    V = np.empty(E, np.float64 if is_real.value else np.int64)
    vi = None if is_real.value else V.ctypes.data
    vf = V.ctypes.data if is_real.value else None
    rc = lib.sev_mtx_read(str(path).encode(), I.ctypes.data, J.ctypes.data, vi, vf)
    if rc != 0:
        raise RuntimeError(lib.sev_last_error().decode())
    return (rows.value, cols.value), I, J, V


def _write(path, text, gz=False):
    data = text.encode()
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        path.write_bytes(data)
    return path


def _check_vs_scipy(lib, path):
    shape, I, J, V = _entries(lib, path)
    A = scipy.io.mmread(str(path))
    assert shape == A.shape
    np.testing.assert_array_equal(I - 1, A.row)
    np.testing.assert_array_equal(J - 1, A.col)
    np.testing.assert_array_equal(V, A.data)


def test_small_file_with_comments_duplicates_zeros(lib, tmp_path):
    text = (HEADER_INT + "% a comment\n%\n\n3 4 6\n3 1 2\n1 1 5\n3\t1  1\n2 4 0\n1 2 7\n"
            "2 3 -4\n")
    _check_vs_scipy(lib, _write(tmp_path / "a.mtx", text))


def test_signed_values(lib, tmp_path):
    # Julia's parse(Int64, x) takes a leading '+' (scipy's reader does not)
    shape, I, J, V = _entries(lib, _write(tmp_path / "s.mtx", HEADER_INT + "2 2 2\n+1 2 +7\n2 1 -3\n"))
    np.testing.assert_array_equal(I, [1, 2])
    np.testing.assert_array_equal(V, [7, -3])


def test_crlf_and_trailing_lines_ignored(lib, tmp_path):
    # readMM reads exactly `entries` lines after the size line
    text = (HEADER_INT + "2 2 2\r\n1 1 3\r\n2 2 4\r\n% ignored\nnot an entry\n")
    shape, I, J, V = _entries(lib, _write(tmp_path / "b.mtx", text))
    assert shape == (2, 2)
    np.testing.assert_array_equal(I, [1, 2])
    np.testing.assert_array_equal(J, [1, 2])
    np.testing.assert_array_equal(V, [3, 4])


def test_no_final_newline(lib, tmp_path):
    shape, I, J, V = _entries(lib, _write(tmp_path / "c.mtx", HEADER_INT + "1 1 1\n1 1 9"))
    np.testing.assert_array_equal(V, [9])


def test_real_field_and_gzip(lib, tmp_path):
    rng = np.random.default_rng(0)
    r = rng.integers(1, 51, 3000)
    c = rng.integers(1, 41, 3000)
    v = rng.standard_normal(3000).round(6)
    body = "".join(f"{a} {b} {float(x)!r}\n" for a, b, x in zip(r, c, v))
    text = HEADER_REAL + f"50 40 {len(r)}\n" + body
    _check_vs_scipy(lib, _write(tmp_path / "d.mtx.gz", text, gz=True))
    _check_vs_scipy(lib, _write(tmp_path / "d.mtx", text))


def test_many_lines_split_over_threads(lib, tmp_path):
    # enough text that several host threads each parse a range of lines
    rng = np.random.default_rng(1)
    n = 400_000
    r = rng.integers(1, 2001, n)
    c = rng.integers(1, 30001, n)
    v = rng.integers(0, 1000, n)
    body = "\n".join(f"{a} {b} {x}" for a, b, x in zip(r, c, v)) + "\n"
    path = _write(tmp_path / "e.mtx", HEADER_INT + f"2000 30000 {n}\n" + body)
    shape, I, J, V = _entries(lib, path)
    np.testing.assert_array_equal(I, r)
    np.testing.assert_array_equal(J, c)
    np.testing.assert_array_equal(V, v)


def test_read_mm_matches_scipy_csc(sev, tmp_path):
    text = HEADER_INT + "3 4 6\n3 1 2\n1 1 5\n3 1 1\n2 4 0\n1 2 7\n2 3 -4\n"
    path = _write(tmp_path / "f.mtx", text)
    X = sev.read_mm(str(path))
    A = scipy.io.mmread(str(path)).tocsc()
    A.sort_indices()
    np.testing.assert_array_equal(X.indptr, A.indptr)
    np.testing.assert_array_equal(X.indices, A.indices)
    np.testing.assert_array_equal(X.data, A.data)
    assert X.nnz == 5  # the duplicate (3, 1) summed, the stored zero kept


@pytest.mark.parametrize("text, msg", [
    ("%%MatrixMarket matrix coordinate integer\n1 1 0\n", "Malformed MatrixMarket header"),
    ("%MatrixMarket matrix coordinate integer general\n1 1 0\n", "Not a valid MatrixMarket"),
    ("%%MatrixMarket vector coordinate integer general\n1 1 0\n", "Unknown MatrixMarket data type"),
    ("%%MatrixMarket matrix coordinate pattern general\n1 1 0\n", "Unsupported field pattern"),
    ("%%MatrixMarket matrix coordinate integer symmetric\n1 1 0\n", "Unknown matrix symmetry"),
    ("%%MatrixMarket matrix array integer general\n1 1\n", "Unsupported matrix format"),
    (HEADER_INT + "2 2\n", "Could not read in matrix dimensions"),
    (HEADER_INT + "2 2 3\n1 1 1\n2 2 2\n", "Could not read matrix entry"),
    (HEADER_INT + "2 2 1\n1 1\n", "Could not read matrix entry"),
    (HEADER_INT + "2 2 1\n1 1 1.5\n", "Could not parse matrix entry"),
])
def test_errors(lib, tmp_path, text, msg):
    with pytest.raises(RuntimeError, match=msg):
        _entries(lib, _write(tmp_path / "bad.mtx", text))


def test_make_unique(sev):
    names = ["a", "b", "a", "a.1", "a", "b"]
    # make_unique! (src/utils.jl): "a.1" is already a name, so the first
    # repeated "a" becomes "a.2", the next "a.3"
    assert sev.make_unique(names) == ["a", "b", "a.2", "a.1", "a.3", "b.1"]
