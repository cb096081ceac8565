"""10X MatrixMarket input on the device (src/input.jl:78-115 readMM and
:199-208 _read_10X): the cells x genes counts built on the GPU against
scipy.io.mmread(path).T.tocsc() (duplicates summed, stored zeros kept), on
small files with the edge cases and on files large enough for the staged
upload, plain and gzip; read_10X's v2 and v3 directory layouts."""
import gzip
import os

import numpy as np
import pandas as pd
import pytest
import scipy.io

pytestmark = pytest.mark.gpu

HEADER_INT = "%%MatrixMarket matrix coordinate integer general\n"
HEADER_REAL = "%%MatrixMarket matrix coordinate real general\n"


def _write_mtx(path, shape, r, c, v, real=False, gz=False, comments=("% written by a test",)):
    head = (HEADER_REAL if real else HEADER_INT) + "".join(x + "\n" for x in comments)
    head += f"{shape[0]} {shape[1]} {len(r)}\n"
    body = pd.DataFrame({"r": r, "c": c, "v": v}).to_csv(sep=" ", header=False, index=False)
    data = head.encode() + body.encode()
    if gz:
        with gzip.open(path, "wb", compresslevel=1) as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)
    return str(path)


def _expected(path):
    A = scipy.io.mmread(path).T.tocsc()
    A.sort_indices()
    return A


def _check(sev, path):
    p = sev.Pipeline()
    try:
        m, n, nnz = p.load_mtx(path)
        X = p.get_counts()
    finally:
        p.close()
    A = _expected(path)
    print(f"{os.path.basename(path)}: {m} cells x {n} genes, nnz {nnz}", flush=True)
    assert (m, n) == A.shape and nnz == A.nnz
    np.testing.assert_array_equal(X.indptr, A.indptr)
    np.testing.assert_array_equal(X.indices, A.indices)
    np.testing.assert_array_equal(X.data, A.data)


def test_small_edge_cases(sev, tmp_path):
    # duplicates (summed), a stored zero, unsorted entries, an empty gene and
    # an empty cell, comments and a blank line before the size line
    r = [3, 1, 3, 2, 1, 2, 3]
    c = [1, 1, 1, 4, 2, 3, 5]
    v = [2, 5, 1, 0, 7, 4, 9]
    path = _write_mtx(tmp_path / "small.mtx", (4, 6), r, c, v, comments=("% a", "%", ""))
    _check(sev, path)


def test_random_staged(sev, tmp_path):
    # more than 64 MB of keys: the entries go up through the pinned buffers
    rng = np.random.default_rng(7)
    genes, cells, E = 3000, 20000, 4_500_000
    r = rng.integers(1, genes + 1, E)
    c = rng.integers(1, cells + 1, E)
    v = rng.integers(0, 50, E)
    _check(sev, _write_mtx(tmp_path / "big.mtx", (genes, cells), r, c, v))


def test_gzip_and_real_field(sev, tmp_path):
    rng = np.random.default_rng(8)
    genes, cells, E = 500, 4000, 300_000
    r = rng.integers(1, genes + 1, E)
    c = rng.integers(1, cells + 1, E)
    v = rng.integers(1, 20, E)
    _check(sev, _write_mtx(tmp_path / "m.mtx.gz", (genes, cells), r, c, v, gz=True))
    # a real field holding integers (convert_counts converts it to Int32)
    _check(sev, _write_mtx(tmp_path / "real.mtx", (genes, cells), r, c, v.astype(np.float64),
                           real=True))


def test_errors(sev, tmp_path):
    p = sev.Pipeline()
    try:
        path = _write_mtx(tmp_path / "frac.mtx", (2, 2), [1], [1], [1.5], real=True)
        with pytest.raises(sev.SeveroError, match="not an integer"):
            p.load_mtx(path)
        path = _write_mtx(tmp_path / "big.mtx", (2, 2), [1, 1], [1, 1],
                          [2**31 - 1, 1])
        with pytest.raises(sev.SeveroError, match="Int32"):
            p.load_mtx(path)
        path = _write_mtx(tmp_path / "oob.mtx", (2, 2), [3], [1], [1])
        with pytest.raises(sev.SeveroError, match="out of range"):
            p.load_mtx(path)
    finally:
        p.close()


@pytest.mark.parametrize("v3", [False, True])
def test_read_10X(sev, tmp_path, v3):
    rng = np.random.default_rng(9)
    genes, cells, E = 40, 300, 3000
    r = rng.integers(1, genes + 1, E)
    c = rng.integers(1, cells + 1, E)
    v = rng.integers(1, 9, E)
    d = tmp_path / ("v3" if v3 else "v2")
    d.mkdir()
    feats = [f"ENSG{i:05d}\tGene_{i % 35}\tGene Expression" for i in range(genes)]
    bars = [f"AAAC{i:06d}-1" for i in range(cells)]

    def put(name, lines):
        data = ("\n".join(lines) + "\n").encode()
        if v3:
            with gzip.open(d / (name + ".gz"), "wb") as f:
                f.write(data)
        else:
            (d / name).write_bytes(data)

    put("features.tsv" if v3 else "genes.tsv", feats)
    put("barcodes.tsv", bars)
    mtx = _write_mtx(d / ("matrix.mtx.gz" if v3 else "matrix.mtx"), (genes, cells), r, c, v,
                     gz=v3)
    C = sev.read_10X(str(d))
    A = _expected(mtx)
    np.testing.assert_array_equal(C.X.indptr, A.indptr)
    np.testing.assert_array_equal(C.X.indices, A.indices)
    np.testing.assert_array_equal(C.X.data, A.data)
    assert C.barcodes == bars
    # column 2, '_' -> '-', made unique (Gene-0 .. Gene-34, then Gene-0.1 ...)
    assert C.features[:2] == ["Gene-0", "Gene-1"]
    assert C.features[35] == "Gene-0.1"
    assert len(set(C.features)) == genes
    # into a pipeline: the counts stay on the device
    p = sev.Pipeline()
    try:
        C2 = sev.read_10X(str(d), pipeline=p)
        assert C2.X is None and p.counts_info()[:2] == (cells, genes)
    finally:
        p.close()
