"""The C-ABI library loads and exports every symbol include/severo_hip.h
declares; host-side layout helpers.  CPU only (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "severo_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_path_entry_points():
    names = declared_functions()
    for must in ("sev_normalize_cells", "sev_scale_features", "sev_centered_mul",
                 "sev_irlba_centered_csc", "FindNeighboursEuclidean32",
                 "FindNeighboursEuclidean64", "sev_knn", "sev_snn_jaccard",
                 "sev_pipeline_snn", "irlba"):
        assert must in names


def test_library_exports_every_declared_symbol(sev):
    lib = ctypes.CDLL(sev._lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_bindings_cover_header(sev):
    assert set(declared_functions()) == set(sev._lib.SIGNATURES)


def test_no_cpu_fallback_without_device(sev):
    """Without a GPU the library refuses to compute (no silent CPU path)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = sev.load()
    colptr = np.array([1, 2], np.int64)
    rowval = np.array([1], np.int64)
    counts = np.array([3], np.int32)
    out = np.zeros(1, np.float32)
    rc = L.sev_normalize_cells(1, 1, colptr.ctypes.data, rowval.ctypes.data, counts.ctypes.data,
                               0, 0, 1.0, 0, out.ctypes.data)
    assert rc == -5
    assert b"device" in L.sev_last_error()


def test_synth_tables_calibration(sev):
    t = sev.nb_tables(20000, seed=0, density=0.05)
    assert abs(t.expected_density() - 0.05) < 1e-6
    t2 = sev.nb_tables(33538, seed=0, density=0.04, boost_frac=0.05)
    assert abs(t2.expected_density() - 0.04) < 1e-6
    f = sev.top_mean_features(t, 2000)
    assert len(set(f.tolist())) == 2000
    assert np.all(np.diff(t.gene_mean[f]) <= 0)


@pytest.mark.parametrize("name,dt", [("FindNeighboursEuclidean32", np.float32),
                                     ("FindNeighboursCosine64", np.float64)])
@pytest.mark.parametrize("k", [65, 5])
def test_dropin_knn_failure_is_flagged(sev, name, dt, k):
    """The libcell-signature searches return nothing, so a failure (here k
    beyond the longest list, or no device) leaves indices that Julia's
    sparse() rejects (0) and NaN distances instead of undefined memory."""
    import torch
    L = sev.load()
    n, d = 40, 3
    X = np.asfortranarray(np.random.default_rng(0).standard_normal((n, d)).astype(dt))
    idx = np.full((n, k), 7, np.int32, order="F")
    dist = np.zeros((n, k), dt, order="F")
    if k <= 64 and torch.cuda.is_available():
        pytest.skip("a GPU is present: the k <= 64 call succeeds")
    getattr(L, name)(X.ctypes.data, n, d, n, k, 0, 1, 0, idx.ctypes.data, dist.ctypes.data)
    assert np.all(idx == 0) and np.all(np.isnan(dist))
    if k > 64:
        assert b"k must be in [1, 64]" in L.sev_last_error()
