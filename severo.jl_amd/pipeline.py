"""Device-resident pipeline over the ``sev_pipeline_*`` C ABI.

One handle keeps the count matrix and every intermediate on the GPU:
normalize_cells -> Y[:, features] -> scale_features -> PCA (IRLBA)
-> exact kNN -> Jaccard SNN.  Getters copy results back in Severo's layout
(1-based Int64 CSC, column-major dense, 1-based neighbour indices).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check

PRECISION = {"fp32": _lib.SEV_PREC_FP32, "bf16": _lib.SEV_PREC_BF16,
             "exact": _lib.SEV_PREC_EXACT, "fp8": _lib.SEV_PREC_FP8}


def _ptr(a):
    return None if a is None else a.ctypes.data


class Pipeline:
    def __init__(self):
        self._L = _lib.load()
        h = C.c_void_p()
        check(self._L.sev_pipeline_create(C.byref(h)))
        self._h = h
        self.k = 0
        self.npcs = 0
        self.nfeat = 0
        self.cells = None  # cells per sample of a batch

    def close(self):
        if self._h:
            self._L.sev_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(self._L.sev_pipeline_synchronize(self._h))

    def profile(self, enable=True, only=None):
        """Bracket every kernel of this handle with HIP events, or with
        only="name" just the regions of that name (clears the previous
        records)."""
        if enable and only:
            check(self._L.sev_pipeline_profile_only(self._h, only.encode()))
        else:
            check(self._L.sev_pipeline_profile(self._h, int(enable)))

    def profile_read(self, max_entries=64):
        """{name: (launches, total_ms, algorithmic_bytes, algorithmic_flops)}"""
        names = C.create_string_buffer(32 * max_entries)
        launches = np.zeros(max_entries, np.int64)
        ms = np.zeros(max_entries)
        by = np.zeros(max_entries)
        fl = np.zeros(max_entries)
        cnt = C.c_int()
        check(self._L.sev_pipeline_profile_read(self._h, max_entries, names, _ptr(launches),
                                                _ptr(ms), _ptr(by), _ptr(fl), C.byref(cnt)))
        raw = names.raw
        out = {}
        for i in range(min(cnt.value, max_entries)):
            name = raw[32 * i:32 * i + 32].split(b"\0", 1)[0].decode()
            out[name] = (int(launches[i]), float(ms[i]), float(by[i]), float(fl[i]))
        return out

    # -- counts ------------------------------------------------------------
    def load_counts(self, X):
        X = sp.csc_matrix(X)
        X.sort_indices()
        colptr = np.ascontiguousarray(X.indptr, np.int64) + 1
        rowval = np.ascontiguousarray(X.indices, np.int64) + 1
        vals = np.ascontiguousarray(X.data, np.int32)
        check(self._L.sev_pipeline_load_counts(self._h, X.shape[0], X.shape[1], _ptr(colptr),
                                               _ptr(rowval), _ptr(vals)))
        self.cells = None

    def load_mtx(self, path):
        """Load a 10X matrix.mtx[.gz] (genes x cells) as the cells x genes
        counts, built on the device (readMM + sparse + copy(X') of
        _read_10X, src/input.jl:78-115,199-208).  Returns (cells, genes, nnz)."""
        m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(self._L.sev_pipeline_load_mtx(self._h, os.fsencode(path), C.byref(m), C.byref(n),
                                            C.byref(nnz)))
        self.cells = None
        return m.value, n.value, nnz.value

    def load_counts_julia(self, m, n, colptr, rowval, nzval):
        """Upload counts given in Julia's layout as they are: 1-based Int64
        colptr (n + 1) and rowval (nnz), Int32 nzval (nnz)."""
        for a, dt in ((colptr, np.int64), (rowval, np.int64), (nzval, np.int32)):
            if a.dtype != dt or not a.flags.c_contiguous:
                raise TypeError("load_counts_julia needs contiguous int64/int64/int32 arrays")
        check(self._L.sev_pipeline_load_counts(self._h, m, n, _ptr(colptr), _ptr(rowval),
                                               _ptr(nzval)))
        self.cells = None

    def get_counts_julia(self):
        """The device counts in Julia's layout: (m, n, colptr, rowval, nzval),
        1-based Int64 structure and Int32 values."""
        m, n, nnz = self.counts_info()
        colptr = np.zeros(n + 1, np.int64)
        rowval = np.zeros(nnz, np.int64)
        vals = np.zeros(nnz, np.int32)
        check(self._L.sev_pipeline_get_counts(self._h, _ptr(colptr), _ptr(rowval), _ptr(vals)))
        return m, n, colptr, rowval, vals

    def synth_counts(self, n_cells, tables, seed=None):
        seed = tables.seed if seed is None else seed
        check(self._L.sev_pipeline_synth_counts(self._h, n_cells, tables.n_genes, seed, tables.r,
                                                _ptr(tables.p0), _ptr(tables.q),
                                                tables.boost_frac))
        self.cells = None

    # -- batches (samples stacked by rows) -----------------------------------
    def synth_batch(self, cells, tables, seeds=None):
        """Generate one synthetic sample per entry of ``cells`` (with the
        matching ``tables[s]`` and seed) and stack them by rows."""
        cells = np.ascontiguousarray(cells, np.int64)
        S = len(cells)
        if len(tables) != S:
            raise ValueError("one table per sample")
        n = tables[0].n_genes
        if any(t.n_genes != n for t in tables):
            raise ValueError("samples must share the genes")
        seeds = np.ascontiguousarray([t.seed for t in tables] if seeds is None else seeds,
                                     np.uint64)
        r = np.ascontiguousarray([t.r for t in tables], np.float64)
        p0 = np.ascontiguousarray(np.stack([t.p0 for t in tables]), np.float64)
        q = np.ascontiguousarray(np.stack([t.q for t in tables]), np.float64)
        bf = np.ascontiguousarray([t.boost_frac for t in tables], np.float64)
        check(self._L.sev_pipeline_synth_batch(self._h, S, _ptr(cells), n, _ptr(seeds), _ptr(r),
                                               _ptr(p0), _ptr(q), _ptr(bf)))
        self.cells = cells

    def set_batch(self, cells):
        """Declare the loaded counts to be samples of ``cells`` rows each,
        stacked in order (None clears it)."""
        if cells is None:
            check(self._L.sev_pipeline_set_batch(self._h, 0, None))
            self.cells = None
            return
        cells = np.ascontiguousarray(cells, np.int64)
        check(self._L.sev_pipeline_set_batch(self._h, len(cells), _ptr(cells)))
        self.cells = cells

    def select_scale_batch(self, features, scale_max=10.0):
        """features: S x F 0-based gene indices (row s for sample s)."""
        f = np.ascontiguousarray(np.asarray(features, np.int64) + 1)
        if f.ndim != 2 or f.shape[0] != len(self.cells):
            raise ValueError("features must be S x F")
        check(self._L.sev_pipeline_select_scale_batch(self._h, f.shape[1], _ptr(f),
                                                      float(scale_max)))
        self.nfeat = f.shape[1]

    def pca_batch(self, npcs=30, init=None, tol=1e-5, maxit=1000, graph=True, rng=None, work=0):
        """One IRLBA per sample (BF16 operator values, FP32 left bases, FP64
        arithmetic), in lockstep; returns (iterations per sample, total products).  work:
        Lanczos vectors per sample (0: npcs + 7, the reference's choice; at
        most 64, capped by the smallest sample)."""
        S = len(self.cells)
        if init is None:
            rng = np.random.default_rng(0) if rng is None else rng
            init = rng.standard_normal((S, self.nfeat))
        init = np.ascontiguousarray(init, np.float64)
        if init.shape != (S, self.nfeat):
            raise ValueError("init must be S x F")
        it = np.zeros(S, np.int64)
        nm = C.c_int64()
        check(self._L.sev_pipeline_pca_batch(self._h, npcs, int(work), maxit, float(tol), _ptr(init),
                                             int(bool(graph)), _ptr(it), C.byref(nm)))
        self.npcs = min(npcs, int(self.cells.min()), self.nfeat)
        return it, nm.value

    def get_pca_batch(self):
        """s (S x d), U (M x d), V (S x F x d), Z (M x d) for the batch."""
        S = len(self.cells)
        M = int(self.cells.sum())
        d = self.npcs
        s = np.zeros((S, d))
        U = np.zeros((M, d), order="F")
        V = np.zeros((S, d, self.nfeat))
        Z = np.zeros((M, d), np.float32, order="F")
        check(self._L.sev_pipeline_get_pca(self._h, _ptr(s), _ptr(U), _ptr(V), _ptr(Z)))
        return s, U, np.ascontiguousarray(V.transpose(0, 2, 1)), Z

    def counts_info(self):
        m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(self._L.sev_pipeline_counts_info(self._h, C.byref(m), C.byref(n), C.byref(nnz)))
        return m.value, n.value, nnz.value

    def get_counts(self):
        m, n, nnz = self.counts_info()
        colptr = np.zeros(n + 1, np.int64)
        rowval = np.zeros(nnz, np.int64)
        vals = np.zeros(nnz, np.int32)
        check(self._L.sev_pipeline_get_counts(self._h, _ptr(colptr), _ptr(rowval), _ptr(vals)))
        return sp.csc_matrix((vals, rowval - 1, colptr - 1), shape=(m, n))

    def find_variable_features(self, nfeatures=2000, method="vst", loess_span=0.5, num_bins=20,
                               binning_method="width", alpha_thresh=0.1, return_metric=False):
        """find_variable_features on the loaded counts (src/variablefeatures.jl:
        128-161): 0-based genes, best first; methods as the one-shot
        find_variable_features."""
        if isinstance(method, str):
            method = method.lstrip(":")
        if isinstance(binning_method, str):
            binning_method = binning_method.lstrip(":")
        codes = {"dispersion": _lib.SEV_HVG_DISPERSION, "meanvarplot": _lib.SEV_HVG_MEANVARPLOT,
                 "saunders": _lib.SEV_HVG_SAUNDERS}
        if method != "vst" and method not in codes:
            raise ValueError(f"unknown selection method: {method}")
        if binning_method != "width":
            raise ValueError("only width binning is provided")
        _, n, _ = self.counts_info()
        sel = np.zeros(nfeatures, np.int64)
        metric = np.zeros(n)
        if method == "vst":
            check(self._L.sev_pipeline_find_variable_features(self._h, nfeatures,
                                                              float(loess_span), _ptr(sel),
                                                              _ptr(metric)))
        else:
            check(self._L.sev_pipeline_find_variable_features_method(
                self._h, codes[method], nfeatures, int(num_bins), _lib.SEV_BIN_WIDTH,
                float(alpha_thresh), _ptr(sel), _ptr(metric)))
        sel -= 1
        return (sel, metric) if return_metric else sel

    # -- stages ------------------------------------------------------------
    def normalize(self, method="lognormalize", scale_factor=1e4):
        code = {"lognormalize": _lib.SEV_LOGNORMALIZE,
                "relativecounts": _lib.SEV_RELATIVECOUNTS}[method]
        check(self._L.sev_pipeline_normalize(self._h, code, float(scale_factor)))

    def get_normalized(self):
        _, _, nnz = self.counts_info()
        out = np.zeros(nnz, np.float32)
        check(self._L.sev_pipeline_get_normalized(self._h, _ptr(out)))
        return out

    def select_scale(self, features, scale_max=10.0):
        """features: 0-based gene indices in the order to keep."""
        f = np.ascontiguousarray(np.asarray(features, np.int64) + 1)
        check(self._L.sev_pipeline_select_scale(self._h, len(f), _ptr(f), float(scale_max)))
        self.nfeat = len(f)

    def get_scaled(self):
        m, _, _ = self.counts_info()
        nnz = C.c_int64()
        check(self._L.sev_pipeline_scaled_nnz(self._h, C.byref(nnz)))
        colptr = np.zeros(self.nfeat + 1, np.int64)
        rowval = np.zeros(nnz.value, np.int64)
        vals = np.zeros(nnz.value, np.float32)
        S = len(self.cells) if self.cells is not None else 0
        mu = np.zeros((S, self.nfeat) if S else self.nfeat, np.float32)
        check(self._L.sev_pipeline_get_scaled(self._h, _ptr(colptr), _ptr(rowval), _ptr(vals),
                                              _ptr(mu)))
        return sp.csc_matrix((vals, rowval - 1, colptr - 1), shape=(m, self.nfeat)), mu

    def pca(self, npcs=50, init=None, tol=1e-5, maxit=1000, precision="fp32", rng=None, work=0,
            algorithm="irlba"):
        """work: Lanczos vectors (0: npcs + 7, the reference's choice).
        algorithm follows _pca's rule (src/embedding.jl:46-53; see
        api.pca_algorithm): more than half the possible components take the
        dense route."""
        from .api import pca_algorithm
        m, _, _ = self.counts_info()
        _, dense_work = pca_algorithm(m, self.nfeat, min(npcs, m, self.nfeat), algorithm)
        if dense_work is not None:
            work = dense_work
        if init is None:
            rng = np.random.default_rng(0) if rng is None else rng
            init = rng.standard_normal(self.nfeat)
        init = np.ascontiguousarray(init, np.float64)
        it, nm = C.c_int64(), C.c_int64()
        check(self._L.sev_pipeline_pca(self._h, npcs, int(work), maxit, float(tol), _ptr(init),
                                       PRECISION[precision], C.byref(it), C.byref(nm)))
        self.npcs = npcs
        return it.value, nm.value

    def get_pca(self):
        m, _, _ = self.counts_info()
        s = np.zeros(self.npcs)
        U = np.zeros((m, self.npcs), order="F")
        V = np.zeros((self.nfeat, self.npcs), order="F")
        Z = np.zeros((m, self.npcs), np.float32, order="F")
        check(self._L.sev_pipeline_get_pca(self._h, _ptr(s), _ptr(U), _ptr(V), _ptr(Z)))
        return s, U, V, Z

    def set_embedding(self, Z):
        Z = np.asfortranarray(Z, np.float32)
        check(self._L.sev_pipeline_set_embedding(self._h, Z.shape[0], Z.shape[1], _ptr(Z),
                                                 Z.shape[0]))
        self.npcs = Z.shape[1]
        self._m = Z.shape[0]

    def neighbours(self, k=20, include_self=True, ndims=0, precision="fp32", metric="euclidean"):
        """Exact kNN of the embedding (metric "euclidean" or "cosine")."""
        fns = {"euclidean": self._L.sev_pipeline_neighbours,
               "cosine": self._L.sev_pipeline_neighbours_cosine}
        if metric not in fns:
            raise ValueError(f"unknown metric: {metric}")
        nfb = C.c_int64()
        check(fns[metric](self._h, k, int(include_self), ndims, PRECISION[precision],
                          C.byref(nfb)))
        self.k = k
        return nfb.value

    def knn_passes(self):
        """FP8 form: rows left without a certificate after the FP8, BF16 and
        FP32 candidate passes of the last search."""
        out = np.zeros(3, np.int64)
        check(self._L.sev_pipeline_knn_passes(self._h, _ptr(out)))
        return out

    def get_neighbours(self, m):
        idx = np.zeros((m, self.k), np.int32, order="F")
        dist = np.zeros((m, self.k), np.float32, order="F")
        check(self._L.sev_pipeline_get_neighbours(self._h, _ptr(idx), _ptr(dist)))
        return idx, dist

    def snn(self, prune=1.0 / 15, dtype=np.float32):
        nnz = C.c_int64()
        code = _lib.SEV_FLOAT32 if np.dtype(dtype) == np.float32 else _lib.SEV_FLOAT64
        check(self._L.sev_pipeline_snn(self._h, float(prune), code, C.byref(nnz)))
        self._snn = (nnz.value, np.dtype(dtype))
        return nnz.value

    def cluster(self, m, resolution=0.8, nrandomstarts=10, niterations=10,
                group_singletons=True, seed=0, return_modularity=False):
        """cluster(SNN) (src/clustering.jl:59-104) on the resident SNN graph
        of the m cells: 1-based cluster ids, as the one-shot cluster."""
        labels = np.zeros(m, np.int64)
        q = C.c_double(0.0)
        nc = C.c_int64(0)
        check(self._L.sev_pipeline_cluster(self._h, float(resolution), int(nrandomstarts),
                                           int(niterations), int(seed) & (2**64 - 1),
                                           1 if group_singletons else 0, _ptr(labels),
                                           C.byref(q), C.byref(nc)))
        return (labels, q.value) if return_modularity else labels

    def get_snn(self, m):
        nnz, dt = self._snn
        colptr = np.zeros(m + 1, np.int64)
        rowval = np.zeros(nnz, np.int64)
        vals = np.zeros(nnz, dt)
        check(self._L.sev_pipeline_get_snn(self._h, _ptr(colptr), _ptr(rowval), _ptr(vals)))
        return sp.csc_matrix((vals, rowval - 1, colptr - 1), shape=(m, m))
