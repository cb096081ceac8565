"""Severo.jl's embedding-and-graph API over the MI355X library.

The functions mirror the Julia API of the reference (same names, argument
meaning, defaults and error behaviour) with scipy CSC matrices standing in for
``SparseMatrixCSC{T,Int64}`` (cells x genes, one CSC column per gene) and
numpy arrays for dense matrices.  Every computation runs in libsevero_hip.so
on the GPU; nothing here computes results on the CPU beyond re-wrapping them.

Reference files: src/normalize.jl, src/scaling.jl, src/irlba.jl,
src/embedding.jl, src/neighbours.jl.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import check


def _csc_julia(X: sp.spmatrix):
    """1-based Int64 colptr/rowval of a scipy matrix (Julia's layout)."""
    X = sp.csc_matrix(X)
    X.sort_indices()
    colptr = np.ascontiguousarray(X.indptr, np.int64) + 1
    rowval = np.ascontiguousarray(X.indices, np.int64) + 1
    return X, colptr, rowval


def _dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return _lib.SEV_FLOAT32
    if dt == np.float64:
        return _lib.SEV_FLOAT64
    raise TypeError(f"unsupported dtype {dtype}")


def _ptr(a):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# normalize_cells  (src/normalize.jl:40-55, :76-79)
# ---------------------------------------------------------------------------
def normalize_cells(X, method="lognormalize", scale_factor=1.0, dtype=np.float64):
    """Divide each cell's counts by its total, multiply by ``scale_factor``
    and (``:lognormalize``) apply log1p.  Returns a CSC matrix of ``dtype``
    with the structure of ``X``."""
    if isinstance(method, str):
        method = method.lstrip(":")
    codes = {"lognormalize": _lib.SEV_LOGNORMALIZE, "relativecounts": _lib.SEV_RELATIVECOUNTS}
    if method not in codes:
        raise ValueError(f"unknown normalization method: {method}")
    X, colptr, rowval = _csc_julia(X)
    if not np.issubdtype(X.dtype, np.integer):
        raise TypeError("normalize_cells expects an integer count matrix")
    counts = np.ascontiguousarray(X.data, np.int64)
    out = np.empty(X.nnz, np.dtype(dtype))
    m, n = X.shape
    L = _lib.load()
    check(L.sev_normalize_cells(m, n, _ptr(colptr), _ptr(rowval), _ptr(counts), _lib.SEV_INT64,
                                codes[method], float(scale_factor), _dtype_code(dtype), _ptr(out)))
    return sp.csc_matrix((out, X.indices.copy(), X.indptr.copy()), shape=X.shape)


# ---------------------------------------------------------------------------
# find_variable_features  (src/variablefeatures.jl:128-161)
# ---------------------------------------------------------------------------
def find_variable_features(X, nfeatures=2000, method="vst", loess_span=0.5, num_bins=20,
                           binning_method="width", alpha_thresh=0.1, return_metric=False):
    """The ``nfeatures`` genes (0-based, best first) with the largest
    selection metric (src/variablefeatures.jl:128-161); ties go to the lower
    gene index.

    ``:vst`` (default): clipped standardized variance after a loess fit of
    log10 sd on log10 mean (the exact local fit at every gene; Loess.jl
    interpolates between kd-tree vertices).  ``:dispersion``,
    ``:meanvarplot`` (``num_bins`` bins of equal width) and ``:saunders``
    (``alpha_thresh``, 0.1 as in the reference's keyword default) work on
    the relative counts row_norm(counts, 1.0)."""
    if isinstance(method, str):
        method = method.lstrip(":")
    if isinstance(binning_method, str):
        binning_method = binning_method.lstrip(":")
    codes = {"dispersion": _lib.SEV_HVG_DISPERSION, "meanvarplot": _lib.SEV_HVG_MEANVARPLOT,
             "saunders": _lib.SEV_HVG_SAUNDERS}
    if method != "vst" and method not in codes:
        raise ValueError(f"unknown selection method: {method}")
    if binning_method != "width":
        raise ValueError("only width binning is provided: frequency binning puts the smallest "
                         "mean in bin 0, outside the reference's bin statistics")
    X, colptr, rowval = _csc_julia(X)
    if not np.issubdtype(X.dtype, np.integer):
        raise TypeError("find_variable_features expects an integer count matrix")
    counts = np.ascontiguousarray(X.data, np.int32)
    m, n = X.shape
    sel = np.zeros(nfeatures, np.int64)
    metric = np.zeros(n)
    L = _lib.load()
    if method == "vst":
        check(L.sev_find_variable_features(m, n, _ptr(colptr), _ptr(rowval), _ptr(counts),
                                           nfeatures, float(loess_span), _ptr(sel),
                                           _ptr(metric)))
    else:
        check(L.sev_find_variable_features_method(m, n, _ptr(colptr), _ptr(rowval),
                                                  _ptr(counts), codes[method], nfeatures,
                                                  int(num_bins), _lib.SEV_BIN_WIDTH,
                                                  float(alpha_thresh), _ptr(sel), _ptr(metric)))
    sel -= 1
    return (sel, metric) if return_metric else sel


# ---------------------------------------------------------------------------
# CenteredMatrix / scale_features  (src/scaling.jl:219-357)
# ---------------------------------------------------------------------------
class CenteredMatrix:
    """Lazy ``A - 1 mu'`` (src/scaling.jl:219-233): never densified."""

    def __init__(self, A, mu):
        A = sp.csc_matrix(A)
        mu = np.asarray(mu)
        if A.shape[1] != mu.shape[0]:
            raise ValueError("length(mu) must equal size(A, 2)")
        self.A = A
        self.mu = mu

    @property
    def shape(self):
        return self.A.shape

    @property
    def dtype(self):
        return np.promote_types(self.A.dtype, self.mu.dtype)

    def _julia(self, dt=None):
        A, colptr, rowval = _csc_julia(self.A)
        dt = np.dtype(self.dtype if dt is None else dt)
        return (A.shape, colptr, rowval, np.ascontiguousarray(A.data, dt),
                np.ascontiguousarray(self.mu, dt), dt)

    def mul(self, y, x, alpha=1.0, beta=0.0, trans=False):
        """mul!(y, C, x, alpha, beta) (src/scaling.jl:245-250) or, with
        ``trans``, mul!(y, C', x, alpha, beta) (:252-257).  FP64 vectors, or
        FP64 matrices (the matrix forms of :259-271, with the adjoint form's
        centering subtracted as in the vector form)."""
        (m, n), colptr, rowval, vals, mu, dt = self._julia()
        x = np.asarray(x, np.float64)
        if x.ndim == 2:
            return self._mul_mat(y, x, alpha, beta, trans)
        x = np.ascontiguousarray(x)
        y = np.ascontiguousarray(y, np.float64)
        if x.shape[0] != (m if trans else n) or y.shape[0] != (n if trans else m):
            raise ValueError("dimension mismatch")
        check(_lib.load().sev_centered_mul(m, n, _ptr(colptr), _ptr(rowval), _ptr(vals),
                                           _ptr(mu), _dtype_code(dt), b"T" if trans else b"N",
                                           _ptr(x), _ptr(y), float(alpha), float(beta)))
        return y

    def _mul_mat(self, Y, X, alpha, beta, trans):
        (m, n), colptr, rowval, vals, mu, dt = self._julia()
        X = np.asfortranarray(X, np.float64)
        Y = np.asfortranarray(Y, np.float64)
        p = X.shape[1]
        if X.shape[0] != (m if trans else n) or Y.shape != ((n if trans else m), p):
            raise ValueError("dimension mismatch")
        check(_lib.load().sev_centered_mul_mat(
            m, n, _ptr(colptr), _ptr(rowval), _ptr(vals), _ptr(mu), _dtype_code(dt),
            b"T" if trans else b"N", p, _ptr(X), max(X.shape[0], 1), _ptr(Y), max(Y.shape[0], 1),
            float(alpha), float(beta)))
        return Y

    def gram(self, other=None, C=None, alpha=1.0, beta=0.0):
        """mul!(C, Sl', Sr, alpha, beta) (src/scaling.jl:273-294) with Sl =
        self and Sr = other (self when None): the FP64 (nl x nr) matrix
        Sl' Sr, never densifying either operand.  Both are read in the dtype
        promoted over the two."""
        Sr = self if other is None else other
        if Sr.shape[0] != self.shape[0]:
            raise ValueError("dimension mismatch")
        dt = np.promote_types(self.dtype, Sr.dtype)
        (m, nl), cpl, rvl, vl, mul, _ = self._julia(dt)
        nr = Sr.shape[1]
        if C is None:
            C = np.zeros((nl, nr), order="F")
        C = np.asfortranarray(C, np.float64)
        if C.shape != (nl, nr):
            raise ValueError("dimension mismatch")
        if other is None:
            right = (None, None, None, None)
        else:
            _, cpr, rvr, vr, mur, _ = Sr._julia(dt)
            right = (cpr, rvr, vr, mur)
        check(_lib.load().sev_centered_gram(
            m, nl, _ptr(cpl), _ptr(rvl), _ptr(vl), _ptr(mul), nr, *(_ptr(a) for a in right),
            _dtype_code(dt), _ptr(C), max(nl, 1), float(alpha), float(beta)))
        return C

    def __matmul__(self, x):
        x = np.asarray(x)
        if x.ndim == 2:
            return self.mul(np.zeros((self.shape[0], x.shape[1]), order="F"), x)
        return self.mul(np.zeros(self.shape[0]), x)

    def rmatvec(self, x):
        return self.mul(np.zeros(self.shape[1]), x, trans=True)

    def to_dense(self):
        """convert(Matrix, C) (src/scaling.jl:298-304)."""
        return self.A.toarray() - self.mu[None, :]


def scale_features(X, scale_max=np.inf, dtype=None, features=None):
    """Standardise each feature (column) with the sequential Welford
    mean/variance of src/scaling.jl:18-34, divide by sigma and clip at
    ``scale_max + mu/sigma``; returns the lazy CenteredMatrix(B, mu/sigma).
    ``features`` are 0-based column indices (Julia's are 1-based)."""
    X = sp.csc_matrix(X)
    if features is not None:
        X = X[:, np.asarray(features)]
    if dtype is None:
        dtype = X.dtype if np.issubdtype(X.dtype, np.floating) else np.float64
    X, colptr, rowval = _csc_julia(X)
    vals = np.ascontiguousarray(X.data, np.dtype(dtype))
    out = np.empty_like(vals)
    mu = np.empty(X.shape[1], np.dtype(dtype))
    m, n = X.shape
    check(_lib.load().sev_scale_features(m, n, _ptr(colptr), _ptr(rowval), _ptr(vals),
                                         _dtype_code(dtype), float(scale_max), _ptr(out),
                                         _ptr(mu)))
    B = sp.csc_matrix((out, X.indices.copy(), X.indptr.copy()), shape=X.shape)
    return CenteredMatrix(B, mu)


# ---------------------------------------------------------------------------
# irlba / pca / embedding  (src/irlba.jl:47-99, src/embedding.jl:46-94)
# ---------------------------------------------------------------------------
@dataclass
class SVD:
    U: np.ndarray
    S: np.ndarray
    V: np.ndarray
    iters: int = 0
    nmult: int = 0


def _host_operator(A):
    """((m, n), A x, A' y) of an operator applied on the host."""
    if isinstance(A, CenteredMatrix):
        B = sp.csc_matrix(A.A, dtype=np.float64)
        mu = np.asarray(A.mu, np.float64)
        return B.shape, (lambda x: B @ x - mu @ x), (lambda y: B.T @ y - mu * y.sum())
    if hasattr(A, "matvec") and hasattr(A, "rmatvec"):
        return A.shape, A.matvec, A.rmatvec
    return A.shape, (lambda x: A @ x), (lambda y: A.T @ y)


def irlba_operator(A, nu: int, S: SVD | None = None, tol=1e-5, maxit=1000, init=None,
                   rng=None, work=None) -> SVD:
    """irlba through libcell's callback interface (src/irlba.jl:22-99): the
    library runs the Lanczos process on the GPU and calls back for every
    product, which the operator computes wherever it lives (a LinearOperator,
    numpy or scipy matrix, or a CenteredMatrix applied as B x - (mu'x) 1 and
    B'y - mu sum(y)).  ``S`` (a previous SVD with d <= nu triplets) warm-starts
    the bases, as ``irlba(A, nu, S)`` does (src/irlba.jl:87-99)."""
    (m, n), matvec, rmatvec = _host_operator(A)
    if init is None:
        rng = np.random.default_rng() if rng is None else rng
        init = rng.standard_normal(n)
    init = np.ascontiguousarray(init, np.float64)
    s = np.zeros(nu)
    U = np.zeros((m, nu), order="F")
    V = np.zeros((n, nu), order="F")
    restart = 0
    if S is not None:
        restart = len(S.S)
        s[:restart] = S.S
        U[:, :restart] = S.U[:, :restart]
        V[:, :restart] = S.V[:, :restart]
    work = nu + 7 if work is None else work
    work = min(work, m, n)
    errors = []

    def _mm(yp, trans, xp, tp, data):
        try:
            if trans == b"T":
                np.ctypeslib.as_array(yp, (n,))[:] = rmatvec(np.ctypeslib.as_array(xp, (m,)))
            else:
                np.ctypeslib.as_array(yp, (m,))[:] = matvec(np.ctypeslib.as_array(xp, (n,)))
        except Exception as e:  # surfaced after the call returns
            errors.append(e)

    cb = _lib.MATMUL_FN(_mm)
    rc = _lib.load().irlba(m, n, nu, work, maxit, restart, float(tol), _ptr(init), _ptr(s),
                           _ptr(U), _ptr(V), None, C.cast(cb, C.c_void_p), None)
    if errors:
        raise errors[0]
    check(rc)
    return SVD(U, s, V)


def irlba(A, nu: int, tol=1e-5, maxit=1000, init=None, rng=None, precision="fp32",
          work=None) -> SVD:
    """Truncated SVD by augmented restarted Lanczos bidiagonalisation on the
    GPU (src/irlba.jl:47-85).  ``A`` is a CenteredMatrix or any sparse/dense
    matrix, held on the device for the whole run; work defaults to nu + 7
    (src/irlba.jl:50); ``init`` defaults to a standard-normal vector
    (src/irlba.jl:63).  Raises ConvergenceError like the reference's
    ``error("convergence failed")``."""
    mu = None
    if isinstance(A, CenteredMatrix):
        mu, A = A.mu, A.A
    Am = sp.csc_matrix(A)
    dt = np.float32 if (Am.dtype == np.float32 and (mu is None or mu.dtype == np.float32)) \
        else np.float64
    Am, colptr, rowval = _csc_julia(Am)
    m, n = Am.shape
    if init is None:
        rng = np.random.default_rng() if rng is None else rng
        init = rng.standard_normal(n)
    init = np.ascontiguousarray(init, np.float64)
    vals = np.ascontiguousarray(Am.data, dt)
    mu = None if mu is None else np.ascontiguousarray(mu, dt)
    s = np.zeros(nu)
    U = np.zeros((m, nu), order="F")
    V = np.zeros((n, nu), order="F")
    it = C.c_int64(0)
    nm = C.c_int64(0)
    prec = {"fp32": _lib.SEV_PREC_FP32, "bf16": _lib.SEV_PREC_BF16}[precision]
    work = nu + 7 if work is None else work
    check(_lib.load().sev_irlba_centered_csc(m, n, _ptr(colptr), _ptr(rowval), _ptr(vals),
                                             _ptr(mu), _dtype_code(dt), nu, work, maxit,
                                             float(tol), _ptr(init), _ptr(s), _ptr(U), _ptr(V),
                                             prec, C.byref(it), C.byref(nm)))
    return SVD(U, s, V, it.value, nm.value)


@dataclass
class LinearEmbedding:
    """src/embedding.jl:20-25."""
    coordinates: np.ndarray  # C x npcs FP64 (U * diag(S))
    stdev: np.ndarray
    basis: np.ndarray  # F x npcs (loadings)


DENSE_MAX = 128  # largest min(m, n) of the dense route (a full Lanczos basis)


def pca_algorithm(m, n, npcs, algorithm="arpack"):
    """The algorithm _pca runs (src/embedding.jl:46-53): more than half the
    possible components switch :arpack and :irlba to the dense SVD, with the
    reference's warning.  Returns (algorithm, work): :arpack and :irlba both
    run the device IRLBA (svds and irlba compute the same truncated SVD to
    tolerance), work None meaning npcs + 7; :svd runs it with a full basis
    of min(m, n) Lanczos vectors, which is the Golub-Kahan bidiagonalisation
    followed by the SVD of the bidiagonal -- a dense SVD to working
    precision -- for min(m, n) <= 128.  :tssvd is not provided."""
    import warnings
    if isinstance(algorithm, str):
        algorithm = algorithm.lstrip(":")
    if algorithm not in ("arpack", "irlba", "svd", "tssvd"):
        raise ValueError(f"unknown algorithm: {algorithm}")
    if npcs > 0.5 * min(m, n) and algorithm in ("arpack", "irlba"):
        warnings.warn("Computing too large a percentage of principal components, using "
                      "standard svd instead")
        algorithm = "svd"
    if algorithm == "tssvd":
        raise NotImplementedError("algorithm=:tssvd is not provided on the device")
    if algorithm == "svd":
        if min(m, n) > DENSE_MAX:
            raise NotImplementedError(f"the dense SVD route needs min(m, n) <= {DENSE_MAX}")
        return "svd", min(m, n)
    return algorithm, None


def pca(X: CenteredMatrix, npcs: int = 50, tol=1e-5, maxit=1000, init=None, rng=None,
        precision="fp32", algorithm="arpack", work=None) -> LinearEmbedding:
    """_pca (src/embedding.jl:46-76): the truncated SVD of the algorithm
    pca_algorithm picks, Z = U diag(S), stdev = S / sqrt(max(1, m - 1)),
    basis = V."""
    m, n = X.shape
    npcs = min(min(m, n), npcs)
    algorithm, dense_work = pca_algorithm(m, n, npcs, algorithm)
    S = irlba(X, npcs, tol=tol, maxit=maxit, init=init, rng=rng, precision=precision,
              work=dense_work if dense_work is not None else work)
    Z = S.U[:, :npcs] * S.S[:npcs][None, :]
    stdev = S.S[:npcs] / np.sqrt(max(1, m - 1))
    return LinearEmbedding(Z, stdev, S.V[:, :npcs])


def embedding(X, ncomponents: int = 50, method="pca", **kw) -> LinearEmbedding:
    """src/embedding.jl:202-212 (algorithm= passes on to _pca)."""
    if isinstance(method, str):
        method = method.lstrip(":")
    if method != "pca":
        raise ValueError(f"unknown reduction method: {method}")
    return pca(X, ncomponents, **kw)


# ---------------------------------------------------------------------------
# ann / nearest_neighbours / jaccard_index / shared_nearest_neighbours
# (src/neighbours.jl:19-108, 127-349)
# ---------------------------------------------------------------------------
def _coords(X, dims):
    if isinstance(X, LinearEmbedding):
        X = X.coordinates
    X = np.asarray(X)
    if dims is not None:
        X = X[:, dims]
    return X


def ann(X, k: int, include_self=True, ntables=None, precision="fp32", return_fallback=False,
        metric="euclidean"):
    """Exact kNN (the search behind ann!/FindNeighboursEuclidean and
    FindNeighboursCosine, src/neighbours.jl:19-72).  Returns (nn_index n x k
    1-based Int32, distances n x k), column-major like the reference.
    ``precision`` selects how candidates are proposed ("fp32" / "bf16"
    matrix cores with an exact re-rank, or "exact" exhaustive); the result is
    the same.  ``metric="cosine"`` (Distances.CosineDist) searches the rows
    scaled to unit length and reports D / 2 of the unit rows (1 - cos in
    exact arithmetic)."""
    X = np.asfortranarray(X)
    if X.dtype not in (np.float32, np.float64):
        X = X.astype(np.float64, order="F")
    n, d = X.shape
    idx = np.zeros((n, k), np.int32, order="F")
    dist = np.zeros((n, k), X.dtype, order="F")
    nfb = C.c_int64(0)
    prec = {"fp32": _lib.SEV_PREC_FP32, "bf16": _lib.SEV_PREC_BF16,
            "exact": _lib.SEV_PREC_EXACT, "fp8": _lib.SEV_PREC_FP8}[precision]
    metrics = {"euclidean": "sev_knn", "cosine": "sev_knn_cosine"}
    if metric not in metrics:
        raise ValueError(f"unknown metric: {metric}")
    fn = getattr(_lib.load(), metrics[metric])
    check(fn(_ptr(X), n, d, n, _dtype_code(X.dtype), k, int(include_self), prec, _ptr(idx),
             _ptr(dist), C.byref(nfb)))
    if return_fallback:
        return idx, dist, nfb.value
    return idx, dist


def _nn_matrix(nn_index):
    """sparse(vec(nn_index'), repeat(1:n, inner=k), true) (src/neighbours.jl:79):
    column j holds N(j)."""
    n, k = nn_index.shape
    rows = np.asarray(nn_index, np.int64).ravel(order="C") - 1
    cols = np.repeat(np.arange(n), k)
    M = sp.csc_matrix((np.ones(n * k, bool), (rows, cols)), shape=(n, n))
    M.sort_indices()
    return M


def nearest_neighbours(X, k: int, dims=None, include_self=True, ntables=None):
    """k-NN indicator graph, column j holding N(j) (src/neighbours.jl:76-86)."""
    idx, _ = ann(_coords(X, dims), k, include_self=include_self)
    return _nn_matrix(idx)


def _snn_from_index(nn_index, k, prune, dtype):
    n = nn_index.shape[0]
    nn_index = np.asfortranarray(nn_index, np.int32)
    nnz = C.c_int64(0)
    L = _lib.load()
    dt = _dtype_code(dtype)
    check(L.sev_snn_jaccard(n, k, _ptr(nn_index), float(prune), dt, C.byref(nnz), None, None,
                            None))
    colptr = np.zeros(n + 1, np.int64)
    rowval = np.zeros(nnz.value, np.int64)
    vals = np.zeros(nnz.value, np.dtype(dtype))
    check(L.sev_snn_jaccard(n, k, _ptr(nn_index), float(prune), dt, C.byref(nnz), _ptr(colptr),
                            _ptr(rowval), _ptr(vals)))
    return sp.csc_matrix((vals, rowval - 1, colptr - 1), shape=(n, n))


def jaccard_index(nn, prune=1.0 / 15, dtype=np.float64):
    """Jaccard graph of a k-NN indicator matrix (src/neighbours.jl:96-108,
    :127-131); k is taken from the column counts, which must all be equal."""
    nn = sp.csc_matrix(nn)
    counts = np.diff(nn.indptr)
    if len(counts) == 0 or not np.all(counts == counts[0]):
        raise ValueError("jaccard_index: every column of nn must hold k neighbours")
    k = int(counts[0])
    n = nn.shape[1]
    nn.sort_indices()
    idx = (nn.indices.reshape(n, k) + 1).astype(np.int32)
    return _snn_from_index(np.asfortranarray(idx), k, prune, dtype)


def shared_nearest_neighbours(X, k: int, dims=None, include_self=True, ntables=None,
                              prune=1.0 / 15):
    """kNN then Jaccard graph with weights in eltype(X) (src/neighbours.jl:272-278)."""
    Xc = _coords(X, dims)
    idx, _ = ann(Xc, k, include_self=include_self)
    return _snn_from_index(idx, k, prune, Xc.dtype)


# ---------------------------------------------------------------------------
# cluster  (src/clustering.jl:59-104, src/modularity.jl:165-380)
# ---------------------------------------------------------------------------
def cluster(SNN, algorithm="louvain", resolution=0.8, nrandomstarts=10, niterations=10,
            group_singletons=True, seed=0, return_modularity=False):
    """Louvain clusters of a neighbour graph (square sparse, weights >= 0;
    self-links ignored), Severo's cluster(SNN): 1-based cluster ids in order
    of first occurrence.  Local moving runs in parallel sweeps on the device,
    so the labels are one outcome of the reference's randomised search;
    ``seed`` fixes it.  ``return_modularity`` adds the best start's
    modularity (before singleton grouping)."""
    if isinstance(algorithm, str):
        algorithm = algorithm.lstrip(":")
    if algorithm != "louvain":
        raise ValueError(f"unknown clustering algorithm: {algorithm}")
    if SNN.shape[0] != SNN.shape[1]:
        raise ValueError(f"matrix is not square: dimensions are {SNN.shape}")
    X, colptr, rowval = _csc_julia(SNN)
    dt = X.dtype if X.dtype in (np.float32, np.float64) else np.dtype(np.float64)
    vals = np.ascontiguousarray(X.data, dt)
    n = X.shape[0]
    labels = np.zeros(n, np.int64)
    q = C.c_double(0.0)
    nc = C.c_int64(0)
    check(_lib.load().sev_cluster_louvain(n, _ptr(colptr), _ptr(rowval), _ptr(vals),
                                          _dtype_code(dt), float(resolution), int(nrandomstarts),
                                          int(niterations), int(seed) & (2**64 - 1),
                                          1 if group_singletons else 0, _ptr(labels),
                                          C.byref(q), C.byref(nc)))
    return (labels, q.value) if return_modularity else labels


def copy_probe() -> float:
    """Achievable HBM copy rate of the device in GB/s (1 GiB read + 1 GiB
    written per launch, best of 5)."""
    r = C.c_double(0.0)
    check(_lib.load().sev_selftest(3, C.byref(r)))
    return r.value


# ---------------------------------------------------------------------------
# 10X input  (src/input.jl:26-115 readMM, :199-208 _read_10X, :634-646
# read_10X, :776-790 convert_counts, src/utils.jl make_unique!)
# ---------------------------------------------------------------------------
def read_mm(fname):
    """readMM: the MatrixMarket coordinate file (``.gz`` inflated) as a CSC
    matrix of its own orientation, built as Julia's ``sparse(I, J, V, rows,
    cols)`` builds it (duplicates summed, stored zeros kept).  Int64 values
    for the integer field, Float64 for real.  The text is parsed by the
    library's host threads."""
    import os
    path = os.fsencode(fname)
    L = _lib.load()
    rows, cols, entries, is_real = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
    check(L.sev_mtx_info(path, C.byref(rows), C.byref(cols), C.byref(entries),
                         C.byref(is_real)))
    E = entries.value
    I = np.empty(E, np.int64)
    J = np.empty(E, np.int64)
    V = np.empty(E, np.float64 if is_real.value else np.int64)
    check(L.sev_mtx_read(path, _ptr(I), _ptr(J), None if is_real.value else _ptr(V),
                         _ptr(V) if is_real.value else None))
    if E and (I.min() < 1 or I.max() > rows.value or J.min() < 1 or J.max() > cols.value):
        raise ValueError("matrix entry index out of range")
    X = sp.coo_matrix((V, (I - 1, J - 1)), shape=(rows.value, cols.value)).tocsc()
    X.sort_indices()
    return X


def make_unique(names, sep="."):
    """make_unique!: the second and later copies of a name get ``sep`` and a
    counter, skipping names already present."""
    n = len(names)
    seen = {}
    dup = [False] * n
    for i, x in enumerate(names):
        if x not in seen:
            seen[x] = 1
        else:
            dup[i] = True
    out = list(names)
    for i, x in enumerate(names):
        if dup[i]:
            cnt = seen[x]
            y = f"{x}{sep}{cnt}"
            while y in seen and cnt <= n:
                cnt += 1
                y = f"{x}{sep}{cnt}"
            out[i] = y
            seen[x] = cnt + 1
            seen[y] = 1
    return out


@dataclass
class NamedCounts:
    """A cells x genes count matrix with its barcodes and feature names
    (Severo's NamedCountMatrix).  ``X`` is None when the counts were loaded
    into a pipeline and left on the device."""
    X: sp.csc_matrix | None
    barcodes: list
    features: list


def _read_lines(fname):
    import gzip
    opener = gzip.open if str(fname).endswith(".gz") else open
    with opener(fname, "rt") as f:
        return [ln.rstrip("\r\n") for ln in f]


def read_10X(dirname, gene_column: int = 2, unique_features: bool = True, pipeline=None):
    """read_10X: matrix.mtx, genes.tsv and barcodes.tsv (or the v3 names
    matrix.mtx.gz, features.tsv.gz, barcodes.tsv.gz) of a 10X directory.
    The counts are parsed by host threads and transposed to cells x genes on
    the GPU; with ``pipeline`` they stay there (X is None), otherwise they are
    copied back.  Feature names take column ``gene_column`` with '_' replaced
    by '-' and, with ``unique_features``, made unique."""
    import os
    if not os.path.isdir(dirname):
        raise ValueError(f"Directory {dirname} does not exist")
    v3 = os.path.isfile(os.path.join(dirname, "features.tsv.gz"))
    feature_file = os.path.join(dirname, "features.tsv.gz" if v3 else "genes.tsv")
    barcodes_file = os.path.join(dirname, "barcodes.tsv.gz" if v3 else "barcodes.tsv")
    matrix_file = os.path.join(dirname, "matrix.mtx.gz" if v3 else "matrix.mtx")
    for f in (feature_file, barcodes_file, matrix_file):
        if not os.path.isfile(f):
            raise ValueError(f"Couldn't locate all components: {feature_file}, "
                             f"{barcodes_file}, {matrix_file}")
    barcodes = _read_lines(barcodes_file)
    features = [ln.split("\t")[gene_column - 1] for ln in _read_lines(feature_file)]
    features = [f.replace("_", "-") for f in features]
    if unique_features:
        features = make_unique(features)
    from .pipeline import Pipeline
    p = pipeline if pipeline is not None else Pipeline()
    try:
        m, n, _ = p.load_mtx(matrix_file)
        if len(barcodes) != m or len(features) != n:
            raise ValueError(f"matrix is {n} x {m} (genes x cells) but there are "
                             f"{len(features)} features and {len(barcodes)} barcodes")
        X = None if pipeline is not None else p.get_counts()
    finally:
        if pipeline is None:
            p.close()
    return NamedCounts(X, barcodes, features)
