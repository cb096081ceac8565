"""ctypes front end of the CPU oracle (``libsevero_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and bench.py's ``cpu_baseline`` leg, always as the checker.  The product
library never calls into it.

Every function takes and returns numpy arrays in the layouts Severo.jl uses
(cells x genes CSC, column-major dense matrices, 1-based neighbour indices)
except that CSC row indices are 0-based int32 here, as on the device.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64 = C.c_int64
_i32 = C.c_int32
_u64 = C.c_uint64
_dbl = C.c_double
_flt = C.c_float
_p = C.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsevero_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.so_u01.restype = _dbl
        L.so_u01.argtypes = [_u64, C.c_uint32, C.c_uint32]
        L.so_synth_nb_csc.restype = _i64
        L.so_synth_nb_csc.argtypes = [_i64, _i64, _u64, _dbl, _p, _p, _dbl, _p, _p, _p]
        L.so_normalize_f32.argtypes = [_i64, _i64, _p, _p, _p, _flt, C.c_int, _p]
        L.so_normalize_f64.argtypes = [_i64, _i64, _p, _p, _p, _dbl, C.c_int, _p]
        L.so_gather_cols.restype = _i64
        L.so_gather_cols.argtypes = [_i64, _p, _p, _p, _p, _p, _p, _p]
        L.so_scale_f32.argtypes = [_i64, _i64, _p, _p, _flt, _p, _p, _p]
        L.so_scale_f64.argtypes = [_i64, _i64, _p, _p, _dbl, _p, _p, _p]
        L.so_centered_mul.argtypes = [_i64, _i64, _p, _p, _p, C.c_int, _p, C.c_int, _p, _p,
                                      _dbl, _dbl]
        L.so_centered_tmul_mat.argtypes = [_i64, _i64, _p, _p, _p, _p, _p, _i64, _p, C.c_int]
        L.so_centered_gram.restype = C.c_int
        L.so_centered_gram.argtypes = [_i64, _i64, _p, _p, _p, _p, _i64, _p, _p, _p, _p, C.c_int,
                                       _p, _dbl, _dbl]
        L.so_svd_jacobi.argtypes = [_i64, _p, _p, _p, _p]
        L.so_irlba_csc.restype = C.c_int
        L.so_irlba_csc.argtypes = [_i64, _i64, _p, _p, _p, C.c_int, _p, _i64, _i64, _i64, _dbl,
                                   _dbl, _p, _p, _p, _p, _p, _p]
        L.so_irlba_dense.restype = C.c_int
        L.so_irlba_dense.argtypes = [_i64, _i64, _p, _p, _i64, _i64, _i64, _dbl, _dbl, _p, _p,
                                     _p, _p, _p, _p]
        for name in ("so_knn_f32", "so_knn_f64"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [_p, _i64, _i64, _i64, _i64, C.c_int, _p, _i64, _p, _p, C.c_int]
        L.so_snn_jaccard.restype = _i64
        L.so_snn_jaccard.argtypes = [_i64, _i64, _p, _dbl, C.c_int, _p, _p, _p]
        L.so_count_mean_std.argtypes = [_i64, _i64, _p, _p, _p, _p]
        L.so_rel_mean_var.argtypes = [_i64, _i64, _p, _p, _p, _p, _p]
        L.so_loess_fit.restype = C.c_int
        L.so_loess_fit.argtypes = [_i64, _p, _p, _dbl, _p]
        L.so_standardized_var.argtypes = [_i64, _i64, _p, _p, _p, _p, _p]
        L.so_louvain.restype = _i64
        L.so_louvain.argtypes = [_i64, _p, _p, _p, _dbl, _i64, _i64, _dbl, _u64, _p, _p]
        L.so_modularity.restype = _dbl
        L.so_modularity.argtypes = [_i64, _p, _p, _p, _p, _dbl]
        L.so_group_singletons.restype = _i64
        L.so_group_singletons.argtypes = [_i64, _p, _p, _p, _p]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data


def u01(seed, stream, index):
    return lib().so_u01(seed, stream, index)


def synth_csc(n_cells, tables, seed=None):
    """cells x genes CSC of NB counts: (colptr int64, rowval int32 0-based, counts int32)."""
    seed = tables.seed if seed is None else seed
    G = tables.n_genes
    colptr = np.zeros(G + 1, np.int64)
    L = lib()
    nnz = L.so_synth_nb_csc(n_cells, G, seed, tables.r, _ptr(tables.p0), _ptr(tables.q),
                            tables.boost_frac, _ptr(colptr), None, None)
    rowval = np.empty(nnz, np.int32)
    counts = np.empty(nnz, np.int32)
    L.so_synth_nb_csc(n_cells, G, seed, tables.r, _ptr(tables.p0), _ptr(tables.q),
                      tables.boost_frac, _ptr(colptr), _ptr(rowval), _ptr(counts))
    return colptr, rowval, counts


def normalize(m, colptr, rowval, counts, scale_factor=1e4, log=True, dtype=np.float32):
    n = len(colptr) - 1
    colptr = np.ascontiguousarray(colptr, np.int64)
    rowval = np.ascontiguousarray(rowval, np.int32)
    counts = np.ascontiguousarray(counts, np.int32)
    out = np.empty(len(rowval), dtype)
    if dtype == np.float32:
        lib().so_normalize_f32(m, n, _ptr(colptr), _ptr(rowval), _ptr(counts), scale_factor,
                               int(log), _ptr(out))
    else:
        lib().so_normalize_f64(m, n, _ptr(colptr), _ptr(rowval), _ptr(counts), scale_factor,
                               int(log), _ptr(out))
    return out


def gather(feats, colptr, rowval, vals):
    feats = np.ascontiguousarray(feats, np.int64)
    colptr = np.ascontiguousarray(colptr, np.int64)
    rowval = np.ascontiguousarray(rowval, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    out_ptr = np.zeros(len(feats) + 1, np.int64)
    L = lib()
    nnz = L.so_gather_cols(len(feats), _ptr(feats), _ptr(colptr), None, None, _ptr(out_ptr),
                           None, None)
    r = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float32)
    L.so_gather_cols(len(feats), _ptr(feats), _ptr(colptr), _ptr(rowval), _ptr(vals),
                     _ptr(out_ptr), _ptr(r), _ptr(v))
    return out_ptr, r, v


def scale(m, colptr, vals, scale_max=np.inf):
    """Returns (B values, mu/sigma, sigma) in the dtype of vals."""
    n = len(colptr) - 1
    colptr = np.ascontiguousarray(colptr, np.int64)
    vals = np.ascontiguousarray(vals)
    out = np.empty_like(vals)
    mu = np.empty(n, vals.dtype)
    sd = np.empty(n, vals.dtype)
    if vals.dtype == np.float32:
        lib().so_scale_f32(m, n, _ptr(colptr), _ptr(vals), scale_max, _ptr(out), _ptr(mu),
                           _ptr(sd))
    else:
        lib().so_scale_f64(m, n, _ptr(colptr), _ptr(vals), scale_max, _ptr(out), _ptr(mu),
                           _ptr(sd))
    return out, mu, sd


def centered_mul(m, colptr, rowval, vals, mu, x, trans=False, y=None, alpha=1.0, beta=0.0):
    n = len(colptr) - 1
    vals = np.ascontiguousarray(vals)
    f64 = int(vals.dtype == np.float64)
    if mu is not None:
        mu = np.ascontiguousarray(mu, vals.dtype)
    x = np.ascontiguousarray(x, np.float64)
    if y is None:
        y = np.zeros(n if trans else m, np.float64)
    else:
        y = np.array(y, np.float64, copy=True)
    colptr = np.ascontiguousarray(colptr, np.int64)
    rowval = np.ascontiguousarray(rowval, np.int32)
    lib().so_centered_mul(m, n, _ptr(colptr), _ptr(rowval), _ptr(vals), f64, _ptr(mu),
                          int(trans), _ptr(x), _ptr(y), alpha, beta)
    return y


def centered_mul_mat(m, colptr, rowval, vals, mu, X, trans=False, Y=None, alpha=1.0, beta=0.0):
    """mul!(Y, S, X, alpha, beta) / mul!(Y, S', X, alpha, beta) with a matrix
    X (src/scaling.jl:259-271): the vector form column by column (the
    reference's matrix form runs the same SparseArrays loop per column of X
    and subtracts the same centering term for S X; for S' X the vector
    form's -mu (1'x) is used, the reference's matrix form adds it)."""
    n = len(colptr) - 1
    X = np.asarray(X, np.float64)
    p = X.shape[1]
    out = np.zeros(((n if trans else m), p)) if Y is None else np.array(Y, np.float64)
    for c in range(p):
        out[:, c] = centered_mul(m, colptr, rowval, vals, mu, X[:, c], trans=trans, y=out[:, c],
                                 alpha=alpha, beta=beta)
    return out


def centered_gram(m, cpl, rvl, vl, mul, cpr=None, rvr=None, vr=None, mur=None, C=None,
                  alpha=1.0, beta=0.0):
    """mul!(C, Sl', Sr, alpha, beta) (src/scaling.jl:273-294); cpr None: Sr = Sl."""
    if cpr is None:
        cpr, rvr, vr, mur = cpl, rvl, vl, mul
    vl = np.ascontiguousarray(vl)
    f64 = int(vl.dtype == np.float64)
    vr = np.ascontiguousarray(vr, vl.dtype)
    mul = None if mul is None else np.ascontiguousarray(mul, vl.dtype)
    mur = None if mur is None else np.ascontiguousarray(mur, vl.dtype)
    nl, nr = len(cpl) - 1, len(cpr) - 1
    out = np.zeros((nl, nr), order="F") if C is None else np.array(C, np.float64, order="F")
    args = [np.ascontiguousarray(a, t) for a, t in ((cpl, np.int64), (rvl, np.int32),
                                                   (cpr, np.int64), (rvr, np.int32))]
    rc = lib().so_centered_gram(m, nl, _ptr(args[0]), _ptr(args[1]), _ptr(vl), _ptr(mul), nr,
                                _ptr(args[2]), _ptr(args[3]), _ptr(vr), _ptr(mur), f64,
                                _ptr(out), alpha, beta)
    if rc != 0:
        raise MemoryError("so_centered_gram")
    return out


def centered_tmul_mat(m, colptr, rowval, vals, mu, U, nthreads=None):
    """A'U for the centered matrix A = B - 1 mu' (Float32 B and mu), U m x k
    FP64; returns n x k FP64 (threaded over the columns of B)."""
    n = len(colptr) - 1
    U = np.ascontiguousarray(U, np.float64)
    k = U.shape[1]
    Y = np.zeros((n, k))
    colptr = np.ascontiguousarray(colptr, np.int64)
    rowval = np.ascontiguousarray(rowval, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    mu = None if mu is None else np.ascontiguousarray(mu, np.float32)
    lib().so_centered_tmul_mat(m, n, _ptr(colptr), _ptr(rowval), _ptr(vals), _ptr(mu), _ptr(U),
                               k, _ptr(Y), int(nthreads or os.cpu_count() or 1))
    return Y


def svd_jacobi(A):
    A = np.asfortranarray(A, np.float64)
    n = A.shape[0]
    U = np.zeros((n, n), order="F")
    V = np.zeros((n, n), order="F")
    s = np.zeros(n)
    lib().so_svd_jacobi(n, _ptr(A), _ptr(U), _ptr(s), _ptr(V))
    return U, s, V


def irlba_csc(m, colptr, rowval, vals, mu, nu, init, work=None, maxit=1000, tol=1e-5,
              svtol=None):
    n = len(colptr) - 1
    work = nu + 7 if work is None else work
    work = min(work, m, n)
    svtol = tol if svtol is None else svtol
    vals = np.ascontiguousarray(vals)
    f64 = int(vals.dtype == np.float64)
    if mu is not None:
        mu = np.ascontiguousarray(mu, vals.dtype)
    s = np.zeros(nu)
    U = np.zeros((m, nu), order="F")
    V = np.zeros((n, nu), order="F")
    it = np.zeros(1, np.int64)
    nm = np.zeros(1, np.int64)
    init = np.ascontiguousarray(init, np.float64)
    colptr = np.ascontiguousarray(colptr, np.int64)
    rowval = np.ascontiguousarray(rowval, np.int32)
    rc = lib().so_irlba_csc(m, n, _ptr(colptr), _ptr(rowval), _ptr(vals), f64, _ptr(mu), nu, work, maxit, tol, svtol, _ptr(init), _ptr(s), _ptr(U),
                            _ptr(V), _ptr(it), _ptr(nm))
    return rc, s, U, V, int(it[0]), int(nm[0])


def irlba_dense(A, nu, init, mu=None, work=None, maxit=1000, tol=1e-5, svtol=None):
    A = np.asfortranarray(A, np.float64)
    m, n = A.shape
    work = nu + 7 if work is None else work
    work = min(work, m, n)
    svtol = tol if svtol is None else svtol
    if mu is not None:
        mu = np.ascontiguousarray(mu, np.float64)
    s = np.zeros(nu)
    U = np.zeros((m, nu), order="F")
    V = np.zeros((n, nu), order="F")
    it = np.zeros(1, np.int64)
    nm = np.zeros(1, np.int64)
    init = np.ascontiguousarray(init, np.float64)
    rc = lib().so_irlba_dense(m, n, _ptr(A), _ptr(mu), nu, work, maxit, tol, svtol, _ptr(init),
                              _ptr(s), _ptr(U), _ptr(V), _ptr(it), _ptr(nm))
    return rc, s, U, V, int(it[0]), int(nm[0])


def knn(X, k, include_self=True, qidx=None, nthreads=None):
    """Exact kNN of the rows of X (n x d).  Returns (nn_index n x k 1-based
    int32, distances n x k), both column-major like Severo's ann()."""
    X = np.asfortranarray(X)
    n, d = X.shape
    nthreads = nthreads or os.cpu_count() or 1
    nrow = n if qidx is None else len(qidx)
    if qidx is not None:
        qidx = np.ascontiguousarray(qidx, np.int64)
    idx = np.zeros((nrow, k), np.int32, order="F")
    dist = np.zeros((nrow, k), X.dtype, order="F")
    f = lib().so_knn_f32 if X.dtype == np.float32 else lib().so_knn_f64
    rc = f(_ptr(X), n, d, n, k, int(include_self), _ptr(qidx), nrow, _ptr(idx), _ptr(dist),
           int(nthreads))
    if rc != 0:
        raise ValueError(f"oracle knn failed ({rc})")
    return idx, dist


def unit_rows(X):
    """Rows of X scaled to unit length in X's precision: sum of squares in
    column order (each square and sum rounded), IEEE sqrt and division; an
    all-zero row stays zero.  The cosine form of ann! (CosineDist,
    src/neighbours.jl:41-50, :63-72) is defined here on these rows."""
    X = np.asfortranarray(X)
    dt = X.dtype.type
    s = np.zeros(X.shape[0], X.dtype)
    for t in range(X.shape[1]):
        v = X[:, t]
        s = (s + v * v).astype(X.dtype)
    nrm = np.sqrt(s).astype(X.dtype)
    U = X.copy(order="F")
    nz = s > dt(0)
    U[nz, :] = (X[nz, :] / nrm[nz, None]).astype(X.dtype)
    return U


def half_pair_dist(U, nn_index):
    """D / 2 of the unit rows for each (i, nn_index[i, t]) (1-based), with D
    the Euclidean definition of knn(): squares of differences, accumulated in
    column order, no FMA."""
    U = np.asfortranarray(U)
    n, k = nn_index.shape
    j = np.asarray(nn_index, np.int64) - 1
    D = np.zeros((n, k), U.dtype)
    rows = np.arange(n)[:, None]
    for t in range(U.shape[1]):
        diff = (U[rows, t] - U[j, t]).astype(U.dtype)
        D = (D + diff * diff).astype(U.dtype)
    return (U.dtype.type(0.5) * D).astype(U.dtype)


def knn_cosine(X, k, include_self=True, nthreads=None):
    """Exact cosine kNN: the exact Euclidean search on unit_rows(X), ties to
    the lower index, distance D / 2 (1 - cos in exact arithmetic)."""
    U = unit_rows(X)
    idx, _ = knn(U, k, include_self=include_self, nthreads=nthreads)
    return idx, np.asfortranarray(half_pair_dist(U, idx))


def snn(nn_index, k, prune=1.0 / 15, dtype=np.float64):
    """Jaccard SNN graph as CSC (colptr, rowval 1-based int64, nzval)."""
    nn_index = np.asfortranarray(nn_index, np.int32)
    n = nn_index.shape[0]
    f64 = int(dtype == np.float64)
    L = lib()
    nnz = L.so_snn_jaccard(n, k, _ptr(nn_index), prune, f64, None, None, None)
    if nnz < 0:
        raise ValueError("bad neighbour index")
    colptr = np.zeros(n + 1, np.int64)
    rowval = np.zeros(nnz, np.int64)
    nzval = np.zeros(nnz, dtype)
    L.so_snn_jaccard(n, k, _ptr(nn_index), prune, f64, _ptr(colptr), _ptr(rowval), _ptr(nzval))
    return colptr, rowval, nzval


def count_mean_std(m, colptr, counts):
    """mean_std(Float64, A) of an Int32 count CSC (src/scaling.jl:119-130)."""
    n = len(colptr) - 1
    colptr = np.ascontiguousarray(colptr, np.int64)
    counts = np.ascontiguousarray(counts, np.int32)
    mu = np.empty(n)
    sd = np.empty(n)
    lib().so_count_mean_std(m, n, _ptr(colptr), _ptr(counts), _ptr(mu), _ptr(sd))
    return mu, sd


def rel_mean_var(m, colptr, rows, counts):
    """mean_var(Float64, row_norm(A, 1.0)) of an Int32 count CSC: per gene
    Welford in FP64 over x / s_row (src/scaling.jl:18-34, src/normalize.jl)."""
    n = len(colptr) - 1
    colptr = np.ascontiguousarray(colptr, np.int64)
    rows = np.ascontiguousarray(rows, np.int32)
    counts = np.ascontiguousarray(counts, np.int32)
    mu = np.empty(n)
    var = np.empty(n)
    lib().so_rel_mean_var(m, n, _ptr(colptr), _ptr(rows), _ptr(counts), _ptr(mu), _ptr(var))
    return mu, var


def _nan2zero(v):
    return np.where(np.isnan(v), 0.0, v)


def _libm(fn, v):
    """fn (a C library function through Python's math module) per element;
    log of 0 is -inf and of a negative or NaN argument NaN, as in C."""
    import math

    def one(x):
        if x != x or x < 0.0:
            return float("nan")
        if x == 0.0 and fn in (math.log, math.log10):
            return float("-inf")
        if x == float("inf"):
            return x
        return fn(x)
    return np.array([one(float(x)) for x in np.asarray(v, np.float64)])


def hvg_dispersion(mu, var):
    """select_dispersion (src/variablefeatures.jl:74-77): log(var / mu) with
    NaN -> 0 (log_VMR, src/scaling.jl:190-192)."""
    import math
    with np.errstate(divide="ignore", invalid="ignore"):
        return _nan2zero(_libm(math.log, var / mu))


def width_breaks(v, nbins):
    """cut(v, nbins; method=:width) breaks (src/utils.jl:140-148): nbins + 1
    evenly spaced points from min to max, widened by (max - min)/1000 at
    both ends (interior points from an extended-precision interpolation)."""
    lo, hi = float(np.min(v)), float(np.max(v))
    dx = hi - lo
    i = np.arange(nbins + 1, dtype=np.longdouble)
    br = ((np.longdouble(lo) * (nbins - i) + np.longdouble(hi) * i) / nbins).astype(np.float64)
    br[0], br[-1] = lo, hi
    br[0] -= dx / 1000
    br[-1] += dx / 1000
    return br


def cut_labels(v, breaks):
    """_cut! with right=true (src/utils.jl:109-125): searchsortedlast, minus
    one when v equals that break; 1-based labels."""
    idx = np.searchsorted(breaks, v, side="right")  # searchsortedlast, 1-based
    on = breaks[np.clip(idx - 1, 0, len(breaks) - 1)] == v
    return np.where(on, idx - 1, idx)


def hvg_meanvarplot(mu, var, nbins=20):
    """select_meanvarplot (src/variablefeatures.jl:79-91): log1p(mu) and
    log(var/mu) (NaN -> 0), width bins of log1p(mu), then the z-score of
    the dispersion in its bin; per-bin mean and sd by the labelled Welford
    in gene order (src/scaling.jl:90-112), NaN -> 0."""
    import math
    with np.errstate(divide="ignore", invalid="ignore"):
        lmu = _nan2zero(_libm(math.log1p, mu))
        disp = _nan2zero(_libm(math.log, var / mu))
        lab = cut_labels(lmu, width_breaks(lmu, nbins))
        bm = np.zeros(nbins + 1)
        bv = np.zeros(nbins + 1)
        bn = np.zeros(nbins + 1, np.int64)
        for v, k in zip(disp.tolist(), lab.tolist()):
            bn[k] += 1
            delta = v - bm[k]
            bm[k] = bm[k] + delta / bn[k]
            bv[k] = bv[k] + delta * (v - bm[k])
        bs = np.sqrt(bv / (bn - 1))
        z = (disp - bm[lab]) / bs[lab]
    return _nan2zero(z)


def norm_quantile_upper(tail):
    """x with P(Z > x) = tail for a standard normal Z, by Newton steps on
    erfc (the quantile at 1 - tail)."""
    import math
    x = math.sqrt(-2.0 * math.log(tail))  # start above the root
    for _ in range(100):
        f = 0.5 * math.erfc(x / math.sqrt(2.0)) - tail
        dfdx = -math.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)
        step = f / dfdx
        x -= step
        if abs(step) <= 1e-16 * abs(x):
            break
    return x


def hvg_saunders(m, colptr, rows, counts, mu, var, alpha=0.1):
    """select_features_saunders (src/variablefeatures.jl:60-72): the Nolan
    constant mean(1 ./ cell totals) (summed in cell order here), the
    Bonferroni-corrected upper bound of a Poisson-like gene mean, and
    log10(var) - log10(mu * nolan) for genes above it, else 0."""
    G = len(colptr) - 1
    tot = np.zeros(m, np.int64)
    np.add.at(tot, np.asarray(rows, np.int64), np.asarray(counts, np.int64))
    acc = 0.0
    for t in tot.tolist():
        acc += 1.0 / t
    nolan = acc / m
    a = alpha / G
    p = 1.0 - a / 2.0
    q = norm_quantile_upper(1.0 - p)
    upper = mu + q * np.sqrt(mu * nolan / m)
    import math
    with np.errstate(divide="ignore", invalid="ignore"):
        base = _libm(math.log10, mu * nolan)
        J = (var / nolan) > upper
        metric = np.zeros(G)
        metric[J] = _libm(math.log10, var[J]) - base[J]
    return metric


def top_features(metric, nfeatures):
    """partialsortperm(metric, 1:nfeatures, rev=true), ties to the lower
    gene (0-based)."""
    order = np.lexsort((np.arange(len(metric)), -np.asarray(metric)))
    return order[:nfeatures]


def loess_fit(x, y, span):
    """Exact degree-2 loess at every point of x (ascending)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    fit = np.empty(len(x))
    if lib().so_loess_fit(len(x), _ptr(x), _ptr(y), span, _ptr(fit)) != 0:
        raise ValueError("loess needs at least one point")
    return fit


def standardized_var(m, colptr, counts, mu, sd):
    n = len(colptr) - 1
    colptr = np.ascontiguousarray(colptr, np.int64)
    counts = np.ascontiguousarray(counts, np.int32)
    out = np.empty(n)
    lib().so_standardized_var(m, n, _ptr(colptr), _ptr(counts), _ptr(np.ascontiguousarray(mu)),
                              _ptr(np.ascontiguousarray(sd)), _ptr(out))
    return out


def vst(m, colptr, counts, nfeatures=2000, loess_span=0.5):
    """find_variable_features(:vst) (src/variablefeatures.jl:34-50,128-161):
    returns (selected 0-based genes, metric, mu, sd, expected sd).  The
    non-constant genes are ordered by (log10 mu, gene) for the fit; ties in
    the metric go to the lower gene index."""
    mu, sd = count_mean_std(m, colptr, counts)
    nc = np.nonzero(sd > 0)[0]
    xs, ys = np.log10(mu[nc]), np.log10(sd[nc])
    order = np.lexsort((nc, xs))
    fit = np.empty(len(nc))
    fit[order] = loess_fit(xs[order], ys[order], loess_span)
    esd = sd.copy()
    esd[nc] = 10.0 ** fit
    esd[np.isnan(esd)] = 0.0
    metric = standardized_var(m, colptr, counts, mu, esd)
    sel = np.lexsort((np.arange(len(metric)), -metric))[:nfeatures]
    return sel, metric, mu, sd, esd


def _graph(colptr, rows, w):
    return (np.ascontiguousarray(colptr, np.int64), np.ascontiguousarray(rows, np.int32),
            np.ascontiguousarray(w, np.float64))


def louvain(n, colptr, rows, w, resolution=0.8, nstarts=10, max_iter=10, min_mod=1e-4,
            seed=0):
    """Multi-start Louvain of the reference's cluster(SNN) in FP64 on a
    square CSC graph (src/clustering.jl:59-77, src/modularity.jl:255-380);
    node orders from a seeded shuffle.  Returns (0-based labels in order of
    first occurrence, modularity)."""
    colptr, rows, w = _graph(colptr, rows, w)
    lab = np.empty(n, np.int64)
    q = C.c_double()
    lib().so_louvain(n, _ptr(colptr), _ptr(rows), _ptr(w), resolution, nstarts, max_iter,
                     min_mod, seed, _ptr(lab), C.byref(q))
    return lab, q.value


def modularity(n, colptr, rows, w, labels, resolution=0.8):
    """modularity of compact 0-based labels on Network(SNN) (self-links
    dropped; src/modularity.jl:125-148,165-194)."""
    colptr, rows, w = _graph(colptr, rows, w)
    labels = np.ascontiguousarray(labels, np.int64)
    return lib().so_modularity(n, _ptr(colptr), _ptr(rows), _ptr(w), _ptr(labels), resolution)


def group_singletons(n, colptr, rows, w, labels):
    """group_singletons! then relabel! (src/clustering.jl:79-98,
    src/utils.jl:91-106) on compact 0-based labels; None when every cluster
    is a singleton (the reference's findmax over an empty set)."""
    colptr, rows, w = _graph(colptr, rows, w)
    lab = np.array(labels, np.int64)
    r = lib().so_group_singletons(n, _ptr(colptr), _ptr(rows), _ptr(w), _ptr(lab))
    return None if r < 0 else lab
