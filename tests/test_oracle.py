"""Pins the CPU oracle against the reference's fixtures, closed forms, library
routines (numpy/LAPACK) and brute force, so that the oracle can serve as the
checker of the GPU path.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp


# ---------------------------------------------------------------------------
# synthetic generator
# ---------------------------------------------------------------------------
def test_u01_moments(orc):
    u = np.array([orc.u01(5, 3, i) for i in range(20000)])
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003


def test_synth_matches_negative_binomial(sev, orc):
    from scipy import stats
    t = sev.nb_tables(40, seed=3, density=0.3)
    C = 20000
    colptr, rowval, counts = orc.synth_csc(C, t)
    D = sp.csc_matrix((counts, rowval, colptr), shape=(C, 40)).toarray()
    for g in range(40):
        mu = t.size_factor * t.gene_mean[g]
        p = t.r / (t.r + mu)
        # zero fraction and mean against the NB(r, p) law
        p0 = stats.nbinom.pmf(0, t.r, p)
        assert abs((D[:, g] == 0).mean() - p0) < 5 * np.sqrt(p0 * (1 - p0) / C) + 1e-3
        sd = np.sqrt(stats.nbinom.var(t.r, p) / C)
        assert abs(D[:, g].mean() - mu) < 6 * sd + 1e-3
    # achieved density close to the calibrated expectation
    assert abs(len(counts) / (C * 40) - t.expected_density()) < 0.01


def test_synth_boost_class(sev, orc):
    t = sev.nb_tables(200, seed=1, density=0.04, boost_frac=0.05)
    colptr, rowval, counts = orc.synth_csc(4000, t)
    X = sp.csc_matrix((counts, rowval, colptr), shape=(4000, 200))
    lib = np.asarray(X.sum(axis=1)).ravel()
    cls = np.array([orc.u01(t.seed, 0x7FFFFFFF, c) < 0.05 for c in range(4000)])
    assert 0.03 < cls.mean() < 0.07
    ratio = lib[cls].mean() / lib[~cls].mean()
    assert 2.5 < ratio < 3.5


# ---------------------------------------------------------------------------
# normalize_cells  (src/normalize.jl)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt,tag", [(np.float32, "f32"), (np.float64, "f64")])
@pytest.mark.parametrize("sf", [1.0, 10.0, 1e4])
@pytest.mark.parametrize("log", [False, True])
def test_normalize_golden(orc, golden, dt, tag, sf, log):
    out = orc.normalize(10, golden["colptr"], golden["rowval"], golden["nzval"], sf, log, dt)
    exp = golden[f"norm_{tag}_sf{sf:g}_{'log' if log else 'rel'}"]
    assert out.dtype == dt and np.array_equal(out, exp)


def test_normalize_row_sum_invariants(orc, golden):
    # test/test_input.jl:64,74: relative counts sum to the scale factor per cell
    for sf in (1.0, 10.0):
        out = orc.normalize(10, golden["colptr"], golden["rowval"], golden["nzval"], sf, False,
                            np.float64)
        X = sp.csc_matrix((out, golden["rowval"], golden["colptr"]), shape=(10, 5))
        rs = np.asarray(X.sum(axis=1)).ravel()
        nonempty = np.asarray((golden["counts"] != 0).any(axis=1))
        assert np.allclose(rs[nonempty], sf, rtol=1e-14)
    # log path is log1p of the relative counts (test/test_input.jl:70)
    rel = orc.normalize(10, golden["colptr"], golden["rowval"], golden["nzval"], 1.0, False,
                        np.float64)
    lg = orc.normalize(10, golden["colptr"], golden["rowval"], golden["nzval"], 1.0, True,
                       np.float64)
    assert np.array_equal(lg, np.log1p(rel))


def test_normalize_operation_order(orc):
    """sf*x/s (left to right) differs from x/s*sf in Float32 on some inputs;
    the oracle must follow the reference's order (src/normalize.jl:27)."""
    rng = np.random.default_rng(0)
    C, G = 400, 50
    D = rng.integers(0, 40, size=(C, G)) * (rng.random((C, G)) < 0.3)
    D[:, 0] += 1
    X = sp.csc_matrix(D)
    out = orc.normalize(C, X.indptr, X.indices, X.data, 1e4, False, np.float32)
    tot = D.sum(axis=1).astype(np.float32)
    rows = X.indices
    x = X.data.astype(np.float32)
    ref = (np.float32(1e4) * x) / tot[rows]
    alt = (x / tot[rows]) * np.float32(1e4)
    assert np.array_equal(out, ref)
    assert not np.array_equal(out, alt)


# ---------------------------------------------------------------------------
# column selection Y[:, features]  (docs/src/pbmc.md:121, src/scaling.jl:350-352;
# the order of the features is the order of src/variablefeatures.jl:159)
# ---------------------------------------------------------------------------
def _sparse_with_empty_columns(seed, m=300, n=60):
    rng = np.random.default_rng(seed)
    X = sp.random(m, n, density=0.1, format="csc", random_state=rng, dtype=np.float32)
    X.data[:] = rng.uniform(0.5, 5.0, X.nnz).astype(np.float32)
    X = X.tolil()
    for j in (0, 7, 31, n - 1):
        X[:, j] = 0
    X = X.tocsc()
    X.sort_indices()
    return X


@pytest.mark.parametrize("order", ["unsorted", "descending", "single", "empty_cols", "none"])
def test_gather_vs_scipy_column_indexing(orc, order):
    X = _sparse_with_empty_columns(3)
    n = X.shape[1]
    feats = {
        "unsorted": np.random.default_rng(9).permutation(n)[:25],
        "descending": np.arange(n - 1, -1, -3),
        "single": np.array([17]),
        "empty_cols": np.array([31, 2, 0, n - 1, 7, 40]),
        "none": np.array([], np.int64),
    }[order]
    cp, rows, vals = orc.gather(feats, X.indptr, X.indices, X.data)
    ref = X[:, feats]  # scipy's own column indexing keeps the given order
    ref = ref.tocsc()
    ref.sort_indices()
    assert np.array_equal(cp, ref.indptr)
    assert np.array_equal(rows, ref.indices) and np.array_equal(vals, ref.data)
    # dense cross-check through a second independent route
    D = X.toarray()
    G = sp.csc_matrix((vals, rows, cp), shape=(X.shape[0], len(feats))).toarray()
    assert np.array_equal(G, D[:, feats])


def test_gather_planted_mistakes_differ(orc):
    X = _sparse_with_empty_columns(4)
    feats = np.array([40, 3, 22, 9, 55, 1])
    cp, rows, vals = orc.gather(feats, X.indptr, X.indices, X.data)
    ref = X[:, feats].toarray()
    G = sp.csc_matrix((vals, rows, cp), shape=(X.shape[0], len(feats))).toarray()
    assert np.array_equal(G, ref)
    # a gather that sorts the features, or that is off by one gene, fails
    for wrong in (np.sort(feats), feats + 1):
        cpw, rw, vw = orc.gather(wrong, X.indptr, X.indices, X.data)
        Gw = sp.csc_matrix((vw, rw, cpw), shape=(X.shape[0], len(feats))).toarray()
        assert not np.array_equal(Gw, ref)


def _brute_top(metric, nfeatures):
    """Selection by repeated scans: the largest remaining value, the lowest
    index among equal values (partialsortperm(metric, 1:n, rev=true) with
    ties in index order)."""
    left = list(range(len(metric)))
    out = []
    for _ in range(nfeatures):
        best = left[0]
        for i in left[1:]:
            if metric[i] > metric[best]:
                best = i
        out.append(best)
        left.remove(best)
    return np.array(out, np.int64)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_top_features_vs_brute_force_with_ties(orc, seed):
    rng = np.random.default_rng(seed)
    metric = rng.integers(0, 12, 150).astype(np.float64)  # many planted ties
    metric[rng.choice(150, 10, replace=False)] = 0.0
    for nf in (1, 7, 40, 150):
        got = orc.top_features(metric, nf)
        assert np.array_equal(got, _brute_top(metric, nf))
    # a selection that breaks ties towards the higher index differs
    rev = np.lexsort((-np.arange(150), -metric))[:40]
    assert not np.array_equal(rev, _brute_top(metric, 40))
    # continuous values: the order equals a plain descending sort
    v = rng.standard_normal(300)
    assert np.array_equal(orc.top_features(v, 50), np.argsort(-v, kind="stable")[:50])


# ---------------------------------------------------------------------------
# scale_features  (src/scaling.jl)
# ---------------------------------------------------------------------------
def test_scale_golden_closed_form(orc, golden):
    # test/test_scaling.jl:41-44 on the 10x5 fixture (Float64)
    vals = golden["nzval"].astype(np.float64)
    out, mu, sd = orc.scale(10, golden["colptr"], vals)
    assert np.allclose(mu, golden["scale_mu_over_sd"], rtol=1e-13)
    B = sp.csc_matrix((out, golden["rowval"], golden["colptr"]), shape=(10, 5)).toarray()
    assert np.allclose(B - mu[None, :], golden["scale_dense"], rtol=1e-12, atol=1e-13)


def _welford_py(col_vals, m):
    """Pure-Python Float32 restatement of src/scaling.jl:18-34 (small cases)."""
    f = np.float32
    count = m - len(col_vals)
    mu = f(0.0)
    s = f(0.0)
    for v in col_vals:
        v = f(v)
        count += 1
        delta = f(v - mu)
        mu = f(mu + f(delta / f(count)))
        s = f(s + f(delta * f(v - mu)))
    var = f(s / f(m - 1))
    return mu, f(np.sqrt(var))


def test_scale_welford_independent_loop(orc):
    rng = np.random.default_rng(4)
    m, n = 3000, 6
    D = (rng.random((m, n)) < 0.2) * rng.gamma(2.0, 1.5, (m, n))
    X = sp.csc_matrix(D.astype(np.float32))
    out, mu, sd = orc.scale(m, X.indptr, X.data, np.float32(10.0))
    for j in range(n):
        v = X.data[X.indptr[j]:X.indptr[j + 1]]
        mu_j, sd_j = _welford_py(v, m)
        assert sd[j] == sd_j
        assert mu[j] == np.float32(mu_j / sd_j)
        # close to the exact statistics (catches a wrong variance formula)
        full = D[:, j].astype(np.float64)
        assert abs(sd_j - full.std(ddof=1)) < 1e-5 * full.std(ddof=1)


def test_scale_clip_upper_only(orc):
    rng = np.random.default_rng(5)
    m = 500
    d = np.zeros(m, np.float32)
    d[:40] = rng.gamma(1.0, 1.0, 40)
    d[0] = 500.0  # an outlier to clip
    X = sp.csc_matrix(d[:, None])
    out, mu, sd = orc.scale(m, X.indptr, X.data, np.float32(3.0))
    v = X.data / sd[0]
    clip = np.float32(3.0) + mu[0]
    assert np.array_equal(out, np.where(v > clip, clip, v))
    assert out.max() == clip and (v > clip).sum() >= 1
    out_inf, _, _ = orc.scale(m, X.indptr, X.data, np.float32(np.inf))
    assert np.array_equal(out_inf, v)


# ---------------------------------------------------------------------------
# CenteredMatrix mul!  (src/scaling.jl:245-257, known answers of
# test/test_scaling.jl:70-113)
# ---------------------------------------------------------------------------
def test_centered_mul_known_answers(orc, golden):
    A = sp.csc_matrix(golden["cm_A"])
    mu, r, y0 = golden["cm_mu"], golden["cm_r"], golden["cm_y0"]
    Qr = orc.centered_mul(4, A.indptr, A.indices, A.data, mu, r)
    assert np.allclose(Qr, golden["cm_Qr"], rtol=1e-14)
    y1 = orc.centered_mul(4, A.indptr, A.indices, A.data, mu, r, y=y0, alpha=2.0, beta=1.0)
    assert np.allclose(y1, golden["cm_y1"], rtol=1e-14)
    Qty = orc.centered_mul(4, A.indptr, A.indices, A.data, mu, y1, trans=True)
    assert np.allclose(Qty, golden["cm_Qty1"], rtol=1e-14)
    r1 = orc.centered_mul(4, A.indptr, A.indices, A.data, mu, y1, trans=True, y=r, alpha=2.0,
                          beta=1.0)
    assert np.allclose(r1, golden["cm_r1"], rtol=1e-14)


def _centered(rng, m, n, dens, dt):
    A = sp.random(m, n, density=dens, format="csc", random_state=rng, dtype=np.float64)
    A.data = (A.data * 4.0 - 1.0).astype(dt)
    A = A.astype(dt)
    A.sort_indices()
    mu = (rng.standard_normal(n) * 0.7).astype(dt)
    return A, mu


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_centered_mul_mat_dense_closed_form(orc, dt):
    """mul! of a CenteredMatrix with matrices (src/scaling.jl:259-271) against
    (A - 1 mu') X and (A - 1 mu')' X of the densified operator."""
    rng = np.random.default_rng(5)
    m, n, p = 120, 30, 7
    A, mu = _centered(rng, m, n, 0.2, dt)
    D = A.toarray().astype(np.float64) - mu.astype(np.float64)[None, :]
    X = rng.standard_normal((n, p))
    W = rng.standard_normal((m, p))
    Y0 = rng.standard_normal((m, p))
    Z0 = rng.standard_normal((n, p))
    got = orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, X, Y=Y0, alpha=2.0, beta=0.5)
    assert np.allclose(got, 2.0 * D @ X + 0.5 * Y0, rtol=1e-12, atol=1e-12)
    got = orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, W, trans=True, Y=Z0,
                               alpha=-1.5, beta=2.0)
    assert np.allclose(got, -1.5 * D.T @ W + 2.0 * Z0, rtol=1e-12, atol=1e-12)
    # a planted sign error in the adjoint centering term is caught
    wrong = -1.5 * (A.toarray().T @ W + np.outer(mu, W.sum(0))) + 2.0 * Z0
    assert not np.allclose(got, wrong, rtol=1e-6)


@pytest.mark.parametrize("same", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_centered_gram_dense_closed_form(orc, same, dt):
    """mul!(C, Sl', Sr, alpha, beta) (src/scaling.jl:273-294) against
    (Al - 1 mul')'(Ar - 1 mur') of the densified operands, for Sl === Sr,
    two different matrices and an uncentered right operand."""
    rng = np.random.default_rng(6 + same)
    m, nl, nr = 150, 25, 25 if same else 18
    Al, mul = _centered(rng, m, nl, 0.15, dt)
    if same:
        Ar, mur = Al, mul
        got = orc.centered_gram(m, Al.indptr, Al.indices, Al.data, mul, alpha=0.5)
    else:
        Ar, mur = _centered(rng, m, nr, 0.3, dt)
        C0 = rng.standard_normal((nl, nr))
        got = orc.centered_gram(m, Al.indptr, Al.indices, Al.data, mul, Ar.indptr, Ar.indices,
                                Ar.data, mur, C=C0, alpha=0.5, beta=-1.0)
    Dl = Al.toarray().astype(np.float64) - mul.astype(np.float64)[None, :]
    Dr = Ar.toarray().astype(np.float64) - mur.astype(np.float64)[None, :]
    ref = 0.5 * Dl.T @ Dr + (0.0 if same else -1.0 * C0)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    if not same:
        got = orc.centered_gram(m, Al.indptr, Al.indices, Al.data, mul, Ar.indptr, Ar.indices,
                                Ar.data, None)
        assert np.allclose(got, Dl.T @ Ar.toarray().astype(np.float64), rtol=1e-11,
                           atol=1e-11 * np.abs(got).max())


# ---------------------------------------------------------------------------
# small SVD and IRLBA  (test/test_irlba.jl cases against LAPACK)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,rank", [(7, 7), (30, 30), (20, 12)])
def test_svd_jacobi_vs_lapack(orc, n, rank):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, rank)) @ rng.standard_normal((rank, n))
    U, s, V = orc.svd_jacobi(A)
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(s, ref, rtol=1e-12, atol=1e-12 * ref[0])
    assert np.allclose(U @ np.diag(s) @ V.T, A, atol=1e-11 * ref[0])
    assert np.allclose(U.T @ U, np.eye(n), atol=1e-12)


def _rel_err(X, U, s, V):
    return np.linalg.norm(X - (U * s) @ V.T)


@pytest.mark.parametrize("m,n,nu,tol", [(50, 50, 20, 1e-5), (100, 50, 20, 1e-5)])
def test_irlba_dense(orc, m, n, nu, tol):
    rng = np.random.default_rng(m + n)
    X = rng.standard_normal((m, n))
    rc, s, U, V, it, nm = orc.irlba_dense(X, nu, rng.standard_normal(n), tol=tol)
    assert rc == 0
    ss = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(s, ss[:nu], rtol=1e-8)
    assert np.linalg.norm(X.T @ U - V * s) / np.linalg.norm(X) < tol
    Us, sv, Vt = np.linalg.svd(X)
    assert np.isclose(_rel_err(X, U, s, V), _rel_err(X, Us[:, :nu], sv[:nu], Vt[:nu].T),
                      rtol=1e-6)


def test_irlba_centered_sparse(orc):
    A = sp.random(2000, 400, 0.1, format="csc", random_state=3, dtype=np.float64)
    mu = np.asarray(A.mean(axis=0)).ravel()
    Q = A.toarray() - mu
    rng = np.random.default_rng(0)
    rc, s, U, V, it, nm = orc.irlba_csc(2000, A.indptr, A.indices, A.data, mu, 2,
                                        rng.standard_normal(400), tol=1e-9)
    assert rc == 0
    ss = np.linalg.svd(Q, compute_uv=False)
    assert np.allclose(s, ss[:2], rtol=1e-10)
    assert np.linalg.norm(Q.T @ U - V * s) / np.linalg.norm(Q) < 1e-9


def test_irlba_tall_skinny_transpose(orc):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((10000, 10))
    rc1, s1, U1, V1, _, _ = orc.irlba_dense(X, 2, rng.standard_normal(10), tol=1e-9)
    rc2, s2, U2, V2, _, _ = orc.irlba_dense(X.T.copy(), 2, rng.standard_normal(10000), tol=1e-9)
    assert rc1 == 0 and rc2 == 0
    assert np.allclose(s1, s2, rtol=1e-10)
    # svd_flip! (src/utils.jl:215-228): align signs, then compare vectors
    for j in range(2):
        a = U1[:, j] * np.sign(U1[np.argmax(np.abs(U1[:, j])), j])
        b = V2[:, j] * np.sign(V2[np.argmax(np.abs(V2[:, j])), j])
        assert np.allclose(a, b, atol=1e-7)


def test_irlba_count_matrix(orc):
    # test/test_irlba.jl:107-115: Poisson(10) counts, centered by row means,
    # transposed operator
    rng = np.random.default_rng(11)
    X = sp.random(3000, 2000, 0.01, format="csc", random_state=11,
                  data_rvs=lambda k: rng.poisson(10, k).astype(float))
    mu = np.asarray(X.mean(axis=1)).ravel()
    At = sp.csc_matrix(X.T)
    rc, s, U, V, it, nm = orc.irlba_csc(2000, At.indptr, At.indices, At.data, mu, 10,
                                        rng.standard_normal(3000))
    assert rc == 0
    ss = np.linalg.svd(At.toarray() - mu[None, :], compute_uv=False)
    assert np.allclose(s, ss[:10], rtol=1e-6)


def test_irlba_mutation_is_caught(orc):
    """A dropped centering term (planted in the oracle's input, on the CPU)
    is detected by the LAPACK comparison above."""
    A = sp.random(500, 80, 0.2, format="csc", random_state=2, dtype=np.float64)
    mu = np.asarray(A.mean(axis=0)).ravel()
    Q = A.toarray() - mu
    rc, s, U, V, _, _ = orc.irlba_csc(500, A.indptr, A.indices, A.data, None, 3,
                                      np.ones(80))
    ss = np.linalg.svd(Q, compute_uv=False)[:3]
    assert not np.allclose(s, ss, rtol=1e-3)


# ---------------------------------------------------------------------------
# exact kNN
# ---------------------------------------------------------------------------
def _knn_numpy(Z, k, include_self=True):
    """Independent restatement: float32 sequential-t distances, stable argsort."""
    n, d = Z.shape
    D = np.zeros((n, n), Z.dtype)
    for t in range(d):
        diff = Z[:, None, t] - Z[None, :, t]
        D = D + diff * diff
    if not include_self:
        D[np.arange(n), np.arange(n)] = np.inf
    idx = np.argsort(D, axis=1, kind="stable")[:, :k]
    return idx + 1, np.sqrt(np.take_along_axis(D, idx, 1))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("include_self", [True, False])
def test_knn_bruteforce(orc, dtype, include_self):
    rng = np.random.default_rng(12)
    Z = rng.random((300, 10)).astype(dtype)
    idx, dist = orc.knn(Z, 7, include_self=include_self)
    ridx, rdist = _knn_numpy(Z, 7, include_self)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(dist, rdist.astype(dtype))


def test_knn_ties_lower_index_wins(orc):
    rng = np.random.default_rng(13)
    base = rng.random((40, 5)).astype(np.float32)
    Z = np.concatenate([base, base, base[:10]])  # exact duplicate points
    idx, dist = orc.knn(Z, 4)
    ridx, _ = _knn_numpy(Z, 4)
    assert np.array_equal(idx, ridx)
    # duplicates at distance 0: the lowest index comes first
    assert idx[40, 0] == 1 and idx[40, 1] == 41 and idx[80, 0] == 1


@pytest.mark.parametrize("dtype,tol", [(np.float32, 3e-7), (np.float64, 1e-15)])
def test_knn_cosine_vs_float64_cosine(orc, dtype, tol):
    """The cosine form ranks like 1 - cos of a float64 brute force (points
    without near ties) and reports 1 - cos to the rounding of dtype."""
    rng = np.random.default_rng(15)
    X = (rng.standard_normal((400, 12)) * rng.uniform(0.1, 10.0, (400, 1))).astype(dtype)
    idx, dist = orc.knn_cosine(X, 8)
    Xd = X.astype(np.float64)
    U = Xd / np.linalg.norm(Xd, axis=1)[:, None]
    C = 1.0 - U @ U.T
    ridx = np.stack([np.lexsort((np.arange(400), C[i]))[:8] for i in range(400)]) + 1
    assert np.array_equal(idx, ridx)
    assert np.max(np.abs(dist - np.take_along_axis(C, idx - 1, 1))) < tol * 4
    assert dist.dtype == dtype


def test_knn_cosine_scale_invariance_and_zero_row(orc):
    """Positive multiples of a row are at distance 0 (lower index first);
    an all-zero row stays zero, at D / 2 = |unit row|^2 / 2 (0.5 up to
    rounding) from every unit row."""
    rng = np.random.default_rng(16)
    base = rng.standard_normal((50, 6)).astype(np.float32)
    X = np.concatenate([base, 4.0 * base[:10], np.zeros((1, 6), np.float32)])
    idx, dist = orc.knn_cosine(X, 3)
    assert idx[50, 0] == 1 and idx[50, 1] == 51 and dist[50, 1] <= 1e-7
    assert np.all(np.abs(dist[60, 1:] - 0.5) < 1e-6)  # |unit row|^2 / 2, rounded
    # a planted mutation (no scaling to unit length) ranks differently
    ridx, _ = orc.knn(X, 3)
    assert not np.array_equal(idx, ridx)


def test_knn_reference_recall_property(orc):
    # test/test_nn.jl:31-38: 30th percentile Jaccard recall vs brute force == 1
    rng = np.random.default_rng(14)
    X = rng.random((100, 10))
    idx, _ = orc.knn(X, 4)
    D = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    j = []
    for i in range(100):
        nn = set(np.argsort(D[:, i], kind="stable")[:4].tolist())
        x = len(nn & set((idx[i] - 1).tolist()))
        j.append(x / (8 - x))
    assert np.quantile(j, 0.3) == 1.0


# ---------------------------------------------------------------------------
# Jaccard SNN
# ---------------------------------------------------------------------------
def _snn_sets(idx, k, prune, dt):
    n = idx.shape[0]
    N = [set(r) for r in idx.tolist()]
    cols = []
    for j in range(n):
        rows, vals = [], []
        for i in range(n):
            x = len(N[i] & N[j])
            if x == 0:
                continue
            xf = dt(x)
            w = xf / (dt(k) + (dt(k) - xf))
            if abs(w) > dt(prune):
                rows.append(i + 1)
                vals.append(w)
        cols.append((rows, vals))
    return cols


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("k,prune", [(5, 1 / 15), (4, 0.0), (6, 0.3)])
def test_snn_sets(orc, dt, k, prune):
    rng = np.random.default_rng(k)
    Z = rng.random((150, 6)).astype(np.float32)
    idx, _ = orc.knn(Z, k)
    cp, rv, nz = orc.snn(idx, k, prune, dt)
    cols = _snn_sets(idx, k, prune, dt)
    for j, (rows, vals) in enumerate(cols):
        assert list(rv[cp[j] - 1:cp[j + 1] - 1]) == rows
        assert np.array_equal(nz[cp[j] - 1:cp[j + 1] - 1], np.array(vals, dt))


def test_snn_symmetric_with_unit_diagonal(orc):
    rng = np.random.default_rng(3)
    Z = rng.random((400, 8)).astype(np.float32)
    idx, _ = orc.knn(Z, 10)
    cp, rv, nz = orc.snn(idx, 10, 1 / 15)
    S = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(400, 400))
    assert abs(S - S.T).max() == 0
    assert np.all(S.diagonal() == 1.0)


def test_welford_quotient_identity():
    """The device Welford step computes delta / Float32(count) as
    Float32(Float64(delta) * (1 / Float64(Float32(count)))); for normal
    deltas this equals the IEEE Float32 division bit for bit (checked here on
    random deltas over a wide exponent range and counts up to 2^26, where the
    Float32 count itself is rounded)."""
    rng = np.random.default_rng(7)
    for hi in (2**6, 2**12, 2**20, 2**26 - 1):
        N = 400_000
        e = rng.integers(-100, 100, N)
        d = (np.sign(rng.random(N) - 0.5) * (rng.random(N) + 1.0) *
             np.exp2(e.astype(np.float64))).astype(np.float32)
        d[: N // 8] = rng.integers(-2**24, 2**24, N // 8).astype(np.float32)
        cf = rng.integers(1, hi, N).astype(np.float32)
        ref = d / cf
        fast = (d.astype(np.float64) * (1.0 / cf.astype(np.float64))).astype(np.float32)
        assert np.array_equal(fast.view(np.uint32), ref.view(np.uint32))


# ---------------------------------------------------------------------------
# find_variable_features(:vst) pieces (src/variablefeatures.jl:21-50)
# ---------------------------------------------------------------------------
def _count_csc(seed, m=400, n=60):
    rng = np.random.default_rng(seed)
    lam = np.exp(rng.normal(-1.0, 1.5, n))
    X = rng.poisson(lam[None, :] * rng.uniform(0.3, 2.0, (m, 1))).astype(np.int32)
    X[:, 5] = 0          # an all-zero (constant) gene
    X[:, 7] = 3          # a constant nonzero gene
    return sp.csc_matrix(X), X


def test_count_mean_std_vs_numpy(orc):
    S, X = _count_csc(1)
    mu, sd = orc.count_mean_std(X.shape[0], S.indptr, S.data)
    assert np.allclose(mu, X.mean(axis=0), rtol=1e-13, atol=0)
    assert np.allclose(sd, X.std(axis=0, ddof=1), rtol=1e-12, atol=1e-15)
    assert sd[5] == 0.0 and sd[7] == 0.0


def test_loess_reproduces_quadratics(orc):
    rng = np.random.default_rng(2)
    x = np.sort(np.concatenate([rng.uniform(-2, 3, 300), np.full(20, 0.5)]))  # with ties
    y = 1.5 - 0.7 * x + 0.3 * x * x
    for span in (0.1, 0.5, 0.9):
        assert np.abs(orc.loess_fit(x, y, span) - y).max() < 1e-10


def _loess_bruteforce(x, y, span):
    """The definition, written independently: the q-th smallest distance,
    tricube weights, numpy weighted least squares per point."""
    N = len(x)
    q = max(min(N, 3), min(N, int(np.floor(N * span))))
    out = np.empty(N)
    for i in range(N):
        d = np.abs(x - x[i])
        dmax = np.sort(d)[q - 1]
        r = d / dmax
        w = np.where(r < 1, (1 - r ** 3) ** 3, 0.0)
        u = (x - x[i]) / dmax
        A = np.stack([np.ones(N), u, u * u], axis=1) * np.sqrt(w)[:, None]
        coef = np.linalg.lstsq(A, y * np.sqrt(w), rcond=None)[0]
        out[i] = coef[0]
    return out


def test_loess_vs_weighted_least_squares(orc):
    rng = np.random.default_rng(3)
    x = np.sort(np.round(rng.normal(0, 1, 250), 2))  # rounded: many tied x
    y = np.sin(2 * x) + 0.1 * rng.standard_normal(len(x))
    for span in (0.2, 0.5):
        f = orc.loess_fit(x, y, span)
        assert np.allclose(f, _loess_bruteforce(x, y, span), rtol=1e-8, atol=1e-10)
    # a wrong neighbourhood (one point fewer) is detected
    f_bad = _loess_bruteforce(x, y, 0.5 - 1.0 / len(x))
    assert not np.allclose(orc.loess_fit(x, y, 0.5), f_bad, rtol=1e-8, atol=1e-10)


def test_standardized_var_closed_form(orc):
    S, X = _count_csc(4)
    m = X.shape[0]
    mu = X.mean(axis=0)
    esd = np.abs(np.random.default_rng(5).normal(1.0, 0.3, X.shape[1])) + 0.05
    esd[5] = 0.0
    v = orc.standardized_var(m, S.indptr, S.data, mu, esd)
    Z = np.minimum((X - mu[None, :]) / np.where(esd > 0, esd, 1)[None, :], np.sqrt(m))
    ref = (Z ** 2).sum(axis=0) / (m - 1)
    ref[5] = 0.0
    assert np.allclose(v, ref, rtol=1e-12, atol=0)
    # the clip at sqrt(m) matters for a rare, large count
    assert (((X - mu) / np.where(esd > 0, esd, 1)) > np.sqrt(m)).any()


def test_vst_selection_order(orc):
    S, X = _count_csc(6, m=500, n=80)
    sel, metric, mu, sd, esd = orc.vst(X.shape[0], S.indptr, S.data, nfeatures=20)
    assert np.array_equal(sel, np.lexsort((np.arange(80), -metric))[:20])
    assert metric[5] == 0.0 and metric[7] == 0.0
    assert np.all(np.diff(metric[sel]) <= 0)


# ---------------------------------------------------------------------------
# find_variable_features(:dispersion / :meanvarplot / :saunders) pieces
# (src/variablefeatures.jl:60-91, src/scaling.jl:18-34,90-112,190-192)
# ---------------------------------------------------------------------------
def _rel(X):
    tot = X.sum(axis=1).astype(np.float64)
    return X / np.where(tot == 0, 1.0, tot)[:, None]


def test_rel_mean_var_vs_numpy(orc):
    S, X = _count_csc(11, m=500, n=70)
    mu, var = orc.rel_mean_var(X.shape[0], S.indptr, S.indices, S.data)
    R = _rel(X)
    assert np.allclose(mu, R.mean(axis=0), rtol=1e-13, atol=1e-18)
    assert np.allclose(var, R.var(axis=0, ddof=1), rtol=1e-11, atol=1e-20)
    assert mu[5] == 0.0 and var[5] == 0.0


def test_dispersion_closed_form(orc):
    S, X = _count_csc(12, m=500, n=70)
    mu, var = orc.rel_mean_var(X.shape[0], S.indptr, S.indices, S.data)
    d = orc.hvg_dispersion(mu, var)
    R = _rel(X)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.log(R.var(axis=0, ddof=1) / R.mean(axis=0))
    ok = np.isfinite(ref)
    assert np.allclose(d[ok], ref[ok], rtol=1e-11)
    assert d[5] == 0.0  # 0/0 -> NaN -> 0


def test_width_bins_right_closed(orc):
    v = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 2.0])
    br = orc.width_breaks(v, 4)
    assert br[0] < 0.0 and br[-1] > 4.0 and np.allclose(br[1:-1], [1.0, 2.0, 3.0])
    lab = orc.cut_labels(v, br)
    # (b0, 1], (1, 2], (2, 3], (3, b4]: a value on an interior break goes down
    assert lab.tolist() == [1, 1, 2, 3, 4, 2]
    # planted mutation: left-closed bins differ
    assert np.searchsorted(br, v, side="right").tolist() != lab.tolist()


def test_meanvarplot_vs_numpy_bins(orc):
    S, X = _count_csc(13, m=800, n=150)
    mu, var = orc.rel_mean_var(X.shape[0], S.indptr, S.indices, S.data)
    z = orc.hvg_meanvarplot(mu, var, 20)
    with np.errstate(divide="ignore", invalid="ignore"):
        lmu = np.nan_to_num(np.log1p(mu), nan=0.0)
        disp = np.log(var / mu)
        disp = np.where(np.isnan(disp), 0.0, disp)
    lab = orc.cut_labels(lmu, orc.width_breaks(lmu, 20))
    ref = np.zeros_like(z)
    for k in np.unique(lab):
        sel = lab == k
        if sel.sum() < 2:
            continue  # sd of one value is NaN: z -> 0
        ref[sel] = (disp[sel] - disp[sel].mean()) / disp[sel].std(ddof=1)
    ref = np.where(np.isnan(ref), 0.0, ref)
    fin = np.isfinite(disp)
    assert np.allclose(z[fin], ref[fin], rtol=1e-9, atol=1e-12)


def test_saunders_closed_form(orc):
    import scipy.stats as st
    S, X = _count_csc(14, m=700, n=90)
    X[X.sum(axis=1) == 0, 0] = 1  # every cell has a count
    S = sp.csc_matrix(X)
    m = X.shape[0]
    mu, var = orc.rel_mean_var(m, S.indptr, S.indices, S.data)
    met = orc.hvg_saunders(m, S.indptr, S.indices, S.data, mu, var, alpha=0.1)
    nolan = np.mean(1.0 / X.sum(axis=1))
    q = st.norm.ppf(1 - 0.1 / 90 / 2)
    assert abs(orc.norm_quantile_upper(0.1 / 90 / 2) - q) <= 1e-14 * q
    with np.errstate(divide="ignore", invalid="ignore"):
        upper = mu + q * np.sqrt(mu * nolan / m)
        J = var / nolan > upper
        ref = np.where(J, np.log10(var) - np.log10(mu * nolan), 0.0)
    assert np.allclose(met, ref, rtol=1e-11, atol=0)
    assert (met > 0).sum() > 0 and met[5] == 0.0


def test_centered_tmul_mat_vs_dense(orc):
    """A'U of the centered matrix for several vectors at once (the checker
    of the IRLBA residual at full size) against the dense product; a
    planted mistake (the mu term added instead of subtracted) differs."""
    rng = np.random.default_rng(21)
    X = sp.random(700, 60, 0.15, format="csc", random_state=2, dtype=np.float64)
    X = X.astype(np.float32)
    mu = rng.random(60).astype(np.float32)
    U = rng.standard_normal((700, 7))
    A = X.toarray().astype(np.float64) - mu.astype(np.float64)
    Y = orc.centered_tmul_mat(700, X.indptr, X.indices, X.data, mu, U, nthreads=3)
    assert np.allclose(Y, A.T @ U, rtol=1e-12, atol=1e-11)
    wrong = X.toarray().astype(np.float64).T @ U + np.outer(mu, U.sum(axis=0))
    assert not np.allclose(Y, wrong, rtol=1e-6)
    # one column equals the single-vector product
    y0 = orc.centered_mul(700, X.indptr, X.indices, X.data, mu, U[:, 3], trans=True)
    assert np.allclose(Y[:, 3], y0, rtol=1e-13, atol=1e-12)
