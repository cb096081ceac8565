"""GPU parity: CenteredMatrix products and IRLBA against the oracle and
LAPACK (the cases of test/test_irlba.jl), plus the pipeline PCA against an
FP64 oracle of the centered Gram matrix.

Tolerances: products rtol 1e-12 (FP64 vectors, different summation order);
singular values within 1e-5 relative of LAPACK in FP32 form (1e-4 in BF16
form); residual ||A'U - V S||_F/||A||_F below the requested tol; singular
subspaces within 1e-3 rad (largest principal angle, sign-free)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def test_centered_mul_known_answers(sev, golden):
    A = sp.csc_matrix(golden["cm_A"])
    Cm = sev.CenteredMatrix(A, golden["cm_mu"])
    r, y0 = golden["cm_r"], golden["cm_y0"]
    assert np.allclose(Cm @ r, golden["cm_Qr"], rtol=1e-14)
    y1 = Cm.mul(y0.copy(), r, 2.0, 1.0)
    assert np.allclose(y1, golden["cm_y1"], rtol=1e-14)
    assert np.allclose(Cm.rmatvec(y1), golden["cm_Qty1"], rtol=1e-14)
    r1 = Cm.mul(r.copy(), y1, 2.0, 1.0, trans=True)
    assert np.allclose(r1, golden["cm_r1"], rtol=1e-14)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_centered_mul_vs_oracle(sev, orc, dt):
    rng = np.random.default_rng(1)
    m, n = 5000, 300  # several row slabs
    A = sp.random(m, n, 0.1, format="csc", random_state=1, dtype=np.float64).astype(dt)
    mu = rng.random(n).astype(dt)
    Cm = sev.CenteredMatrix(A, mu)
    x = rng.standard_normal(n)
    w = rng.standard_normal(m)
    y = Cm @ x
    z = Cm.rmatvec(w)
    assert np.allclose(y, orc.centered_mul(m, A.indptr, A.indices, A.data, mu, x), rtol=1e-12,
                       atol=1e-12)
    assert np.allclose(z, orc.centered_mul(m, A.indptr, A.indices, A.data, mu, w, trans=True),
                       rtol=1e-12, atol=1e-11)


def _angle(U1, U2):
    q1, _ = np.linalg.qr(U1)
    q2, _ = np.linalg.qr(U2)
    s = np.linalg.svd(q1.T @ q2, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1, 1)))


@pytest.mark.parametrize("m,n,nu", [(50, 50, 20), (100, 50, 20)])
def test_irlba_dense(sev, m, n, nu):
    rng = np.random.default_rng(m)
    X = rng.standard_normal((m, n))
    S = sev.irlba(sp.csc_matrix(X), nu, tol=1e-5, init=rng.standard_normal(n))
    ss = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(S.S, ss[:nu], rtol=1e-8)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X) < 1e-5


@pytest.mark.parametrize("nu", [55, 57, 58, 90, 121])
def test_irlba_wide_basis(sev, nu):
    """work = nu + 7 at the top of each projected-SVD kernel's range: 62
    (8 lanes per column pair), 64 (16 lanes per pair, both arrays in LDS),
    then 65, 97 and 128 (one array in LDS; restart products in several
    column groups)."""
    rng = np.random.default_rng(nu)
    m, n = 400, 160
    X = rng.standard_normal((m, n)) * (1.0 + 3.0 * np.exp(-np.arange(n) / 20.0))
    S = sev.irlba(sp.csc_matrix(X), nu, tol=1e-6, init=rng.standard_normal(n))
    ss = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(S.S, ss[:nu], rtol=1e-8)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X) < 1e-6


def test_irlba_graded_spectrum(sev):
    """Singular values spread over eight decades: the projected matrices are
    too ill-conditioned for U = B V / s, and the device SVD takes its
    accumulated-rotation path."""
    rng = np.random.default_rng(12)
    m, n, nu = 300, 120, 10
    Q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = 10.0 ** (-0.5 * np.arange(n))
    X = (Q1 * sig) @ Q2.T
    S = sev.irlba(sp.csc_matrix(X), nu, tol=1e-8, init=rng.standard_normal(n))
    assert np.allclose(S.S, sig[:nu], rtol=1e-7)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X) < 1e-7
    assert _angle(S.U, Q1[:, :nu]) < 1e-6 and _angle(S.V, Q2[:, :nu]) < 1e-6


@pytest.mark.parametrize("nu", [10, 60])
def test_irlba_graded_spectrum_wide(sev, nu):
    """The graded spectrum with work = nu + 7 up to 67: beyond 64 the device
    SVD hands ill-conditioned projected matrices to the host SVD."""
    rng = np.random.default_rng(13)
    m, n = 300, 120
    Q1, _ = np.linalg.qr(rng.standard_normal((m, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = 10.0 ** (-0.1 * np.arange(n))
    X = (Q1 * sig) @ Q2.T
    S = sev.irlba(sp.csc_matrix(X), nu, tol=1e-8, init=rng.standard_normal(n))
    assert np.allclose(S.S, sig[:nu], rtol=1e-7)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X) < 1e-7


def test_irlba_work_limit_reported(sev):
    """work above 128 is refused with the cause in the message (the libcell
    entry also prints it, since Severo only sees a nonzero status)."""
    rng = np.random.default_rng(2)
    X = sp.csc_matrix(rng.standard_normal((300, 200)))
    with pytest.raises(sev.SeveroError, match="128"):
        sev.irlba(X, 122, tol=1e-5, init=rng.standard_normal(200))


def test_irlba_sparse_two(sev):
    X = sp.random(2000, 400, 0.1, format="csc", random_state=7, dtype=np.float64)
    S = sev.irlba(X, 2, tol=1e-9, init=np.random.default_rng(1).standard_normal(400))
    ss = np.linalg.svd(X.toarray(), compute_uv=False)
    assert np.allclose(S.S, ss[:2], rtol=1e-10)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X.toarray()) < 1e-9


def test_irlba_centered_sparse(sev):
    X = sp.random(2000, 400, 0.1, format="csc", random_state=3, dtype=np.float64)
    mu = np.asarray(X.mean(axis=0)).ravel()
    Q = X.toarray() - mu
    S = sev.irlba(sev.CenteredMatrix(X, mu), 2, tol=1e-9,
                  init=np.random.default_rng(0).standard_normal(400))
    ss = np.linalg.svd(Q, compute_uv=False)
    assert np.allclose(S.S, ss[:2], rtol=1e-10)
    assert np.linalg.norm(Q.T @ S.U - S.V * S.S) / np.linalg.norm(Q) < 1e-9


def test_irlba_centered_dense(sev):
    rng = np.random.default_rng(4)
    X = rng.standard_normal((20, 10))
    mu = X.mean(axis=0)
    Q = X - mu
    S = sev.irlba(sev.CenteredMatrix(sp.csc_matrix(X), mu), 3, tol=1e-5,
                  init=rng.standard_normal(10))
    ss = np.linalg.svd(Q, compute_uv=False)
    assert np.allclose(S.S, ss[:3], rtol=1e-8)
    assert np.linalg.norm(Q.T @ S.U - S.V * S.S) / np.linalg.norm(Q) < 1e-9


def test_irlba_count_matrix(sev):
    rng = np.random.default_rng(11)
    X = sp.random(3000, 2000, 0.01, format="csc", random_state=11,
                  data_rvs=lambda k: rng.poisson(10, k).astype(float))
    mu = np.asarray(X.mean(axis=1)).ravel()
    At = sp.csc_matrix(X.T)
    S = sev.irlba(sev.CenteredMatrix(At, mu), 10, init=rng.standard_normal(3000))
    ss = np.linalg.svd(At.toarray() - mu[None, :], compute_uv=False)
    assert np.allclose(S.S, ss[:10], rtol=1e-6)


def test_irlba_maxit_raises(sev):
    rng = np.random.default_rng(2)
    X = sp.csc_matrix(rng.standard_normal((300, 100)))
    with pytest.raises(sev.ConvergenceError):
        sev.irlba(X, 20, tol=1e-14, maxit=1, init=rng.standard_normal(100))


def test_irlba_deterministic(sev):
    X = sp.random(5000, 300, 0.1, format="csc", random_state=5, dtype=np.float64)
    init = np.random.default_rng(3).standard_normal(300)
    a = sev.irlba(X, 10, init=init)
    b = sev.irlba(X, 10, init=init)
    assert np.array_equal(a.S, b.S) and np.array_equal(a.U, b.U)


def _gram_oracle(B, mu):
    """FP64 singular values / right vectors of the centered matrix from its
    Gram matrix (src/scaling.jl:274-296 form), no Lanczos involved."""
    Bd = B.astype(np.float64)
    m = B.shape[0]
    mu = mu.astype(np.float64)
    q = np.asarray(Bd.sum(axis=0)).ravel()
    G = (Bd.T @ Bd).toarray() - np.outer(mu, q) - np.outer(q, mu) + m * np.outer(mu, mu)
    w, V = np.linalg.eigh(G)
    order = np.argsort(w)[::-1]
    return np.sqrt(np.maximum(w[order], 0)), V[:, order]


@pytest.mark.parametrize("precision,tol_s", [("fp32", 1e-5), ("bf16", 1e-4)])
def test_pipeline_pca_vs_gram_oracle(sev, precision, tol_s):
    t = sev.nb_tables(5000, seed=0, density=0.05)
    C = 5000
    p = sev.Pipeline()
    p.synth_counts(C, t)
    p.normalize("lognormalize", 1e4)
    feats = sev.top_mean_features(t, 1000)
    p.select_scale(feats, 10.0)
    B, mu = p.get_scaled()
    it, nm = p.pca(30, init=np.random.default_rng(0).standard_normal(1000), precision=precision)
    s, U, V, Z = p.get_pca()
    sg, Vg = _gram_oracle(B, mu)
    assert np.max(np.abs(s - sg[:30]) / sg[:30]) < tol_s
    if precision == "fp32":
        assert _angle(V[:, :20], Vg[:, :20]) < 1e-3
    assert np.allclose(Z, (U * s).astype(np.float32), rtol=1e-6, atol=1e-6)
    # residual invariant of test/test_irlba.jl on the FP64 centered matrix
    Q = B.toarray().astype(np.float64) - mu.astype(np.float64)
    res = np.linalg.norm(Q.T @ U - V * s) / np.linalg.norm(Q)
    assert res < (1e-5 if precision == "fp32" else 1e-3)


def test_pipeline_pca_fp32_basis_matches_fp64_basis(sev):
    """The BF16 pipeline form stores its left Lanczos basis in FP32; the
    centered IRLBA entry point on the same BF16 operator keeps it in FP64.
    With the same matrix, start vector, basis size and tolerance the two
    agree: singular values to 1e-7 relative, the leading 5-dimensional
    right singular space to 1e-4 rad."""
    t = sev.nb_tables(5000, seed=0, density=0.05)
    C = 5000
    p = sev.Pipeline()
    p.synth_counts(C, t)
    p.normalize("lognormalize", 1e4)
    feats = sev.top_mean_features(t, 1000)
    p.select_scale(feats, 10.0)
    B, mu = p.get_scaled()
    init = np.random.default_rng(0).standard_normal(1000)
    p.pca(30, init=init, precision="bf16", work=60)
    s, U, V, Z = p.get_pca()
    ref = sev.irlba(sev.CenteredMatrix(B, mu), 30, init=init, precision="bf16", work=60)
    assert np.max(np.abs(s - ref.S) / ref.S) < 1e-7
    assert _angle(V[:, :5], ref.V[:, :5]) < 1e-4
    # the residual invariant on the FP64 centered matrix, as for the FP64 basis
    Q = B.toarray().astype(np.float64) - mu.astype(np.float64)
    res = np.linalg.norm(Q.T @ U - V * s) / np.linalg.norm(Q)
    res_ref = np.linalg.norm(Q.T @ ref.U - ref.V * ref.S) / np.linalg.norm(Q)
    assert res < 2 * res_ref + 1e-9


def _clustered_counts(C, G, ncl, seed):
    """Negative-binomial counts of cells from ncl clusters whose gene means
    differ (log-normal, 60% of genes shifted per cluster): the PCA of such
    data has ncl - 1 leading components well separated from each other and
    from the noise."""
    rng = np.random.default_rng(seed)
    base = np.exp(rng.normal(-1.0, 1.2, G))
    means = np.stack([base * np.where(rng.random(G) < 0.6, np.exp(rng.normal(0, 1.0, G)), 1.0)
                      for _ in range(ncl)])
    lab = rng.integers(0, ncl, C)
    size = rng.lognormal(0.0, 0.2, C)[:, None]
    mu = means[lab] * size * 0.6
    r = 1.0 / 0.3
    X = rng.negative_binomial(r, r / (r + mu)).astype(np.int32)
    X[:, 0] += 1  # no empty cell
    return sp.csc_matrix(X)


@pytest.mark.parametrize("precision,tol_s,tol_angle",
                         [("fp32", 1e-5, 1e-4), ("bf16", 1e-4, 1e-2)])
def test_pipeline_pca_singular_vectors_svd_flip(sev, precision, tol_s, tol_angle):
    """Right singular vectors of well-separated components, sign-aligned by
    svd_flip! (src/utils.jl:214-228) on both sides, against the FP64 Gram
    eigenvectors: the angle of each component whose singular value lies 1%
    or more from its neighbours is below 1e-4 rad (FP32 form) / 1e-2 rad
    (BF16 form, whose operator is the BF16-rounded matrix), and the signs
    agree; the leading 5-dimensional subspace is within the same bound."""
    import _gpu_ref as ref
    X = _clustered_counts(6000, 1500, 7, seed=4)
    p = sev.Pipeline()
    p.load_counts(X)
    p.normalize("lognormalize", 1e4)
    p.select_scale(np.arange(1500), 10.0)
    B, mu = p.get_scaled()
    p.pca(20, init=np.random.default_rng(2).standard_normal(1500), precision=precision,
          tol=1e-8 if precision == "fp32" else 1e-6)
    s, U, V, _ = p.get_pca()
    p.close()
    sg, Vg = ref.gram_svd(B, mu, 21)
    assert np.max(np.abs(s - sg[:20]) / sg[:20]) < tol_s
    keep, ang, sign = ref.svd_flip_angles(V, Vg, sg)
    assert len(keep) >= 4, keep
    assert ang.max() < tol_angle and sign.all(), (keep, ang)
    assert ref.angle(V[:, :5], Vg[:, :5]) < tol_angle


@pytest.mark.parametrize("m,n,npcs", [(300, 60, 40), (80, 120, 70), (500, 128, 65)])
def test_pca_dense_route(sev, m, n, npcs):
    """_pca with more than half of the possible components switches to the
    dense SVD (src/embedding.jl:50-53), with the reference's warning: here
    the full Golub-Kahan basis of min(m, n) Lanczos vectors.  All npcs
    singular values match numpy's SVD of the centered matrix (1e-9), the
    embedding is U diag(S) and the loadings span the same subspace."""
    import _gpu_ref as ref
    rng = np.random.default_rng(m + n)
    X = sp.random(m, n, 0.3, format="csc", random_state=m, dtype=np.float64)
    mu = np.asarray(X.mean(axis=0)).ravel()
    Cm = sev.CenteredMatrix(X, mu)
    with pytest.warns(UserWarning, match="standard svd"):
        E = sev.pca(Cm, npcs, init=rng.standard_normal(n))
    Q = X.toarray() - mu
    U, s, Vt = np.linalg.svd(Q, full_matrices=False)
    sd = E.stdev * np.sqrt(max(1, m - 1))
    assert np.allclose(sd, s[:npcs], rtol=1e-9)
    assert ref.angle(E.basis, Vt[:npcs].T) < 1e-6
    assert np.allclose(E.coordinates, Q @ E.basis, rtol=1e-8, atol=1e-9)
    # explicitly requested: the same route without the warning
    E2 = sev.embedding(Cm, 10, algorithm=":svd", init=rng.standard_normal(n))
    assert np.allclose(E2.stdev * np.sqrt(max(1, m - 1)), s[:10], rtol=1e-9)


@pytest.mark.parametrize("work", [0, 64, 80, 128])
def test_pipeline_pca_work_sizes(sev, work):
    """A larger Lanczos basis than the reference's npcs + 7 converges to the
    same tolerance: singular values within 1e-5 of the FP64 Gram oracle, the
    leading right vectors within 1e-3 rad, and the residual below tol."""
    t = sev.nb_tables(4000, seed=3, density=0.05)
    C = 6000
    p = sev.Pipeline()
    p.synth_counts(C, t)
    p.normalize("lognormalize", 1e4)
    feats = sev.top_mean_features(t, 900)
    p.select_scale(feats, 10.0)
    B, mu = p.get_scaled()
    it, nm = p.pca(50, init=np.random.default_rng(1).standard_normal(900), work=work)
    s, U, V, Z = p.get_pca()
    sg, Vg = _gram_oracle(B, mu)
    assert np.max(np.abs(s - sg[:50]) / sg[:50]) < 1e-5
    assert _angle(V[:, :20], Vg[:, :20]) < 1e-3
    Q = B.toarray().astype(np.float64) - mu.astype(np.float64)
    assert np.linalg.norm(Q.T @ U - V * s) / np.linalg.norm(Q) < 1e-5
    p.close()


@pytest.mark.parametrize("m,n,dens", [(3000, 600, 0.5), (9000, 400, 0.3), (4000, 900, 0.08),
                                      (2500, 8000, 0.02), (70000, 50, 0.2)])
def test_centered_mul_layouts(sev, orc, m, n, dens):
    """Row lengths that select every CSR lane-group width, a feature count
    above the LDS limit of A v, and more slabs than column groups."""
    rng = np.random.default_rng(m + n)
    A = sp.random(m, n, dens, format="csc", random_state=m, dtype=np.float64).astype(np.float32)
    mu = rng.random(n).astype(np.float32)
    Cm = sev.CenteredMatrix(A, mu)
    x = rng.standard_normal(n)
    w = rng.standard_normal(m)
    assert np.allclose(Cm @ x, orc.centered_mul(m, A.indptr, A.indices, A.data, mu, x),
                       rtol=1e-12, atol=1e-11)
    assert np.allclose(Cm.rmatvec(w), orc.centered_mul(m, A.indptr, A.indices, A.data, mu, w,
                                                       trans=True), rtol=1e-12, atol=1e-10)


# ---------------------------------------------------------------------------
# matrix and Gram forms of mul! on CenteredMatrix (src/scaling.jl:259-294)
# ---------------------------------------------------------------------------
def _close(got, ref, rtol=1e-12):
    scale = max(np.abs(ref).max(), 1.0)
    assert np.allclose(got, ref, rtol=rtol, atol=rtol * scale)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("m,n,p,dens", [(300, 40, 7, 0.2), (20000, 500, 70, 0.03),
                                        (5000, 3, 1, 0.5)])
def test_centered_mul_mat_vs_oracle(sev, orc, dt, m, n, p, dens):
    """S X and S' X with dense matrices; p = 70 takes two lane passes, an
    empty column and an empty row are planted."""
    rng = np.random.default_rng(m + p)
    A = sp.random(m, n, dens, format="lil", random_state=m, dtype=np.float64)
    A[:, 0] = 0.0
    A[1, :] = 0.0
    A = sp.csc_matrix(A).astype(dt)
    mu = rng.standard_normal(n).astype(dt)
    Cm = sev.CenteredMatrix(A, mu)
    X = rng.standard_normal((n, p))
    W = rng.standard_normal((m, p))
    Y0 = rng.standard_normal((m, p))
    Z0 = rng.standard_normal((n, p))
    _close(Cm @ X, orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, X))
    _close(Cm.mul(Y0.copy(order="F"), X, 2.0, -0.5),
           orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, X, Y=Y0, alpha=2.0, beta=-0.5))
    _close(Cm.mul(np.zeros((n, p), order="F"), W, trans=True),
           orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, W, trans=True))
    _close(Cm.mul(Z0.copy(order="F"), W, 0.5, 3.0, trans=True),
           orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, W, trans=True, Y=Z0,
                                alpha=0.5, beta=3.0))


def test_centered_mul_mat_leading_dimensions(sev, orc):
    """The C ABI with ldx, ldy larger than the rows (submatrices of Julia
    arrays): the padding rows are neither read nor written."""
    import ctypes as C
    rng = np.random.default_rng(3)
    m, n, p = 400, 30, 5
    A = sp.random(m, n, 0.2, format="csc", random_state=3, dtype=np.float64)
    A.sort_indices()
    mu = rng.standard_normal(n)
    Xp = np.asfortranarray(rng.standard_normal((n + 9, p)))
    Yp = np.asfortranarray(np.full((m + 5, p), 7.0))
    cp = (A.indptr.astype(np.int64) + 1)
    rv = (A.indices.astype(np.int64) + 1)
    vals = np.ascontiguousarray(A.data)
    sev._lib.check(sev.load().sev_centered_mul_mat(
        m, n, cp.ctypes.data, rv.ctypes.data, vals.ctypes.data, mu.ctypes.data, 1, b"N", p,
        Xp.ctypes.data, n + 9, Yp.ctypes.data, m + 5, C.c_double(1.0), C.c_double(0.0)))
    _close(Yp[:m], orc.centered_mul_mat(m, A.indptr, A.indices, A.data, mu, Xp[:n]))
    assert np.all(Yp[m:] == 7.0)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("case", ["self", "pair", "wide", "uncentered"])
def test_centered_gram_vs_oracle(sev, orc, dt, case):
    """Sl' Sr for the Gram matrix of one operand, two operands with the same
    rows, a right operand wider than one LDS pass (5000 columns) and an
    uncentered right operand."""
    rng = np.random.default_rng(len(case))
    m = 30000
    Al = sp.random(m, 200, 0.05, format="csc", random_state=1, dtype=np.float64).astype(dt)
    mul = rng.standard_normal(200).astype(dt)
    Sl = sev.CenteredMatrix(Al, mul)
    if case == "self":
        got = Sl.gram(alpha=0.25)
        ref = orc.centered_gram(m, Al.indptr, Al.indices, Al.data, mul, alpha=0.25)
    else:
        nr = 5000 if case == "wide" else 120
        Ar = sp.random(m, nr, 0.004 if case == "wide" else 0.1, format="csc", random_state=2,
                       dtype=np.float64).astype(dt)
        mur = rng.standard_normal(nr).astype(dt)
        if case == "uncentered":
            mur = np.zeros(nr, dt)
        C0 = rng.standard_normal((200, nr))
        got = Sl.gram(sev.CenteredMatrix(Ar, mur), C=C0.copy(order="F"), alpha=1.5, beta=-0.5)
        ref = orc.centered_gram(m, Al.indptr, Al.indices, Al.data, mul, Ar.indptr, Ar.indices,
                                Ar.data, mur, C=C0, alpha=1.5, beta=-0.5)
    _close(got, ref, 1e-11)


# ---------------------------------------------------------------------------
# libcell's callback irlba entry point (src/irlba.jl:22-99)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,n,nu", [(50, 50, 20), (100, 50, 20)])
def test_irlba_callback_dense(sev, m, n, nu):
    """The dense cases of test/test_irlba.jl through host matmul callbacks."""
    from scipy.sparse.linalg import aslinearoperator
    rng = np.random.default_rng(m + 1)
    X = rng.standard_normal((m, n))
    S = sev.irlba_operator(aslinearoperator(X), nu, tol=1e-5, init=rng.standard_normal(n))
    ss = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(S.S, ss[:nu], rtol=1e-8)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X) < 1e-5


def test_irlba_callback_nu58(sev):
    """nu = 58 (work 65) through libcell's callback entry point, which
    Severo's irlba!/pca reach with any npcs (src/irlba.jl:47-76)."""
    from scipy.sparse.linalg import aslinearoperator
    rng = np.random.default_rng(58)
    m, n, nu = 500, 200, 58
    X = rng.standard_normal((m, n)) * (1.0 + 2.0 * np.exp(-np.arange(n) / 30.0))
    S = sev.irlba_operator(aslinearoperator(X), nu, tol=1e-6, init=rng.standard_normal(n))
    ss = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(S.S, ss[:nu], rtol=1e-8)
    assert np.linalg.norm(X.T @ S.U - S.V * S.S) / np.linalg.norm(X) < 1e-6


def test_irlba_callback_matches_device_operator(sev):
    """Same centered matrix and start: the callback path and the device
    operator agree to the FP64 product rounding."""
    X = sp.random(3000, 400, 0.1, format="csc", random_state=21, dtype=np.float64)
    mu = np.asarray(X.mean(axis=0)).ravel()
    Cm = sev.CenteredMatrix(X, mu)
    init = np.random.default_rng(5).standard_normal(400)
    a = sev.irlba(Cm, 8, tol=1e-9, init=init)
    b = sev.irlba_operator(Cm, 8, tol=1e-9, init=init)
    assert np.allclose(a.S, b.S, rtol=1e-11)
    Q = X.toarray() - mu
    assert np.linalg.norm(Q.T @ b.U - b.V * b.S) / np.linalg.norm(Q) < 1e-9
    assert _angle(a.U, b.U) < 1e-6


def test_irlba_callback_restart(sev):
    """irlba(A, nu, S) warm start (test/test_irlba.jl "restart"; the
    reference marks the third value @test_broken, here it must hold)."""
    rng = np.random.default_rng(12)
    X = rng.standard_normal((20, 10))
    ss = np.linalg.svd(X, compute_uv=False)
    S2 = sev.irlba_operator(X, 2, tol=1e-5, init=rng.standard_normal(10))
    assert np.allclose(S2.S, ss[:2], rtol=1e-8)
    S3 = sev.irlba_operator(X, 3, S2, tol=1e-5, init=rng.standard_normal(10))
    assert np.allclose(S3.S, ss[:3], rtol=1e-8)
    Y = sp.random(3000, 300, 0.05, format="csc", random_state=4, dtype=np.float64)
    sy = np.linalg.svd(Y.toarray(), compute_uv=False)
    S5 = sev.irlba_operator(Y, 5, tol=1e-9, init=rng.standard_normal(300))
    S12 = sev.irlba_operator(Y, 12, S5, tol=1e-9, init=rng.standard_normal(300))
    assert np.allclose(S12.S, sy[:12], rtol=1e-8)
    Yd = Y.toarray()
    assert np.linalg.norm(Yd.T @ S12.U - S12.V * S12.S) / np.linalg.norm(Yd) < 1e-8


def test_irlba_callback_randv_and_maxit(sev):
    """init = NULL draws the start from randv once; maxit exhaustion returns
    the convergence status."""
    import ctypes as C
    lib = sev._lib
    rng = np.random.default_rng(3)
    X = np.asfortranarray(rng.standard_normal((60, 30)))
    m, n, nu = 60, 30, 4
    calls = []

    def mm(yp, trans, xp, tp, data):
        if trans == b"T":
            np.ctypeslib.as_array(yp, (n,))[:] = X.T @ np.ctypeslib.as_array(xp, (m,))
        else:
            np.ctypeslib.as_array(yp, (m,))[:] = X @ np.ctypeslib.as_array(xp, (n,))

    def rv(yp, k, data):
        calls.append(k)
        np.ctypeslib.as_array(yp, (k,))[:] = rng.standard_normal(k)

    mcb, rcb = lib.MATMUL_FN(mm), lib.RANDV_FN(rv)
    s = np.zeros(nu)
    U = np.zeros((m, nu), order="F")
    V = np.zeros((n, nu), order="F")
    L = sev.load()
    rc = L.irlba(m, n, nu, nu + 7, 1000, 0, 1e-8, None, s.ctypes.data, U.ctypes.data,
                 V.ctypes.data, C.cast(rcb, C.c_void_p), C.cast(mcb, C.c_void_p), None)
    assert rc == 0 and calls == [n]
    assert np.allclose(s, np.linalg.svd(X, compute_uv=False)[:nu], rtol=1e-8)
    init = rng.standard_normal(n)
    s2 = np.zeros(20)
    U2 = np.zeros((m, 20), order="F")
    V2 = np.zeros((n, 20), order="F")
    rc = L.irlba(m, n, 20, 27, 1, 0, 1e-14, init.ctypes.data, s2.ctypes.data, U2.ctypes.data,
                 V2.ctypes.data, None, C.cast(mcb, C.c_void_p), None)
    assert rc == -3


def _bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("m,n,dens", [(3000, 600, 0.5), (9000, 400, 0.3), (4000, 900, 0.08),
                                      (2500, 8000, 0.02), (2000, 15000, 0.01)])
def test_irlba_bf16_form_layouts(sev, m, n, dens):
    """The BF16 form decomposes the BF16-rounded centered matrix with FP32
    products: singular values within 1e-5 relative of LAPACK on that matrix,
    for row lengths that select every lane width and feature counts above the
    LDS copy of v."""
    rng = np.random.default_rng(m * 3 + n)
    A = sp.random(m, n, dens, format="csc", random_state=m + 1, dtype=np.float64)
    A = A.astype(np.float32)
    mu = np.asarray(A.mean(axis=0), np.float32).ravel()
    nu = 6
    S = sev.irlba(sev.CenteredMatrix(A, mu), nu, tol=1e-7, init=rng.standard_normal(n),
                  precision="bf16")
    Ab = A.copy().astype(np.float64)
    Ab.data = _bf16_round(A.data)
    Q = Ab.toarray() - mu.astype(np.float64)
    ss = np.linalg.svd(Q, compute_uv=False)
    assert np.max(np.abs(S.S - ss[:nu]) / ss[:nu]) < 1e-5
