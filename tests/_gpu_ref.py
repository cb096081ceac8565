"""FP64 references for the GPU parity tests at full size, computed on the
host and independent of the library's kernels: the centered Gram matrix
(dense row blocks through numpy's BLAS) and the IRLBA residual of
test/test_irlba.jl (||A'U - V S||_F / ||A||_F) through the oracle's threaded
A'U, for A = B - 1 mu' with B a scipy CSC matrix (cells x features)."""
import numpy as np


def centered_fro2(B, mu):
    """||A||_F^2 = ||B||^2 - 2 mu.q + m ||mu||^2 with q the column sums."""
    mu = np.asarray(mu, np.float64)
    d = B.data.astype(np.float64)
    q = np.zeros(B.shape[1])
    nz = np.diff(B.indptr) > 0
    if B.nnz:
        q[nz] = np.add.reduceat(d, B.indptr[:-1][nz])
    return float(d @ d - 2.0 * mu @ q + B.shape[0] * (mu @ mu))


def residual(orc, B, mu, s, U, V):
    """||A'U - V diag(s)||_F / ||A||_F (test/test_irlba.jl:30)."""
    AtU = orc.centered_tmul_mat(B.shape[0], B.indptr, B.indices, B.data, mu, U)
    R = AtU - V * s[None, :]
    return float(np.linalg.norm(R) / np.sqrt(centered_fro2(B, mu)))


def gram(B, mu, rows_per_block=32768):
    """Centered Gram matrix A'A in FP64 from dense row blocks of B
    (B'B - mu q' - q mu' + m mu mu', src/scaling.jl:274-296)."""
    m, F = B.shape
    Bc = B.tocsr()
    G = np.zeros((F, F))
    q = np.zeros(F)
    for r0 in range(0, m, rows_per_block):
        D = Bc[r0:r0 + rows_per_block].astype(np.float64).toarray()
        G += D.T @ D
        q += D.sum(axis=0)
        del D
    mu = np.asarray(mu, np.float64)
    return G - np.outer(mu, q) - np.outer(q, mu) + m * np.outer(mu, mu)


def gram_svd(B, mu, npcs):
    """Singular values and right singular vectors of A from its Gram matrix
    (FP64 eigendecomposition, no Lanczos involved)."""
    w, V = np.linalg.eigh(gram(B, mu))
    order = np.argsort(w)[::-1][:npcs]
    return np.sqrt(np.maximum(w[order], 0.0)), V[:, order]


def angle(U1, U2):
    """Largest principal angle between the column spaces (sign-free)."""
    q1, _ = np.linalg.qr(U1)
    q2, _ = np.linalg.qr(U2)
    c = np.linalg.svd(q1.T @ q2, compute_uv=False)
    return float(np.arccos(np.clip(c.min(), -1.0, 1.0)))


def svd_flip_angles(V, Vref, s_ref, min_gap=1e-2):
    """Per-component angles between right singular vectors sign-aligned by
    svd_flip!(S; u_based_decision=false) (src/utils.jl:214-228: the entry of
    largest magnitude of each column made positive), for the components whose
    reference singular value lies at least min_gap (relative) from both
    neighbours.  Returns (indices, angles, signs agree)."""
    n = V.shape[1]
    s = np.asarray(s_ref, np.float64)
    up = np.full(n, np.inf)
    up[1:] = (s[:n - 1] - s[1:n]) / s[1:n]
    down = np.full(n, np.inf)
    if len(s) > n:
        down[:] = (s[:n] - s[1:n + 1]) / s[:n]
    else:
        down[:n - 1] = (s[:n - 1] - s[1:n]) / s[:n - 1]
    gap = np.minimum(up, down)

    def flip(M):
        i = np.argmax(np.abs(M), axis=0)
        return M * np.sign(M[i, np.arange(M.shape[1])])[None, :]
    a, b = flip(np.asarray(V, np.float64)), flip(np.asarray(Vref, np.float64)[:, :n])
    keep = np.nonzero(gap > min_gap)[0]
    cos = np.array([a[:, i] @ b[:, i] / (np.linalg.norm(a[:, i]) * np.linalg.norm(b[:, i]))
                    for i in keep])
    return keep, np.arccos(np.clip(np.abs(cos), -1.0, 1.0)), cos > 0
