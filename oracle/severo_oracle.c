/*
 * severo_oracle.c -- CPU oracle for the Severo embedding-and-graph path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it.  The product
 * (libsevero_hip.so) never links, calls or falls back to it.
 *
 * Each function restates the reference algorithm, citing the Severo.jl
 * file:line it follows.  Built with -ffp-contract=off so that every FP32
 * operation is rounded exactly where the Julia source rounds it (Julia never
 * contracts a*b+c into an FMA).
 *
 * Pinning (what checks each function independently of this file):
 *   so_synth_*          bit-identical to the device generator; distribution
 *                       checked against the negative-binomial pmf (tests).
 *   so_normalize_*      10x5 fixture of test/test_input.jl:47-58, row-sum
 *                       invariants (:64,:74), numpy float32 closed form.
 *   so_scale_*          closed form (X-mu)/sigma of test/test_scaling.jl:41-44,
 *                       a pure-Python float32 Welford loop, numpy clip.
 *   so_centered_mul     known-answer vectors of test/test_scaling.jl:70-113
 *                       (dense closed form computed by numpy).
 *   so_centered_gram    numpy (A - 1 mu')'(B - 1 nu') of the densified
 *                       operands, alpha/beta, uncentered and self cases.
 *   so_irlba_*          numpy/LAPACK dense SVD, residual invariant
 *                       ||A'U - V S||_F / ||A||_F < tol (test/test_irlba.jl).
 *   so_knn_*            numpy float32 brute force with stable argsort,
 *                       crafted duplicate-point ties.
 *   so_snn_jaccard      Python set intersections on small graphs.
 *   so_count_mean_std   numpy mean / var(ddof=1) of the densified columns.
 *   so_rel_mean_var     numpy mean / var(ddof=1) of the densified relative
 *                       counts, including all-zero genes.
 *   so_loess_fit        exact reproduction of quadratics, and a numpy
 *                       weighted least-squares fit (lstsq) per point.
 *   so_standardized_var numpy closed form of the clipped standardized values.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define SO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Counter-based random numbers shared with the device generator.           */
/* ------------------------------------------------------------------------ */

static inline uint64_t so_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* uniform in [0,1) for (seed, stream, index); stream = gene, index = cell */
SO_EXPORT double so_u01(uint64_t seed, uint32_t stream, uint32_t index)
{
    uint64_t h = so_mix64(so_mix64(seed) ^ (((uint64_t)stream << 32) | (uint64_t)index));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

#define SO_CLASS_STREAM 0x7FFFFFFFu
#define SO_NB_XMAX 100000

/* Negative-binomial draw by inverse CDF from one uniform.  p0 = P(X=0),
 * q = mu/(r+mu); P(X=x+1) = P(X=x) * q * (x+r)/(x+1).  Only +,*,/ on
 * doubles, so the device version reproduces it bit for bit. */
static inline int32_t so_nb_draw(double u, double p0, double q, double r)
{
    if (u < p0) return 0;
    double p = p0, cdf = p0;
    int32_t x = 0;
    while (u >= cdf && x < SO_NB_XMAX) {
        p = ((p * q) * ((double)x + r)) / ((double)x + 1.0);
        x += 1;
        cdf = cdf + p;
    }
    return x;
}

/* Synthetic cells x genes CSC count matrix (0-based int32 rows).  Column g
 * draws every cell c with the per-(gene, class) tables p0/q (length 2G,
 * index 2g+class); a cell is in class 1 (boosted library size) when
 * u01(seed, CLASS_STREAM, c) < boost_frac.  With rowval == NULL only colptr
 * is filled.  Returns nnz. */
SO_EXPORT int64_t so_synth_nb_csc(int64_t C, int64_t G, uint64_t seed, double r,
                                  const double *p0, const double *q, double boost_frac,
                                  int64_t *colptr, int32_t *rowval, int32_t *nzval)
{
    unsigned char *cls = (unsigned char *)malloc((size_t)(C > 0 ? C : 1));
    for (int64_t c = 0; c < C; ++c)
        cls[c] = so_u01(seed, SO_CLASS_STREAM, (uint32_t)c) < boost_frac ? 1 : 0;
    int64_t nnz = 0;
    colptr[0] = 0;
    for (int64_t g = 0; g < G; ++g) {
        for (int64_t c = 0; c < C; ++c) {
            int k = cls[c];
            double u = so_u01(seed, (uint32_t)g, (uint32_t)c);
            int32_t x = so_nb_draw(u, p0[2 * g + k], q[2 * g + k], r);
            if (x != 0) {
                if (rowval) {
                    rowval[nnz] = (int32_t)c;
                    nzval[nnz] = x;
                }
                nnz++;
            }
        }
        colptr[g + 1] = nnz;
    }
    free(cls);
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* normalize_cells -> log_norm -> row_norm   (src/normalize.jl:17-55)        */
/* ------------------------------------------------------------------------ */

/* Per-cell totals s = sum(A, dims=2) (src/normalize.jl:24): integer sums. */
static int64_t *so_row_totals(int64_t m, int64_t n, const int64_t *colptr,
                              const int32_t *rowval, const int32_t *counts)
{
    int64_t *s = (int64_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) s[rowval[p]] += counts[p];
    return s;
}

/* FP32 form: nzB = sf * nzA / s[row] (src/normalize.jl:27) evaluated left to
 * right in Float32 (Int32 count and Int64 total promote to Float32), then
 * log1p (src/normalize.jl:36).  log1p(Float32) is defined here as the
 * double-precision log1p rounded to float. */
SO_EXPORT void so_normalize_f32(int64_t m, int64_t n, const int64_t *colptr,
                                const int32_t *rowval, const int32_t *counts, float sf,
                                int do_log, float *out)
{
    int64_t *s = so_row_totals(m, n, colptr, rowval, counts);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            float v = (sf * (float)counts[p]) / (float)s[rowval[p]];
            out[p] = do_log ? (float)log1p((double)v) : v;
        }
    free(s);
}

SO_EXPORT void so_normalize_f64(int64_t m, int64_t n, const int64_t *colptr,
                                const int32_t *rowval, const int32_t *counts, double sf,
                                int do_log, double *out)
{
    int64_t *s = so_row_totals(m, n, colptr, rowval, counts);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            double v = (sf * (double)counts[p]) / (double)s[rowval[p]];
            out[p] = do_log ? log1p(v) : v;
        }
    free(s);
}

/* Column selection Y[:, hvf] keeping hvf order (docs/src/pbmc.md:121,
 * src/scaling.jl:350-352).  feats are 0-based.  With rowval_out == NULL only
 * colptr_out is filled.  Returns nnz of the selection. */
SO_EXPORT int64_t so_gather_cols(int64_t nsel, const int64_t *feats, const int64_t *colptr,
                                 const int32_t *rowval, const float *vals, int64_t *colptr_out,
                                 int32_t *rowval_out, float *vals_out)
{
    colptr_out[0] = 0;
    for (int64_t j = 0; j < nsel; ++j) {
        int64_t g = feats[j], len = colptr[g + 1] - colptr[g];
        if (rowval_out) {
            memcpy(rowval_out + colptr_out[j], rowval + colptr[g], (size_t)len * sizeof(int32_t));
            memcpy(vals_out + colptr_out[j], vals + colptr[g], (size_t)len * sizeof(float));
        }
        colptr_out[j + 1] = colptr_out[j] + len;
    }
    return colptr_out[nsel];
}

/* ------------------------------------------------------------------------ */
/* scale_features -> scale_data -> mean_std -> mean_var                     */
/* (src/scaling.jl:18-34, 36-44, 199-217, 347-357)                           */
/* ------------------------------------------------------------------------ */

/* Float32: Welford over the nonzeros in row order, starting from
 * count = m - nnz (the zeros folded in, :21), var = s/(m-1) (:31),
 * sigma = sqrt(var); mu' = mu/sigma (:207); clip = scale_max + mu' (:209);
 * b = min(x/sigma, clip) with only the upper clip (:210-213). */
SO_EXPORT void so_scale_f32(int64_t m, int64_t n, const int64_t *colptr, const float *x,
                            float scale_max, float *out, float *mu_out, float *sigma_out)
{
    for (int64_t j = 0; j < n; ++j) {
        int64_t count = m - (colptr[j + 1] - colptr[j]);
        float mu = 0.0f, s = 0.0f;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            float v = x[p];
            count += 1;
            float delta = v - mu;
            mu = mu + delta / (float)count;
            s = s + delta * (v - mu);
        }
        float var = s / (float)(m - 1);
        float sd = sqrtf(var);
        float mup = mu / sd;
        float clip = scale_max + mup;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            float v = x[p] / sd;
            out[p] = (v > clip) ? clip : v;
        }
        mu_out[j] = mup;
        if (sigma_out) sigma_out[j] = sd;
    }
}

SO_EXPORT void so_scale_f64(int64_t m, int64_t n, const int64_t *colptr, const double *x,
                            double scale_max, double *out, double *mu_out, double *sigma_out)
{
    for (int64_t j = 0; j < n; ++j) {
        int64_t count = m - (colptr[j + 1] - colptr[j]);
        double mu = 0.0, s = 0.0;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            double v = x[p];
            count += 1;
            double delta = v - mu;
            mu = mu + delta / (double)count;
            s = s + delta * (v - mu);
        }
        double var = s / (double)(m - 1);
        double sd = sqrt(var);
        double mup = mu / sd;
        double clip = scale_max + mup;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            double v = x[p] / sd;
            out[p] = (v > clip) ? clip : v;
        }
        mu_out[j] = mup;
        if (sigma_out) sigma_out[j] = sd;
    }
}

/* ------------------------------------------------------------------------ */
/* CenteredMatrix mul!  (src/scaling.jl:245-257)                              */
/* ------------------------------------------------------------------------ */

/* trans == 0:  y = alpha*(A x - (mu.x) 1) + beta*y    (length m)
 * trans == 1:  y = alpha*(A' x - (sum x) mu) + beta*y (length n)
 * A is m x n CSC with float (f64 == 0) or double (f64 == 1) values; mu may be
 * NULL (plain sparse product).  The CSC loops follow SparseArrays' generic
 * mul! (scatter with x[col]*alpha for A x; per-column dot then *alpha for
 * A' x); dot and sum are sequential here. */
SO_EXPORT void so_centered_mul(int64_t m, int64_t n, const int64_t *colptr,
                               const int32_t *rowval, const void *nzval, int f64,
                               const void *mu, int trans, const double *x, double *y,
                               double alpha, double beta)
{
    const float *vf = (const float *)nzval;
    const double *vd = (const double *)nzval;
    const float *muf = (const float *)mu;
    const double *mud = (const double *)mu;
#define SO_VAL(p) (f64 ? vd[p] : (double)vf[p])
#define SO_MU(i) (f64 ? mud[i] : (double)muf[i])
    if (!trans) {
        for (int64_t i = 0; i < m; ++i) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
        for (int64_t c = 0; c < n; ++c) {
            double ax = x[c] * alpha;
            for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) y[rowval[p]] += SO_VAL(p) * ax;
        }
        if (mu) {
            double d = 0.0;
            for (int64_t c = 0; c < n; ++c) d += SO_MU(c) * x[c];
            double t = alpha * d;
            for (int64_t i = 0; i < m; ++i) y[i] -= t;
        }
    } else {
        for (int64_t c = 0; c < n; ++c) y[c] = (beta == 0.0) ? 0.0 : beta * y[c];
        for (int64_t c = 0; c < n; ++c) {
            double tmp = 0.0;
            for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) tmp += SO_VAL(p) * x[rowval[p]];
            y[c] += tmp * alpha;
        }
        if (mu) {
            double sx = 0.0;
            for (int64_t i = 0; i < m; ++i) sx += x[i];
            double a = -alpha * sx;
            for (int64_t c = 0; c < n; ++c) y[c] += a * SO_MU(c);
        }
    }
#undef SO_VAL
#undef SO_MU
}

/* C (nl x nr, column-major) = alpha (Al - 1 mul')'(Ar - 1 mur') + beta C,
 * in the steps of mul!(C, Sl', Sr, α, β) (src/scaling.jl:273-294):
 *   1. C = β C, then the sparse-sparse loop of src/mul.jl:82-113 on
 *      copy(Al') and Ar: for each column c of Ar and each stored (j, b) of
 *      it (rows ascending), each stored (r, a) of row j of Al:
 *      C[r, c] += (α a) b;
 *   2. q = colsums(Ar); C[r, c] += (-α) (mul[r] q[c]);
 *   3. q' = colsums(Al) (q when Sl is Sr) - m mul; C[r, c] += (-α) (q'[r] mur[c]).
 * Column sums are sequential over the stored entries.  mul or mur NULL: that
 * matrix is uncentered.  The rows of Al are gathered by a counting sort. */
SO_EXPORT int so_centered_gram(int64_t m, int64_t nl, const int64_t *cpl, const int32_t *rvl,
                               const void *vl, const void *mul, int64_t nr, const int64_t *cpr,
                               const int32_t *rvr, const void *vr, const void *mur, int f64,
                               double *C, double alpha, double beta)
{
#define SO_V(v, p) (f64 ? ((const double *)(v))[p] : (double)((const float *)(v))[p])
    const int64_t nnzl = cpl[nl];
    int64_t *rp = (int64_t *)calloc((size_t)m + 1, sizeof(int64_t));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)m + 1));
    int64_t *rc = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnzl > 0 ? nnzl : 1));
    double *ra = (double *)malloc(sizeof(double) * (size_t)(nnzl > 0 ? nnzl : 1));
    double *q = (double *)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1));
    double *ql = (double *)malloc(sizeof(double) * (size_t)(nl > 0 ? nl : 1));
    if (!rp || !fill || !rc || !ra || !q || !ql) {
        free(rp); free(fill); free(rc); free(ra); free(q); free(ql);
        return 1;
    }
    for (int64_t e = 0; e < nnzl; ++e) rp[rvl[e] + 1]++;
    for (int64_t i = 0; i < m; ++i) rp[i + 1] += rp[i];
    memcpy(fill, rp, sizeof(int64_t) * ((size_t)m + 1));
    for (int64_t c = 0; c < nl; ++c)
        for (int64_t e = cpl[c]; e < cpl[c + 1]; ++e) {
            const int64_t u = fill[rvl[e]]++;
            rc[u] = c;
            ra[u] = SO_V(vl, e);
        }
    for (int64_t e = 0; e < nl * nr; ++e) C[e] = beta == 0.0 ? 0.0 : beta * C[e];
    for (int64_t c = 0; c < nr; ++c)
        for (int64_t t = cpr[c]; t < cpr[c + 1]; ++t) {
            const int64_t j = rvr[t];
            const double b = SO_V(vr, t);
            for (int64_t u = rp[j]; u < rp[j + 1]; ++u) C[c * nl + rc[u]] += (alpha * ra[u]) * b;
        }
    for (int64_t c = 0; c < nr; ++c) {
        double a = 0.0;
        for (int64_t t = cpr[c]; t < cpr[c + 1]; ++t) a += SO_V(vr, t);
        q[c] = a;
    }
    for (int64_t r = 0; r < nl; ++r) {
        double a = 0.0;
        for (int64_t t = cpl[r]; t < cpl[r + 1]; ++t) a += SO_V(vl, t);
        ql[r] = a - (mul ? (double)m * SO_V(mul, r) : 0.0);
    }
    if (mul)
        for (int64_t c = 0; c < nr; ++c)
            for (int64_t r = 0; r < nl; ++r) C[c * nl + r] += -alpha * (SO_V(mul, r) * q[c]);
    if (mur)
        for (int64_t c = 0; c < nr; ++c)
            for (int64_t r = 0; r < nl; ++r) C[c * nl + r] += -alpha * (ql[r] * SO_V(mur, c));
    free(rp); free(fill); free(rc); free(ra); free(q); free(ql);
    return 0;
#undef SO_V
}

/* ------------------------------------------------------------------------ */
/* IRLBA: augmented implicitly restarted Lanczos bidiagonalisation           */
/* (Baglama & Reichel, SIAM J. Sci. Comput. 27(1), 2005), in the form the    */
/* native irlba behind src/irlba.jl:47-99 exposes: work = nu+7 Lanczos       */
/* vectors, one classical Gram-Schmidt pass of full reorthogonalisation,     */
/* restart size k = nu + (#converged) capped at work-3, convergence when     */
/* nu Ritz values have |residual| < tol*max(sigma) and relative change       */
/* < svtol between restarts.  FP64 throughout.                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t m, n;
    /* y = A x (trans 0, x length n, y length m) or y = A' x (trans 1) */
    void (*apply)(void *ctx, int trans, const double *x, double *y);
    void *ctx;
} so_op;

static double so_nrm2(int64_t n, const double *x)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}

/* y -= X (X' y) over the first ncol columns of X (leading dim ld). */
static void so_orthog(int64_t len, int64_t ncol, const double *X, int64_t ld, double *y,
                      double *t)
{
    for (int64_t c = 0; c < ncol; ++c) {
        double d = 0.0;
        const double *xc = X + c * ld;
        for (int64_t i = 0; i < len; ++i) d += xc[i] * y[i];
        t[c] = d;
    }
    for (int64_t c = 0; c < ncol; ++c) {
        const double *xc = X + c * ld;
        double tc = t[c];
        for (int64_t i = 0; i < len; ++i) y[i] -= xc[i] * tc;
    }
}

/* Deterministic standard-normal-like vector (Box-Muller over so_u01). */
static void so_randn(int64_t n, uint64_t seed, double *y)
{
    for (int64_t i = 0; i < n; ++i) {
        double u1 = so_u01(seed, 0x51u, (uint32_t)i);
        double u2 = so_u01(seed, 0x52u, (uint32_t)i);
        if (u1 < 1e-300) u1 = 1e-300;
        y[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

/* One-sided (Hestenes) Jacobi SVD of the n x n column-major matrix A.
 * Outputs U (n x n), s (descending), V (n x n). */
SO_EXPORT void so_svd_jacobi(int64_t n, const double *A, double *U, double *s, double *V)
{
    double *W = (double *)malloc((size_t)(n * n) * sizeof(double));
    double *Vt = (double *)malloc((size_t)(n * n) * sizeof(double));
    memcpy(W, A, (size_t)(n * n) * sizeof(double));
    for (int64_t i = 0; i < n * n; ++i) Vt[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) Vt[i * n + i] = 1.0;
    const double eps = 2.220446049250313e-16;
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0.0;
        for (int64_t p = 0; p < n - 1; ++p)
            for (int64_t q = p + 1; q < n; ++q) {
                double a = 0.0, b = 0.0, g = 0.0;
                double *wp = W + p * n, *wq = W + q * n;
                for (int64_t i = 0; i < n; ++i) {
                    a += wp[i] * wp[i];
                    b += wq[i] * wq[i];
                    g += wp[i] * wq[i];
                }
                if (g == 0.0 || a == 0.0 || b == 0.0) continue;
                double rel = fabs(g) / sqrt(a * b);
                if (rel > off) off = rel;
                if (rel <= eps) continue;
                double zeta = (b - a) / (2.0 * g);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                for (int64_t i = 0; i < n; ++i) {
                    double x = wp[i], y = wq[i];
                    wp[i] = c * x - sn * y;
                    wq[i] = sn * x + c * y;
                }
                double *vp = Vt + p * n, *vq = Vt + q * n;
                for (int64_t i = 0; i < n; ++i) {
                    double x = vp[i], y = vq[i];
                    vp[i] = c * x - sn * y;
                    vq[i] = sn * x + c * y;
                }
            }
        if (off <= eps) break;
    }
    /* singular values and left vectors, then sort descending */
    double *sv = (double *)malloc((size_t)n * sizeof(double));
    int64_t *ord = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        sv[i] = so_nrm2(n, W + i * n);
        ord[i] = i;
    }
    for (int64_t i = 1; i < n; ++i) { /* insertion sort, stable, descending */
        int64_t o = ord[i], j = i - 1;
        while (j >= 0 && sv[ord[j]] < sv[o]) {
            ord[j + 1] = ord[j];
            --j;
        }
        ord[j + 1] = o;
    }
    for (int64_t c = 0; c < n; ++c) {
        int64_t o = ord[c];
        s[c] = sv[o];
        memcpy(V + c * n, Vt + o * n, (size_t)n * sizeof(double));
        for (int64_t i = 0; i < n; ++i) U[c * n + i] = sv[o] > 0 ? W[o * n + i] / sv[o] : 0.0;
    }
    /* complete U for zero singular values (Gram-Schmidt on unit vectors) */
    double *t = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t c = 0; c < n; ++c) {
        if (s[c] > 0) continue;
        for (int64_t e = 0; e < n; ++e) {
            double *u = U + c * n;
            for (int64_t i = 0; i < n; ++i) u[i] = (i == e) ? 1.0 : 0.0;
            /* orthogonalise against every other nonzero column (twice) */
            for (int pass = 0; pass < 2; ++pass)
                for (int64_t o = 0; o < n; ++o) {
                    if (o == c || (s[o] <= 0 && o > c)) continue;
                    double d = 0.0;
                    for (int64_t i = 0; i < n; ++i) d += U[o * n + i] * u[i];
                    for (int64_t i = 0; i < n; ++i) u[i] -= d * U[o * n + i];
                }
            double nr = so_nrm2(n, u);
            if (nr > 0.5) {
                for (int64_t i = 0; i < n; ++i) u[i] /= nr;
                break;
            }
        }
    }
    free(t);
    free(sv);
    free(ord);
    free(W);
    free(Vt);
}

/* Returns 0 on success, -1 starting vector is zero, -2 no convergence. */
static int so_irlba_core(const so_op *op, int64_t nu, int64_t work, int64_t maxit, double tol,
                         double svtol, const double *init, double *s, double *U, double *V,
                         int64_t *iters_out, int64_t *nmult_out)
{
    const int64_t m = op->m, n = op->n;
    const double eps = 2.220446049250313e-16;
    double *W = (double *)calloc((size_t)(m * work), sizeof(double));
    double *Vb = (double *)calloc((size_t)(n * work), sizeof(double));
    double *F = (double *)calloc((size_t)n, sizeof(double));
    double *B = (double *)calloc((size_t)(work * work), sizeof(double));
    double *BU = (double *)calloc((size_t)(work * work), sizeof(double));
    double *BV = (double *)calloc((size_t)(work * work), sizeof(double));
    double *BS = (double *)calloc((size_t)work, sizeof(double));
    double *res = (double *)calloc((size_t)work, sizeof(double));
    double *svr = (double *)calloc((size_t)work, sizeof(double));
    double *t = (double *)calloc((size_t)work, sizeof(double));
    int64_t big = m > n ? m : n;
    double *tmp = (double *)calloc((size_t)(big * work), sizeof(double));
    int ret = -2;
    int64_t iter = 0, nmult = 0, k = nu;
    double S = 0.0, RF = 0.0;

    memcpy(Vb, init, (size_t)n * sizeof(double));
    {
        double d = so_nrm2(n, Vb);
        if (d < 2 * eps) {
            ret = -1;
            goto done;
        }
        for (int64_t i = 0; i < n; ++i) Vb[i] /= d;
    }
    while (iter < maxit) {
        int64_t j = (iter == 0) ? 0 : k;
        /* W[:,j] = A V[:,j]; reorthogonalise after a restart */
        op->apply(op->ctx, 0, Vb + j * n, W + j * m);
        nmult++;
        if (iter > 0) so_orthog(m, j, W, m, W + j * m, t);
        S = so_nrm2(m, W + j * m);
        if (S < 2 * eps) {
            ret = -3;
            goto done;
        }
        for (int64_t i = 0; i < m; ++i) W[j * m + i] /= S;
        for (; j < work; ++j) {
            op->apply(op->ctx, 1, W + j * m, F);
            nmult++;
            for (int64_t i = 0; i < n; ++i) F[i] -= S * Vb[j * n + i];
            so_orthog(n, j + 1, Vb, n, F, t);
            if (j + 1 < work) {
                RF = so_nrm2(n, F);
                if (RF < 2 * eps) {
                    so_randn(n, (uint64_t)(iter * work + j) + 17u, F);
                    so_orthog(n, j + 1, Vb, n, F, t);
                    double d = so_nrm2(n, F);
                    for (int64_t i = 0; i < n; ++i) Vb[(j + 1) * n + i] = F[i] / d;
                    RF = 0.0;
                } else {
                    for (int64_t i = 0; i < n; ++i) Vb[(j + 1) * n + i] = F[i] / RF;
                }
                B[j * work + j] = S;        /* B(j,j)   col-major */
                B[(j + 1) * work + j] = RF; /* B(j,j+1) */
                op->apply(op->ctx, 0, Vb + (j + 1) * n, W + (j + 1) * m);
                nmult++;
                for (int64_t i = 0; i < m; ++i) W[(j + 1) * m + i] -= RF * W[j * m + i];
                so_orthog(m, j + 1, W, m, W + (j + 1) * m, t);
                S = so_nrm2(m, W + (j + 1) * m);
                if (S < 2 * eps) {
                    so_randn(m, (uint64_t)(iter * work + j) + 91u, W + (j + 1) * m);
                    so_orthog(m, j + 1, W, m, W + (j + 1) * m, t);
                    double d = so_nrm2(m, W + (j + 1) * m);
                    for (int64_t i = 0; i < m; ++i) W[(j + 1) * m + i] /= d;
                    S = 0.0;
                } else {
                    for (int64_t i = 0; i < m; ++i) W[(j + 1) * m + i] /= S;
                }
            } else {
                B[j * work + j] = S;
            }
        }
        so_svd_jacobi(work, B, BU, BS, BV);
        RF = so_nrm2(n, F);
        if (RF > 0)
            for (int64_t i = 0; i < n; ++i) F[i] /= RF;
        if (RF < 2 * eps) RF = 0.0;
        double smax = 0.0;
        for (int64_t i = 0; i < work; ++i) {
            if (BS[i] > smax) smax = BS[i];
            svr[i] = fabs(svr[i] - BS[i]) / BS[i];
        }
        for (int64_t i = 0; i < work; ++i) res[i] = RF * BU[i * work + (work - 1)];
        int64_t nconv = 0;
        for (int64_t i = 0; i < work; ++i)
            if (fabs(res[i]) < tol * smax && svr[i] < svtol) nconv++;
        iter++;
        if (nconv >= nu || S == 0.0) {
            ret = 0;
            break;
        }
        if (k < nu + nconv) k = nu + nconv;
        if (k > work - 3) k = work - 3;
        if (k < 1) k = 1;
        for (int64_t i = 0; i < work; ++i) svr[i] = BS[i];
        /* restart: V <- V BV[:,0:k], V[:,k] = F; W <- W BU[:,0:k] */
        for (int64_t c = 0; c < k; ++c)
            for (int64_t i = 0; i < n; ++i) {
                double a = 0.0;
                for (int64_t l = 0; l < work; ++l) a += Vb[l * n + i] * BV[c * work + l];
                tmp[c * n + i] = a;
            }
        memcpy(Vb, tmp, (size_t)(k * n) * sizeof(double));
        memcpy(Vb + k * n, F, (size_t)n * sizeof(double));
        for (int64_t c = 0; c < k; ++c)
            for (int64_t i = 0; i < m; ++i) {
                double a = 0.0;
                for (int64_t l = 0; l < work; ++l) a += W[l * m + i] * BU[c * work + l];
                tmp[c * m + i] = a;
            }
        memcpy(W, tmp, (size_t)(k * m) * sizeof(double));
        memset(B, 0, (size_t)(work * work) * sizeof(double));
        for (int64_t c = 0; c < k; ++c) {
            B[c * work + c] = BS[c];
            B[k * work + c] = res[c];
        }
    }
    /* results: s = BS[0:nu], U = W BU[:,0:nu], V = V BV[:,0:nu] */
    for (int64_t c = 0; c < nu; ++c) s[c] = BS[c];
    for (int64_t c = 0; c < nu; ++c)
        for (int64_t i = 0; i < m; ++i) {
            double a = 0.0;
            for (int64_t l = 0; l < work; ++l) a += W[l * m + i] * BU[c * work + l];
            U[c * m + i] = a;
        }
    for (int64_t c = 0; c < nu; ++c)
        for (int64_t i = 0; i < n; ++i) {
            double a = 0.0;
            for (int64_t l = 0; l < work; ++l) a += Vb[l * n + i] * BV[c * work + l];
            V[c * n + i] = a;
        }
done:
    if (iters_out) *iters_out = iter;
    if (nmult_out) *nmult_out = nmult;
    free(W);
    free(Vb);
    free(F);
    free(B);
    free(BU);
    free(BV);
    free(BS);
    free(res);
    free(svr);
    free(t);
    free(tmp);
    return ret;
}

/* Y = A' U for A = B - 1 mu' (B: m x n CSC, Float32 values and mu) and U
 * m x k row-major FP64: Y (n x k row-major) row j = sum over the column's
 * entries of B[p] U[row_p, :] - mu_j (sum of U's rows).  The vector form
 * of src/scaling.jl:252-257 applied to k columns at once (the matrix method
 * at :266-272 adds the mu term with the opposite sign; that is not
 * reproduced here).  Columns of B are split over nthreads OpenMP threads;
 * each output row is summed in entry order. */
SO_EXPORT void so_centered_tmul_mat(int64_t m, int64_t n, const int64_t *colptr,
                                    const int32_t *rowval, const float *nzval, const float *mu,
                                    const double *U, int64_t k, double *Y, int nthreads)
{
    double *colsum = (double *)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
    for (int64_t i = 0; i < m; ++i)
        for (int64_t c = 0; c < k; ++c) colsum[c] += U[i * k + c];
    if (nthreads < 1) nthreads = 1;
    _Pragma("omp parallel for schedule(dynamic, 8) num_threads(nthreads)")
    for (int64_t j = 0; j < n; ++j) {
        double *y = Y + j * k;
        for (int64_t c = 0; c < k; ++c) y[c] = 0.0;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            const double v = (double)nzval[p];
            const double *u = U + (int64_t)rowval[p] * k;
            for (int64_t c = 0; c < k; ++c) y[c] += v * u[c];
        }
        if (mu)
            for (int64_t c = 0; c < k; ++c) y[c] -= (double)mu[j] * colsum[c];
    }
    free(colsum);
}

typedef struct {
    int64_t m, n;
    const int64_t *colptr;
    const int32_t *rowval;
    const void *nzval;
    int f64;
    const void *mu;
} so_csc_ctx;

static void so_csc_apply(void *vctx, int trans, const double *x, double *y)
{
    so_csc_ctx *c = (so_csc_ctx *)vctx;
    so_centered_mul(c->m, c->n, c->colptr, c->rowval, c->nzval, c->f64, c->mu, trans, x, y,
                    1.0, 0.0);
}

typedef struct {
    int64_t m, n;
    const double *A; /* column-major m x n */
    const double *mu; /* may be NULL */
} so_dense_ctx;

static void so_dense_apply(void *vctx, int trans, const double *x, double *y)
{
    so_dense_ctx *c = (so_dense_ctx *)vctx;
    if (!trans) {
        for (int64_t i = 0; i < c->m; ++i) y[i] = 0.0;
        for (int64_t j = 0; j < c->n; ++j)
            for (int64_t i = 0; i < c->m; ++i) y[i] += c->A[j * c->m + i] * x[j];
        if (c->mu) {
            double d = 0.0;
            for (int64_t j = 0; j < c->n; ++j) d += c->mu[j] * x[j];
            for (int64_t i = 0; i < c->m; ++i) y[i] -= d;
        }
    } else {
        for (int64_t j = 0; j < c->n; ++j) {
            double a = 0.0;
            for (int64_t i = 0; i < c->m; ++i) a += c->A[j * c->m + i] * x[i];
            y[j] = a;
        }
        if (c->mu) {
            double sx = 0.0;
            for (int64_t i = 0; i < c->m; ++i) sx += x[i];
            for (int64_t j = 0; j < c->n; ++j) y[j] -= sx * c->mu[j];
        }
    }
}

SO_EXPORT int so_irlba_csc(int64_t m, int64_t n, const int64_t *colptr, const int32_t *rowval,
                           const void *nzval, int f64, const void *mu, int64_t nu, int64_t work,
                           int64_t maxit, double tol, double svtol, const double *init, double *s,
                           double *U, double *V, int64_t *iters, int64_t *nmult)
{
    so_csc_ctx ctx = {m, n, colptr, rowval, nzval, f64, mu};
    so_op op = {m, n, so_csc_apply, &ctx};
    return so_irlba_core(&op, nu, work, maxit, tol, svtol, init, s, U, V, iters, nmult);
}

SO_EXPORT int so_irlba_dense(int64_t m, int64_t n, const double *A, const double *mu, int64_t nu,
                             int64_t work, int64_t maxit, double tol, double svtol,
                             const double *init, double *s, double *U, double *V, int64_t *iters,
                             int64_t *nmult)
{
    so_dense_ctx ctx = {m, n, A, mu};
    so_op op = {m, n, so_dense_apply, &ctx};
    return so_irlba_core(&op, nu, work, maxit, tol, svtol, init, s, U, V, iters, nmult);
}

/* ------------------------------------------------------------------------ */
/* Exact Euclidean kNN (the build's definition of FindNeighboursEuclidean,   */
/* called from src/neighbours.jl:32-63):                                     */
/*   D(i,j) = sum_t (x_it - x_jt)^2, each difference and square rounded in   */
/*   the working precision, accumulated in t order, no FMA;                 */
/*   rank by (D, j) lexicographically; top k; j == i skipped only when       */
/*   include_self == 0; distances returned as sqrt(D).                       */
/* X is n x d column-major with leading dimension ld.  Outputs nn_index and  */
/* distances are n x k column-major, indices 1-based (src/neighbours.jl:79). */
/* Only the queries listed in qidx (0-based, nq of them) are computed when   */
/* qidx != NULL; row r of the outputs then belongs to qidx[r] (ldo = nq).    */
/* ------------------------------------------------------------------------ */

#define SO_KNN_BODY(T, SQRT)                                                              \
    int64_t nrow = qidx ? nq : n;                                                         \
    T *Xr = (T *)malloc((size_t)(n * d) * sizeof(T));                                     \
    for (int64_t i = 0; i < n; ++i)                                                       \
        for (int64_t t = 0; t < d; ++t) Xr[i * d + t] = X[i + t * ld];                    \
    int bad = 0;                                                                          \
    _Pragma("omp parallel for schedule(dynamic, 16) num_threads(nthreads)")               \
    for (int64_t r = 0; r < nrow; ++r) {                                                  \
        int64_t i = qidx ? qidx[r] : r;                                                   \
        T bd[256];                                                                        \
        int64_t bj[256];                                                                  \
        int64_t cnt = 0;                                                                  \
        const T *xi = Xr + i * d;                                                         \
        for (int64_t j = 0; j < n; ++j) {                                                 \
            if (!include_self && j == i) continue;                                        \
            const T *xj = Xr + j * d;                                                     \
            T D = 0;                                                                      \
            for (int64_t t = 0; t < d; ++t) {                                             \
                T diff = xi[t] - xj[t];                                                   \
                D = D + diff * diff;                                                      \
            }                                                                             \
            if (cnt == k && !(D < bd[k - 1] || (D == bd[k - 1] && j < bj[k - 1])))       \
                continue;                                                                 \
            int64_t pos = cnt < k ? cnt++ : k - 1;                                        \
            while (pos > 0 && (bd[pos - 1] > D || (bd[pos - 1] == D && bj[pos - 1] > j))) { \
                bd[pos] = bd[pos - 1];                                                    \
                bj[pos] = bj[pos - 1];                                                    \
                --pos;                                                                    \
            }                                                                             \
            bd[pos] = D;                                                                  \
            bj[pos] = j;                                                                  \
        }                                                                                 \
        if (cnt < k) {                                                                    \
            _Pragma("omp atomic write") bad = 1;                                          \
        }                                                                                 \
        for (int64_t t = 0; t < k; ++t) {                                                 \
            nn_index[r + t * nrow] = t < cnt ? (int32_t)(bj[t] + 1) : 0;                  \
            distances[r + t * nrow] = t < cnt ? SQRT(bd[t]) : (T)INFINITY;                \
        }                                                                                 \
    }                                                                                     \
    free(Xr);                                                                             \
    return bad ? -1 : 0;

SO_EXPORT int so_knn_f32(const float *X, int64_t n, int64_t d, int64_t ld, int64_t k,
                         int include_self, const int64_t *qidx, int64_t nq, int32_t *nn_index,
                         float *distances, int nthreads)
{
    if (k < 1 || k > 256) return -2;
    if (nthreads < 1) nthreads = 1;
    SO_KNN_BODY(float, sqrtf)
}

SO_EXPORT int so_knn_f64(const double *X, int64_t n, int64_t d, int64_t ld, int64_t k,
                         int include_self, const int64_t *qidx, int64_t nq, int32_t *nn_index,
                         double *distances, int nthreads)
{
    if (k < 1 || k > 256) return -2;
    if (nthreads < 1) nthreads = 1;
    SO_KNN_BODY(double, sqrt)
}

/* ------------------------------------------------------------------------ */
/* Jaccard SNN graph  (src/neighbours.jl:79, 88-93)                          */
/*   nn[l, j] = true  <=>  l in N(j)   (nn_index n x k, 1-based, col-major)  */
/*   S = nn' * nn  ->  S[i,j] = |N(i) n N(j)|                                */
/*   w = x / (k + (k - x)) in T; droptol!(S, prune) keeps |w| > prune.       */
/* Output CSC n x n, Int64 colptr (1-based) and rowval (1-based), rows       */
/* sorted within each column.  With colptr == NULL only the count is         */
/* computed.  Returns nnz (or -1 on a bad index).                            */
/* ------------------------------------------------------------------------ */

static int so_cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

SO_EXPORT int64_t so_snn_jaccard(int64_t n, int64_t k, const int32_t *nn_index, double prune,
                                 int f64, int64_t *colptr, int64_t *rowval, void *nzval)
{
    /* neighbour sets N(j) (deduplicated, as sparse() ORs duplicates) */
    int64_t *nb = (int64_t *)malloc((size_t)(n * k) * sizeof(int64_t));
    int64_t *nbn = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t t = 0; t < k; ++t) {
            int64_t l = (int64_t)nn_index[j + t * n] - 1;
            if (l < 0 || l >= n) {
                free(nb);
                free(nbn);
                return -1;
            }
            int dup = 0;
            for (int64_t u = 0; u < nbn[j]; ++u)
                if (nb[j * k + u] == l) dup = 1;
            if (!dup) nb[j * k + nbn[j]++] = l;
        }
    }
    /* reverse lists R(l) = { i : l in N(i) } */
    int64_t *rp = (int64_t *)calloc((size_t)(n + 1), sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t u = 0; u < nbn[j]; ++u) rp[nb[j * k + u] + 1]++;
    for (int64_t l = 0; l < n; ++l) rp[l + 1] += rp[l];
    int64_t *rl = (int64_t *)malloc((size_t)(rp[n] > 0 ? rp[n] : 1) * sizeof(int64_t));
    int64_t *cur = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    memcpy(cur, rp, (size_t)n * sizeof(int64_t));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t u = 0; u < nbn[j]; ++u) rl[cur[nb[j * k + u]]++] = j;
    /* weight table in T:  w[x] = x / (k + (k - x)) */
    const float kf = (float)k, prf = (float)prune;
    const double kd = (double)k;
    int64_t *cnt = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    int64_t *touched = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t nnz = 0;
    if (colptr) colptr[0] = 1;
    for (int64_t j = 0; j < n; ++j) {
        int64_t nt = 0;
        for (int64_t u = 0; u < nbn[j]; ++u) {
            int64_t l = nb[j * k + u];
            for (int64_t p = rp[l]; p < rp[l + 1]; ++p) {
                int64_t i = rl[p];
                if (cnt[i]++ == 0) touched[nt++] = i;
            }
        }
        qsort(touched, (size_t)nt, sizeof(int64_t), so_cmp_i64);
        for (int64_t a = 0; a < nt; ++a) {
            int64_t i = touched[a];
            int64_t x = cnt[i];
            cnt[i] = 0;
            int keep;
            if (f64) {
                double xd = (double)x;
                double w = xd / (kd + (kd - xd));
                keep = fabs(w) > prune;
                if (keep && colptr) {
                    rowval[nnz] = i + 1;
                    ((double *)nzval)[nnz] = w;
                }
            } else {
                float xf = (float)x;
                float w = xf / (kf + (kf - xf));
                keep = fabsf(w) > prf;
                if (keep && colptr) {
                    rowval[nnz] = i + 1;
                    ((float *)nzval)[nnz] = w;
                }
            }
            if (keep) nnz++;
        }
        if (colptr) colptr[j + 1] = nnz + 1;
    }
    free(nb);
    free(nbn);
    free(rp);
    free(rl);
    free(cur);
    free(cnt);
    free(touched);
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* find_variable_features(:vst)  (src/variablefeatures.jl:34-50,128-161)     */
/* ------------------------------------------------------------------------ */

/* mean_std(Float64, A) (src/scaling.jl:18-38,119-130): per column, Welford
 * over the Int32 counts in FP64, count starting at m - nnz; sd = sqrt(s/(m-1)). */
SO_EXPORT void so_count_mean_std(int64_t m, int64_t n, const int64_t *colptr, const int32_t *x,
                                 double *mu_out, double *sd_out)
{
    for (int64_t j = 0; j < n; ++j) {
        int64_t count = m - (colptr[j + 1] - colptr[j]);
        double mu = 0.0, s = 0.0;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            const double v = (double)x[p];
            count += 1;
            const double delta = v - mu;
            mu = mu + delta / (double)count;
            s = s + delta * (v - mu);
        }
        mu_out[j] = mu;
        sd_out[j] = sqrt(s / (double)(m - 1));
    }
}

/* mean_var(Float64, row_norm(A, 1.0)) (src/scaling.jl:18-34,132-142 on
 * src/normalize.jl:24-36): per column, Welford in FP64 over the relative
 * counts v = x / s_row (s_row the Int row total; 1.0 * x is exact), count
 * starting at m - nnz; var = s / (m - 1).  The statistics of the
 * :dispersion, :meanvarplot and :saunders selections
 * (src/variablefeatures.jl:60-91). */
SO_EXPORT void so_rel_mean_var(int64_t m, int64_t n, const int64_t *colptr, const int32_t *row,
                               const int32_t *x, double *mu_out, double *var_out)
{
    int64_t *tot = (int64_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(int64_t));
    for (int64_t p = 0; p < colptr[n]; ++p) tot[row[p]] += x[p];
    for (int64_t j = 0; j < n; ++j) {
        int64_t count = m - (colptr[j + 1] - colptr[j]);
        double mu = 0.0, s = 0.0;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            const double v = (double)x[p] / (double)tot[row[p]];
            count += 1;
            const double delta = v - mu;
            mu = mu + delta / (double)count;
            s = s + delta * (v - mu);
        }
        mu_out[j] = mu;
        var_out[j] = s / (double)(m - 1);
    }
    free(tot);
}

/* Loess (Cleveland's local regression, degree 2, tricube weights) evaluated
 * at every point of x (ascending, N points), the model of
 * `loess(xs, ys, span=...)` + `predict(model, xs)` at :41-45; Loess.jl
 * interpolates between kd-tree vertices, this is the exact local fit.
 *  q = floor(N span), at least min(N, 3); the neighbourhood of x_i is the
 *  window of q consecutive points containing i0 (the first point equal to
 *  x_i) with the smallest dmax = max(x_i - x_s, x_{s+q-1} - x_i) (the
 *  smaller s on a tie);
 *  w_j = (1 - (|x_j - x_i| / dmax)^3)^3 (0 at dmax; 1 when dmax = 0);
 *  u_j = (x_j - x_i) / dmax; weighted least squares y ~ a + b u + c u^2 by
 *  elimination on the normal equations (degree lowered when a pivot
 *  vanishes); fit_i = a. */
static double so_loess_point(int64_t N, const double *x, const double *y, int64_t q, int64_t i)
{
    /* the fit depends on x_i only: tied points use the window of the first
     * of them (its nonzero-weight terms are summed in one order) */
    {
        int64_t a = 0, b = i;
        while (a < b) {
            const int64_t mid = a + (b - a) / 2;
            if (x[mid] < x[i]) a = mid + 1;
            else b = mid;
        }
        i = a;
    }
    int64_t lo = i - q + 1 < 0 ? 0 : i - q + 1;
    int64_t hi = i < N - q ? i : N - q;
    /* smallest s in [lo, hi] with x[s+q-1] - x[i] >= x[i] - x[s] */
    int64_t a = lo, b = hi + 1;
    while (a < b) {
        const int64_t mid = a + (b - a) / 2;
        if (x[mid + q - 1] - x[i] >= x[i] - x[mid]) b = mid;
        else a = mid + 1;
    }
    int64_t s;
    double dmax;
    if (a > hi) {
        s = hi;
        dmax = x[i] - x[hi];
    } else {
        const double dr = x[a + q - 1] - x[i];
        if (a > lo && x[i] - x[a - 1] <= dr) {
            s = a - 1;
            dmax = x[i] - x[a - 1];
        } else {
            s = a;
            dmax = dr;
        }
    }
    double S0 = 0, S1 = 0, S2 = 0, S3 = 0, S4 = 0, T0 = 0, T1 = 0, T2 = 0;
    for (int64_t j = s; j < s + q; ++j) {
        double w, u;
        if (dmax > 0.0) {
            u = (x[j] - x[i]) / dmax;
            const double r = fabs(u);
            const double t = 1.0 - r * r * r;
            w = r < 1.0 ? t * t * t : 0.0;
        } else {
            u = 0.0;
            w = 1.0;
        }
        const double wu = w * u, wu2 = wu * u;
        S0 += w;
        S1 += wu;
        S2 += wu2;
        S3 += wu2 * u;
        S4 += wu2 * u * u;
        T0 += w * y[j];
        T1 += wu * y[j];
        T2 += wu2 * y[j];
    }
    if (!(S0 > 0.0)) return y[i];
    const double l21 = S1 / S0, l31 = S2 / S0;
    const double a22 = S2 - l21 * S1, a23 = S3 - l21 * S2, a33 = S4 - l31 * S2;
    const double b2 = T1 - l21 * T0, b3 = T2 - l31 * T0;
    if (!(a22 > 1e-12 * S2)) return T0 / S0;
    const double l32 = a23 / a22;
    const double a33b = a33 - l32 * a23, b3b = b3 - l32 * b2;
    if (!(a33b > 1e-12 * S4)) {
        const double bl = b2 / a22;
        return (T0 - S1 * bl) / S0;
    }
    const double c = b3b / a33b;
    const double bb = (b2 - a23 * c) / a22;
    return (T0 - S1 * bb - S2 * c) / S0;
}

SO_EXPORT int so_loess_fit(int64_t N, const double *x, const double *y, double span, double *fit)
{
    if (N <= 0) return -1;
    int64_t q = (int64_t)floor((double)N * span);
    const int64_t qmin = N < 3 ? N : 3;
    if (q < qmin) q = qmin;
    if (q > N) q = N;
    for (int64_t i = 0; i < N; ++i) fit[i] = so_loess_point(N, x, y, q, i);
    return 0;
}

/* standardized_var_clipped (src/variablefeatures.jl:21-28): for sd_j > 0,
 * (sum over nonzeros of min((x - mu)/sd, vmax)^2
 *  + (m - nnz) min(-mu/sd, vmax)^2) / (m - 1), vmax = sqrt(m); else 0.
 * The sum runs in row order. */
SO_EXPORT void so_standardized_var(int64_t m, int64_t n, const int64_t *colptr, const int32_t *x,
                                   const double *mu, const double *sd, double *out)
{
    const double vmax = sqrt((double)m);
    for (int64_t j = 0; j < n; ++j) {
        if (sd[j] == 0.0) {
            out[j] = 0.0;
            continue;
        }
        double t = 0.0;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
            double v = ((double)x[p] - mu[j]) / sd[j];
            if (v > vmax) v = vmax;
            t += v * v;
        }
        double z = (0.0 - mu[j]) / sd[j];
        if (z > vmax) z = vmax;
        const int64_t nz = colptr[j + 1] - colptr[j];
        out[j] = (t + (double)(m - nz) * (z * z)) / (double)(m - 1);
    }
}

/* ---- Louvain clustering of the SNN graph (src/clustering.jl, src/modularity.jl) ----
 *
 * A sequential restatement of the reference's multi-start Louvain in FP64.
 * The node visiting order comes from a Fisher-Yates shuffle driven by
 * so_mix64 (Julia's default_rng stream cannot be reproduced here), so a run
 * is one sample of the reference's randomised search, not its exact output.
 */
typedef struct {
    int64_t n;
    int64_t *ptr;   /* n + 1 */
    int64_t *adj;   /* edge targets, self-links excluded */
    double *w;
    double *k;      /* node weights */
    double totw, tot_self;
} so_net;

typedef struct {
    const so_net *net;
    double res;
    int64_t *c;     /* cluster of each node, 0-based */
    double *wtot;   /* length net->n, zero beyond nc */
    int64_t nc;
} so_clu;

static void so_net_free(so_net *g)
{
    free(g->ptr);
    free(g->adj);
    free(g->w);
    free(g->k);
}

/* Network(snn) (src/modularity.jl:165-194): column i's entries are the edges
 * of node i; the self-link (i, i) is skipped and counted nowhere; totw is the
 * sum of the remaining weights, in column order. */
static void so_net_from_csc(int64_t n, const int64_t *colptr, const int32_t *row, const double *w,
                            so_net *g)
{
    g->n = n;
    g->ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    const int64_t nnz = colptr[n];
    g->adj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
    g->w = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    g->k = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    g->totw = 0.0;
    g->tot_self = 0.0;
    int64_t e = 0;
    for (int64_t i = 0; i < n; ++i) {
        g->ptr[i] = e;
        double wi = 0.0;
        for (int64_t p = colptr[i]; p < colptr[i + 1]; ++p) {
            if (row[p] == i) continue;
            g->totw += w[p];
            wi += w[p];
            g->adj[e] = row[p];
            g->w[e] = w[p];
            ++e;
        }
        g->k[i] = wi;
    }
    g->ptr[n] = e;
}

static void so_clu_singletons(const so_net *g, double res, so_clu *cl)
{
    cl->net = g;
    cl->res = res;
    cl->c = (int64_t *)malloc(sizeof(int64_t) * (size_t)(g->n > 0 ? g->n : 1));
    cl->wtot = (double *)malloc(sizeof(double) * (size_t)(g->n > 0 ? g->n : 1));
    for (int64_t i = 0; i < g->n; ++i) {
        cl->c[i] = i;
        cl->wtot[i] = g->k[i];
    }
    cl->nc = g->n;
}

/* modularity(clustering) (src/modularity.jl:125-147): (tot_self + weight of
 * the edges inside clusters) / totw - res * sum_i (w_tot[i] / totw)^2. */
static double so_clu_modularity(const so_clu *cl)
{
    const so_net *g = cl->net;
    double in = g->tot_self;
    for (int64_t v = 0; v < g->n; ++v)
        for (int64_t e = g->ptr[v]; e < g->ptr[v + 1]; ++e)
            if (cl->c[g->adj[e]] == cl->c[v]) in += g->w[e];
    double sq = 0.0;
    for (int64_t i = 0; i < g->n; ++i) {
        const double x = cl->wtot[i] / g->totw;
        sq += x * x;
    }
    return in / g->totw - cl->res * sq;
}

/* local_move! (src/modularity.jl:255-287) with best_local_move (:243-253):
 * one shuffled order per call; each node leaves its cluster and joins the
 * neighbouring cluster with the largest kin - res * w_tot * ki / totw (ties
 * to the smaller label, myfindmax of src/utils.jl:197-212) when that beats
 * staying; sweeps repeat until a sweep gains less than min_mod or moves
 * nothing. */
static double so_local_move(uint64_t *rng, so_clu *cl, double min_mod)
{
    const so_net *g = cl->net;
    const int64_t n = g->n;
    const double totw = g->totw;
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    for (int64_t i = n - 1; i > 0; --i) {
        const uint64_t r = so_mix64((*rng)++);
        const int64_t j = (int64_t)(r % (uint64_t)(i + 1));
        const int64_t t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
    double *kin = (double *)calloc((size_t)n, sizeof(double));
    int64_t *nbr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t nn = 0;
    double delta_mod = min_mod, total = 0.0;
    int stable = 0;
    while (!stable && delta_mod >= min_mod) {
        delta_mod = 0.0;
        stable = 1;
        for (int64_t oi = 0; oi < n; ++oi) {
            const int64_t v = order[oi];
            const double ki = g->k[v];
            const int64_t ci = cl->c[v];
            cl->wtot[ci] -= ki;
            for (int64_t t = 0; t < nn; ++t) kin[nbr[t]] = 0.0;
            nn = 0;
            for (int64_t e = g->ptr[v]; e < g->ptr[v + 1]; ++e) {
                const int64_t cj = cl->c[g->adj[e]];
                if (kin[cj] == 0.0) nbr[nn++] = cj;
                kin[cj] += g->w[e];
            }
            int64_t best = ci;
            double gain = 0.0;
            if (nn > 0) {
                const double init = kin[ci] - cl->res * (cl->wtot[ci] * ki) / totw;
                int64_t m = nbr[0];
                double fm = kin[m] - cl->res * (cl->wtot[m] * ki) / totw;
                for (int64_t t = 1; t < nn; ++t) {
                    const int64_t x = nbr[t];
                    const double fx = kin[x] - cl->res * (cl->wtot[x] * ki) / totw;
                    if (fm < fx || (fm == fx && x < m)) {
                        m = x;
                        fm = fx;
                    }
                }
                gain = 2.0 * (fm - init) / totw;
                best = m;
            }
            if (gain > 0.0) {
                cl->wtot[best] += ki;
                cl->c[v] = best;
                delta_mod += gain;
                stable = 0;
            } else {
                cl->wtot[ci] += ki;
                cl->c[v] = ci;
            }
        }
        total += delta_mod;
    }
    free(order);
    free(kin);
    free(nbr);
    return total;
}

/* renumber! (src/modularity.jl:289-315): labels in order of first
 * occurrence, w_tot carried along. */
static void so_renumber(so_clu *cl)
{
    const int64_t n = cl->net->n;
    int64_t *lab = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    double *w = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int64_t i = 0; i < n; ++i) lab[i] = -1;
    int64_t id = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c = cl->c[i];
        if (lab[c] < 0) {
            lab[c] = id++;
            w[lab[c]] = cl->wtot[c];
        }
        cl->c[i] = lab[c];
    }
    memcpy(cl->wtot, w, sizeof(double) * (size_t)n);
    cl->nc = id;
    free(lab);
    free(w);
}

/* reduced_network (src/modularity.jl:196-241): one node per cluster with
 * the summed node weights; edges between clusters summed (members in node
 * order, their edges in order) and emitted by ascending target; edges
 * inside a cluster added to tot_self. */
static void so_reduce(const so_clu *cl, so_net *r)
{
    const so_net *g = cl->net;
    const int64_t n = g->n, nc = cl->nc;
    int64_t *cnt = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    int64_t *mem = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t v = 0; v < n; ++v) cnt[cl->c[v] + 1]++;
    for (int64_t i = 0; i < nc; ++i) cnt[i + 1] += cnt[i];
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
    for (int64_t i = 0; i < nc; ++i) cur[i] = cnt[i];
    for (int64_t v = 0; v < n; ++v) mem[cur[cl->c[v]]++] = v;
    double *acc = (double *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
    char *seen = (char *)calloc((size_t)(nc > 0 ? nc : 1), 1);
    int64_t *touched = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
    const int64_t cap0 = g->ptr[n] > 0 ? g->ptr[n] : 1;
    r->n = nc;
    r->ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nc + 1));
    r->adj = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap0);
    r->w = (double *)malloc(sizeof(double) * (size_t)cap0);
    r->k = (double *)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1));
    r->totw = g->totw;
    r->tot_self = g->tot_self;
    int64_t ne = 0;
    for (int64_t i = 0; i < nc; ++i) {
        int64_t nt = 0;
        double wi = 0.0;
        for (int64_t t = cnt[i]; t < cnt[i + 1]; ++t) {
            const int64_t v = mem[t];
            wi += g->k[v];
            for (int64_t e = g->ptr[v]; e < g->ptr[v + 1]; ++e) {
                const int64_t j = cl->c[g->adj[e]];
                if (j == i) {
                    r->tot_self += g->w[e];
                } else {
                    if (!seen[j]) {
                        seen[j] = 1;
                        touched[nt++] = j;
                    }
                    acc[j] += g->w[e];
                }
            }
        }
        qsort(touched, (size_t)nt, sizeof(int64_t), so_cmp_i64);
        r->ptr[i] = ne;
        for (int64_t t = 0; t < nt; ++t) {
            const int64_t j = touched[t];
            if (acc[j] != 0.0) {
                r->adj[ne] = j;
                r->w[ne] = acc[j];
                ++ne;
            }
            acc[j] = 0.0;
            seen[j] = 0;
        }
        r->k[i] = wi;
    }
    r->ptr[nc] = ne;
    free(cnt);
    free(mem);
    free(cur);
    free(acc);
    free(seen);
    free(touched);
}

/* louvain! (src/modularity.jl:326-340): local moving, renumbering, and the
 * same on the reduced network, whose clusters are merged back. */
static double so_louvain_level(uint64_t *rng, so_clu *cl, double min_mod)
{
    if (cl->net->n <= 1) return 0.0;
    double gain = so_local_move(rng, cl, min_mod);
    so_renumber(cl);
    if (cl->nc < cl->net->n) {
        so_net r;
        so_reduce(cl, &r);
        so_clu rc;
        so_clu_singletons(&r, cl->res, &rc);
        gain += so_louvain_level(rng, &rc, min_mod);
        for (int64_t v = 0; v < cl->net->n; ++v) cl->c[v] = rc.c[cl->c[v]];
        memset(cl->wtot, 0, sizeof(double) * (size_t)cl->net->n);
        memcpy(cl->wtot, rc.wtot, sizeof(double) * (size_t)r.n);
        cl->nc = rc.nc;
        free(rc.c);
        free(rc.wtot);
        so_net_free(&r);
    }
    return gain;
}

/* cluster(rng, SNN; resolution, nrandomstarts, niterations) with
 * louvain_multi (src/clustering.jl:59-77, src/modularity.jl:344-380): per
 * start, louvain! from singletons and again while it gains at least
 * min_mod, at most max_iter further times; the start with the largest
 * modularity is kept (the first on ties).  labels[n] receive 0-based
 * clusters in order of first occurrence; returns their number. */
SO_EXPORT int64_t so_louvain(int64_t n, const int64_t *colptr, const int32_t *row, const double *w,
                             double res, int64_t nstarts, int64_t max_iter, double min_mod,
                             uint64_t seed, int64_t *labels, double *q_out)
{
    so_net g;
    so_net_from_csc(n, colptr, row, w, &g);
    uint64_t rng = so_mix64(seed);
    double qbest = 0.0;
    int64_t ncbest = 0;
    for (int64_t s = 0; s < nstarts; ++s) {
        so_clu cl;
        so_clu_singletons(&g, res, &cl);
        double gain = so_louvain_level(&rng, &cl, min_mod);
        for (int64_t it = 1; gain >= min_mod && it <= max_iter; ++it)
            gain = so_louvain_level(&rng, &cl, min_mod);
        const double q = so_clu_modularity(&cl);
        if (s == 0 || qbest < q) {
            qbest = q;
            ncbest = cl.nc;
            memcpy(labels, cl.c, sizeof(int64_t) * (size_t)n);
        }
        free(cl.c);
        free(cl.wtot);
    }
    so_net_free(&g);
    if (q_out) *q_out = qbest;
    return ncbest;
}

/* modularity of given 0-based labels (compact, 0..nc-1) on Network(snn):
 * Clustering(network, nodecluster, resolution) (src/modularity.jl:52-66)
 * followed by modularity (:145-148). */
SO_EXPORT double so_modularity(int64_t n, const int64_t *colptr, const int32_t *row,
                               const double *w, const int64_t *labels, double res)
{
    so_net g;
    so_net_from_csc(n, colptr, row, w, &g);
    so_clu cl;
    cl.net = &g;
    cl.res = res;
    cl.c = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    cl.wtot = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    int64_t nc = 0;
    for (int64_t v = 0; v < n; ++v) {
        cl.c[v] = labels[v];
        if (labels[v] + 1 > nc) nc = labels[v] + 1;
    }
    for (int64_t ci = 0; ci < nc; ++ci)
        for (int64_t v = 0; v < n; ++v)
            if (labels[v] == ci) cl.wtot[ci] += g.k[v];
    cl.nc = nc;
    const double q = so_clu_modularity(&cl);
    free(cl.c);
    free(cl.wtot);
    so_net_free(&g);
    return q;
}

/* group_singletons! (src/clustering.jl:79-98) on the SNN itself (self-links
 * included): each member of a cluster of size 1, in node order, moves to the
 * cluster of size > 1 (ascending label order) with the largest mean weight
 * of its entries into that cluster, labels as updated so far.  The mean over
 * no entries is NaN, and findmax returns the first NaN it meets, so the
 * first such cluster without an entry from the node wins.  Labels are then
 * renumbered by first occurrence (relabel!, src/utils.jl:91-106).  labels are
 * 0-based and compact; returns the new number of clusters, or -1 when no
 * cluster has more than one member (findmax over an empty set).
 * The node takes the cluster label that findmax(f, others) returns, as Julia
 * 1.6 (the version the reference's CI pins, .github/workflows/CI.yml:12-14)
 * defines findmax(f, domain) -> (f(x), x).  Julia 1.7 and later return the
 * position in `others` instead, so under those versions src/clustering.jl:89-96
 * assigns a position, not a label; that behaviour is not reproduced (parity
 * unpinned for Julia >= 1.7). */
SO_EXPORT int64_t so_group_singletons(int64_t n, const int64_t *colptr, const int32_t *row,
                                      const double *w, int64_t *labels)
{
    int64_t nc = 0;
    for (int64_t v = 0; v < n; ++v)
        if (labels[v] + 1 > nc) nc = labels[v] + 1;
    int64_t *cnt = (int64_t *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(int64_t));
    for (int64_t v = 0; v < n; ++v) cnt[labels[v]]++;
    int64_t nothers = 0, nsingle = 0;
    for (int64_t c = 0; c < nc; ++c) {
        if (cnt[c] > 1) ++nothers;
        else ++nsingle;
    }
    if (nsingle == 0) {
        free(cnt);
        return nc;
    }
    if (nothers == 0) {
        free(cnt);
        return -1;
    }
    int64_t *others = (int64_t *)malloc(sizeof(int64_t) * (size_t)nothers);
    int64_t *single = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    nothers = 0;
    for (int64_t c = 0; c < nc; ++c)
        if (cnt[c] > 1) others[nothers++] = c;
    int64_t ns = 0;
    for (int64_t v = 0; v < n; ++v)
        if (cnt[labels[v]] <= 1) single[ns++] = v;
    for (int64_t t = 0; t < ns; ++t) {
        const int64_t j = single[t];
        int64_t closest = -1;
        double best = 0.0;
        for (int64_t o = 0; o < nothers; ++o) {
            const int64_t cl = others[o];
            double s = 0.0;
            int64_t c = 0;
            for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p)
                if (labels[row[p]] == cl) {
                    s += w[p];
                    ++c;
                }
            if (c == 0) { /* NaN: the first one wins */
                closest = cl;
                break;
            }
            const double mean = s / (double)c;
            if (closest < 0 || best < mean) {
                closest = cl;
                best = mean;
            }
        }
        labels[j] = closest;
    }
    int64_t *lab = (int64_t *)malloc(sizeof(int64_t) * (size_t)nc);
    for (int64_t c = 0; c < nc; ++c) lab[c] = -1;
    int64_t id = 0;
    for (int64_t v = 0; v < n; ++v) {
        if (lab[labels[v]] < 0) lab[labels[v]] = id++;
        labels[v] = lab[labels[v]];
    }
    free(lab);
    free(cnt);
    free(others);
    free(single);
    return id;
}
