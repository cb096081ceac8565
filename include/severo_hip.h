/*
 * severo_hip.h -- C ABI of libsevero_hip.so, the MI355X (gfx950) engine for
 * Severo.jl's embedding-and-graph path:
 *
 *   normalize_cells -> feature selection -> scale_features -> PCA (IRLBA)
 *     -> nearest_neighbours (exact Euclidean kNN) -> Jaccard SNN graph
 *
 * Conventions
 *   - Every entry point returns an int status (SEV_OK == 0) unless it mirrors
 *     a reference native symbol with another signature; sev_last_error()
 *     returns the message of the calling thread's last failure.
 *   - "Host" entry points take caller-owned host buffers in Severo's Julia
 *     layout (cells x genes SparseMatrixCSC{T,Int64}: Int64 colptr/rowval,
 *     1-based; column-major dense matrices; 1-based neighbour indices) and
 *     write results in place, exactly as the reference's ccall sites do.
 *   - The sev_pipeline_* handle keeps a dataset resident in device memory so
 *     the whole path runs without host round trips.  A handle is not
 *     thread-safe; use one per thread.
 *   - No entry point falls back to CPU code: without a usable gfx950 device
 *     every call fails with SEV_ERR_NODEV.
 */
#ifndef SEVERO_HIP_H
#define SEVERO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEV_OK 0
#define SEV_ERR_ARG (-1)
#define SEV_ERR_HIP (-2)
#define SEV_ERR_CONVERGENCE (-3)
#define SEV_ERR_NOMEM (-4)
#define SEV_ERR_NODEV (-5)
#define SEV_ERR_INTERNAL (-6)

/* dtype codes */
#define SEV_FLOAT32 0
#define SEV_FLOAT64 1
/* count element types */
#define SEV_INT32 0
#define SEV_INT64 1
/* normalisation methods (src/normalize.jl:45-51) */
#define SEV_LOGNORMALIZE 0
#define SEV_RELATIVECOUNTS 1
/* precision forms */
#define SEV_PREC_FP32 0 /* plain form */
#define SEV_PREC_BF16 1 /* BF16 values / BF16 MFMA candidates, exact re-rank */
#define SEV_PREC_EXACT 2 /* kNN only: exhaustive exact search, no MFMA stage */
#define SEV_PREC_FP8 3   /* kNN only: FP8 e4m3 candidates (per-128-row scaling), then
                          * BF16 / FP32 candidates for rows without a certificate */

/* find_variable_features methods besides :vst */
#define SEV_HVG_DISPERSION 1
#define SEV_HVG_MEANVARPLOT 2
#define SEV_HVG_SAUNDERS 3
#define SEV_BIN_WIDTH 0 /* binning_method=:width (the only one supported) */

/* ---- library ----------------------------------------------------------- */
const char *sev_last_error(void);
int sev_version(void);
/* Device self-checks (no counterpart in the reference): what == 0 / 1 / 2
 * / 4 checks the FP32 / BF16 / FP64 / FP8 e4m3 matrix-core lane layouts the
 * kNN and IRLBA kernels rely on; *result receives the largest deviation (0
 * when they hold).  what == 3 runs the HBM copy probe (k_copy_probe: 1 GiB read +
 * 1 GiB written per launch, best of 5) and returns its rate in GB/s; it is
 * the known-size kernel against which HBM byte counters are calibrated. */
int sev_selftest(int what, double *result);
/* Selects the device for the calling thread and checks it is usable. */
int sev_device_init(int device);

/* ---- drop-ins over host buffers --------------------------------------- */

/* normalize_cells (src/normalize.jl:40-55, row_norm :17-32, log_norm :34-38).
 * m cells x n genes; colptr[n+1], rowval[nnz] 1-based Int64; nzval is Int32
 * (count_type SEV_INT32) or Int64 counts.  Per cell c: s_c = sum of its
 * counts; out = sf * x / s_c in dtype, then log1p for SEV_LOGNORMALIZE.
 * out_nzval has nnz elements of dtype (the structure is unchanged). */
int sev_normalize_cells(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                        const void *nzval, int count_type, int method, double scale_factor,
                        int dtype, void *out_nzval);

/* Allocation-free normalize_cells.  No counterpart in the reference: this
 * entry point is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It computes what sev_normalize_cells computes, on a count
 * matrix already in device memory: colptr[n+1] (int64, 0-based), rowval[nnz]
 * (int32, 0-based, below m and strictly increasing within each column) and
 * counts[nnz] (int32, >= 0); out_nzval[nnz] (device, dtype) receives the
 * values.  The work is enqueued on `stream` (a hipStream_t; NULL for the
 * null stream) and the call returns without waiting.  The per-cell totals
 * live in a static device array of 8,388,608 entries (32 MiB on every
 * device the library runs on, held until the process ends); m above that
 * is refused with SEV_ERR_ARG.  Every call on one device shares that array,
 * so calls that overlap on one device (other threads or streams) race.  No
 * hipMalloc, hipFree or other allocating call is made. */
int sev_normalize_cells_noalloc(int64_t m, int64_t n, int64_t nnz, const int64_t *colptr,
                                const int32_t *rowval, const int32_t *counts, int method,
                                double scale_factor, int dtype, void *out_nzval, void *stream);

/* scale_features -> scale_data (src/scaling.jl:199-217, :335-357) with the
 * Welford mean/var of src/scaling.jl:18-34 in dtype.  Writes the clipped
 * scaled values (same structure) and mu' = mu / sigma per column (the
 * CenteredMatrix mu).  scale_max may be +Inf. */
int sev_scale_features(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                       const void *nzval, int dtype, double scale_max, void *out_nzval,
                       void *out_mu);

/* CenteredMatrix mul! (src/scaling.jl:245-257):
 *   trans == 'N':  y = alpha*(A x - (mu.x) 1) + beta*y     (y length m)
 *   trans == 'T':  y = alpha*(A' x - (sum x) mu) + beta*y  (y length n)
 * A is m x n CSC with dtype values (1-based Int64 structure); mu may be NULL
 * for a plain sparse product. */
int sev_centered_mul(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                     const void *nzval, const void *mu, int dtype, char trans, const double *x,
                     double *y, double alpha, double beta);

/* mul!(C, S, X, α, β) and mul!(C, S', X, α, β) of a CenteredMatrix S with a
 * dense matrix X (src/scaling.jl:259-271).  S = A - 1 mu' with A as in
 * sev_centered_mul.  trans 'N': Y (m x p) = alpha S X + beta Y, X n x p;
 * 'T': Y (n x p) = alpha S' X + beta Y, X m x p.  X and Y are column-major
 * FP64 with leading dimensions ldx, ldy; beta == 0 overwrites Y.  The 'T'
 * form subtracts mu (1' X) like the vector form (src/scaling.jl:252-257);
 * the reference's matrix form adds it (src/scaling.jl:268-269). */
int sev_centered_mul_mat(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                         const void *nzval, const void *mu, int dtype, char trans, int64_t p,
                         const double *X, int64_t ldx, double *Y, int64_t ldy, double alpha,
                         double beta);

/* mul!(C, Sl', Sr, α, β) of two CenteredMatrix with m rows
 * (src/scaling.jl:273-294): C (nl x nr, column-major FP64, leading
 * dimension ldc) = alpha Sl' Sr + beta C, computed as the reference does:
 * Al'Ar - mul (1'Ar) - (1'Al - m mul')' mur'.  colptr_r == NULL takes
 * Sr = Sl (the Gram matrix S'S; nr and the other _r arguments are then
 * ignored).  Both matrices have the same dtype. */
int sev_centered_gram(int64_t m, int64_t nl, const int64_t *colptr_l, const int64_t *rowval_l,
                      const void *nzval_l, const void *mu_l, int64_t nr,
                      const int64_t *colptr_r, const int64_t *rowval_r, const void *nzval_r,
                      const void *mu_r, int dtype, double *C, int64_t ldc, double alpha,
                      double beta);

/* IRLBA on the centered matrix held on the device (replaces the libcell
 * irlba call of src/irlba.jl:66-71 for the CenteredMatrix operator of
 * src/scaling.jl:219-257, without per-product callbacks).  B (dtype values,
 * 1-based Int64 structure) and mu (dtype, length n, or NULL for an
 * uncentered matrix) describe A = B - 1 mu'.  work = nu + 7 in the reference
 * (src/irlba.jl:50); init has n entries (src/irlba.jl:63).  Outputs s[nu],
 * U (m x nu), V (n x nu) column-major FP64.  precision: SEV_PREC_FP32 reads
 * B in its dtype, SEV_PREC_BF16 as BF16 (Float32 input only); vectors and
 * reorthogonalisation are FP64 either way.  Returns SEV_ERR_CONVERGENCE
 * where the reference raises "convergence failed". */
int sev_irlba_centered_csc(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                           const void *B, const void *mu, int dtype, int64_t nu, int64_t work,
                           int64_t maxit, double tol, const double *init, double *s, double *U,
                           double *V, int precision, int64_t *iters, int64_t *nmult);

/* libcell's irlba entry point, bound unchanged by src/irlba.jl:66-71
 * (BlasInt = int64 arguments, Cint status): the truncated SVD of any operator
 * the caller applies through `matmul` (src/irlba.jl:22-38): y (n) = A' x (m)
 * when trans == 'T', else y (m) = A x (n); x and y are host arrays and temp
 * holds max(m, n) doubles.  `randv` (src/irlba.jl:40-45) fills a host
 * vector with normal deviates; it is called only when init is NULL.  restart
 * = d > 0 warm-starts from the d singular triplets already in the first d
 * entries / columns of s, U, V (src/irlba.jl:87-99).  The Lanczos vectors,
 * reorthogonalisation and restarts run on the GPU; every product copies one
 * vector to the host and back.  Returns 0, or a nonzero status where the
 * reference raises "convergence failed"; the cause is also printed to stderr.
 * Limits: work <= 128 (nu <= 121 with the reference's work = nu + 7).  The
 * integer arguments are 64-bit: this binds an ILP64 Severo build (BlasInt =
 * Int64, deps/build.jl:17-57); an LP64 build passes 32-bit BlasInt and needs
 * its ccall argument types widened to Int64 (the values are unchanged). */
typedef void (*sev_irlba_randv_fn)(double *y, int32_t n, void *data);
typedef void (*sev_irlba_matmul_fn)(double *y, char trans, double *x, double *temp, void *data);
int irlba(int64_t m, int64_t n, int64_t nu, int64_t work, int64_t maxit, int64_t restart,
          double tol, double *init, double *s, double *U, double *V, sev_irlba_randv_fn randv,
          sev_irlba_matmul_fn matmul, void *data);

/* Exact Euclidean kNN with the native libcell signature
 * (src/neighbours.jl:32-39 and :54-61).  X is n x d column-major with
 * leading dimension ld; nn_index (1-based) and distances are n x k
 * column-major.  ntables and seed are accepted and ignored (the search is
 * exact: D = sum_t (x_it - x_jt)^2 accumulated in t order, ties by lower
 * index, distance sqrt(D)).  Errors are reported via sev_last_error(). */
void FindNeighboursEuclidean32(const float *X, int n, int d, int ld, int k, int ntables,
                               int include_self, int seed, int32_t *nn_index, float *distances);
void FindNeighboursEuclidean64(const double *X, int n, int d, int ld, int k, int ntables,
                               int include_self, int seed, int32_t *nn_index, double *distances);

/* Exact cosine kNN with the native libcell signature (the CosineDist
 * methods of src/neighbours.jl:41-50 and :63-72).  Each row is first scaled
 * to unit length in its precision (sum of squares in t order, IEEE sqrt and
 * division; an all-zero row stays zero); the Euclidean search above then
 * runs on the unit rows, and the distance reported is D / 2 of the unit
 * rows (1 - cos in exact arithmetic). */
void FindNeighboursCosine32(const float *X, int n, int d, int ld, int k, int ntables,
                            int include_self, int seed, int32_t *nn_index, float *distances);
void FindNeighboursCosine64(const double *X, int n, int d, int ld, int k, int ntables,
                            int include_self, int seed, int32_t *nn_index, double *distances);

/* Same search with a status and a precision form: SEV_PREC_FP32 (FP32
 * candidates on MFMA), SEV_PREC_BF16 (BF16 candidates on MFMA) -- both
 * followed by an exact FP32 re-rank and a certificate, so the result equals
 * the exact search -- or SEV_PREC_EXACT (exhaustive).  Float64 input always
 * uses the exhaustive search.  *n_fallback (may be NULL) receives the number
 * of query rows that were recomputed exhaustively. */
int sev_knn(const void *X, int64_t n, int64_t d, int64_t ld, int dtype, int64_t k,
            int include_self, int precision, int32_t *nn_index, void *distances,
            int64_t *n_fallback);
/* The cosine form of sev_knn (unit rows, distance D / 2 of the unit rows),
 * with the same arguments and precision forms. */
int sev_knn_cosine(const void *X, int64_t n, int64_t d, int64_t ld, int dtype, int64_t k,
                   int include_self, int precision, int32_t *nn_index, void *distances,
                   int64_t *n_fallback);

/* Jaccard SNN graph (src/neighbours.jl:79, :88-93): from nn_index (n x k
 * column-major, 1-based), S = nn'nn, w = x/(k + (k - x)) in dtype, entries
 * with w <= prune dropped (droptol!).  Two calls: with colptr == NULL only
 * *nnz is set; then colptr[n+1], rowval[nnz] (1-based Int64, rows sorted)
 * and nzval[nnz] of dtype are filled. */
int sev_snn_jaccard(int64_t n, int64_t k, const int32_t *nn_index, double prune, int dtype,
                    int64_t *nnz, int64_t *colptr, int64_t *rowval, void *nzval);

/* cluster(SNN; algorithm=:louvain, resolution, nrandomstarts, niterations,
 * group_singletons) (src/clustering.jl:59-104, src/modularity.jl:165-380) on
 * a square graph in 1-based CSC (rows sorted within columns; weights of
 * dtype, finite and >= 0; self-links ignored as Network(snn) does).
 * Louvain with nrandomstarts starts from singletons, each repeated while it
 * gains at least 1e-4 modularity, at most niterations further times; the
 * start of largest modularity is kept; group_singletons != 0 then moves each
 * member of a one-node cluster as group_singletons! does.  Local moving runs
 * in parallel sweeps on the device (seed picks the random label orders), so
 * labels are one outcome of the reference's randomised search, not its
 * exact output; the same seed gives the same labels.  labels[n] receive
 * 1-based clusters in order of first occurrence; *modularity (optional) the
 * best start's modularity before singleton grouping (NaN without edges);
 * *nclusters (optional) the number of clusters.  Fails when group_singletons
 * is set and every cluster has one member (the reference's findmax over no
 * clusters). */
int sev_cluster_louvain(int64_t n, const int64_t *colptr, const int64_t *rowval,
                        const void *nzval, int dtype, double resolution, int64_t nrandomstarts,
                        int64_t niterations, uint64_t seed, int group_singletons,
                        int64_t *labels, double *modularity, int64_t *nclusters);

/* find_variable_features(counts, nfeatures; method=:vst, loess_span)
 * (src/variablefeatures.jl:34-50,128-161): per-gene mean and sd of the
 * counts (mean_std(Float64, A), src/scaling.jl:119-130), a degree-2 loess of
 * log10 sd on log10 mean over the non-constant genes (evaluated exactly at
 * every gene; Loess.jl interpolates between kd-tree vertices), the clipped
 * standardized variance per gene (:21-28), and the nfeatures genes with the
 * largest one.  selected[nfeatures] receives 1-based gene indices in
 * descending order (ties to the lower index); metric[n] (optional) the
 * variances.  rowval is accepted for the CSC's shape and not read. */
int sev_find_variable_features(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                               const int32_t *nzval, int64_t nfeatures, double loess_span,
                               int64_t *selected, double *metric);

/* find_variable_features(counts, nfeatures; method=:dispersion,
 * :meanvarplot or :saunders) (src/variablefeatures.jl:60-91,128-161) on
 * norm = row_norm(counts, 1.0): the per-gene FP64 mean and variance of the
 * relative counts (src/scaling.jl:18-34), then
 *   SEV_HVG_DISPERSION  log(var/mean) (NaN -> 0);
 *   SEV_HVG_MEANVARPLOT the z-score of log(var/mean) within num_bins bins of
 *                       equal width of log1p(mean) (binning SEV_BIN_WIDTH;
 *                       the reference's :frequency binning places the
 *                       smallest mean in bin 0, outside its bin statistics,
 *                       and is refused);
 *   SEV_HVG_SAUNDERS    log10(var) - log10(mean nolan) for genes whose
 *                       var/nolan exceeds mean + q sqrt(mean nolan / m),
 *                       nolan = mean(1 ./ cell totals), q the standard
 *                       normal quantile at 1 - alpha_thresh/n/2 (the
 *                       reference's default alpha_thresh is 0.1); 0 else.
 * rowval: 1-based, strictly increasing within each column.  Outputs as in
 * sev_find_variable_features.  num_bins is read for MEANVARPLOT,
 * alpha_thresh for SAUNDERS. */
int sev_find_variable_features_method(int64_t m, int64_t n, const int64_t *colptr,
                                      const int64_t *rowval, const int32_t *nzval, int method,
                                      int64_t nfeatures, int64_t num_bins, int binning,
                                      double alpha_thresh, int64_t *selected, double *metric);

/* ---- MatrixMarket (10X matrix.mtx[.gz]) --------------------------------- */
/* readMM (src/input.jl:26-115) up to its sparse() call: checks the header
 * ("%%MatrixMarket matrix coordinate real|integer general"), skips blank and
 * '%' lines, reads the size line.  rows/cols/entries as written; *is_real is
 * 1 for the real field (values as Float64), 0 for integer (Int64).  A path
 * ending in ".gz" is inflated (GzipDecompressorStream in maybe_gzip). */
int sev_mtx_info(const char *path, int64_t *rows, int64_t *cols, int64_t *entries,
                 int32_t *is_real);
/* The entries in file order: I[i], J[i] 1-based row and column, the value in
 * Vi (integer field) or Vf (real field); the other may be NULL.  Arrays of
 * `entries` elements.  Errors carry readMM's messages (sev_last_error). */
int sev_mtx_read(const char *path, int64_t *I, int64_t *J, int64_t *Vi, double *Vf);

/* ---- device-resident pipeline ----------------------------------------- */
typedef struct sev_pipeline sev_pipeline;

int sev_pipeline_create(sev_pipeline **out);
void sev_pipeline_destroy(sev_pipeline *p);
int sev_pipeline_synchronize(sev_pipeline *p);

/* Load a cells x genes count matrix (1-based Int64 structure, Int32 counts). */
int sev_pipeline_load_counts(sev_pipeline *p, int64_t m, int64_t n, const int64_t *colptr,
                             const int64_t *rowval, const int32_t *nzval);
/* Load a 10X matrix.mtx[.gz] (genes x cells) as the cells x genes counts:
 * readMM, sparse(I, J, V) (duplicates summed, stored zeros kept) and the
 * copy(X') of _read_10X (src/input.jl:199-208), the transpose and the sums
 * built on the device from the parsed entries.  Counts must fit Int32 (a
 * real field must hold integers, as convert_counts converts it).  *out_m
 * (cells), *out_n (genes) and *out_nnz may be NULL. */
int sev_pipeline_load_mtx(sev_pipeline *p, const char *path, int64_t *out_m, int64_t *out_n,
                          int64_t *out_nnz);
/* Generate the seeded negative-binomial matrix on the device: entry (g, c)
 * is 0 when u < p0[2g+class(c)], else drawn by inverse CDF with
 * q[2g+class(c)] and size r; class(c) = 1 for a boost_frac share of cells. */
int sev_pipeline_synth_counts(sev_pipeline *p, int64_t m, int64_t n, uint64_t seed, double r,
                              const double *p0, const double *q, double boost_frac);
/* Batches (the reference runs its path once per sample, e.g. per dataset of
 * an integration, src/embedding.jl:46-94 and src/neighbours.jl:272-278 per
 * call): S samples share the gene columns and are stacked by rows (cells of
 * sample s are rows row0[s] .. row0[s+1]-1, row0 = prefix sums of cells).
 * sev_pipeline_synth_batch generates sample s like sev_pipeline_synth_counts
 * (seeds[s], r[s], p0/q rows s of S x 2n, boost_frac[s]) and stacks them;
 * sev_pipeline_set_batch declares loaded counts to be such a stack (S = 0
 * clears it).  Normalisation is per cell and needs nothing more; the stages
 * below have batch forms that keep every sample separate, and the SNN of a
 * batch's neighbours is the block-diagonal union of the per-sample graphs. */
int sev_pipeline_synth_batch(sev_pipeline *p, int64_t S, const int64_t *cells, int64_t n,
                             const uint64_t *seeds, const double *r, const double *p0,
                             const double *q, const double *boost_frac);
int sev_pipeline_set_batch(sev_pipeline *p, int64_t S, const int64_t *cells);
int sev_pipeline_counts_info(sev_pipeline *p, int64_t *m, int64_t *n, int64_t *nnz);
int sev_pipeline_get_counts(sev_pipeline *p, int64_t *colptr, int64_t *rowval, int32_t *nzval);

/* find_variable_features(:vst) on the loaded counts (as
 * sev_find_variable_features). */
int sev_pipeline_find_variable_features(sev_pipeline *p, int64_t nfeatures, double loess_span,
                                        int64_t *selected, double *metric);

/* find_variable_features(:dispersion | :meanvarplot | :saunders) on the
 * loaded counts (as sev_find_variable_features_method, same codes and
 * errors); src/variablefeatures.jl:128-161. */
int sev_pipeline_find_variable_features_method(sev_pipeline *p, int method, int64_t nfeatures,
                                               int64_t num_bins, int binning, double alpha_thresh,
                                               int64_t *selected, double *metric);

/* Stage 1: normalize_cells in Float32 (method SEV_LOGNORMALIZE or
 * SEV_RELATIVECOUNTS). */
int sev_pipeline_normalize(sev_pipeline *p, int method, double scale_factor);
int sev_pipeline_get_normalized(sev_pipeline *p, float *nzval);

/* Stage 2: Y[:, features] then scale_features(scale_max) in Float32.
 * features are 1-based gene indices in the order to keep. */
int sev_pipeline_select_scale(sev_pipeline *p, int64_t nfeat, const int64_t *features,
                              double scale_max);
/* Batch form: features is S x nfeat (row s = sample s's 1-based genes);
 * column j of sample s is scaled with sample s's own statistics
 * (src/scaling.jl:199-217 over its cells).  The result is one stacked
 * cells x nfeat CSC. */
int sev_pipeline_select_scale_batch(sev_pipeline *p, int64_t nfeat, const int64_t *features,
                                    double scale_max);
/* nnz of the selected matrix, then its 1-based structure, scaled values
 * and mu' (length nfeat; S x nfeat, row per sample, for a batch). */
int sev_pipeline_scaled_nnz(sev_pipeline *p, int64_t *nnz);
int sev_pipeline_get_scaled(sev_pipeline *p, int64_t *colptr, int64_t *rowval, float *nzval,
                            float *mu);

/* Stage 3: embedding(S, npcs; method=:pca, algorithm=:irlba)
 * (src/embedding.jl:46-94): IRLBA, then Z = U diag(s) kept on the device in
 * Float32 for the neighbour search.  work is the number of Lanczos vectors:
 * 0 takes the reference's npcs + 7 (src/irlba.jl:50); a larger basis (at
 * most 128) converges to the same tolerance in fewer restarts.  precision
 * SEV_PREC_BF16 reads the operator as BF16 and stores the left Lanczos
 * basis (cells x work) in FP32; every dot, norm and the recurrence stay
 * FP64.  SEV_PREC_FP32 keeps the FP32 operator and an FP64 basis. */
int sev_pipeline_pca(sev_pipeline *p, int64_t npcs, int64_t work, int64_t maxit, double tol,
                     const double *init, int precision, int64_t *iters, int64_t *nmult);
/* Batch form: one IRLBA per sample (src/irlba.jl:47-85 on the sample's
 * centered matrix), all run in lockstep over the Lanczos steps with BF16
 * operator values, left bases stored in FP32 and FP64 arithmetic; init is
 * S x nfeat (row per sample);
 * iters[S].  use_graph != 0 records each restart round once as a HIP graph
 * and replays it.  npcs is capped by the smallest sample.  work: Lanczos
 * vectors per sample, 0 for the reference's npcs + 7, else npcs + 1 .. 64
 * (capped by the smallest sample and the feature count). */
int sev_pipeline_pca_batch(sev_pipeline *p, int64_t npcs, int64_t work, int64_t maxit, double tol,
                           const double *init, int use_graph, int64_t *iters, int64_t *nmult);
/* s[npcs]; U (m x npcs) and V (nfeat x npcs) column-major FP64; Z (m x npcs
 * column-major Float32).  Any pointer may be NULL.  For a batch, s is
 * S x npcs (row per sample), V holds S blocks of nfeat x npcs, and U, Z
 * keep the stacked rows. */
int sev_pipeline_get_pca(sev_pipeline *p, double *s, double *U, double *V, float *Z);
/* Replace the embedding with a caller-supplied m x d Float32 matrix
 * (column-major, ld >= m), e.g. to search neighbours of given coordinates. */
int sev_pipeline_set_embedding(sev_pipeline *p, int64_t m, int64_t d, const float *Z, int64_t ld);

/* Stage 4: exact kNN on the first ndims embedding dimensions (ndims <= 0:
 * all).  For a batch, the neighbours of a cell are cells of its own sample
 * (indices stay those of the stacked rows). */
int sev_pipeline_neighbours(sev_pipeline *p, int64_t k, int include_self, int64_t ndims,
                            int precision, int64_t *n_fallback);
/* The same with metric=CosineDist() (src/neighbours.jl:159-204): the search
 * runs on the embedding's rows scaled to unit length and the distances are
 * D / 2 of the unit rows, as in FindNeighboursCosine32. */
int sev_pipeline_neighbours_cosine(sev_pipeline *p, int64_t k, int include_self, int64_t ndims,
                                   int precision, int64_t *n_fallback);
int sev_pipeline_get_neighbours(sev_pipeline *p, int32_t *nn_index, float *distances);
/* FP8 form only: rows left without a certificate after the FP8, BF16 and
 * FP32 candidate passes of the last search (passes[3]). */
int sev_pipeline_knn_passes(sev_pipeline *p, int64_t *passes);

/* Stage 5: Jaccard SNN of the current neighbours (dtype of the weights). */
int sev_pipeline_snn(sev_pipeline *p, double prune, int dtype, int64_t *nnz);

/* cluster(SNN; algorithm=:louvain, ...) on the graph built by
 * sev_pipeline_snn, as sev_cluster_louvain (same arguments, outputs and
 * errors; labels[m] for the m cells). */
int sev_pipeline_cluster(sev_pipeline *p, double resolution, int64_t nrandomstarts,
                         int64_t niterations, uint64_t seed, int group_singletons,
                         int64_t *labels, double *modularity, int64_t *nclusters);
int sev_pipeline_get_snn(sev_pipeline *p, int64_t *colptr, int64_t *rowval, void *nzval);

/* Kernel timing (no counterpart in the reference, which has no tracing):
 * with enable != 0 every kernel launched by this handle is bracketed by HIP
 * events on the handle's stream; enabling again clears the records.
 * sev_pipeline_profile_read synchronises the stream and returns up to max
 * aggregated entries: name (32 bytes each, NUL-terminated), launches, total
 * milliseconds, and the algorithmic bytes and flops of those launches. */
int sev_pipeline_profile(sev_pipeline *p, int enable);
/* The same, bracketing only the regions called name (e.g. "spmv_av"): the
 * other regions then add no events to the stream.  Clears the records. */
int sev_pipeline_profile_only(sev_pipeline *p, const char *name);
int sev_pipeline_profile_read(sev_pipeline *p, int max, char *names, int64_t *launches,
                              double *ms, double *bytes, double *flops, int *count);

#ifdef __cplusplus
}
#endif

#endif /* SEVERO_HIP_H */
