// Host-side launchers of the device kernels (internal to libsevero_hip).
#pragma once

#include <hip/hip_bf16.h>

#include "sev_common.hpp"
#include "sev_prof.hpp"

#include <functional>
#include <map>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace sev {

// ---- sev_capi.cpp ---------------------------------------------------------
// Runs fn(ctx) and turns a thrown error into the status and message the C
// ABI reports (sev_last_error).
int guarded_call(void (*fn)(void *), void *ctx);
// guarded_call for a callable
template <class F>
int guarded_fn(F &&f) {
    using G = typename std::remove_reference<F>::type;
    return guarded_call([](void *c) { (*(G *)c)(); }, (void *)&f);
}
// Throws unless a gfx950 device is current.
void require_device();

// ---- sev_mtx.cpp ------------------------------------------------------------
// A MatrixMarket coordinate file (plain or .gz): the header, comment and
// size lines are read on construction; parse() calls emit(i, row, col,
// ivalue, fvalue) for the entries i in [0, entries) (1-based row/col as
// written; ivalue for the integer field, fvalue for the real one) from
// several host threads, and done(first, count, thread) after each thread's
// range.  Errors throw std::runtime_error with readMM's messages.
class MtxReader {
  public:
    explicit MtxReader(const char *path);
    ~MtxReader();
    int64_t rows() const { return rows_; }
    int64_t cols() const { return cols_; }
    int64_t entries() const { return entries_; }
    bool is_real() const { return real_; }
    void parse(const std::function<void(int64_t, int64_t, int64_t, int64_t, double)> &emit,
               const std::function<void(int64_t, int64_t, int)> &done);

  private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
    int64_t rows_ = 0, cols_ = 0, entries_ = 0;
    bool real_ = false;
};

// ---- sev_coo.hip -------------------------------------------------------------
// The cells x genes CSC of the transpose of a genes x cells coordinate list:
// keys[i] = gene << 32 | cell and vals[i] (both overwritten), n entries.
// Entries with equal keys are summed (Julia's sparse(I, J, V) combines
// duplicates with +; stored zeros stay).  Outputs colptr (ngenes + 1), the
// cell rows (ascending in each column) and the int32 sums; *nnz receives
// the number of distinct keys.  Throws if a sum leaves the int32 range.
void coo_to_csc(int64_t ngenes, uint64_t *keys, int64_t *vals, int64_t n, DevBuf<int64_t> &cp,
                DevBuf<int32_t> &row, DevBuf<int32_t> &val, int64_t &nnz, hipStream_t st);

// ---- sev_noalloc.hip -------------------------------------------------------
// Resolves the device address of the allocation-free entry points' static
// scratch array on the current device (once per device).
void noalloc_setup();

// ---- sev_scan.hip --------------------------------------------------------
// out[i] = sum_{j<i} in[j] for i in [0, n]; out has n+1 entries.  Uses the
// caller-provided scratch of scan_scratch_elems(n) int64 values.  in may
// alias out.
size_t scan_scratch_elems(int64_t n);
void scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *scratch,
                        hipStream_t st);
// Same for int32 counts widened to int64 offsets.
void scan_exclusive_i32(const int32_t *in, int64_t *out, int64_t n, int64_t *scratch,
                        hipStream_t st);

// out[p] = in[p] - 1 for 1-based Int64 rows; *bad is set when a row falls
// outside [1, m].
void rows_to_device(const int64_t *in, int64_t len, int64_t m, int32_t *out, int32_t *bad,
                    hipStream_t st);

// ---- sev_prep.hip --------------------------------------------------------
// Synthetic NB count matrix, cells x genes CSC.  Two calls: counts only
// (rowval == nullptr) fills colptr and returns nnz, then the fill.
struct SynthPlan {
    int64_t cells = 0, genes = 0, chunk = 0, nchunk = 0;
};
int64_t synth_count(int64_t C, int64_t G, uint64_t seed, double r, const double *p0_d,
                    const double *q_d, double boost_frac, uint8_t *cls_d, int64_t *chunk_cnt_d,
                    int64_t *chunk_off_d, int64_t *colptr_d, int64_t *scratch, hipStream_t st,
                    SynthPlan *plan);
void synth_fill(const SynthPlan &plan, uint64_t seed, double r, const double *p0_d,
                const double *q_d, const uint8_t *cls_d, const int64_t *chunk_off_d,
                int32_t *rowval_d, int32_t *nzval_d, hipStream_t st);
void synth_classes(int64_t C, uint64_t seed, double boost_frac, uint8_t *cls_d, hipStream_t st);

// Per-cell totals and the per-nonzero map of normalize_cells.
// colptr/row: CSC with strictly increasing rows in every column
void row_totals(const int64_t *colptr, const int32_t *row, const int32_t *cnt, int64_t nnz,
                int64_t ncells, int64_t ngenes, uint32_t *tot, hipStream_t st);
// The same with the totals cleared by a kernel instead of a memset.
void row_totals_fill(const int64_t *colptr, const int32_t *row, const int32_t *cnt, int64_t nnz,
                     int64_t ncells, int64_t ngenes, uint32_t *tot, hipStream_t st);
// cnt[r] = number of stored entries of row r (same CSC requirement)
void row_counts(const int64_t *colptr, const int32_t *row, int64_t nnz, int64_t m, int64_t n,
                int32_t *cnt, hipStream_t st);
// bad |= 1 when some column's rows are not strictly increasing
void check_sorted(int64_t n, const int64_t *colptr, const int32_t *row, int32_t *bad,
                  hipStream_t st);
void lognorm_f32(const int32_t *row, const int32_t *cnt, const uint32_t *tot, int64_t nnz,
                 float sf, int dolog, float *out, hipStream_t st);
void lognorm_f64(const int32_t *row, const int32_t *cnt, const uint32_t *tot, int64_t nnz,
                 double sf, int dolog, double *out, hipStream_t st);

// Column selection (feature gather) keeping the feature order.
void gather_lengths(const int64_t *feat, const int64_t *colptr, int64_t nfeat, int64_t *len,
                    hipStream_t st);
template <class T>
void gather_copy(const int64_t *feat, const int64_t *colptr, const int32_t *row, const T *val,
                 int64_t nfeat, const int64_t *fcolptr, int32_t *orow, T *oval, hipStream_t st);

// scale_data: sequential Welford per column then the clipped map.
template <class T>
void scale_stats(int64_t m, int64_t ncol, const int64_t *colptr, const T *x, T scale_max,
                 T *sd, T *mup, T *clip, hipStream_t st);
template <class T>
void scale_map(int64_t ncol, const int64_t *colptr, const T *x, const T *sd, const T *clip,
               T *out, hipStream_t st);

// mean_std(Float64, A) of Int32 counts per column (FP64 Welford in row order).
void count_mean_std(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *x, double *mu,
                    double *sd, hipStream_t st);

// mean_var(Float64, row_norm(A, 1.0)) per column: FP64 Welford over x / s_row
// (tot = the per-row totals).
void rel_mean_var(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *row,
                  const int32_t *x, const uint32_t *tot, double *mu, double *var, hipStream_t st);

// ---- sev_hvg.hip ---------------------------------------------------------
// find_variable_features(:vst) on device-resident counts (cells x genes CSC,
// rows sorted within columns): selected[0..nfeat) = 0-based genes by
// descending clipped standardized variance (ties to the lower gene);
// metric_out (host, G entries) optional.
void find_variable_features(int64_t m, int64_t G, const int64_t *colptr_d, const int32_t *cnt_d,
                            int64_t nfeat, double span, int64_t *selected, double *metric_out,
                            hipStream_t st);
// The :dispersion (method 1), :meanvarplot (2, width bins) and :saunders (3)
// selections on device-resident counts (row_d: 0-based rows, sorted within
// columns); same outputs.
void find_variable_features_method(int64_t m, int64_t G, const int64_t *colptr_d,
                                   const int32_t *row_d, const int32_t *cnt_d, int method,
                                   int64_t nfeat, int64_t nbins, double alpha, int64_t *selected,
                                   double *metric_out, hipStream_t st);

// ---- sev_louvain.hip -----------------------------------------------------
// cluster(SNN; algorithm=:louvain) on a square CSC (0-based rows, sorted
// within columns; nnz weights >= 0 in FP32 (f32) or FP64): labels_d[n]
// receive 0-based clusters in order of first occurrence, *q_out the
// modularity of the best start (NaN without edges); returns the number of
// clusters.
int64_t louvain_cluster(int64_t n, const int64_t *colptr_d, const int32_t *row_d,
                        const void *val_d, bool f32, int64_t nnz, double res, int64_t nstarts,
                        int64_t niter, double min_mod, uint64_t seed, bool group,
                        int32_t *labels_d, double *q_out, hipStream_t st);

// Batches of S samples stacked by rows of one cells x genes CSC.
// soff[s*G + g] (s = 0..S): first entry of column g in sample >= s.
void sample_offsets(int64_t G, int64_t S, const int32_t *row_sample, const int64_t *colptr,
                    const int32_t *row, int64_t *soff, hipStream_t st);
// Virtual column v = j*S + s is gene feat[s*F + j] restricted to sample s:
// src_b[v] = its first entry, len[v] = its length.
void batch_seg_lengths(int64_t F, int64_t S, int64_t G, const int64_t *feat, const int64_t *soff,
                       int64_t *src_b, int64_t *len, hipStream_t st);
void batch_gather(int64_t nv, const int64_t *src_b, const int64_t *vcp, const int32_t *row,
                  const float *val, int32_t *orow, float *oval, hipStream_t st);
// scale_stats over virtual columns, m of column v = cells of sample v % S.
void batch_scale_stats(int64_t ncol, const int64_t *colptr, const float *x, float scale_max,
                       float *sd, float *mup, float *clip, const int64_t *row0, int64_t S,
                       hipStream_t st);
// out[j] = in[j*S], j = 0..n
void stride_ptr(const int64_t *in, int64_t n, int64_t S, int64_t *out, hipStream_t st);
// mu_all[s*F + j] = mup[j*S + s]
void mu_by_sample(const float *mup, int64_t F, int64_t S, float *mu_all, hipStream_t st);
// Stacking S sample CSCs (G columns each, device column pointers cps_d[s]):
// len[g*S + s] = entries of column g in sample s; after an exclusive scan
// into starts, stack_copy places sample s (rows offset by roff).
void stack_lengths(int64_t G, int64_t S, const int64_t *const *cps_d, int64_t *len,
                   hipStream_t st);
void stack_copy(int64_t G, int64_t S, int64_t s, const int64_t *cp, const int32_t *row,
                const int32_t *cnt, int64_t roff, const int64_t *starts, int32_t *orow,
                int32_t *ocnt, hipStream_t st);

}  // namespace sev

namespace sev {

// ---- sev_irlba.hip -------------------------------------------------------
// Device layouts of the centered, scaled matrix A = B - 1 mu' (C x F) used by
// the two Lanczos products:
//  * A v  : CSR with every row padded to a multiple of 4 entries (16-bit
//           column indices), read one 4-entry vector per lane;
//  * A' w : the CSC itself, cut into slabs of kSlabRows rows: off[s*F + j] is
//           the position of column j's first entry in slab s ((nslab+1) x F,
//           slab-major) and lrow holds 16-bit in-slab rows.
constexpr int64_t kSlabRows = 4096;
// CSC value arrays handed to build_slab_csc must hold kValuePad readable
// elements past nnz (A' w reads 4-aligned covers of each column segment).
constexpr int64_t kValuePad = 8;

enum ValType : int { kValF32 = 0, kValBF16 = 1, kValF64 = 2 };

// Temporaries of the operator build, kept with the operator so that a
// rebuild (every step of a pipeline) reuses them.
struct BuildScratch {
    DevBuf<int64_t> scan, cstart;
    DevBuf<int32_t> cnt, ids, sids;
    DevBuf<uint32_t> srow;
    DevBuf<unsigned char> tmp;
};

struct SlabCsc {
    int64_t m = 0, n = 0, nnz = 0, nslab = 0, csr_nnz = 0;
    DevBuf<int64_t> off;      // (nslab + 1) * n
    DevBuf<uint16_t> lrow;    // nnz
    DevBuf<uint16_t> vbf16;   // nnz bf16 values (precision form only)
    DevBuf<uint32_t> lrow12;  // precision form: lrow 12-bit packed (8 entries / 12 bytes)
    DevBuf<int64_t> rowptr;   // m + 1 (padded CSR)
    DevBuf<uint16_t> ccol;    // csr_nnz
    DevBuf<uint32_t> ccol12;  // precision form, n <= 4096: ccol 12-bit packed
    bool col12 = false;       // A v reads ccol12
    DevBuf<unsigned char> cval;  // csr_nnz values of vtype
    int lanes_per_row = 64;       // A v: lanes per CSR row (8/16/32/64)
    int atw_groups = 0;           // A' w: column groups per slab (0: heuristic)
    int atw_lanes = 64;           // A' w: lanes per column segment (4..64)
    const void *vals = nullptr;  // CSC values read by A' w (not owned unless
                                 // they point into vbf16)
    const void *mu = nullptr;    // mu' (not owned), same float type as the
                                 // source values; nullptr = uncentered
    int vtype = kValF32;         // type of vals / cval
    bool mu64 = false;           // mu is double
    BuildScratch build;          // temporaries of build_slab_csc
};

// val is float (src_f64 == false) or double; bf16 converts float values.
void build_slab_csc(SlabCsc &op, int64_t m, int64_t n, const int64_t *colptr,
                    const int32_t *row, const void *val, bool src_f64, const void *mu, bool bf16,
                    hipStream_t st);

// Times the A v lane widths and A' w column-group counts on this operator
// and keeps the fastest (cached per shape; small operators keep the
// heuristics).
void autotune_operator(SlabCsc &op, hipStream_t st);

// ---- sev_centered.hip --------------------------------------------------------
// A (m x n) on the device as CSC: cp (n + 1, 0-based), row (int32), val
// (double when f64, else float); mu (n doubles) centres it, nullptr: none.
struct DevCsc {
    int64_t m = 0, n = 0, nnz = 0;
    const int64_t *cp = nullptr;
    const int32_t *row = nullptr;
    const void *val = nullptr;
    bool f64 = false;
    const double *mu = nullptr;
};
// Y = alpha S X + beta Y (Y m x p, X n x p) or, with trans, alpha S' X +
// beta Y (Y n x p, X m x p), S = A - 1 mu'; X and Y host column-major FP64.
void centered_mul_mat(const DevCsc &A, bool trans, int64_t p, const double *X, int64_t ldx,
                      double *Y, int64_t ldy, double alpha, double beta, hipStream_t st);
// C = alpha Sl' Sr + beta C (C nl x nr, host column-major); Ar null: Sr = Sl.
void centered_gram(const DevCsc &Al, const DevCsc *Ar, double *C, int64_t ldc, double alpha,
                   double beta, hipStream_t st);

// y = A x (trans == 0, y length m) or y = A' x (trans == 1, y length n) on
// the device, A = B - 1 mu'.  Uses w.part / w.part2 as scratch.
struct IrlbaWork;
void op_apply(const SlabCsc &op, int trans, const double *x, double *y, IrlbaWork &w,
              hipStream_t st);

struct IrlbaParams {
    int64_t nu = 50, work = 57, maxit = 1000;
    double tol = 1e-5, svtol = 1e-5;
    int reorth_passes = 1;  // CGS passes on the W (left) side
    // Left basis W stored in FP32 (products, dots and the recurrence stay
    // FP64; the device operator only): halves the reorthogonalisation reads
    int basis32 = 0;
    // Warm start from `restart` previous singular triplets (host arrays,
    // column-major, leading dimensions m and n): their vectors seed the first
    // columns of the Krylov bases and B starts as diag(rs).
    int64_t restart = 0;
    const double *rU = nullptr, *rs = nullptr, *rV = nullptr;
};

// An operator applied by the host: matmul(y, 'N', x, temp, data) sets
// y (m) = A x (n); matmul(y, 'T', x, temp, data) sets y (n) = A' x (m).  x and
// y are host arrays; temp holds max(m, n) doubles of scratch.  This is the
// callback pair of the reference's libcell irlba (src/irlba.jl:22-45).
struct HostOperator {
    void (*matmul)(double *y, char trans, double *x, double *temp, void *data) = nullptr;
    void (*randv)(double *y, int32_t n, void *data) = nullptr;
    void *data = nullptr;
};

struct IrlbaOut {
    int64_t iters = 0, nmult = 0, svd_sweeps = 0;
};

struct IrlbaWork {
    DevBuf<double> W, W2, V, V2, B, BU, BV, BS, svr, T, part, part2, red, Fv, scal, rnd, rpart;
    DevBuf<float> Wf, Wf2;        // the left basis in FP32 (IrlbaParams::basis32)
    DevBuf<double> ycur, wprev;  // with Wf: the current product, the last w_j
    DevBuf<int> flags, info;
};

// Runs the Lanczos bidiagonalisation on the device operator.  init (host,
// length n) is the starting vector.  On return s (host, nu), U (device,
// m x nu column-major) and V (device, n x nu column-major) hold the
// singular triplets.  Throws Error(kErrConvergence) like the reference's
// "convergence failed".
//
// With `hop` set, the products are the host operator's and `op` only gives
// the shape (m, n, nslab; mu must be null).
void irlba_device(const SlabCsc &op, const IrlbaParams &p, const double *init, IrlbaWork &w,
                  double *s, double *U, double *V, hipStream_t st, IrlbaOut *out,
                  const HostOperator *hop = nullptr);

// ---- batched IRLBA ---------------------------------------------------------
// S independent problems whose rows are stacked in one matrix (sample s owns
// rows row0[s] .. row0[s+1]-1; all share the n feature columns, each with its
// own centering).  The operator is a SlabCsc built on the stacked matrix with
// slabs that never straddle a sample (BatchLayout::slab_row0).
struct BatchLayout {
    int64_t S = 0, M = 0;
    std::vector<int64_t> row0;              // S+1 (host)
    std::vector<int64_t> slab_row0;          // nslab+1 (host)
    std::vector<int32_t> slab_sample;        // nslab
    std::vector<int64_t> sample_slab0;       // S+1
    std::vector<int32_t> blk_sample;         // A v row blocks
    std::vector<int64_t> blk_r0;             // nblk+1
    DevBuf<int64_t> row0_d, slab_row0_d, sample_slab0_d, blk_r0_d;
    DevBuf<int32_t> row_sample_d, slab_sample_d, blk_sample_d;
};
// Partition rows (cells per sample) into slabs of at most kSlabRows rows and
// A v blocks of at most 256 rows inside each sample; uploads the tables.
void batch_layout(BatchLayout &L, const std::vector<int64_t> &cells, hipStream_t st);
// Rebuild op.off / op.lrow / op.nslab of a built operator for the layout's
// slabs (A' w partials per slab never mix samples).
void batch_slabs(SlabCsc &op, const BatchLayout &L, const int64_t *colptr, const int32_t *row,
                 hipStream_t st);

struct BatchIrlbaWork {
    DevBuf<double> W[2], V[2], B, BU, BV, BS, svr, T, part, part2, red, Fy, F, scal, rpart;
    DevBuf<float> Wf[2];         // the left bases in FP32 (IrlbaParams::basis32)
    DevBuf<double> ycur, wprev;  // with Wf: the current product, the last w_j
    DevBuf<int> flags, info, state;  // state: j0, done, iters, fail per sample
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (first step, buffer parity)
    std::vector<uint64_t> sig;  // pointers and sizes the graphs were recorded with
    ~BatchIrlbaWork();
};
// Batched IRLBA (mu_all: S x n centering).  init (host, S x n) are the
// starting vectors.  Outputs: s (host, S x nu), U (device, M x nu column-major
// over stacked rows), V (device, S x n x nu), iters (host, S).  graph: record
// each restart round once as a hipGraph and replay it.
void irlba_batch(const SlabCsc &op, const BatchLayout &L, const float *mu_all,
                 const IrlbaParams &p, const double *init, BatchIrlbaWork &w, double *s,
                 double *U, double *V, int64_t *iters, int64_t *nmult, bool graph,
                 hipStream_t st);
// Z[r*rs + c*cs] = float(U[c*M + r] * s[sample(r)*d + c]) for c < d.
void batch_embedding(const double *U, const double *s_dev, const BatchLayout &L, int64_t d,
                     float *Z, int64_t rs, int64_t cs, hipStream_t st);

// Z[r*rs + c*cs] = float(U[c*m + r] * s[c]) for c < d: the embedding
// U*diag(s) (src/embedding.jl:67) rounded to Float32.
void embedding_f32(const double *U, const double *s_dev, int64_t m, int64_t d, float *Z,
                   int64_t rs, int64_t cs, hipStream_t st);

// Host SVD of a small square matrix (sev_svd.cpp): A = U diag(s) V',
// column-major, s descending.
void small_svd(int64_t n, const double *A, double *U, double *s, double *V);

}  // namespace sev

namespace sev {

// ---- sev_knn.hip ---------------------------------------------------------
// Exact Euclidean kNN, the definition FindNeighboursEuclidean32/64
// (src/neighbours.jl:32-63) is given here: D(i,j) = sum_t (x_it - x_jt)^2 in
// T, each difference and square rounded, accumulated in t order (no FMA);
// rank by (D, j); j == i skipped only when include_self == 0; distance
// sqrt(D).  X is column-major (x_it at X[t*ld + i]).  Queries are rows
// qidx[r] (or r when qidx == nullptr); outputs row-major nq x k (0-based
// indices).
constexpr int kMaxK = 64;
template <class T>
void knn_exact(const T *X, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
               const int64_t *qidx, int64_t nq, int32_t *idx, T *dist, hipStream_t st);
// Cosine form: X (n x d, column-major, ld = n) scaled in place to unit rows
// (zero rows stay zero); then D / 2 of the pairs (i, idx[i*k + t]) into dist.
template <class T>
void knn_unit_rows(T *X, int64_t n, int64_t d, hipStream_t st);
template <class T>
void knn_half_pair_dist(const T *X, int64_t n, int64_t d, const int32_t *idx, int64_t k, T *dist,
                        hipStream_t st);
// Same, but result row r is written at row qidx[r] of idx/dist (n x k).
template <class T>
void knn_exact_rows(const T *X, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                    const int64_t *qidx, int64_t nq, int32_t *idx, T *dist, hipStream_t st);

// Batches: queries qidx (scattered into idx/dist rows) search only the rows
// of their own sample (seg = row0 of the samples, row_sample = sample of each
// row); qidx lists every query of a 16-row block from one sample.
void knn_exact_rows_batch(const float *X, int64_t n, int64_t d, int64_t ld, int64_t k,
                          int include_self, const int64_t *qidx, int64_t nq, const int64_t *seg,
                          const int32_t *row_sample, int32_t *idx, float *dist, hipStream_t st);

// ---- sev_knn_mfma.hip ----------------------------------------------------
// The same exact kNN (FP32 definition) by MFMA candidate search (FP32 or
// BF16 operands), exact FP32 re-rank and a certificate; uncertified rows are
// recomputed exhaustively.  Z is n x d column-major (ld).  Outputs row-major
// n x k, 0-based.  *n_fallback = rows recomputed exhaustively.
struct KnnWork {
    DevBuf<unsigned char> zq, zc, zcp, stmp;
    DevBuf<float> zr, nhat, eps, cthr, pval, pthr, scf, hn;
    DevBuf<unsigned> gmax, bmax, skey, skey2;
    DevBuf<uint64_t> skey64, skey64b;
    DevBuf<int32_t> cidx, ccnt, nfail, pidx, pcnt, perm, sval;
    DevBuf<int64_t> fail, fail2, qpad;
};
void knn_mfma(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
              int bf16, KnnWork &w, int32_t *idx, float *dist, hipStream_t st,
              int64_t *n_fallback);
// FP8 (e4m3, per-128-row power-of-two scaling) candidates; rows without a
// certificate go through BF16, then FP32 candidates, then the exhaustive
// kernel.  passes[0..2] (optional) = rows left after each candidate pass.
void knn_mfma_fp8(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                  KnnWork &w, int32_t *idx, float *dist, hipStream_t st, int64_t *n_fallback,
                  int64_t *passes);
// The same per sample of a batch (every row's neighbours are rows of its own
// sample; indices stay global, 0-based).
void knn_mfma_batch(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k,
                    int include_self, int bf16, const BatchLayout &L, KnnWork &w, int32_t *idx,
                    float *dist, hipStream_t st, int64_t *n_fallback);
// Largest deviation of the MFMA lane-layout model from scalar arithmetic
// (mode 0: 16x16x4 FP32, 1: 16x16x32 BF16); 0 when the model holds.
double mfma_selftest(int mode);  // mode 2: 16x16x4 FP64, mode 3: 16x16x32 FP8 e4m3

// Copies 1 GiB between two device buffers with 16-byte loads and stores
// (k_copy_probe) and returns the best rate of 5 launches in GB/s (2 GiB of
// traffic per launch).
double hbm_copy_probe();

// ---- sev_snn.hip ---------------------------------------------------------
// Jaccard SNN graph from 0-based row-major neighbour lists (n x k).
// wtab holds w[x] = x / (k + (k - x)) in T for x = 0..k; entries with
// x < xmin are pruned.  Two phases: count (colptr filled, nnz returned),
// then fill (rows 0-based, values).
struct SnnWork {
    DevBuf<int32_t> deg, cur, rev, cnt, heavy, nheavy, staged;
    DevBuf<unsigned char> stage;  // sorted kept slots of the wave-path columns
    DevBuf<int64_t> rptr, scratch, htab_off, heavyP;
    DevBuf<uint64_t> htab;
    HostPinned<int64_t> hP, hoff;     // host copies: heavy columns' P, table offsets
    HostPinned<int32_t> hcol, horder;  // heavy columns as found, and by class
    int64_t nheavy_h = 0;
    int64_t nheavy_cls[3] = {0, 0, 0};  // heavy columns per table class
};
template <class T>
int64_t snn_count(const int32_t *nn, int64_t n, int64_t k, int xmin, SnnWork &w,
                  int64_t *colptr, hipStream_t st);
template <class T>
void snn_fill(const int32_t *nn, int64_t n, int64_t k, int xmin, const T *wtab, SnnWork &w,
              const int64_t *colptr, int32_t *rows, T *vals, hipStream_t st);

}  // namespace sev
