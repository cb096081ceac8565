// Products of the lazy centered matrix S = A - 1 mu' with dense matrices and
// with another centered matrix: the matrix forms of mul! on CenteredMatrix
// (src/scaling.jl:259-294).  A is m x n CSC on the device; products and
// results are FP64, every sum in a fixed order (no atomics), so a call is
// reproducible bit for bit.
//
//   S X   = A X - 1 (mu' X)                       rows of A (CSR built here)
//   S' X  = A' X - mu (1' X)                      column segments of A
//   Sl'Sr = Al'Ar - mul (1'Ar) - (1'Al - m mul') ' mur'
//                                                 rows of Ar per column of Al
//
// The adjoint matrix form of the reference adds the centering term
// (mul!(C, mu, s, α, 1.0) with s = sum(v, dims=1), src/scaling.jl:268-269);
// S' X is A' X - mu (1' X), the same operator as the vector form
// (src/scaling.jl:252-257), and that is what is computed here.
#include "sev_kernels.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace sev {

namespace {

constexpr int kBlock = 256;
constexpr int kNW = kBlock / kWave;
constexpr int64_t kSegChunk = 2048;  // entries of a column per partial sum of S' X
constexpr int kGramW = 4096;         // output columns per pass of the Gram (per wave)

template <bool F64>
__device__ __forceinline__ double load_val(const void *v, int64_t e) {
    if constexpr (F64)
        return ((const double *)v)[e];
    else
        return (double)((const float *)v)[e];
}

unsigned grid_for(int64_t work, int per_block = kBlock) {
    const int64_t want = (work + per_block - 1) / per_block;
    const int64_t cap = (int64_t)device_cu_count() * 8;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, cap));
}

// fixed-order block sum (every thread's value in the tree of the block)
__device__ __forceinline__ double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---- CSR of a CSC matrix -----------------------------------------------------
// keys = the row of every entry, ids = the entry, cols = its column
__global__ void __launch_bounds__(kBlock)
k_entry_keys(int64_t n, const int64_t *cp, const int32_t *row, uint32_t *key, int32_t *id,
             int32_t *col) {
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x)
        for (int64_t e = cp[j] + threadIdx.x; e < cp[j + 1]; e += kBlock) {
            key[e] = (uint32_t)row[e];
            id[e] = (int32_t)e;
            col[e] = (int32_t)j;
        }
}

// rowptr[i] = number of entries with row < i (sorted keys)
__global__ void __launch_bounds__(kBlock)
k_rowptr(const uint32_t *skey, int64_t nnz, int64_t m, int64_t *rowptr) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i <= m;
         i += (int64_t)gridDim.x * kBlock) {
        int64_t lo = 0, hi = nnz;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if ((int64_t)skey[mid] < i)
                lo = mid + 1;
            else
                hi = mid;
        }
        rowptr[i] = lo;
    }
}

template <bool F64>
__global__ void __launch_bounds__(kBlock)
k_csr_gather(int64_t nnz, const int32_t *sid, const int32_t *ecol, const void *val,
             int32_t *col, double *cval) {
    for (int64_t u = (int64_t)blockIdx.x * kBlock + threadIdx.x; u < nnz;
         u += (int64_t)gridDim.x * kBlock) {
        const int32_t e = sid[u];
        col[u] = ecol[e];
        cval[u] = load_val<F64>(val, e);
    }
}

// ---- S X ---------------------------------------------------------------------
// one wavefront per row, lanes over the columns of X (chunks of 64); Xr is
// X row-major (n x p); Y column-major m x p
__global__ void __launch_bounds__(kBlock)
k_spmm_rows(int64_t m, int64_t p, const int64_t *rowptr, const int32_t *col, const double *cval,
            const double *Xr, const double *muX, double alpha, double beta, double *Y) {
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * kNW;
    for (int64_t i = (int64_t)blockIdx.x * kNW + threadIdx.x / kWave; i < m; i += nw) {
        const int64_t b = rowptr[i], e = rowptr[i + 1];
        for (int64_t c = lane; c < p; c += kWave) {
            double acc = 0.0;
            for (int64_t t = b; t < e; ++t) acc += cval[t] * Xr[(int64_t)col[t] * p + c];
            const double y = alpha * (acc - muX[c]);
            double *yo = Y + c * m + i;
            *yo = beta == 0.0 ? y : y + beta * *yo;
        }
    }
}

// muX[c] = mu' X[:, c]  (block per column)
__global__ void __launch_bounds__(kBlock)
k_mu_dot(int64_t n, int64_t p, const double *mu, const double *Xr, double *muX) {
    __shared__ double sh[kBlock];
    for (int64_t c = blockIdx.x; c < p; c += gridDim.x) {
        double s = 0.0;
        if (mu)
            for (int64_t j = threadIdx.x; j < n; j += kBlock) s += mu[j] * Xr[j * p + c];
        const double r = block_sum(s, sh);
        if (threadIdx.x == 0) muX[c] = r;
    }
}

// ---- S' X --------------------------------------------------------------------
// column sums of a dense column-major m x p matrix (block per column)
__global__ void __launch_bounds__(kBlock)
k_dense_colsum(int64_t m, int64_t p, const double *X, double *s) {
    __shared__ double sh[kBlock];
    for (int64_t c = blockIdx.x; c < p; c += gridDim.x) {
        double a = 0.0;
        for (int64_t i = threadIdx.x; i < m; i += kBlock) a += X[c * m + i];
        const double r = block_sum(a, sh);
        if (threadIdx.x == 0) s[c] = r;
    }
}

// Xr (m x p row-major) from X (m x p column-major)
__global__ void __launch_bounds__(kBlock)
k_to_rowmajor(int64_t m, int64_t p, const double *X, double *Xr) {
    const int64_t tot = m * p;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t i = e / p, c = e - i * p;
        Xr[e] = X[c * m + i];
    }
}

// one wavefront per chunk of a column: part[chunk][c] = sum of val * X[row, c]
template <bool F64>
__global__ void __launch_bounds__(kBlock)
k_spmm_seg(int64_t nchunk, int64_t p, const int64_t *cbeg, const int64_t *cend, const int32_t *row,
           const void *val, const double *Xr, double *part) {
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * kNW;
    for (int64_t q = (int64_t)blockIdx.x * kNW + threadIdx.x / kWave; q < nchunk; q += nw) {
        const int64_t b = cbeg[q], e = cend[q];
        for (int64_t c = lane; c < p; c += kWave) {
            double acc = 0.0;
            for (int64_t t = b; t < e; ++t) acc += load_val<F64>(val, t) * Xr[(int64_t)row[t] * p + c];
            part[q * p + c] = acc;
        }
    }
}

// Y[j, c] = alpha (sum of the chunks of column j - mu_j s_c) + beta Y[j, c]
__global__ void __launch_bounds__(kBlock)
k_spmm_seg_final(int64_t n, int64_t p, const int64_t *cchunk, const double *part, const double *mu,
                 const double *s, double alpha, double beta, double *Y) {
    const int64_t tot = n * p;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t c = e / n, j = e - c * n;
        double acc = 0.0;
        for (int64_t q = cchunk[j]; q < cchunk[j + 1]; ++q) acc += part[q * p + c];
        const double y = alpha * (acc - (mu ? mu[j] : 0.0) * s[c]);
        Y[e] = beta == 0.0 ? y : y + beta * Y[e];
    }
}

// ---- Sl' Sr ------------------------------------------------------------------
// column sums of a CSC matrix (wavefront per column, lanes then a fixed
// butterfly)
template <bool F64>
__global__ void __launch_bounds__(kBlock)
k_csc_colsum(int64_t n, const int64_t *cp, const void *val, double *s) {
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * kNW;
    for (int64_t j = (int64_t)blockIdx.x * kNW + threadIdx.x / kWave; j < n; j += nw) {
        double a = 0.0;
        for (int64_t e = cp[j] + lane; e < cp[j + 1]; e += kWave) a += load_val<F64>(val, e);
        for (int o = kWave / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, kWave);
        if (lane == 0) s[j] = a;
    }
}

// AA[j, k0 + kk] (column-major nl x nr) = sum over the rows i of column j of
// Al of Al[i, j] Ar[i, k0 + kk]: each wavefront takes every kNW-th entry of
// the column and adds the row of Ar into its own LDS accumulator (the lanes
// of a row hold distinct columns); the kNW accumulators are then added in
// wavefront order.
template <bool F64>
__global__ void __launch_bounds__(kBlock)
k_gram_cols(int64_t nl, const int64_t *cpl, const int32_t *rowl, const void *vall,
            const int64_t *rpr, const int32_t *colr, const double *valr, int64_t k0, int kw,
            double *AA) {
    extern __shared__ double acc[];  // kNW * kw
    const int w = threadIdx.x / kWave, lane = lane_id();
    double *mine = acc + (int64_t)w * kw;
    for (int64_t j = blockIdx.x; j < nl; j += gridDim.x) {
        for (int e = threadIdx.x; e < kNW * kw; e += kBlock) acc[e] = 0.0;
        __syncthreads();
        for (int64_t t = cpl[j] + w; t < cpl[j + 1]; t += kNW) {
            const int64_t i = rowl[t];
            const double a = load_val<F64>(vall, t);
            for (int64_t u = rpr[i] + lane; u < rpr[i + 1]; u += kWave) {
                const int64_t k = (int64_t)colr[u] - k0;
                if (k >= 0 && k < kw) mine[k] += a * valr[u];
            }
        }
        __syncthreads();
        for (int kk = threadIdx.x; kk < kw; kk += kBlock) {
            double s = 0.0;
            for (int v = 0; v < kNW; ++v) s += acc[v * kw + kk];
            AA[(k0 + kk) * nl + j] = s;
        }
        __syncthreads();
    }
}

// C = ((alpha AA + beta C) - alpha mul_j qr_k) - alpha ql_j mur_k, the
// reference's order of the three terms (src/scaling.jl:280-292)
__global__ void __launch_bounds__(kBlock)
k_gram_final(int64_t nl, int64_t nr, const double *AA, const double *mul, const double *qr,
             const double *ql, const double *mur, double alpha, double beta, double *C) {
    const int64_t tot = nl * nr;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t k = e / nl, j = e - k * nl;
        double c = alpha * AA[e];
        if (beta != 0.0) c += beta * C[e];
        if (mul) c -= alpha * (mul[j] * qr[k]);
        if (mur) c -= alpha * (ql[j] * mur[k]);
        C[e] = c;
    }
}

// ql = colsums(Al) - m mul
__global__ void __launch_bounds__(kBlock)
k_sub_mmu(int64_t n, int64_t m, const double *mu, double *q) {
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * kBlock)
        if (mu) q[j] -= (double)m * mu[j];
}

// CSR (rows of A in column order) of the device CSC
struct DevCsr {
    DevBuf<int64_t> rowptr;
    DevBuf<int32_t> col;
    DevBuf<double> val;
};

void csc_to_csr(const DevCsc &A, DevCsr &R, hipStream_t st) {
    SEV_REQUIRE(A.nnz < ((int64_t)1 << 31) - 1, "more than 2^31 - 2 entries");
    const int64_t nnz = A.nnz;
    R.rowptr.ensure((size_t)A.m + 1);
    R.col.ensure((size_t)nnz);
    R.val.ensure((size_t)nnz);
    if (nnz == 0) {
        SEV_HIP(hipMemsetAsync(R.rowptr.get(), 0, sizeof(int64_t) * (A.m + 1), st));
        return;
    }
    DevBuf<uint32_t> key, skey;
    DevBuf<int32_t> id, sid, ecol;
    key.ensure((size_t)nnz);
    skey.ensure((size_t)nnz);
    id.ensure((size_t)nnz);
    sid.ensure((size_t)nnz);
    ecol.ensure((size_t)nnz);
    hipLaunchKernelGGL(k_entry_keys, dim3(grid_for(A.n, 1)), dim3(kBlock), 0, st, A.n, A.cp,
                       A.row, key.get(), id.get(), ecol.get());
    SEV_LAUNCH_CHECK();
    int end_bit = 1;
    while (end_bit < 32 && (A.m >> end_bit) != 0) ++end_bit;
    size_t tb = 0;
    SEV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.get(), skey.get(), id.get(),
                                               sid.get(), (int)nnz, 0, end_bit, st));
    DevBuf<unsigned char> tmp;
    tmp.ensure(tb);
    SEV_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tb, key.get(), skey.get(), id.get(),
                                               sid.get(), (int)nnz, 0, end_bit, st));
    hipLaunchKernelGGL(k_rowptr, dim3(grid_for(A.m + 1)), dim3(kBlock), 0, st, skey.get(), nnz,
                       A.m, R.rowptr.get());
    SEV_LAUNCH_CHECK();
    if (A.f64)
        hipLaunchKernelGGL(k_csr_gather<true>, dim3(grid_for(nnz)), dim3(kBlock), 0, st, nnz,
                           sid.get(), ecol.get(), A.val, R.col.get(), R.val.get());
    else
        hipLaunchKernelGGL(k_csr_gather<false>, dim3(grid_for(nnz)), dim3(kBlock), 0, st, nnz,
                           sid.get(), ecol.get(), A.val, R.col.get(), R.val.get());
    SEV_LAUNCH_CHECK();
    SEV_HIP(hipStreamSynchronize(st));  // the temporaries are freed on return
}

// dense host column-major (ld) <-> device column-major (rows x cols, ld rows)
void upload_dense(DevBuf<double> &d, const double *h, int64_t rows, int64_t cols, int64_t ld,
                  hipStream_t st) {
    d.ensure((size_t)(rows * cols));
    if (rows * cols == 0) return;
    SEV_HIP(hipMemcpy2DAsync(d.get(), sizeof(double) * rows, h, sizeof(double) * ld,
                             sizeof(double) * rows, cols, hipMemcpyHostToDevice, st));
}

void download_dense(double *h, int64_t ld, const double *d, int64_t rows, int64_t cols,
                    hipStream_t st) {
    if (rows * cols == 0) return;
    SEV_HIP(hipMemcpy2DAsync(h, sizeof(double) * ld, d, sizeof(double) * rows,
                             sizeof(double) * rows, cols, hipMemcpyDeviceToHost, st));
}

}  // namespace

void centered_mul_mat(const DevCsc &A, bool trans, int64_t p, const double *X, int64_t ldx,
                      double *Y, int64_t ldy, double alpha, double beta, hipStream_t st) {
    const int64_t lx = trans ? A.m : A.n, ly = trans ? A.n : A.m;
    SEV_REQUIRE(p >= 0, "negative column count");
    SEV_REQUIRE(p == 0 || (X && Y), "NULL matrix");
    SEV_REQUIRE(ldx >= std::max<int64_t>(1, lx) && ldy >= std::max<int64_t>(1, ly),
                "leading dimension smaller than the rows");
    if (p == 0 || ly == 0) return;
    DevBuf<double> dX, dXr, dY, aux;
    upload_dense(dX, X, lx, p, ldx, st);
    dY.ensure((size_t)(ly * p));
    if (beta != 0.0) upload_dense(dY, Y, ly, p, ldy, st);
    dXr.ensure((size_t)(lx * p));
    hipLaunchKernelGGL(k_to_rowmajor, dim3(grid_for(lx * p)), dim3(kBlock), 0, st, lx, p,
                       dX.get(), dXr.get());
    SEV_LAUNCH_CHECK();
    aux.ensure((size_t)p);
    if (!trans) {
        hipLaunchKernelGGL(k_mu_dot, dim3(grid_for(p, 1)), dim3(kBlock), 0, st, A.n, p, A.mu,
                           dXr.get(), aux.get());
        SEV_LAUNCH_CHECK();
        DevCsr R;
        csc_to_csr(A, R, st);
        hipLaunchKernelGGL(k_spmm_rows, dim3(grid_for(A.m, kNW)), dim3(kBlock), 0, st, A.m, p,
                           R.rowptr.get(), R.col.get(), R.val.get(), dXr.get(), aux.get(), alpha,
                           beta, dY.get());
        SEV_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_dense_colsum, dim3(grid_for(p, 1)), dim3(kBlock), 0, st, A.m, p,
                           dX.get(), aux.get());
        SEV_LAUNCH_CHECK();
        // every column in chunks of kSegChunk entries (one chunk for an
        // empty column)
        std::vector<int64_t> cp((size_t)A.n + 1), cb, ce, cc((size_t)A.n + 1);
        SEV_HIP(hipMemcpyAsync(cp.data(), A.cp, sizeof(int64_t) * (A.n + 1),
                               hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        for (int64_t j = 0; j < A.n; ++j) {
            cc[(size_t)j] = (int64_t)cb.size();
            int64_t b = cp[(size_t)j];
            do {
                const int64_t e = std::min(cp[(size_t)j + 1], b + kSegChunk);
                cb.push_back(b);
                ce.push_back(e);
                b = e;
            } while (b < cp[(size_t)j + 1]);
        }
        cc[(size_t)A.n] = (int64_t)cb.size();
        const int64_t nchunk = (int64_t)cb.size();
        DevBuf<int64_t> dcb, dce, dcc;
        DevBuf<double> part;
        dcb.ensure((size_t)nchunk);
        dce.ensure((size_t)nchunk);
        dcc.ensure((size_t)A.n + 1);
        part.ensure((size_t)(nchunk * p));
        SEV_HIP(hipMemcpyAsync(dcb.get(), cb.data(), sizeof(int64_t) * nchunk,
                               hipMemcpyHostToDevice, st));
        SEV_HIP(hipMemcpyAsync(dce.get(), ce.data(), sizeof(int64_t) * nchunk,
                               hipMemcpyHostToDevice, st));
        SEV_HIP(hipMemcpyAsync(dcc.get(), cc.data(), sizeof(int64_t) * (A.n + 1),
                               hipMemcpyHostToDevice, st));
        if (A.f64)
            hipLaunchKernelGGL(k_spmm_seg<true>, dim3(grid_for(nchunk, kNW)), dim3(kBlock), 0, st,
                               nchunk, p, dcb.get(), dce.get(), A.row, A.val, dXr.get(),
                               part.get());
        else
            hipLaunchKernelGGL(k_spmm_seg<false>, dim3(grid_for(nchunk, kNW)), dim3(kBlock), 0,
                               st, nchunk, p, dcb.get(), dce.get(), A.row, A.val, dXr.get(),
                               part.get());
        SEV_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_spmm_seg_final, dim3(grid_for(A.n * p)), dim3(kBlock), 0, st, A.n,
                           p, dcc.get(), part.get(), A.mu, aux.get(), alpha, beta, dY.get());
        SEV_LAUNCH_CHECK();
        download_dense(Y, ldy, dY.get(), ly, p, st);
        SEV_HIP(hipStreamSynchronize(st));
        return;
    }
    download_dense(Y, ldy, dY.get(), ly, p, st);
    SEV_HIP(hipStreamSynchronize(st));
}

void centered_gram(const DevCsc &Al, const DevCsc *Ar_, double *C, int64_t ldc, double alpha,
                   double beta, hipStream_t st) {
    const DevCsc &Ar = Ar_ ? *Ar_ : Al;
    SEV_REQUIRE(Al.m == Ar.m, "the two matrices need the same number of rows");
    const int64_t nl = Al.n, nr = Ar.n;
    SEV_REQUIRE(C != nullptr || nl * nr == 0, "NULL output");
    SEV_REQUIRE(ldc >= std::max<int64_t>(1, nl), "leading dimension smaller than the rows");
    if (nl * nr == 0) return;
    DevCsr R;
    csc_to_csr(Ar, R, st);
    DevBuf<double> qr, ql, AA, dC;
    qr.ensure((size_t)nr);
    ql.ensure((size_t)nl);
    if (Ar.f64)
        hipLaunchKernelGGL(k_csc_colsum<true>, dim3(grid_for(nr, kNW)), dim3(kBlock), 0, st, nr,
                           Ar.cp, Ar.val, qr.get());
    else
        hipLaunchKernelGGL(k_csc_colsum<false>, dim3(grid_for(nr, kNW)), dim3(kBlock), 0, st, nr,
                           Ar.cp, Ar.val, qr.get());
    SEV_LAUNCH_CHECK();
    if (Al.f64)
        hipLaunchKernelGGL(k_csc_colsum<true>, dim3(grid_for(nl, kNW)), dim3(kBlock), 0, st, nl,
                           Al.cp, Al.val, ql.get());
    else
        hipLaunchKernelGGL(k_csc_colsum<false>, dim3(grid_for(nl, kNW)), dim3(kBlock), 0, st, nl,
                           Al.cp, Al.val, ql.get());
    SEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sub_mmu, dim3(grid_for(nl)), dim3(kBlock), 0, st, nl, Al.m, Al.mu,
                       ql.get());
    SEV_LAUNCH_CHECK();
    AA.ensure((size_t)(nl * nr));
    const int kw = (int)std::min<int64_t>(nr, kGramW);
    const size_t lds = sizeof(double) * kNW * (size_t)kw;
    if (lds > 65536) {
        // set on every call: the attribute belongs to the current device
        const void *fn = Al.f64 ? (const void *)k_gram_cols<true> : (const void *)k_gram_cols<false>;
        SEV_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(double) * kNW * kGramW)));
    }
    const unsigned g = (unsigned)std::min<int64_t>(nl, (int64_t)device_cu_count() * 4);
    for (int64_t k0 = 0; k0 < nr; k0 += kw) {
        const int w = (int)std::min<int64_t>(kw, nr - k0);
        if (Al.f64)
            hipLaunchKernelGGL(k_gram_cols<true>, dim3(g), dim3(kBlock), lds, st, nl, Al.cp,
                               Al.row, Al.val, R.rowptr.get(), R.col.get(), R.val.get(), k0, w,
                               AA.get());
        else
            hipLaunchKernelGGL(k_gram_cols<false>, dim3(g), dim3(kBlock), lds, st, nl, Al.cp,
                               Al.row, Al.val, R.rowptr.get(), R.col.get(), R.val.get(), k0, w,
                               AA.get());
        SEV_LAUNCH_CHECK();
    }
    dC.ensure((size_t)(nl * nr));
    if (beta != 0.0) upload_dense(dC, C, nl, nr, ldc, st);
    hipLaunchKernelGGL(k_gram_final, dim3(grid_for(nl * nr)), dim3(kBlock), 0, st, nl, nr,
                       AA.get(), Al.mu, qr.get(), ql.get(), Ar.mu, alpha, beta, dC.get());
    SEV_LAUNCH_CHECK();
    download_dense(C, ldc, dC.get(), nl, nr, st);
    SEV_HIP(hipStreamSynchronize(st));
}

}  // namespace sev
