// Device IRLBA for the centered scaled matrix A = B - 1 mu' (cells x features).
//
// Replaces the native irlba behind src/irlba.jl:47-99 together with the
// CenteredMatrix products of src/scaling.jl:245-257 that it calls back into
// for every Lanczos step.  The operator never leaves the device and is never
// densified: centering is folded into each product (A v = B v - (mu'.v) 1,
// A'w = B'w - (sum w) mu').
//
// Algorithm: augmented implicitly restarted Lanczos bidiagonalisation
// (Baglama & Reichel, SIAM J. Sci. Comput. 27(1), 2005) in the form the
// reference's native library exposes (work = nu + 7 vectors, one classical
// Gram-Schmidt pass of full reorthogonalisation on each side, restart size
// k = nu + #converged capped at work - 3, convergence when nu Ritz values
// have |residual| < tol * max(sigma) and relative change < svtol).  All
// vector arithmetic is FP64; B is read as FP32 (plain form) or BF16
// (precision form) values with FP64 accumulation.
//
// Every reduction has a fixed order (wave butterflies read back from lane 0,
// per-slab partials summed in slab order), so results are bit-for-bit
// reproducible run to run.
#include "sev_kernels.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace sev {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kMaxWork = 64;
constexpr int64_t kMaxFL = 3584;  // largest feature count kept in LDS
constexpr int kRedChunks = 16;    // first-level chunks of the slab partials
constexpr int kDotBlocks = 512;   // blocks of the W' y partial products

__device__ __forceinline__ double ldval(const float *v, int64_t p) { return (double)v[p]; }
__device__ __forceinline__ double ldval(const double *v, int64_t p) { return v[p]; }
__device__ __forceinline__ double ldmu(const float *mu, int64_t i) { return (double)mu[i]; }
__device__ __forceinline__ double ldmu(const double *mu, int64_t i) { return mu[i]; }
__device__ __forceinline__ double ldval(const uint16_t *v, int64_t p) {
    return (double)__uint_as_float((uint32_t)v[p] << 16);
}

__device__ __forceinline__ int wave_index() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

// Block-wide sum in fixed order (waves' lane-0 values summed in wave order).
__device__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) scratch[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s += scratch[i];
    __syncthreads();
    return s;
}

// --------------------------------------------------------------------------
// Layout build: slab offsets and 16-bit in-slab rows.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_slab_offsets(int64_t n, int64_t nslab, const int64_t *colptr, const int32_t *row,
               int64_t *off, uint16_t *lrow) {
    const int64_t j = blockIdx.x;
    const int64_t b = colptr[j], e = colptr[j + 1];
    if (b == e) {
        for (int64_t s = threadIdx.x; s <= nslab; s += kBlock) off[s * n + j] = b;
        return;
    }
    for (int64_t p = b + threadIdx.x; p < e; p += kBlock) {
        const int64_t r = row[p];
        const int64_t sc = r / kSlabRows;
        const int64_t sp = (p == b) ? -1 : (int64_t)row[p - 1] / kSlabRows;
        for (int64_t s = sp + 1; s <= sc; ++s) off[s * n + j] = p;
        if (p == e - 1)
            for (int64_t s = sc + 1; s <= nslab; ++s) off[s * n + j] = e;
        lrow[p] = (uint16_t)(r - sc * kSlabRows);
    }
}

__global__ void __launch_bounds__(kBlock) k_to_bf16(const float *in, uint16_t *out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        // round to nearest even
        uint32_t u = __float_as_uint(in[i]);
        uint32_t lsb = (u >> 16) & 1u;
        u += 0x7FFFu + lsb;
        out[i] = (uint16_t)(u >> 16);
    }
}

// --------------------------------------------------------------------------
// y = B v - (mu'.v) 1 - RF * wprev          (one block per slab)
// Each wave owns a private FP64 accumulator over the slab's rows and walks
// the columns j = wave, wave + kWaves, ...; within one column the rows are
// distinct, so the read-modify-write needs no atomics.  The four wave
// accumulators are summed in wave order.
// --------------------------------------------------------------------------
template <class VT, class MT>
__global__ void __launch_bounds__(kBlock)
k_av(int64_t m, int64_t n, const int64_t *off, const uint16_t *lrow, const VT *val,
     const MT *mu, const double *v, const double *rf, const double *wprev, double *y,
     int use_lds) {
    extern __shared__ double smem[];
    double *acc = smem;                       // kWaves * kSlabRows
    double *vl = smem + kWaves * kSlabRows;   // n (if use_lds)
    __shared__ double dmu_s;
    const int64_t s = blockIdx.x;
    const int64_t r0 = s * kSlabRows;
    const int64_t nrows = (m - r0) < kSlabRows ? (m - r0) : kSlabRows;
    const int w = wave_index(), lane = lane_id();

    if (use_lds)
        for (int64_t i = threadIdx.x; i < n; i += kBlock) vl[i] = v[i];
    for (int64_t i = threadIdx.x; i < kWaves * kSlabRows; i += kBlock) acc[i] = 0.0;
    __syncthreads();
    const double *vs = use_lds ? vl : v;
    if (w == 0) {
        double d = 0.0;
        if (mu)
            for (int64_t i = lane; i < n; i += kWave) d += ldmu(mu, i) * vs[i];
        d = wave_sum(d);
        if (lane == 0) dmu_s = d;
    }
    double *aw = acc + w * kSlabRows;
    for (int64_t j = w; j < n; j += kWaves) {
        const int64_t a = off[s * n + j], b = off[(s + 1) * n + j];
        const double vj = vs[j];
        for (int64_t p = a + lane; p < b; p += kWave) {
            const int lr = lrow[p];
            aw[lr] = fma(ldval(val, p), vj, aw[lr]);
        }
    }
    __syncthreads();
    const double dmu = dmu_s;
    const double rfv = rf ? *rf : 0.0;
    for (int64_t lr = threadIdx.x; lr < nrows; lr += kBlock) {
        double t = acc[lr];
        for (int q = 1; q < kWaves; ++q) t += acc[q * kSlabRows + lr];
        t -= dmu;
        if (wprev) t -= rfv * wprev[r0 + lr];
        y[r0 + lr] = t;
    }
}

// --------------------------------------------------------------------------
// Reorthogonalisation partial products: part[b*kMaxWork + i] = W_i' y over
// the rows of block b, for i < jc.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_dot_w(int64_t m, const double *W, int jc, const double *y, double *part) {
    __shared__ double red[kWaves][32];
    const int64_t per = (m + gridDim.x - 1) / gridDim.x;
    const int64_t rb = (int64_t)blockIdx.x * per;
    const int64_t re = rb + per < m ? rb + per : m;
    for (int c0 = 0; c0 < jc; c0 += 32) {
        double acc[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc[i] = 0.0;
        for (int64_t r = rb + threadIdx.x; r < re; r += kBlock) {
            const double yr = y[r];
#pragma unroll
            for (int i = 0; i < 32; ++i)
                if (c0 + i < jc) acc[i] = fma(W[(int64_t)(c0 + i) * m + r], yr, acc[i]);
        }
        const int w = threadIdx.x / kWave;
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            double t = wave_sum(acc[i]);
            if (lane_id() == 0) red[w][i] = t;
        }
        __syncthreads();
        if (threadIdx.x < 32 && c0 + (int)threadIdx.x < jc) {
            double t = 0.0;
            for (int q = 0; q < kWaves; ++q) t += red[q][threadIdx.x];
            part[(int64_t)blockIdx.x * kMaxWork + c0 + threadIdx.x] = t;
        }
        __syncthreads();
    }
}

// T[i] = sum_b part[b*kMaxWork + i] in block order (single block, 1024 threads:
// 16 threads per column each summing a strided subset, combined in order).
__global__ void __launch_bounds__(1024) k_red_t(int nblk, int jc, const double *part, double *T) {
    const int i = threadIdx.x / 16, q = threadIdx.x % 16;
    double t = 0.0;
    if (i < jc)
        for (int b = q; b < nblk; b += 16) t += part[(int64_t)b * kMaxWork + i];
    // fixed-order combine of the 16 partials (lanes q = 0..15 of a group)
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        double u = __shfl_down(t, o, 16);
        if ((q & (2 * o - 1)) == 0) t += u;
    }
    if (q == 0 && i < jc) T[i] = t;
}

// --------------------------------------------------------------------------
// Fused: y -= W[:,0:jc] T (reorthogonalisation apply), slab partials of
// ||y||^2 and sum(y), and the slab partials of B' y (one block per slab).
// part row s holds n column dots followed by ||y||^2 and sum(y).
// --------------------------------------------------------------------------
template <class VT>
__global__ void __launch_bounds__(kBlock)
k_upd_atw(int64_t m, int64_t n, const int64_t *off, const uint16_t *lrow, const VT *val,
          const double *W, int jc, const double *T, double *y, double *part) {
    __shared__ double wl[kSlabRows];
    __shared__ double tl[kMaxWork];
    __shared__ double scratch[kWaves];
    const int64_t s = blockIdx.x;
    const int64_t r0 = s * kSlabRows;
    const int64_t nrows = (m - r0) < kSlabRows ? (m - r0) : kSlabRows;
    const int w = wave_index(), lane = lane_id();
    if (threadIdx.x < jc) tl[threadIdx.x] = T[threadIdx.x];
    __syncthreads();
    double nrm = 0.0, sm = 0.0;
    for (int64_t lr = threadIdx.x; lr < kSlabRows; lr += kBlock) {
        double yv = 0.0;
        if (lr < nrows) {
            const int64_t r = r0 + lr;
            yv = y[r];
            if (jc > 0) {
                double d = 0.0;
                for (int i = 0; i < jc; ++i) d = fma(W[(int64_t)i * m + r], tl[i], d);
                yv -= d;
                y[r] = yv;
            }
            nrm = fma(yv, yv, nrm);
            sm += yv;
        }
        wl[lr] = yv;
    }
    nrm = block_sum(nrm, scratch);
    sm = block_sum(sm, scratch);
    double *prow = part + s * (n + 2);
    if (threadIdx.x == 0) {
        prow[n] = nrm;
        prow[n + 1] = sm;
    }
    for (int64_t j = w; j < n; j += kWaves) {
        const int64_t a = off[s * n + j], b = off[(s + 1) * n + j];
        double acc = 0.0;
        for (int64_t p = a + lane; p < b; p += kWave) acc = fma(ldval(val, p), wl[lrow[p]], acc);
        acc = wave_sum(acc);
        if (lane == 0) prow[j] = acc;
    }
}

// First-level reduction of slab partials: part2[q*ncol + c] = sum over the
// slabs of chunk q (slab order).
__global__ void __launch_bounds__(kBlock)
k_red_slabs(int64_t nslab, int64_t ncol, const double *part, double *part2) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t q = blockIdx.y;
    if (c >= ncol) return;
    const int64_t per = (nslab + kRedChunks - 1) / kRedChunks;
    const int64_t sb = q * per, se = sb + per < nslab ? sb + per : nslab;
    double t = 0.0;
    for (int64_t s = sb; s < se; ++s) t += part[s * ncol + c];
    part2[q * ncol + c] = t;
}

// y = sum_q part2[q][i] - (sum x) mu'_i  (finishes A' x after k_upd_atw)
template <class MT>
__global__ void __launch_bounds__(kBlock)
k_finish_at(int64_t n, const double *part2, const MT *mu, double *y) {
    const int64_t ncol = n + 2;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, t = 0.0;
    for (int q = 0; q < kRedChunks; ++q) {
        t += part2[q * ncol + i];
        sx += part2[q * ncol + n + 1];
    }
    if (mu) t -= sx * ldmu(mu, i);
    y[i] = t;
}

__device__ double hash_normal(uint64_t seed, uint32_t i) {
    const uint64_t sm = mix64(seed);
    double u1 = u01(sm, 0x51u, i), u2 = u01(sm, 0x52u, i);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// --------------------------------------------------------------------------
// Right-side step (single block): finishes A' w_j from the partials, forms
// F = A' w_j - S v_j, reorthogonalises it against V[:,0:j+1], and writes
// v_{j+1} = F/||F|| plus the bidiagonal entries B(j,j) = S, B(j,j+1) = RF.
// scal: [0] S, [1] RF (next step), [2] ||F|| after the last step.
// flags[0] is set when ||w_j|| underflows (invariant subspace).
// --------------------------------------------------------------------------
template <class MT>
__global__ void __launch_bounds__(1024)
k_right_step(int64_t n, int j, int work, const double *part2, const MT *mu, double *V,
             double *Fv, double *B, double *scal, int *flags, uint64_t seed) {
    __shared__ double scratch[16];
    __shared__ double tsh[kMaxWork];
    __shared__ double sc[2];
    const int64_t ncol = n + 2;
    const double eps2 = 2.0 * 2.220446049250313e-16;
    if (threadIdx.x < 2) {
        double t = 0.0;
        for (int q = 0; q < kRedChunks; ++q) t += part2[q * ncol + n + threadIdx.x];
        sc[threadIdx.x] = t;
    }
    __syncthreads();
    double S = sqrt(sc[0]);
    const double sumy = sc[1];
    if (!(S >= eps2)) {
        if (threadIdx.x == 0) flags[0] = 1;
        S = 1.0;
    }
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double t = 0.0;
        for (int q = 0; q < kRedChunks; ++q) t += part2[q * ncol + i];
        if (mu) t -= sumy * ldmu(mu, i);
        Fv[i] = t / S - S * V[(int64_t)j * n + i];
    }
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int nw = blockDim.x / kWave;
    auto orthog = [&]() {
        for (int c = w; c <= j; c += nw) {
            double d = 0.0;
            for (int64_t i = lane; i < n; i += kWave) d = fma(V[(int64_t)c * n + i], Fv[i], d);
            d = wave_sum(d);
            if (lane == 0) tsh[c] = d;
        }
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
            double d = 0.0;
            for (int c = 0; c <= j; ++c) d = fma(V[(int64_t)c * n + i], tsh[c], d);
            Fv[i] -= d;
        }
        __syncthreads();
    };
    orthog();
    double nf = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) nf = fma(Fv[i], Fv[i], nf);
    const double RF = sqrt(block_sum(nf, scratch));
    if (j + 1 < work) {
        double RFs = RF, div = RF;
        if (RF < eps2) {
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
                Fv[i] = hash_normal(seed, (uint32_t)i);
            __syncthreads();
            orthog();
            double nn = 0.0;
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x) nn = fma(Fv[i], Fv[i], nn);
            div = sqrt(block_sum(nn, scratch));
            RFs = 0.0;
        }
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
            V[(int64_t)(j + 1) * n + i] = Fv[i] / div;
        if (threadIdx.x == 0) {
            B[(int64_t)j * work + j] = S;
            B[(int64_t)(j + 1) * work + j] = RFs;
            scal[0] = S;
            scal[1] = RFs;
        }
    } else if (threadIdx.x == 0) {
        B[(int64_t)j * work + j] = S;
        scal[0] = S;
        scal[2] = RF;
    }
}

__global__ void __launch_bounds__(kBlock) k_scale_col(double *y, int64_t m, const double *scal) {
    const double S = scal[0];
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * kBlock)
        y[r] = y[r] / S;
}

// Y[:, c] = X[:, 0:work] Cf[:, c] for c < kout (X, Y column-major, ld = m);
// optionally Y[:, kout] = extra / scal[2] (the normalised residual vector).
__global__ void __launch_bounds__(kBlock)
k_gemm_small(int64_t m, int work, const double *X, const double *Cf, int kout, double *Y,
             const double *extra, const double *scal) {
    __shared__ double cf[kMaxWork * kMaxWork];
    for (int i = threadIdx.x; i < work * kout; i += kBlock) cf[i] = Cf[(i / work) * work + i % work];
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * kBlock) {
        double x[kMaxWork];
#pragma unroll
        for (int l = 0; l < kMaxWork; ++l) x[l] = l < work ? X[(int64_t)l * m + r] : 0.0;
        for (int c = 0; c < kout; ++c) {
            double a = 0.0;
#pragma unroll
            for (int l = 0; l < kMaxWork; ++l)
                if (l < work) a = fma(x[l], cf[c * work + l], a);
            Y[(int64_t)c * m + r] = a;
        }
        if (extra) Y[(int64_t)kout * m + r] = extra[r] / scal[2];
    }
}

__global__ void __launch_bounds__(kBlock)
k_embed(const double *U, const double *s, int64_t m, int64_t d, float *Z, int64_t rs,
        int64_t cs) {
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * kBlock)
        for (int64_t c = 0; c < d; ++c) Z[r * rs + c * cs] = (float)(U[c * m + r] * s[c]);
}

unsigned grid_rows(int64_t m, int per_cu = 4) {
    int64_t want = ceil_div(m > 0 ? m : 1, kBlock);
    int64_t cap = (int64_t)device_cu_count() * per_cu;
    return (unsigned)(want < cap ? want : cap);
}

}  // namespace

void build_slab_csc(SlabCsc &op, int64_t m, int64_t n, const int64_t *colptr,
                    const int32_t *row, const void *val, bool src_f64, const void *mu, bool bf16,
                    hipStream_t st) {
    SEV_REQUIRE(m > 0 && n > 0, "empty operator");
    SEV_REQUIRE(!(bf16 && src_f64), "bf16 operator needs Float32 source values");
    op.m = m;
    op.n = n;
    op.nslab = ceil_div(m, kSlabRows);
    SEV_HIP(hipMemcpyAsync(&op.nnz, colptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    op.off.ensure((size_t)((op.nslab + 1) * n));
    op.lrow.ensure((size_t)op.nnz);
    op.mu = mu;
    op.mu64 = src_f64;
    op.vals = val;
    op.vtype = src_f64 ? kValF64 : kValF32;
    ProfScope ps("slab_build", st, 6.0 * op.nnz + 8.0 * (op.nslab + 1) * n);
    hipLaunchKernelGGL(k_slab_offsets, dim3((unsigned)n), dim3(kBlock), 0, st, n, op.nslab,
                       colptr, row, op.off.get(), op.lrow.get());
    SEV_LAUNCH_CHECK();
    if (bf16) {
        op.vbf16.ensure((size_t)op.nnz);
        if (op.nnz > 0) {
            hipLaunchKernelGGL(k_to_bf16, dim3(grid_rows(op.nnz, 16)), dim3(kBlock), 0, st,
                               (const float *)val, op.vbf16.get(), op.nnz);
            SEV_LAUNCH_CHECK();
        }
        op.vals = op.vbf16.get();
        op.vtype = kValBF16;
    }
}

namespace {


int val_bytes(const SlabCsc &op) {
    return op.vtype == kValBF16 ? 2 : (op.vtype == kValF64 ? 8 : 4);
}

template <class VT, class MT>
void launch_av_t(const SlabCsc &op, size_t smem, int use_lds, const double *v, const double *rf,
                 const double *wprev, double *y, hipStream_t st) {
    ProfScope ps("spmv_av", st,
                 (double)op.nnz * (val_bytes(op) + 2) + 8.0 * op.m * (wprev ? 2 : 1) + 8.0 * op.n);
    hipLaunchKernelGGL((k_av<VT, MT>), dim3((unsigned)op.nslab), dim3(kBlock), smem, st, op.m, op.n,
                       op.off.get(), op.lrow.get(), (const VT *)op.vals, (const MT *)op.mu, v, rf,
                       wprev, y, use_lds);
    SEV_LAUNCH_CHECK();
}

void launch_av(const SlabCsc &op, size_t smem, int use_lds, const double *v, const double *rf,
               const double *wprev, double *y, hipStream_t st) {
    if (op.vtype == kValBF16)
        launch_av_t<uint16_t, float>(op, smem, use_lds, v, rf, wprev, y, st);
    else if (op.vtype == kValF64)
        launch_av_t<double, double>(op, smem, use_lds, v, rf, wprev, y, st);
    else
        launch_av_t<float, float>(op, smem, use_lds, v, rf, wprev, y, st);
}

void launch_upd_atw(const SlabCsc &op, const double *W, int jc, const double *T, double *y,
                    double *part, hipStream_t st) {
    ProfScope ps("spmv_atw", st,
                 (double)op.nnz * (val_bytes(op) + 2) + 8.0 * op.m * (jc > 0 ? jc + 2 : 1) +
                     8.0 * op.n);
    if (op.vtype == kValBF16)
        hipLaunchKernelGGL(k_upd_atw<uint16_t>, dim3((unsigned)op.nslab), dim3(kBlock), 0, st,
                           op.m, op.n, op.off.get(), op.lrow.get(), (const uint16_t *)op.vals, W,
                           jc, T, y, part);
    else if (op.vtype == kValF64)
        hipLaunchKernelGGL(k_upd_atw<double>, dim3((unsigned)op.nslab), dim3(kBlock), 0, st, op.m,
                           op.n, op.off.get(), op.lrow.get(), (const double *)op.vals, W, jc, T,
                           y, part);
    else
        hipLaunchKernelGGL(k_upd_atw<float>, dim3((unsigned)op.nslab), dim3(kBlock), 0, st, op.m,
                           op.n, op.off.get(), op.lrow.get(), (const float *)op.vals, W, jc, T, y,
                           part);
    SEV_LAUNCH_CHECK();
}

void launch_right_step(const SlabCsc &op, int j, int work, const double *part2, double *V,
                       double *Fv, double *B, double *scal, int *flags, uint64_t seed,
                       hipStream_t st) {
    ProfScope ps("right_step", st, 16.0 * op.n * (j + 1));
    if (op.mu64)
        hipLaunchKernelGGL(k_right_step<double>, dim3(1), dim3(1024), 0, st, op.n, j, work, part2,
                           (const double *)op.mu, V, Fv, B, scal, flags, seed);
    else
        hipLaunchKernelGGL(k_right_step<float>, dim3(1), dim3(1024), 0, st, op.n, j, work, part2,
                           (const float *)op.mu, V, Fv, B, scal, flags, seed);
    SEV_LAUNCH_CHECK();
}

size_t av_smem_bytes(int64_t n, int use_lds) {
    return sizeof(double) * (size_t)(kWaves * kSlabRows + (use_lds ? n : 0));
}

}  // namespace

void irlba_device(const SlabCsc &op, const IrlbaParams &p, const double *init, IrlbaWork &w,
                  double *s, double *U, double *V, hipStream_t st, IrlbaOut *out) {
    const int64_t m = op.m, n = op.n;
    const int64_t nu = p.nu;
    int64_t workl = p.work;
    if (workl < nu) workl = nu + 1;
    if (workl > m) workl = m;
    if (workl > n) workl = n;
    const int work = (int)workl;
    SEV_REQUIRE(nu >= 1 && nu <= work, "irlba: need 1 <= nu <= min(m, n)");
    SEV_REQUIRE(work <= kMaxWork, "irlba: work (nu + 7) exceeds 64 on the device path");
    const double eps = 2.220446049250313e-16;

    w.W.ensure((size_t)(m * work));
    w.W2.ensure((size_t)(m * work));
    w.V.ensure((size_t)(n * work));
    w.V2.ensure((size_t)(n * work));
    w.B.ensure((size_t)(work * work));
    w.BU.ensure((size_t)(work * work));
    w.BV.ensure((size_t)(work * work));
    w.T.ensure(kMaxWork);
    w.part.ensure((size_t)(op.nslab * (n + 2)));
    w.part2.ensure((size_t)(kRedChunks * (n + 2)));
    w.red.ensure((size_t)kDotBlocks * kMaxWork);
    w.Fv.ensure((size_t)n);
    w.scal.ensure(4);
    w.flags.ensure(4);

    // starting vector v_0 = init / ||init||
    std::vector<double> v0(init, init + n);
    double nr = 0.0;
    for (int64_t i = 0; i < n; ++i) nr += v0[i] * v0[i];
    nr = std::sqrt(nr);
    if (nr < 2 * eps) fail(kErrConvergence, "irlba: starting vector is zero");
    for (int64_t i = 0; i < n; ++i) v0[i] /= nr;
    SEV_HIP(hipMemcpyAsync(w.V.get(), v0.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    SEV_HIP(hipMemsetAsync(w.B.get(), 0, sizeof(double) * work * work, st));
    SEV_HIP(hipMemsetAsync(w.flags.get(), 0, sizeof(int) * 4, st));
    SEV_HIP(hipMemsetAsync(w.scal.get(), 0, sizeof(double) * 4, st));

    const int use_lds = n <= kMaxFL ? 1 : 0;
    const size_t av_smem = av_smem_bytes(n, use_lds);
    const int dot_blocks = (int)std::min<int64_t>(kDotBlocks, ceil_div(m, kBlock));

    std::vector<double> Bh((size_t)work * work), BUh((size_t)work * work),
        BVh((size_t)work * work), BSh(work), svr(work, 0.0), res(work);
    double scal_h[4];
    int flags_h[4];
    int64_t iter = 0, nmult = 0, k = nu;
    bool converged = false;

    while (iter < p.maxit) {
        const int j0 = (iter == 0) ? 0 : (int)k;
        for (int j = j0; j < work; ++j) {
            double *Wd = w.W.get(), *Vd = w.V.get();
            double *y = Wd + (int64_t)j * m;
            // W[:,j] = A v_j - RF w_{j-1}
            launch_av(op, av_smem, use_lds, Vd + (int64_t)j * n, j > j0 ? w.scal.get() + 1 : nullptr,
                      j > j0 ? Wd + (int64_t)(j - 1) * m : nullptr, y, st);
            nmult++;
            // reorthogonalise against W[:,0:j]
            const int jc = j;
            if (jc > 0) {
                ProfScope ps("reorth_dot", st, 8.0 * m * (jc + 1));
                hipLaunchKernelGGL(k_dot_w, dim3(dot_blocks), dim3(kBlock), 0, st, m, Wd, jc, y,
                                   w.red.get());
                SEV_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_red_t, dim3(1), dim3(1024), 0, st, dot_blocks, jc,
                                   w.red.get(), w.T.get());
                SEV_LAUNCH_CHECK();
            }
            launch_upd_atw(op, Wd, jc, w.T.get(), y, w.part.get(), st);
            nmult++;
            {
                ProfScope ps("reduce_slabs", st, 8.0 * op.nslab * (n + 2));
                hipLaunchKernelGGL(k_red_slabs, dim3((unsigned)ceil_div(n + 2, kBlock), kRedChunks),
                                   dim3(kBlock), 0, st, op.nslab, n + 2, w.part.get(),
                                   w.part2.get());
                SEV_LAUNCH_CHECK();
            }
            launch_right_step(op, j, work, w.part2.get(), Vd, w.Fv.get(), w.B.get(),
                              w.scal.get(), w.flags.get(), (uint64_t)(iter * work + j) + 17u, st);
            ProfScope ps("scale_col", st, 16.0 * m);
            hipLaunchKernelGGL(k_scale_col, dim3(grid_rows(m)), dim3(kBlock), 0, st, y, m,
                               w.scal.get());
            SEV_LAUNCH_CHECK();
        }
        SEV_HIP(hipMemcpyAsync(Bh.data(), w.B.get(), sizeof(double) * work * work,
                               hipMemcpyDeviceToHost, st));
        SEV_HIP(hipMemcpyAsync(scal_h, w.scal.get(), sizeof scal_h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipMemcpyAsync(flags_h, w.flags.get(), sizeof flags_h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        if (flags_h[0]) fail(kErrConvergence, "irlba: invariant subspace encountered (rank-deficient operator)");
        small_svd(work, Bh.data(), BUh.data(), BSh.data(), BVh.data());
        double RF = scal_h[2];
        const double RFn = RF;
        if (RF < 2 * eps) RF = 0.0;
        double smax = 0.0;
        for (int i = 0; i < work; ++i) {
            if (BSh[i] > smax) smax = BSh[i];
            svr[i] = std::fabs(svr[i] - BSh[i]) / BSh[i];
        }
        for (int i = 0; i < work; ++i) res[i] = RF * BUh[(size_t)i * work + (work - 1)];
        int64_t nconv = 0;
        for (int i = 0; i < work; ++i)
            if (std::fabs(res[i]) < p.tol * smax && svr[i] < p.svtol) nconv++;
        iter++;
        if (nconv >= nu) {
            converged = true;
            break;
        }
        if (iter >= p.maxit) break;
        if (k < nu + nconv) k = nu + nconv;
        if (k > work - 3) k = work - 3;
        if (k < 1) k = 1;
        for (int i = 0; i < work; ++i) svr[i] = BSh[i];
        // restart: V <- [V BV[:,0:k], F/||F||], W <- W BU[:,0:k]
        SEV_HIP(hipMemcpyAsync(w.BU.get(), BUh.data(), sizeof(double) * work * work,
                               hipMemcpyHostToDevice, st));
        SEV_HIP(hipMemcpyAsync(w.BV.get(), BVh.data(), sizeof(double) * work * work,
                               hipMemcpyHostToDevice, st));
        if (RFn <= 0.0) fail(kErrConvergence, "irlba: zero residual vector at restart");
        hipLaunchKernelGGL(k_gemm_small, dim3(grid_rows(n)), dim3(kBlock), 0, st, n, work,
                           w.V.get(), w.BV.get(), (int)k, w.V2.get(), w.Fv.get(), w.scal.get());
        SEV_LAUNCH_CHECK();
        w.V.swap(w.V2);
        ProfScope ps("restart_w", st, 8.0 * m * (work + k), 2.0 * m * work * k);
        hipLaunchKernelGGL(k_gemm_small, dim3(grid_rows(m)), dim3(kBlock), 0, st, m, work,
                           w.W.get(), w.BU.get(), (int)k, w.W2.get(), (const double *)nullptr,
                           w.scal.get());
        SEV_LAUNCH_CHECK();
        w.W.swap(w.W2);
        std::fill(Bh.begin(), Bh.end(), 0.0);
        for (int64_t cc = 0; cc < k; ++cc) {
            Bh[(size_t)cc * work + cc] = BSh[cc];
            Bh[(size_t)k * work + cc] = res[cc];
        }
        SEV_HIP(hipMemcpyAsync(w.B.get(), Bh.data(), sizeof(double) * work * work,
                               hipMemcpyHostToDevice, st));
    }
    if (out) {
        out->iters = iter;
        out->nmult = nmult;
    }
    if (!converged) fail(kErrConvergence, "irlba: convergence failed (maxit reached)");
    for (int64_t i = 0; i < nu; ++i) s[i] = BSh[i];
    SEV_HIP(hipMemcpyAsync(w.BU.get(), BUh.data(), sizeof(double) * work * work,
                           hipMemcpyHostToDevice, st));
    SEV_HIP(hipMemcpyAsync(w.BV.get(), BVh.data(), sizeof(double) * work * work,
                           hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gemm_small, dim3(grid_rows(m)), dim3(kBlock), 0, st, m, work, w.W.get(),
                       w.BU.get(), (int)nu, U, (const double *)nullptr, w.scal.get());
    SEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gemm_small, dim3(grid_rows(n)), dim3(kBlock), 0, st, n, work, w.V.get(),
                       w.BV.get(), (int)nu, V, (const double *)nullptr, w.scal.get());
    SEV_LAUNCH_CHECK();
    SEV_HIP(hipStreamSynchronize(st));
}

void op_apply(const SlabCsc &op, int trans, const double *x, double *y, IrlbaWork &w,
              hipStream_t st) {
    const int64_t m = op.m, n = op.n;
    if (!trans) {
        const int use_lds = n <= kMaxFL ? 1 : 0;
        launch_av(op, av_smem_bytes(n, use_lds), use_lds, x, nullptr, nullptr, y, st);
        return;
    }
    w.part.ensure((size_t)(op.nslab * (n + 2)));
    w.part2.ensure((size_t)(kRedChunks * (n + 2)));
    // k_upd_atw with jc == 0 leaves its y argument unchanged
    launch_upd_atw(op, nullptr, 0, nullptr, const_cast<double *>(x), w.part.get(), st);
    hipLaunchKernelGGL(k_red_slabs, dim3((unsigned)ceil_div(n + 2, kBlock), kRedChunks),
                       dim3(kBlock), 0, st, op.nslab, n + 2, w.part.get(), w.part2.get());
    SEV_LAUNCH_CHECK();
    if (op.mu64)
        hipLaunchKernelGGL(k_finish_at<double>, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock),
                           0, st, n, w.part2.get(), (const double *)op.mu, y);
    else
        hipLaunchKernelGGL(k_finish_at<float>, dim3((unsigned)ceil_div(n, kBlock)), dim3(kBlock),
                           0, st, n, w.part2.get(), (const float *)op.mu, y);
    SEV_LAUNCH_CHECK();
    (void)m;
}

void embedding_f32(const double *U, const double *s_dev, int64_t m, int64_t d, float *Z,
                   int64_t rs, int64_t cs, hipStream_t st) {
    hipLaunchKernelGGL(k_embed, dim3(grid_rows(m)), dim3(kBlock), 0, st, U, s_dev, m, d, Z, rs, cs);
    SEV_LAUNCH_CHECK();
}

}  // namespace sev
