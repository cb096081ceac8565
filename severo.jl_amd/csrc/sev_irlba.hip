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
// vector arithmetic is FP64; B is read as FP32 (plain form), BF16 (precision
// form) or FP64 values with FP64 accumulation.
//
// Every reduction has a fixed order (wave butterflies read back from one
// lane, per-slab partials summed in slab order), so results are bit-for-bit
// reproducible run to run.
#include "sev_kernels.hpp"

#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstdlib>

#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>
#include <tuple>
#include <type_traits>
#include <type_traits>
#include <vector>

namespace sev {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kMaxWork = 128;  // Lanczos vectors (single operator)
constexpr int kMaxWorkB = 64;  // Lanczos vectors per sample of a batch
constexpr int64_t kMaxFL = 7168;     // largest feature count kept in LDS (A v, FP64)
constexpr int64_t kMaxFL32 = 14336;  // the same for the FP32 copy of the BF16 form
constexpr int kRedChunks = 16;       // first-level chunks of the slab partials

__device__ __forceinline__ double ldval(const float *v, int64_t p) { return (double)v[p]; }
__device__ __forceinline__ double ldval(const double *v, int64_t p) { return v[p]; }
__device__ __forceinline__ double ldval(const uint16_t *v, int64_t p) {
    return (double)__uint_as_float((uint32_t)v[p] << 16);
}
__device__ __forceinline__ double ldmu(const float *mu, int64_t i) { return (double)mu[i]; }
__device__ __forceinline__ double ldmu(const double *mu, int64_t i) { return mu[i]; }

// four consecutive values starting at a multiple of 4
__device__ __forceinline__ void ld4(const float *v, int64_t p, double o[4]) {
    const float4 x = *(const float4 *)(v + p);
    o[0] = x.x;
    o[1] = x.y;
    o[2] = x.z;
    o[3] = x.w;
}
__device__ __forceinline__ void ld4(const uint16_t *v, int64_t p, double o[4]) {
    const ushort4 x = *(const ushort4 *)(v + p);
    o[0] = (double)__uint_as_float((uint32_t)x.x << 16);
    o[1] = (double)__uint_as_float((uint32_t)x.y << 16);
    o[2] = (double)__uint_as_float((uint32_t)x.z << 16);
    o[3] = (double)__uint_as_float((uint32_t)x.w << 16);
}
__device__ __forceinline__ void ld4(const double *v, int64_t p, double o[4]) {
    const double2 a = *(const double2 *)(v + p);
    const double2 b = *(const double2 *)(v + p + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
}

__device__ __forceinline__ float cvt_store(float x, float *) { return x; }
__device__ __forceinline__ uint16_t cvt_store(float x, uint16_t *) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ double cvt_store(double x, double *) { return x; }

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

// Reduce-scatter of 32 per-lane values over the 64 lanes of a wavefront by a
// butterfly: each step halves the values a lane keeps and adds its partner's
// copy of them.  On return v[0] of lane l holds the sum over all lanes of
// the input v[l >> 1] (lanes 2t and 2t+1 hold the same sum; the order of the
// additions is fixed).
template <int HALF>
__device__ __forceinline__ void butterfly_step(double (&v)[32], int lane) {
    const bool hi = (lane & (2 * HALF)) != 0;
#pragma unroll
    for (int i = 0; i < HALF; ++i) {
        const double send = hi ? v[i] : v[HALF + i];
        const double keep = hi ? v[HALF + i] : v[i];
        v[i] = keep + __shfl_xor(send, 2 * HALF, kWave);
    }
}

__device__ __forceinline__ double butterfly32(double (&v)[32]) {
    const int lane = lane_id();
    butterfly_step<16>(v, lane);
    butterfly_step<8>(v, lane);
    butterfly_step<4>(v, lane);
    butterfly_step<2>(v, lane);
    butterfly_step<1>(v, lane);
    return v[0] + __shfl_xor(v[0], 1, kWave);
}

// --------------------------------------------------------------------------
// Layout build
// --------------------------------------------------------------------------

// off[s*n + j] = first position of column j with row >= s*R (s = 0..nslab);
// lrow[p] = row - slab base (when lrow != nullptr).
__global__ void __launch_bounds__(kBlock)
k_slab_offsets(int64_t n, int64_t nslab, int64_t R, const int64_t *colptr, const int32_t *row,
               int64_t *off, uint16_t *lrow) {
    // column blockIdx.x; its entries are shared by the gridDim.y blocks of
    // that column (each entry's writes are its own)
    const int64_t j = blockIdx.x;
    const int64_t b = colptr[j], e = colptr[j + 1];
    if (b == e) {
        if (blockIdx.y == 0)
            for (int64_t s = threadIdx.x; s <= nslab; s += kBlock) off[s * n + j] = b;
        return;
    }
    for (int64_t p = b + (int64_t)blockIdx.y * kBlock + threadIdx.x; p < e;
         p += (int64_t)gridDim.y * kBlock) {
        const int64_t r = row[p];
        const int64_t sc = r / R;
        const int64_t sp = (p == b) ? -1 : (int64_t)row[p - 1] / R;
        for (int64_t s = sp + 1; s <= sc; ++s) off[s * n + j] = p;
        if (p == e - 1)
            for (int64_t s = sc + 1; s <= nslab; ++s) off[s * n + j] = e;
        if (lrow) lrow[p] = (uint16_t)(r - sc * R);
    }
}

__global__ void __launch_bounds__(kBlock) k_to_bf16(const float *in, uint16_t *out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = cvt_store(in[i], (uint16_t *)nullptr);
}

// row lengths rounded up to a multiple of R (a power of two)
__global__ void __launch_bounds__(kBlock) k_round_up(int32_t *cnt, int64_t m, int R) {
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * kBlock)
        cnt[r] = (cnt[r] + R - 1) & ~(R - 1);
}

// Padded CSR placement: the entries sorted by row (a stable radix sort keeps
// each row's entries in CSC, i.e. column, order) are copied to
// rowptr[row] + (their rank within the row); each entry's column is the last
// j with colptr[j] <= e.  Consecutive sorted entries land in consecutive
// slots, so the writes are contiguous.  Padding slots were zeroed before.
template <class ST, class DT>
__global__ void __launch_bounds__(kBlock)
k_csr_place(int64_t nnz, int64_t n, const int64_t *colptr, const uint32_t *srow,
            const int32_t *sid, const int64_t *cstart, const int64_t *rowptr, const ST *val,
            uint16_t *ccol, DT *cval) {
    for (int64_t u = (int64_t)blockIdx.x * kBlock + threadIdx.x; u < nnz;
         u += (int64_t)gridDim.x * kBlock) {
        const int64_t r = srow[u];
        const int64_t e = sid[u];
        int64_t lo = 0, hi = n;  // colptr[lo] <= e < colptr[hi]
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (colptr[mid] <= e)
                lo = mid;
            else
                hi = mid;
        }
        const int64_t pos = rowptr[r] + (u - cstart[r]);
        ccol[pos] = (uint16_t)lo;
        cval[pos] = cvt_store(val[e], (DT *)nullptr);
    }
}

// BF16 operators: one 32-bit word per entry, column << 16 | bf16(value), so
// the sorted words hold everything the CSR needs
__global__ void __launch_bounds__(kBlock)
k_csr_words(int64_t n, const int64_t *colptr, const float *val, uint32_t *word) {
    // a column's entries are shared by the gridDim.y blocks of its x index
    const int64_t step = (int64_t)gridDim.y * kBlock;
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x)
        for (int64_t e = colptr[j] + (int64_t)blockIdx.y * kBlock + threadIdx.x;
             e < colptr[j + 1]; e += step)
            word[e] = ((uint32_t)j << 16) | cvt_store(val[e], (uint16_t *)nullptr);
}

__global__ void __launch_bounds__(kBlock)
k_csr_place_words(int64_t nnz, const uint32_t *srow, const uint32_t *sword, const int64_t *cstart,
                  const int64_t *rowptr, uint16_t *ccol, uint16_t *cval) {
    for (int64_t u = (int64_t)blockIdx.x * kBlock + threadIdx.x; u < nnz;
         u += (int64_t)gridDim.x * kBlock) {
        const int64_t r = srow[u];
        const uint32_t wd = sword[u];
        const int64_t pos = rowptr[r] + (u - cstart[r]);
        ccol[pos] = (uint16_t)(wd >> 16);
        cval[pos] = (uint16_t)(wd & 0xFFFFu);
    }
}

__global__ void __launch_bounds__(kBlock) k_iota32(int32_t *v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        v[i] = (int32_t)i;
}

// --------------------------------------------------------------------------
// y = B v - (mu'.v) 1 - RF * wprev   (padded CSR; G lanes per row, every
// lane reads 4 entries per step; the G partial sums are combined by a fixed
// butterfly and the group leader's value is used)
// --------------------------------------------------------------------------
template <class VT, class MT, int G>
__global__ void __launch_bounds__(kBlock)
k_av_csr(int64_t m, int64_t n, const int64_t *rowptr, const uint16_t *col, const VT *val,
         const MT *mu, const double *v, const double *rf, const double *wprev, double *y,
         int use_lds) {
    extern __shared__ double vl[];
    __shared__ double dmu_s;
    const int w = wave_index(), lane = lane_id();
    if (use_lds)
        for (int64_t i = threadIdx.x; i < n; i += kBlock) vl[i] = v[i];
    __syncthreads();
    const double *vs = use_lds ? vl : v;
    if (w == 0) {
        double d = 0.0;
        if (mu)
            for (int64_t i = lane; i < n; i += kWave) d = fma(ldmu(mu, i), vs[i], d);
        d = wave_sum(d);
        if (lane == 0) dmu_s = d;
    }
    __syncthreads();
    const double dmu = dmu_s;
    const double rfv = rf ? *rf : 0.0;
    constexpr int RPW = kWave / G;  // rows per wave
    const int gl = lane % G, gi = lane / G;
    const int64_t stride = (int64_t)gridDim.x * kWaves * RPW;
    for (int64_t r0 = ((int64_t)blockIdx.x * kWaves + w) * RPW; r0 < m; r0 += stride) {
        const int64_t r = r0 + gi;
        double acc = 0.0;
        if (r < m) {
            const int64_t b = rowptr[r], e = rowptr[r + 1];
            for (int64_t p = b + 4 * gl; p < e; p += 4 * G) {
                const ushort4 c4 = *(const ushort4 *)(col + p);
                double x[4];
                ld4(val, p, x);
                acc = fma(x[0], vs[c4.x], acc);
                acc = fma(x[1], vs[c4.y], acc);
                acc = fma(x[2], vs[c4.z], acc);
                acc = fma(x[3], vs[c4.w], acc);
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        acc = __shfl(acc, gi * G, kWave);
        if (gl == 0 && r < m) {
            double t = acc - dmu;
            if (wprev) t -= rfv * wprev[r];
            y[r] = t;
        }
    }
}

// --------------------------------------------------------------------------
// BF16 form of A v: the same product with BF16 matrix values and FP32
// arithmetic (v staged in LDS as FP32, FP32 accumulation); the centering and
// recurrence terms are applied in FP64.  Rows are padded to multiples of 8, so
// each lane reads 8 columns and 8 values with two 16-byte loads; two
// accumulators alternate and are combined in order.
// --------------------------------------------------------------------------
__device__ __forceinline__ float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }

// 12-bit packed indices (columns of A v when n <= 4096, in-slab rows of A' w):
// entries 8g .. 8g+7 share three 32-bit words, entry e at bits 12e .. 12e+11
// of the little-endian 96-bit group.
__device__ __forceinline__ void unpack12(const uint3 w, uint32_t c[8]) {
    c[0] = w.x & 0xFFFu;
    c[1] = (w.x >> 12) & 0xFFFu;
    c[2] = (w.x >> 24) | ((w.y & 0xFu) << 8);
    c[3] = (w.y >> 4) & 0xFFFu;
    c[4] = (w.y >> 16) & 0xFFFu;
    c[5] = (w.y >> 28) | ((w.z & 0xFFu) << 4);
    c[6] = (w.z >> 8) & 0xFFFu;
    c[7] = w.z >> 20;
}

// the 8 indices of entries p .. p+7 (p a multiple of 8)
template <bool C12>
__device__ __forceinline__ void load_idx8(const void *idx, int64_t p, uint32_t c[8]) {
    if constexpr (C12) {
        unpack12(((const uint3 *)idx)[p >> 3], c);
    } else {
        const uint4 c8 = *(const uint4 *)((const uint16_t *)idx + p);
        c[0] = c8.x & 0xFFFFu;
        c[1] = c8.x >> 16;
        c[2] = c8.y & 0xFFFFu;
        c[3] = c8.y >> 16;
        c[4] = c8.z & 0xFFFFu;
        c[5] = c8.z >> 16;
        c[6] = c8.w & 0xFFFFu;
        c[7] = c8.w >> 16;
    }
}

__global__ void __launch_bounds__(kBlock)
k_pack12(const uint16_t *in, int64_t nent, int64_t ngroups, uint32_t *out) {
    for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < ngroups;
         g += (int64_t)gridDim.x * kBlock) {
        uint32_t c[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) c[e] = 8 * g + e < nent ? (in[8 * g + e] & 0xFFFu) : 0u;
        out[3 * g] = c[0] | (c[1] << 12) | (c[2] << 24);
        out[3 * g + 1] = (c[2] >> 8) | (c[3] << 4) | (c[4] << 16) | (c[5] << 28);
        out[3 * g + 2] = (c[5] >> 4) | (c[6] << 8) | (c[7] << 20);
    }
}

template <int G, bool C12>
__global__ void __launch_bounds__(kBlock)
k_av_bf16(int64_t m, int64_t n, const int64_t *rowptr, const void *col, const uint16_t *val,
          const float *mu, const double *v, const double *rf, const double *wprev, double *y) {
    extern __shared__ float vf[];
    __shared__ double dmu_s;
    const int w = wave_index(), lane = lane_id();
    for (int64_t i = threadIdx.x; i < n; i += kBlock) vf[i] = (float)v[i];
    if (w == 0) {
        double d = 0.0;
        if (mu)
            for (int64_t i = lane; i < n; i += kWave) d = fma((double)mu[i], v[i], d);
        d = wave_sum(d);
        if (lane == 0) dmu_s = d;
    }
    __syncthreads();
    const double dmu = dmu_s;
    const double rfv = rf ? *rf : 0.0;
    constexpr int RPW = kWave / G;  // rows per wave
    const int gl = lane % G, gi = lane / G;
    const int64_t stride = (int64_t)gridDim.x * kWaves * RPW;
    for (int64_t r0 = ((int64_t)blockIdx.x * kWaves + w) * RPW; r0 < m; r0 += stride) {
        const int64_t r = r0 + gi;
        float a0 = 0.0f, a1 = 0.0f;
        if (r < m) {
            const int64_t b = rowptr[r], e = rowptr[r + 1];
            for (int64_t p = b + 8 * gl; p < e; p += 8 * G) {
                uint32_t c[8];
                load_idx8<C12>(col, p, c);
                const uint4 v8 = *(const uint4 *)(val + p);
                a0 = fmaf(bf_lo(v8.x), vf[c[0]], a0);
                a1 = fmaf(bf_hi(v8.x), vf[c[1]], a1);
                a0 = fmaf(bf_lo(v8.y), vf[c[2]], a0);
                a1 = fmaf(bf_hi(v8.y), vf[c[3]], a1);
                a0 = fmaf(bf_lo(v8.z), vf[c[4]], a0);
                a1 = fmaf(bf_hi(v8.z), vf[c[5]], a1);
                a0 = fmaf(bf_lo(v8.w), vf[c[6]], a0);
                a1 = fmaf(bf_hi(v8.w), vf[c[7]], a1);
            }
        }
        float acc = a0 + a1;
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        acc = __shfl(acc, gi * G, kWave);
        if (gl == 0 && r < m) {
            double t = (double)acc - dmu;
            if (wprev) t -= rfv * wprev[r];
            y[r] = t;
        }
    }
}

// --------------------------------------------------------------------------
// Reorthogonalisation partial products: part[i*nblk + b] = W_i' y over the
// rows of slab b (kSlabRows rows, one block of kDotThreads), for i < jc:
// columns in chunks of 32, each chunk reduced across the wavefront by the
// butterfly and across the block's waves in wave order.
// --------------------------------------------------------------------------
constexpr int kDotThreads = 512;

template <class TW>
__global__ void __launch_bounds__(kDotThreads)
k_dot_w(int64_t m, const TW *W, int jc, const double *y, double *part) {
    constexpr int NW = kDotThreads / kWave;
    __shared__ double red[NW][32];
    const int64_t nblk = gridDim.x;
    const int64_t rb = (int64_t)blockIdx.x * kSlabRows;
    const int64_t re = rb + kSlabRows < m ? rb + kSlabRows : m;
    const int w = threadIdx.x / kWave, lane = lane_id();
    for (int c0 = 0; c0 < jc; c0 += 32) {
        double acc[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc[i] = 0.0;
        for (int64_t r = rb + threadIdx.x; r < re; r += kDotThreads) {
            const double yr = y[r];
            // every load is issued (columns past jc read column jc - 1; their
            // sums are never stored), so the 32 loads of a row are in flight
            // together
            double wv[32];
#pragma unroll
            for (int i = 0; i < 32; ++i)
                wv[i] = (double)W[(int64_t)(c0 + i < jc ? c0 + i : jc - 1) * m + r];
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the FMAs
#pragma unroll
            for (int i = 0; i < 32; ++i) acc[i] = fma(wv[i], yr, acc[i]);
        }
        const double t = butterfly32(acc);
        if ((lane & 1) == 0) red[w][lane >> 1] = t;
        __syncthreads();
        if (threadIdx.x < 32 && c0 + (int)threadIdx.x < jc) {
            double u = 0.0;
            for (int q = 0; q < NW; ++q) u += red[q][threadIdx.x];
            part[(int64_t)(c0 + threadIdx.x) * nblk + blockIdx.x] = u;
        }
        __syncthreads();
    }
}

// T[i] = sum_b part[i*nblk + b]: block i sums column i (each thread a
// strided subset with independent loads, then the fixed-order block sum).
__global__ void __launch_bounds__(kBlock) k_red_t(int nblk, const double *part, double *T) {
    __shared__ double scratch[kWaves];
    const double *pc = part + (int64_t)blockIdx.x * nblk;
    double t0 = 0.0, t1 = 0.0;
    int b = threadIdx.x;
    for (; b + kBlock < nblk; b += 2 * kBlock) {
        t0 += pc[b];
        t1 += pc[b + kBlock];
    }
    if (b < nblk) t0 += pc[b];
    const double t = block_sum(t0 + t1, scratch);
    if (threadIdx.x == 0) T[blockIdx.x] = t;
}

// --------------------------------------------------------------------------
// y' = y - W[:,0:jc] T (reorthogonalisation apply, one pass over W) into
// yout, with the slab partials of ||y'||^2 and sum(y') (one block per slab;
// part row s holds n column dots followed by these two).
// --------------------------------------------------------------------------
constexpr int kApplyThreads = 1024;

template <class TW>
__global__ void __launch_bounds__(kApplyThreads)
k_apply(int64_t m, int64_t n, const TW *W, int jc, const double *T, const double *y,
        double *yout, double *part) {
    __shared__ double tl[kMaxWork];
    __shared__ double scratch[kApplyThreads / kWave];
    const int64_t s = blockIdx.x;
    const int64_t r0 = s * kSlabRows;
    const int64_t nrows = (m - r0) < kSlabRows ? (m - r0) : kSlabRows;
    if ((int)threadIdx.x < jc) tl[threadIdx.x] = T[threadIdx.x];
    __syncthreads();
    double nrm = 0.0, sm = 0.0;
    for (int64_t lr = threadIdx.x; lr < nrows; lr += kApplyThreads) {
        const int64_t r = r0 + lr;
        double yv = y[r];
        if (jc > 0) {
            double d = 0.0;
#pragma unroll 8
            for (int i = 0; i < jc; ++i) d = fma((double)W[(int64_t)i * m + r], tl[i], d);
            yv -= d;
        }
        yout[r] = yv;
        nrm = fma(yv, yv, nrm);
        sm += yv;
    }
    nrm = block_sum(nrm, scratch);
    sm = block_sum(sm, scratch);
    if (threadIdx.x == 0) {
        part[s * (n + 2) + n] = nrm;
        part[s * (n + 2) + n + 1] = sm;
    }
}

// Workgroups are dealt to the 8 XCDs in turn (linear id mod 8).  A' w maps
// block x to the slab that puts consecutive slabs of a column group on one
// XCD and 8 dispatches apart: the cache lines at the boundary of a column's
// consecutive slab segments, which both slabs read, are then found in that
// XCD's L2.  A bijection of [0, nslab).
__device__ __forceinline__ int64_t xcd_slab(int64_t x, int64_t nslab) {
    const int64_t xcd = x & 7, i = x >> 3;
    const int64_t q = nslab >> 3, r = nslab & 7;
    return xcd * q + (xcd < r ? xcd : r) + i;
}

// Sum over the G lanes of each lane group (G a power of two); lane 0 of the
// group holds the result.  G = 64 is the wave sum (result in every lane).
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
    if constexpr (G == kWave) {
        return wave_sum(v);
    } else {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        return v;
    }
}

// --------------------------------------------------------------------------
// Slab partials of B' y: block (slab s, column group g) stages the slab of y
// in LDS and computes the dots of the columns of its group; each column
// segment is summed by a group of G lanes (kWave / G segments per wavefront
// at once, G set from the mean segment length when the operator is built).
// Lane u of a group reads 4 consecutive entries per step from the 4-aligned
// cover of the segment (entries outside it are masked; the value and row
// arrays are padded so the cover stays in bounds); two accumulators
// alternate and are combined in order, then the group is summed.
// --------------------------------------------------------------------------
template <class VT, int G>
__global__ void __launch_bounds__(kBlock)
k_atw(int64_t m, int64_t n, const int64_t *off, const uint16_t *lrow, const VT *val,
      const double *y, double *part) {
    constexpr int C = kWave / G;  // segments per wavefront
    __shared__ double wl[kSlabRows];
    const int64_t s = xcd_slab(blockIdx.x, gridDim.x);
    const int g = blockIdx.y, ng = gridDim.y;
    const int64_t r0 = s * kSlabRows;
    const int64_t nrows = (m - r0) < kSlabRows ? (m - r0) : kSlabRows;
    const int w = wave_index(), lane = lane_id();
    const int u = lane % G, c = lane / G;
    for (int64_t lr = threadIdx.x; lr < kSlabRows; lr += kBlock)
        wl[lr] = lr < nrows ? y[r0 + lr] : 0.0;
    __syncthreads();
    double *prow = part + s * (n + 2);
    for (int64_t j0 = ((int64_t)g * kWaves + w) * C; j0 < n; j0 += (int64_t)ng * kWaves * C) {
        const int64_t j = j0 + c;
        double acc0 = 0.0, acc1 = 0.0;
        if (j < n) {
            const int64_t a = off[s * n + j], b = off[(s + 1) * n + j];
            for (int64_t p = (a & ~(int64_t)3) + 4 * u; p < b; p += 4 * G) {
                const ushort4 l4 = *(const ushort4 *)(lrow + p);
                double x[4];
                ld4(val, p, x);
                const bool v0 = p >= a, v1 = p + 1 >= a && p + 1 < b,
                           v2 = p + 2 >= a && p + 2 < b, v3 = p + 3 < b;
                acc0 = fma(v0 ? x[0] : 0.0, wl[v0 ? l4.x : 0], acc0);
                acc1 = fma(v1 ? x[1] : 0.0, wl[v1 ? l4.y : 0], acc1);
                acc0 = fma(v2 ? x[2] : 0.0, wl[v2 ? l4.z : 0], acc0);
                acc1 = fma(v3 ? x[3] : 0.0, wl[v3 ? l4.w : 0], acc1);
            }
        }
        const double acc = group_sum<G>(acc0 + acc1);
        if (u == 0 && j < n) prow[j] = acc;
    }
}

// BF16 form of the slab partials of B' y: y staged in LDS as FP32, FP32
// accumulation per column segment (the slab sums are reduced in FP64).  The
// in-slab rows are 12-bit packed; each lane reads the 8-aligned groups of 8
// entries covering the segment [a, b) (entries outside it are masked; the
// value and index arrays are padded so the cover stays in bounds).
// (Lane u of a group of G; the whole segment for G = kWave, u = lane.)
template <int G = kWave>
__device__ __forceinline__ float atw_seg_bf16(int64_t a, int64_t b, const uint32_t *lrow12,
                                              const uint16_t *val, const float *wl, int u) {
    float acc0 = 0.0f, acc1 = 0.0f;
    for (int64_t p = (a & ~(int64_t)7) + 8 * u; p < b; p += 8 * G) {
        uint32_t l[8];
        load_idx8<true>(lrow12, p, l);
        const uint4 v8 = *(const uint4 *)(val + p);
        const uint32_t vw[4] = {v8.x, v8.y, v8.z, v8.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const bool ok = p + e >= a && p + e < b;
            const float v = (e & 1) ? bf_hi(vw[e >> 1]) : bf_lo(vw[e >> 1]);
            const float t = ok ? v : 0.0f;
            if (e & 1)
                acc1 = fmaf(t, wl[ok ? l[e] : 0], acc1);
            else
                acc0 = fmaf(t, wl[ok ? l[e] : 0], acc0);
        }
    }
    return acc0 + acc1;
}

template <int G>
__global__ void __launch_bounds__(kBlock)
k_atw_bf16(int64_t m, int64_t n, const int64_t *off, const uint32_t *lrow12, const uint16_t *val,
           const double *y, double *part) {
    constexpr int C = kWave / G;  // segments per wavefront
    __shared__ float wl[kSlabRows];
    const int64_t s = xcd_slab(blockIdx.x, gridDim.x);
    const int g = blockIdx.y, ng = gridDim.y;
    const int64_t r0 = s * kSlabRows;
    const int64_t nrows = (m - r0) < kSlabRows ? (m - r0) : kSlabRows;
    const int w = wave_index(), lane = lane_id();
    const int u = lane % G, c = lane / G;
    for (int64_t lr = threadIdx.x; lr < kSlabRows; lr += kBlock)
        wl[lr] = lr < nrows ? (float)y[r0 + lr] : 0.0f;
    __syncthreads();
    double *prow = part + s * (n + 2);
    for (int64_t j0 = ((int64_t)g * kWaves + w) * C; j0 < n; j0 += (int64_t)ng * kWaves * C) {
        const int64_t j = j0 + c;
        float acc = 0.0f;
        if (j < n) {
            const int64_t a = off[s * n + j], b = off[(s + 1) * n + j];
            acc = atw_seg_bf16<G>(a, b, lrow12, val, wl, u);
        }
        acc = group_sum<G>(acc);
        if (u == 0 && j < n) prow[j] = (double)acc;
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
    // 16 slabs at a time: all loads issued first (indices clamped, values
    // past the chunk multiplied by zero), then added in slab order
    double t = 0.0;
    for (int64_t s0 = sb; s0 < se; s0 += 16) {
        double v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = part[(s0 + i < se ? s0 + i : se - 1) * ncol + c];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 16; ++i) t += v[i] * (s0 + i < se ? 1.0 : 0.0);
    }
    part2[q * ncol + c] = t;
}

// y = sum_q part2[q][i] - (sum x) mu'_i  (finishes A' x after k_atw)
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
// Right-side step, three launches over the n-vectors (n = features):
//  A: every block forms F = (B'y - sum(y) mu')/S - S v_j for its rows
//     (S = ||y|| from the slab partials) and the block partials of
//     V[:,0:j+1]' F;
//  B: every block sums those partials in block order (T), applies
//     F -= V[:,0:j+1] T to its rows and writes the block partial of ||F||^2;
//  C: one block: RF = ||F||, v_{j+1} = F / RF, B(j,j) = S, B(j,j+1) = RF
//     (a deterministic pseudo-random direction replaces F when it vanishes).
// scal: [0] S, [1] RF (next step), [2] ||F|| after the last step; flags[0]
// is set when ||y|| underflows (invariant subspace).
// --------------------------------------------------------------------------
constexpr int kRsBlock = 256;

// Block bid of the rows; column groups cg, cg + ncg, ... of 32 (every group
// recomputes f for its rows, group 0 stores it).
template <class MT>
__device__ __forceinline__ void right_a_body(int64_t n, int j, const double *part2, const MT *mu, const double *V,
                             double *Fv, double *rpart, unsigned bid, int cg = 0,
                             int ncg = 1) {
    __shared__ double red[kRsBlock / kWave][kMaxWork];
    const int64_t ncol = n + 2;
    double nrm = 0.0, sumy = 0.0;
    for (int q = 0; q < kRedChunks; ++q) {
        nrm += part2[q * ncol + n];
        sumy += part2[q * ncol + n + 1];
    }
    double S = sqrt(nrm);
    if (!(S >= 2.0 * 2.220446049250313e-16)) S = 1.0;
    const int64_t i = (int64_t)bid * kRsBlock + threadIdx.x;
    double f = 0.0;
    if (i < n) {
        double t = 0.0;
        for (int q = 0; q < kRedChunks; ++q) t += part2[q * ncol + i];
        if (mu) t -= sumy * ldmu(mu, i);
        f = t / S - S * V[(int64_t)j * n + i];
        if (cg == 0) Fv[i] = f;
    }
    const int w = threadIdx.x / kWave, lane = lane_id();
    const int64_t ic = i < n ? i : n - 1;
    for (int c0 = 32 * cg; c0 <= j; c0 += 32 * ncg) {
        double v[32];
        // unconditional (clamped) loads: all 32 in flight together
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            const bool in = i < n && c0 + c <= j;
            const double vv = V[(int64_t)(c0 + c <= j ? c0 + c : j) * n + ic];
            v[c] = in ? vv * f : 0.0;
        }
        const double d = butterfly32(v);
        if ((lane & 1) == 0 && c0 + (lane >> 1) <= j) red[w][c0 + (lane >> 1)] = d;
    }
    __syncthreads();
    for (int c = threadIdx.x; c <= j; c += kRsBlock) {
        if ((c / 32) % ncg != cg) continue;
        double t = 0.0;
        for (int q = 0; q < kRsBlock / kWave; ++q) t += red[q][c];
        rpart[(int64_t)bid * kMaxWork + c] = t;
    }
}

template <class MT>
__global__ void __launch_bounds__(kRsBlock)
k_right_a(int64_t n, int j, const double *part2, const MT *mu, const double *V, double *Fv,
          double *rpart) {
    right_a_body<MT>(n, j, part2, mu, V, Fv, rpart, blockIdx.x, blockIdx.y, gridDim.y);
}

__device__ __forceinline__ void right_b_body(int64_t n, int j, int nblk, const double *rpart, const double *V,
                             double *Fv, double *npart, unsigned bid) {
    __shared__ double T[kMaxWork];
    __shared__ double scratch[kRsBlock / kWave];
    for (int c = threadIdx.x; c <= j; c += kRsBlock) {
        double t = 0.0;
        for (int b = 0; b < nblk; ++b) t += rpart[(int64_t)b * kMaxWork + c];
        T[c] = t;
    }
    __syncthreads();
    const int64_t i = (int64_t)bid * kRsBlock + threadIdx.x;
    double f = 0.0;
    if (i < n) {
        // 16 columns at a time: loads first (clamped column index, the extra
        // terms multiplied by zero), then the FMA chain in column order
        double d = 0.0;
        for (int c0 = 0; c0 <= j; c0 += 16) {
            double v[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) v[c] = V[(int64_t)(c0 + c <= j ? c0 + c : j) * n + i];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const double t = T[c0 + c <= j ? c0 + c : j];
                d = fma(v[c], c0 + c <= j ? t : 0.0, d);
            }
        }
        f = Fv[i] - d;
        Fv[i] = f;
    }
    const double s2 = block_sum(f * f, scratch);
    if (threadIdx.x == 0) npart[bid] = s2;
}

__global__ void __launch_bounds__(kRsBlock)
k_right_b(int64_t n, int j, int nblk, const double *rpart, const double *V, double *Fv,
          double *npart) {
    right_b_body(n, j, nblk, rpart, V, Fv, npart, blockIdx.x);
}

__device__ __forceinline__ void right_c_body(int64_t n, int j, int work, int nblk, const double *part2,
                             const double *npart, double *V, double *Fv, double *B, double *scal,
                             int *flags, uint64_t seed) {
    __shared__ double scratch[16];
    __shared__ double tsh[kMaxWork];
    const int64_t ncol = n + 2;
    const double eps2 = 2.0 * 2.220446049250313e-16;
    double nrm = 0.0;
    for (int q = 0; q < kRedChunks; ++q) nrm += part2[q * ncol + n];
    double S = sqrt(nrm);
    if (!(S >= eps2)) {
        if (threadIdx.x == 0) flags[0] = 1;
        S = 1.0;
    }
    double nf = 0.0;
    for (int b = 0; b < nblk; ++b) nf += npart[b];
    const double RF = sqrt(nf);
    if (j + 1 < work) {
        double RFs = RF, div = RF;
        if (RF < eps2) {
            // replace F by a pseudo-random direction orthogonal to V[:,0:j+1]
            const int w = threadIdx.x / kWave, lane = lane_id();
            const int nw = blockDim.x / kWave;
            for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
                Fv[i] = hash_normal(seed, (uint32_t)i);
            __syncthreads();
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

template <class MT>
__global__ void __launch_bounds__(1024)
k_right_c(int64_t n, int j, int work, int nblk, const double *part2, const double *npart,
          double *V, double *Fv, double *B, double *scal, int *flags, uint64_t seed) {
    right_c_body(n, j, work, nblk, part2, npart, V, Fv, B, scal, flags, seed);
}

// y -= RF w_prev (the recurrence term of A v_j when the product came from a
// host operator; same rounding as the fused form in k_av_csr)
__global__ void __launch_bounds__(kBlock)
k_sub_rf(double *y, int64_t m, const double *rf, const double *wprev) {
    const double r = *rf;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * kBlock)
        y[i] -= r * wprev[i];
}

// dst = src / S (the normalised Lanczos vector w_j = y / ||y||)
__global__ void __launch_bounds__(kBlock)
k_scale_col(const double *src, double *dst, int64_t m, const double *scal) {
    const double S = scal[0];
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * kBlock)
        dst[r] = src[r] / S;
}

// The same into a basis stored in FP32 (dst) and an FP64 copy (dst64) that
// the next product's recurrence term reads.
__global__ void __launch_bounds__(kBlock)
k_scale_col32(const double *src, float *dst, double *dst64, int64_t m, const double *scal) {
    const double S = scal[0];
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * kBlock) {
        const double v = src[r] / S;
        dst64[r] = v;
        dst[r] = (float)v;
    }
}

// Y[:, c0 + c] = X[:, 0:work] Cf[:, c0 + c] for c < kout (X, Y column-major,
// ld = m), computed as the 16x16 tiles of D = Cf' X' on
// v_mfma_f64_16x16x4_f64: one wavefront per 32 rows (two 16-row tiles sharing
// every A operand), lanes (n = row, k-group) read 16 consecutive rows of one
// column of X, and each lane writes 16 consecutive rows of an output column.
// In this instruction lane l holds D[l/16 + 4i][l%16], i = 0..3.  NS =
// ceil(work / 4) k-steps (the launch picks the instantiation): the whole
// 32 x work slice of X is loaded before the MFMA chain, so its 2 NS loads per
// lane are in flight together.  One launch covers at most gemm_cols(NS)
// output columns (their coefficients sit in LDS, rows padded to an odd
// stride); wider outputs take several launches.  Optionally Y[:, c0 + kout] =
// extra / scal[2] (the normalised residual).
__host__ __device__ constexpr int gemm_cols(int ns) { return ns <= 31 ? 64 : 48; }

template <int NS, class TX = double, class TY = double>
__global__ void __launch_bounds__(kBlock)
k_gemm_mfma(int64_t m, int work, const TX *X, const double *Cf, int c0, int kout, TY *Y,
            const double *extra, const double *scal) {
    constexpr int LD = 4 * NS + 1;
    constexpr int GC = gemm_cols(NS);
    __shared__ double cfl[GC * LD];  // cfl[c*LD + l] = Cf[(c0 + c)*work + l], zero padded
    for (int i = threadIdx.x; i < GC * (LD - 1); i += kBlock) {
        const int c = i / (LD - 1), l = i % (LD - 1);
        cfl[c * LD + l] = (c < kout && l < work) ? Cf[(int64_t)(c0 + c) * work + l] : 0.0;
    }
    __syncthreads();
    const int w = wave_index(), lane = lane_id();
    const int nr = lane & 15, kq = lane >> 4;
    const int ncb = (kout + 15) / 16;
    const int64_t nrb = (m + 31) / 32;
    TY *Yc = Y + (int64_t)c0 * m;
    for (int64_t rb = (int64_t)blockIdx.x * kWaves + w; rb < nrb;
         rb += (int64_t)gridDim.x * kWaves) {
        const int64_t row0 = rb * 32 + nr, row1 = row0 + 16;
        // indices clamped into range, out-of-range values multiplied by zero
        // (a product, not a select: the loads then stay unconditional)
        const int64_t rc0 = row0 < m ? row0 : m - 1, rc1 = row1 < m ? row1 : m - 1;
        double bv0[NS], bv1[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int k = 4 * s + kq;
            const int64_t kc = (int64_t)(k < work ? k : work - 1) * m;
            bv0[s] = (double)X[kc + rc0] * ((k < work && row0 < m) ? 1.0 : 0.0);
            bv1[s] = (double)X[kc + rc1] * ((k < work && row1 < m) ? 1.0 : 0.0);
        }
#pragma unroll
        for (int cb = 0; cb < GC / 16; ++cb) {
            if (cb >= ncb) break;
            v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double av = cfl[(cb * 16 + nr) * LD + 4 * s + kq];
                acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv0[s], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv1[s], acc1, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c = cb * 16 + kq + 4 * i;
                if (c < kout) {
                    if (row0 < m) Yc[(int64_t)c * m + row0] = (TY)acc0[i];
                    if (row1 < m) Yc[(int64_t)c * m + row1] = (TY)acc1[i];
                }
            }
        }
        if (extra && kq == 0) {
            if (row0 < m) Yc[(int64_t)kout * m + row0] = (TY)(extra[row0] / scal[2]);
            if (row1 < m) Yc[(int64_t)kout * m + row1] = (TY)(extra[row1] / scal[2]);
        }
    }
}

// --------------------------------------------------------------------------
// SVD of the projected matrix B (work x work, column-major) on the device:
// one-sided Jacobi on G = B' with the round-robin (circle) pair schedule, so
// that the ne/2 pairs of a round touch disjoint columns and run in parallel,
// one 16-lane group per pair.  Rotations accumulate in J; at convergence
// B' J = G, the singular values are the column norms of G, U = J and
// V = G diag(1/s) (ordered by descending s, ties by column index).
// status[0] = 1 if a singular value vanished (V then needs completing; the
// caller redoes that restart on the host).
// --------------------------------------------------------------------------
constexpr int kSvdThreads = 512;  // 32 groups of 16 lanes: one pair each
// A pair is rotated while its cosine exceeds 1e-12.  Further rotations would
// move the singular values by under 1e-24 (relative) and the vectors'
// orthogonality by under 1e-12, far below what a restart needs (each new
// Lanczos vector is reorthogonalised in FP64); against the eps * sqrt(n)
// level this saves about one sweep in ten at 112 columns.
constexpr double kJacobiCos2 = 1e-24;

// 1/x and 1/sqrt(x) from the hardware approximations with two Newton steps
// (relative error at the FP64 rounding level).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = fma(y, fma(-h * y, y, 0.5), y);
    y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}

// DPP lane exchange of a double within rows of 16 lanes.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Sum over a row of 16 lanes (quad xor 1, quad xor 2, half-row mirror, row
// mirror).  Every step adds two equal-order partial sums, so all 16 lanes
// end with bit-identical totals.
__device__ __forceinline__ double sum16(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    return v;
}

__device__ __forceinline__ void small_svd_body(int n, const double *B, double *U, double *s, double *V,
                               int *status) {
    extern __shared__ double sm[];
    double *G = sm;
    // second array: B itself (for U = B V diag(1/s)), or the accumulated
    // rotations when B is too ill-conditioned for that product
    double *J = sm + n * n;
    __shared__ int rotated, need_j, nsweeps;
    __shared__ double sv[kMaxWork];
    const int tid = threadIdx.x;
    const int ne = n + (n & 1);
    const int pr = tid >> 4, gl = tid & 15;
    const double tol2 = kJacobiCos2;
    if (tid == 0) {
        need_j = 0;
        nsweeps = 0;
    }
    auto load = [&](bool withj) {
        for (int e = tid; e < n * n; e += kSvdThreads) {
            const int i = e % n, j = e / n;
            G[e] = B[(int64_t)i * n + j];
            J[e] = withj ? (i == j ? 1.0 : 0.0) : B[e];
        }
        __syncthreads();
    };
    auto sweeps = [&](bool withj) {
        for (int sweep = 0; sweep < 100; ++sweep) {
            if (tid == 0) rotated = 0;
            __syncthreads();
            for (int r = 0; r < ne - 1; ++r) {
                int p = r, q = ne - 1;
                if (pr > 0) {
                    p = (r + pr) % (ne - 1);
                    q = (r - pr + 2 * (ne - 1)) % (ne - 1);
                }
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                const bool valid = pr < ne / 2 && q < n;
                double *gp = G + p * n, *gq = G + q * n;
                double x[4], y[4];
                double app = 0.0, aqq = 0.0, apq = 0.0;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int i = gl + 16 * e;
                    const bool in = valid && i < n;
                    x[e] = in ? gp[i] : 0.0;
                    y[e] = in ? gq[i] : 0.0;
                    app = fma(x[e], x[e], app);
                    aqq = fma(y[e], y[e], aqq);
                    apq = fma(x[e], y[e], apq);
                }
                app = sum16(app);
                aqq = sum16(aqq);
                apq = sum16(apq);
                if (valid && apq != 0.0 && apq * apq > tol2 * app * aqq) {
                    if (gl == 0) rotated = 1;
                    // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), c = 1/sqrt(1 + t^2),
                    // with Newton-refined hardware reciprocals (the rotation
                    // needs c^2 + s^2 = 1 to working precision, not rounding)
                    const double zeta = (aqq - app) * rcp_nr(2.0 * apq);
                    const double az = fabs(zeta);
                    double t;
                    if (az > 1e150) {
                        t = 0.5 * rcp_nr(zeta);
                    } else {
                        const double xx = fma(az, az, 1.0);
                        t = copysign(rcp_nr(az + xx * rsq_nr(xx)), zeta);
                    }
                    const double c = rsq_nr(fma(t, t, 1.0));
                    const double sn = c * t;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int i = gl + 16 * e;
                        if (i < n) {
                            gp[i] = c * x[e] - sn * y[e];
                            gq[i] = sn * x[e] + c * y[e];
                        }
                    }
                    if (withj) {
                        double *jp = J + p * n, *jq = J + q * n;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int i = gl + 16 * e;
                            if (i < n) {
                                const double u = jp[i], v = jq[i];
                                jp[i] = c * u - sn * v;
                                jq[i] = sn * u + c * v;
                            }
                        }
                    }
                }
                __syncthreads();
            }
            const int any = rotated;
            __syncthreads();
            if (tid == 0) nsweeps = nsweeps + 1;
            if (!any) break;
        }
    };
    auto norms = [&]() {
        if (tid < n) {
            double t = 0.0;
            for (int i = 0; i < n; ++i) t += G[tid * n + i] * G[tid * n + i];
            sv[tid] = sqrt(t);
        }
        __syncthreads();
    };
    load(false);
    sweeps(false);
    norms();
    // U = B V diag(1/s) is accurate to eps s_max / s_j: without the
    // accumulated rotations only while s_min >= 1e-5 s_max
    if (tid == 0) {
        double mx = 0.0, mn = __builtin_huge_val();
        for (int i = 0; i < n; ++i) {
            mx = fmax(mx, sv[i]);
            mn = fmin(mn, sv[i]);
        }
        need_j = !(mn >= 1e-5 * mx);
    }
    __syncthreads();
    const bool withj = need_j != 0;
    if (withj) {
        load(true);
        sweeps(true);
        norms();
    }
    if (tid == 0) {
        status[0] = 0;
        status[1] = nsweeps;
    }
    __syncthreads();
    for (int j = tid >> 6; j < n; j += kSvdThreads / 64) {
        const int l = tid & 63;
        const double sj = sv[j];
        int rk = 0;
        for (int i = 0; i < n; ++i) rk += (sv[i] > sj || (sv[i] == sj && i < j)) ? 1 : 0;
        if (l == 0) {
            s[rk] = sj;
            if (!(sj > 0.0)) status[0] = 1;
        }
        for (int i = l; i < n; i += 64) {
            if (withj) {
                U[(int64_t)rk * n + i] = J[j * n + i];
            } else {
                // (B g_j)_i / s_j^2, g_j = s_j v_j the column of G
                double t = 0.0;
                for (int c = 0; c < n; ++c) t = fma(J[i + c * n], G[j * n + c], t);
                U[(int64_t)rk * n + i] = t / (sj * sj);
            }
            V[(int64_t)rk * n + i] = sj > 0.0 ? G[j * n + i] / sj : 0.0;
        }
    }
}

__global__ void __launch_bounds__(kSvdThreads)
k_small_svd(int n, const double *B, double *U, double *s, double *V, int *status) {
    small_svd_body(n, B, U, s, V, status);
}

// The same SVD with 8 lanes per column pair (ne/2 <= 31 groups in four
// wavefronts): columns of G and J sit in LDS with stride kSvdLd, rows n..63
// held at zero, and lane l of a group holds rows 16k + 2l + h (k < 4, h < 2),
// so that each of its four 16-byte LDS reads per column is one contiguous
// 128-byte segment across the group.  Half the lanes of the 16-lane form for
// the same rounds: the per-pair reductions and rotation formulas, repeated on
// every lane of a group, are the bulk of a round.  For work <= kSvd8MaxWork
// (the two arrays fit the 64 KiB of LDS a block may allocate by default).
constexpr int kSvd8Threads = 256;
constexpr int kSvdLd = 64;
constexpr int kSvd8MaxWork = 62;

// Sum over each group of 8 lanes (quad xor 1, quad xor 2, half-row mirror):
// commutative steps, so the 8 lanes end with bit-identical totals.
__device__ __forceinline__ double sum8(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    return v;
}

__device__ __forceinline__ void small_svd8_body(int n, const double *B, double *U, double *s,
                                                double *V, int *status) {
    extern __shared__ double2 svd8_lds[];  // 16-byte aligned for the paired reads
    const int ne = n + (n & 1);
    double *G = (double *)svd8_lds;
    // B itself (for U = B V diag(1/s)), or the accumulated rotations when B
    // is too ill-conditioned for that product
    double *J = G + ne * kSvdLd;
    __shared__ int rotated, need_j, nsweeps;
    __shared__ double sv[kMaxWork];
    const int tid = threadIdx.x;
    const int pr = tid >> 3, gl = tid & 7;
    const int npair = ne / 2;
    const double tol2 = kJacobiCos2;
    if (tid == 0) {
        need_j = 0;
        nsweeps = 0;
    }
    auto load = [&](bool withj) {
        for (int e = tid; e < ne * kSvdLd; e += kSvd8Threads) {
            const int c = e / kSvdLd, i = e % kSvdLd;
            const bool in = c < n && i < n;
            G[e] = in ? B[(int64_t)i * n + c] : 0.0;  // G(i, c) = B(c, i)
            J[e] = withj ? (in && i == c ? 1.0 : 0.0) : (in ? B[(int64_t)c * n + i] : 0.0);
        }
        __syncthreads();
    };
    auto sweeps = [&](bool withj) {
        for (int sweep = 0; sweep < 100; ++sweep) {
            if (tid == 0) rotated = 0;
            __syncthreads();
            // circle schedule: round r pairs (r + pr, r - pr) mod (ne - 1),
            // group 0 pairs r with the fixed column ne - 1
            int pp = pr, qq = ne - 1 - pr;
            for (int r = 0; r < ne - 1; ++r) {
                if (pr < npair) {
                    const int p = pp < qq ? pp : qq, q = pp < qq ? qq : pp;
                    const bool valid = q < n;
                    double2 *gp = (double2 *)(G + p * kSvdLd) + gl;
                    double2 *gq = (double2 *)(G + q * kSvdLd) + gl;
                    double2 x[4], y[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        x[k] = gp[8 * k];
                        y[k] = gq[8 * k];
                    }
                    double app = 0.0, aqq = 0.0, apq = 0.0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        app = fma(x[k].x, x[k].x, app);
                        aqq = fma(y[k].x, y[k].x, aqq);
                        apq = fma(x[k].x, y[k].x, apq);
                        app = fma(x[k].y, x[k].y, app);
                        aqq = fma(y[k].y, y[k].y, aqq);
                        apq = fma(x[k].y, y[k].y, apq);
                    }
                    app = sum8(app);
                    aqq = sum8(aqq);
                    apq = sum8(apq);
                    if (valid && apq != 0.0 && apq * apq > tol2 * app * aqq) {
                        if (gl == 0) rotated = 1;
                        // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), c = 1/sqrt(1 + t^2)
                        const double zeta = (aqq - app) * rcp_nr(2.0 * apq);
                        const double az = fabs(zeta);
                        double t;
                        if (az > 1e150) {
                            t = 0.5 * rcp_nr(zeta);
                        } else {
                            const double xx = fma(az, az, 1.0);
                            t = copysign(rcp_nr(az + xx * rsq_nr(xx)), zeta);
                        }
                        const double c = rsq_nr(fma(t, t, 1.0));
                        const double sn = c * t;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            gp[8 * k] = make_double2(c * x[k].x - sn * y[k].x, c * x[k].y - sn * y[k].y);
                            gq[8 * k] = make_double2(sn * x[k].x + c * y[k].x, sn * x[k].y + c * y[k].y);
                        }
                        if (withj) {
                            double2 *jp = (double2 *)(J + p * kSvdLd) + gl;
                            double2 *jq = (double2 *)(J + q * kSvdLd) + gl;
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const double2 u = jp[8 * k], v = jq[8 * k];
                                jp[8 * k] = make_double2(c * u.x - sn * v.x, c * u.y - sn * v.y);
                                jq[8 * k] = make_double2(sn * u.x + c * v.x, sn * u.y + c * v.y);
                            }
                        }
                    }
                }
                __syncthreads();
                pp = pp + 1 == ne - 1 ? 0 : pp + 1;
                if (pr > 0) qq = qq + 1 == ne - 1 ? 0 : qq + 1;
            }
            const int any = rotated;
            __syncthreads();
            if (tid == 0) nsweeps = nsweeps + 1;
            if (!any) break;
        }
    };
    auto norms = [&]() {
        if (tid < n) {
            double t = 0.0;
            for (int i = 0; i < n; ++i) t += G[tid * kSvdLd + i] * G[tid * kSvdLd + i];
            sv[tid] = sqrt(t);
        }
        __syncthreads();
    };
    load(false);
    sweeps(false);
    norms();
    // U = B V diag(1/s) is accurate to eps s_max / s_j: without the
    // accumulated rotations only while s_min >= 1e-5 s_max
    if (tid == 0) {
        double mx = 0.0, mn = __builtin_huge_val();
        for (int i = 0; i < n; ++i) {
            mx = fmax(mx, sv[i]);
            mn = fmin(mn, sv[i]);
        }
        need_j = !(mn >= 1e-5 * mx);
    }
    __syncthreads();
    const bool withj = need_j != 0;
    if (withj) {
        load(true);
        sweeps(true);
        norms();
    }
    if (tid == 0) {
        status[0] = 0;
        status[1] = nsweeps;
    }
    __syncthreads();
    for (int j = tid >> 6; j < n; j += kSvd8Threads / 64) {
        const int l = tid & 63;
        const double sj = sv[j];
        int rk = 0;
        for (int i = 0; i < n; ++i) rk += (sv[i] > sj || (sv[i] == sj && i < j)) ? 1 : 0;
        if (l == 0) {
            s[rk] = sj;
            if (!(sj > 0.0)) status[0] = 1;
        }
        for (int i = l; i < n; i += 64) {
            if (withj) {
                U[(int64_t)rk * n + i] = J[j * kSvdLd + i];
            } else {
                // (B g_j)_i / s_j^2, g_j = s_j v_j the column of G
                double t = 0.0;
                for (int c = 0; c < n; ++c) t = fma(J[c * kSvdLd + i], G[j * kSvdLd + c], t);
                U[(int64_t)rk * n + i] = t / (sj * sj);
            }
            V[(int64_t)rk * n + i] = sj > 0.0 ? G[j * kSvdLd + i] / sj : 0.0;
        }
    }
}

__global__ void __launch_bounds__(kSvd8Threads)
k_small_svd8(int n, const double *B, double *U, double *s, double *V, int *status) {
    small_svd8_body(n, B, U, s, V, status);
}

// The same SVD for 64 < work <= kMaxWork with 8 lanes per column pair (64
// groups in eight wavefronts, circle schedule): lane l of a group holds rows
// 16k + 2l + h (k < 8, h < 2) of its two columns, read and written as
// 16-byte pairs; only G = B' sits in LDS (columns of kSvdLdB doubles, rows
// n.. held at zero), so U = B V diag(1/s) reads B from global memory.  When
// B is too ill-conditioned for that product (s_min < 1e-5 s_max) the kernel
// sets status[0] and the caller redoes the SVD on the host (the accumulated
// rotations would need a second n x n array).  Per round the per-pair
// reductions and rotation formulas, repeated on every lane of a group, are
// half those of 16 lanes per pair.
constexpr int kSvdBigThreads = 512;
constexpr int kSvdLdB = 130;  // column stride (doubles): 128 rows + 16 bytes

__global__ void __launch_bounds__(kSvdBigThreads)
k_small_svd_big(int n, const double *B, double *U, double *s, double *V, int *status) {
    extern __shared__ double2 svdb_lds[];
    double *G = (double *)svdb_lds;  // G(i, c) = B(c, i) at G[c * kSvdLdB + i]
    __shared__ int rotated, nsweeps, ill;
    __shared__ double sv[kMaxWork];
    const int tid = threadIdx.x;
    const int ne = n + (n & 1);
    const int npair = ne / 2;
    const int pr = tid >> 3, gl = tid & 7;
    const double tol2 = kJacobiCos2;
    if (tid == 0) {
        nsweeps = 0;
        ill = 0;
    }
    for (int e = tid; e < ne * kSvdLdB; e += kSvdBigThreads) {
        const int c = e / kSvdLdB, i = e % kSvdLdB;
        G[e] = (c < n && i < n) ? B[(int64_t)i * n + c] : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 100; ++sweep) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        // circle schedule: round r pairs (r + pr, r - pr) mod (ne - 1), group
        // 0 pairs r with the fixed column ne - 1
        int pp = pr, qq = ne - 1 - pr;
        for (int r = 0; r < ne - 1; ++r) {
            if (pr < npair) {
                const int p = pp < qq ? pp : qq, q = pp < qq ? qq : pp;
                const bool valid = q < n;
                double2 *gp = (double2 *)(G + p * kSvdLdB) + gl;
                double2 *gq = (double2 *)(G + q * kSvdLdB) + gl;
                double2 x[8], y[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    x[k] = gp[8 * k];
                    y[k] = gq[8 * k];
                }
                double app = 0.0, aqq = 0.0, apq = 0.0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    app = fma(x[k].x, x[k].x, app);
                    aqq = fma(y[k].x, y[k].x, aqq);
                    apq = fma(x[k].x, y[k].x, apq);
                    app = fma(x[k].y, x[k].y, app);
                    aqq = fma(y[k].y, y[k].y, aqq);
                    apq = fma(x[k].y, y[k].y, apq);
                }
                app = sum8(app);
                aqq = sum8(aqq);
                apq = sum8(apq);
                if (valid && apq != 0.0 && apq * apq > tol2 * app * aqq) {
                    if (gl == 0) rotated = 1;
                    // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), c = 1/sqrt(1 + t^2)
                    const double zeta = (aqq - app) * rcp_nr(2.0 * apq);
                    const double az = fabs(zeta);
                    double t;
                    if (az > 1e150) {
                        t = 0.5 * rcp_nr(zeta);
                    } else {
                        const double xx = fma(az, az, 1.0);
                        t = copysign(rcp_nr(az + xx * rsq_nr(xx)), zeta);
                    }
                    const double c = rsq_nr(fma(t, t, 1.0));
                    const double sn = c * t;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        gp[8 * k] = make_double2(c * x[k].x - sn * y[k].x, c * x[k].y - sn * y[k].y);
                        gq[8 * k] = make_double2(sn * x[k].x + c * y[k].x, sn * x[k].y + c * y[k].y);
                    }
                }
            }
            __syncthreads();
            pp = pp + 1 == ne - 1 ? 0 : pp + 1;
            if (pr > 0) qq = qq + 1 == ne - 1 ? 0 : qq + 1;
        }
        const int any = rotated;
        __syncthreads();
        if (tid == 0) nsweeps = nsweeps + 1;
        if (!any) break;
    }
    if (tid < n) {
        double t = 0.0;
        for (int i = 0; i < n; ++i) t += G[tid * kSvdLdB + i] * G[tid * kSvdLdB + i];
        sv[tid] = sqrt(t);
    }
    __syncthreads();
    if (tid == 0) {
        double mx = 0.0, mn = __builtin_huge_val();
        for (int i = 0; i < n; ++i) {
            mx = fmax(mx, sv[i]);
            mn = fmin(mn, sv[i]);
        }
        ill = !(mn >= 1e-5 * mx);
        status[0] = ill;
        status[1] = nsweeps;
    }
    __syncthreads();
    if (ill) return;
    // U[:, rank(j)] = B g_j / s_j^2 (g_j = s_j v_j, the column of G) for the
    // columns j = w + 8q of wavefront w: lane l holds rows l and l + 64 of B,
    // loaded in blocks of 16 columns with every load in flight, and applies
    // each block to all the wavefront's j (the sums run over c in order)
    constexpr int kJw = kMaxWork / (kSvdBigThreads / 64);
    const int w = tid >> 6, l = tid & 63;
    for (int ih = 0; ih * 64 < n; ++ih) {
        const int i = l + 64 * ih;
        double acc[kJw];
#pragma unroll
        for (int q = 0; q < kJw; ++q) acc[q] = 0.0;
        for (int c0 = 0; c0 < n; c0 += 16) {
            double bv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
                bv[u] = (c0 + u < n && i < n) ? B[(int64_t)(c0 + u) * n + i] : 0.0;
#pragma unroll
            for (int q = 0; q < kJw; ++q) {
                const int j = w + 8 * q;
                if (j < n) {
                    // rows n.. of G are zero
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        acc[q] = fma(bv[u], G[j * kSvdLdB + c0 + u], acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kJw; ++q) {
            const int j = w + 8 * q;
            if (j >= n) break;
            const double sj = sv[j];
            int rk = 0;
            for (int r = 0; r < n; ++r) rk += (sv[r] > sj || (sv[r] == sj && r < j)) ? 1 : 0;
            if (ih == 0 && l == 0) s[rk] = sj;
            if (i < n) {
                U[(int64_t)rk * n + i] = acc[q] / (sj * sj);
                V[(int64_t)rk * n + i] = G[j * kSvdLdB + i] / sj;
            }
        }
    }
}

// LDS bytes and launch of the projected SVD for work columns
size_t svd_lds_bytes(int work) {
    const size_t ne = (size_t)(work + (work & 1));
    if (work <= kSvd8MaxWork) return sizeof(double) * 2 * ne * kSvdLd;
    if (work <= 64) return sizeof(double) * 2 * (size_t)work * (size_t)work;
    return sizeof(double) * ne * kSvdLdB;
}

// The projected SVD of the single-operator IRLBA for any work <= kMaxWork.
void launch_small_svd(int work, const double *B, double *BU, double *BS, double *BV, int *status,
                      hipStream_t st) {
    const size_t lds = svd_lds_bytes(work);
    if (work <= kSvd8MaxWork) {
        hipLaunchKernelGGL(k_small_svd8, dim3(1), dim3(kSvd8Threads), lds, st, work, B, BU, BS,
                           BV, status);
    } else if (work <= 64) {
        hipLaunchKernelGGL(k_small_svd, dim3(1), dim3(kSvdThreads), lds, st, work, B, BU, BS, BV,
                           status);
    } else {
        static bool attr = false;
        if (!attr) {
            SEV_HIP(hipFuncSetAttribute((const void *)k_small_svd_big,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)(sizeof(double) * kMaxWork * kSvdLdB)));
            attr = true;
        }
        hipLaunchKernelGGL(k_small_svd_big, dim3(1), dim3(kSvdBigThreads), lds, st, work, B, BU,
                           BS, BV, status);
    }
    SEV_LAUNCH_CHECK();
}

// --------------------------------------------------------------------------
// Restart decision after the SVD of B (one thread; the rules of the
// augmented restart): residuals res_i = ||F|| BU(work-1, i), Ritz-value
// change svr_i = |s_i^prev - s_i| / s_i, converged when nu pairs have
// |res_i| < tol max(s) and svr_i < svtol; otherwise the restart size
// k = max(k, nu + nconv) capped to [1, work - 3] and the new B =
// [diag(s_0..s_k-1) res_0..k-1 in column k].  info: [0] converged,
// [1] k, [2] ||F|| <= 0 (cannot restart), [3] nconv, [4] SVD needs the host.
// --------------------------------------------------------------------------
__device__ __forceinline__ void restart_check_body(int work, int nu, double tol, double svtol, const double *BU,
                                   const double *BS, const double *scal, const int *status,
                                   double *svr, double *B, int *info, int last) {
    __shared__ int s_go;  // rebuild B for a restart
    __shared__ int s_k;
    __shared__ double s_rf;
    const int tid = threadIdx.x, lane = lane_id();
    if (tid < kWave) {
        const int st = status[0];
        const double eps = 2.220446049250313e-16;
        double RF = scal[2];
        const double RFn = RF;
        if (RF < 2 * eps) RF = 0.0;
        // lane l looks after Ritz values l and l + 64
        const int i0 = lane, i1 = lane + kWave;
        const double bs0 = i0 < work ? BS[i0] : 0.0, bs1 = i1 < work ? BS[i1] : 0.0;
        double smax = fmax(bs0, bs1);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) smax = fmax(smax, __shfl_xor(smax, o, kWave));
        bool conv0 = false, conv1 = false;
        if (i0 < work) {
            const double r = RF * BU[(int64_t)i0 * work + (work - 1)];
            conv0 = fabs(r) < tol * smax && fabs(svr[i0] - bs0) / bs0 < svtol;
        }
        if (i1 < work) {
            const double r = RF * BU[(int64_t)i1 * work + (work - 1)];
            conv1 = fabs(r) < tol * smax && fabs(svr[i1] - bs1) / bs1 < svtol;
        }
        const int nconv = __popcll(__ballot(conv0)) + __popcll(__ballot(conv1));
        int go = 0, k = info[1];
        if (!st && nconv < nu && !last) {
            go = 1;
            if (k < nu + nconv) k = nu + nconv;
            if (k > work - 3) k = work - 3;
            if (k < 1) k = 1;
        }
        if (lane == 0) {
            info[5] += status[1];
            info[4] = st ? 1 : 0;
            if (!st) {
                info[3] = nconv;
                info[2] = RFn <= 0.0 ? 1 : 0;
                info[0] = nconv >= nu ? 1 : 0;
                info[1] = k;
            }
            s_go = go;
            s_k = k;
            s_rf = RF;
        }
    }
    __syncthreads();
    if (!s_go) return;
    const int k = s_k;
    const double RF = s_rf;
    // B = [diag(s_0..s_k-1), residuals in column k]; Ritz values kept for
    // the next convergence test
    for (int e = tid; e < work * work; e += kBlock) B[e] = 0.0;
    __syncthreads();
    for (int c = tid; c < k; c += kBlock) {
        B[(int64_t)c * work + c] = BS[c];
        B[(int64_t)k * work + c] = RF * BU[(int64_t)c * work + (work - 1)];
    }
    for (int i = tid; i < work; i += kBlock) svr[i] = BS[i];
}

__global__ void __launch_bounds__(kBlock)
k_restart_check(int work, int nu, double tol, double svtol, const double *BU, const double *BS,
                const double *scal, const int *status, double *svr, double *B, int *info,
                int last) {
    restart_check_body(work, nu, tol, svtol, BU, BS, scal, status, svr, B, info, last);
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

int val_bytes(int vtype) { return vtype == kValBF16 ? 2 : (vtype == kValF64 ? 8 : 4); }

template <class ST, class DT>
void launch_csr_place(int64_t nnz, int64_t n, const int64_t *colptr, const uint32_t *srow,
                      const int32_t *sid, const int64_t *cstart, const int64_t *rowptr,
                      const void *val, uint16_t *ccol, void *cval, hipStream_t st) {
    if (nnz == 0) return;
    hipLaunchKernelGGL((k_csr_place<ST, DT>), dim3(grid_rows(nnz)), dim3(kBlock), 0, st, nnz, n,
                       colptr, srow, sid, cstart, rowptr, (const ST *)val, ccol, (DT *)cval);
    SEV_LAUNCH_CHECK();
}

}  // namespace

// 12-bit copy of a 16-bit index array (nent entries, readable through the
// 8-entry group covering nent + kValuePad).
static void pack12(const uint16_t *in, int64_t nent, DevBuf<uint32_t> &out, hipStream_t st) {
    const int64_t ngroups = ceil_div(nent + kValuePad, (int64_t)8);
    out.ensure((size_t)(3 * ngroups));
    hipLaunchKernelGGL(k_pack12, dim3(grid_rows(ngroups)), dim3(kBlock), 0, st, in, nent,
                       ngroups, out.get());
    SEV_LAUNCH_CHECK();
}

void build_slab_csc(SlabCsc &op, int64_t m, int64_t n, const int64_t *colptr,
                    const int32_t *row, const void *val, bool src_f64, const void *mu, bool bf16,
                    hipStream_t st) {
    SEV_REQUIRE(m > 0 && n > 0, "empty operator");
    SEV_REQUIRE(n <= 65535, "more than 65535 features");
    SEV_REQUIRE(!(bf16 && src_f64), "bf16 operator needs Float32 source values");
    op.m = m;
    op.n = n;
    op.nslab = ceil_div(m, kSlabRows);
    SEV_HIP(hipMemcpyAsync(&op.nnz, colptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    op.off.ensure((size_t)((op.nslab + 1) * n));
    op.lrow.ensure((size_t)op.nnz + kValuePad);
    op.mu = mu;
    op.mu64 = src_f64;
    op.vals = val;
    op.vtype = src_f64 ? kValF64 : kValF32;
    {
        ProfScope ps("slab_build", st, 6.0 * op.nnz + 8.0 * (op.nslab + 1) * n);
        hipLaunchKernelGGL(k_slab_offsets, dim3((unsigned)n, 8), dim3(kBlock), 0, st, n,
                           op.nslab, kSlabRows, colptr, row, op.off.get(), op.lrow.get());
        SEV_LAUNCH_CHECK();
    }
    if (bf16) {
        op.vbf16.ensure((size_t)op.nnz + kValuePad);
        if (op.nnz > 0) {
            hipLaunchKernelGGL(k_to_bf16, dim3(grid_rows(op.nnz, 16)), dim3(kBlock), 0, st,
                               (const float *)val, op.vbf16.get(), op.nnz);
            SEV_LAUNCH_CHECK();
        }
        op.vals = op.vbf16.get();
        op.vtype = kValBF16;
        pack12(op.lrow.get(), op.nnz, op.lrow12, st);
    }
    // padded CSR for A v
    ProfScope ps("csr_build", st, 2.0 * 10.0 * op.nnz);
    BuildScratch &b = op.build;
    SEV_REQUIRE(op.nnz < ((int64_t)1 << 31) - 1, "more than 2^31 - 2 selected entries");
    b.cnt.ensure((size_t)m);
    b.scan.ensure(scan_scratch_elems(m));
    b.cstart.ensure((size_t)m + 1);
    op.rowptr.ensure((size_t)m + 1);
    row_counts(colptr, row, op.nnz, m, n, b.cnt.get(), st);
    // the entries' offsets among the sorted entries, then the padded rows
    scan_exclusive_i32(b.cnt.get(), b.cstart.get(), m, b.scan.get(), st);
    // BF16 rows are padded to 8 entries (16-byte loads of values and columns)
    const int ralign = bf16 ? 8 : 4;
    hipLaunchKernelGGL(k_round_up, dim3(grid_rows(m)), dim3(kBlock), 0, st, b.cnt.get(), m,
                       ralign);
    SEV_LAUNCH_CHECK();
    scan_exclusive_i32(b.cnt.get(), op.rowptr.get(), m, b.scan.get(), st);
    SEV_HIP(hipMemcpyAsync(&op.csr_nnz, op.rowptr.get() + m, sizeof(int64_t),
                           hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    const size_t vb = (size_t)val_bytes(op.vtype);
    op.ccol.ensure((size_t)op.csr_nnz + 8);
    op.cval.ensure(((size_t)op.csr_nnz + 8) * vb);
    SEV_HIP(hipMemsetAsync(op.ccol.get(), 0, sizeof(uint16_t) * (size_t)op.csr_nnz, st));
    SEV_HIP(hipMemsetAsync(op.cval.get(), 0, vb * (size_t)op.csr_nnz, st));
    const bool words = op.vtype == kValBF16;
    if (op.nnz > 0) {
        // entries by row: keys = the CSC rows (non-negative int32); values =
        // the entry's column and BF16 value (BF16 operators) or its id
        b.ids.ensure((size_t)op.nnz);
        b.sids.ensure((size_t)op.nnz);
        b.srow.ensure((size_t)op.nnz);
        if (words) {
            hipLaunchKernelGGL(k_csr_words, dim3((unsigned)std::min<int64_t>(n, 65535), 8),
                               dim3(kBlock), 0, st, n, colptr, (const float *)val,
                               (uint32_t *)b.ids.get());
        } else {
            hipLaunchKernelGGL(k_iota32, dim3(grid_rows(op.nnz)), dim3(kBlock), 0, st,
                               b.ids.get(), op.nnz);
        }
        SEV_LAUNCH_CHECK();
        int end_bit = 1;
        while (end_bit < 32 && (m >> end_bit) != 0) ++end_bit;
        size_t tb = 0;
        const uint32_t *keys = (const uint32_t *)row;
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, b.srow.get(), b.ids.get(),
                                                   b.sids.get(), (int)op.nnz, 0, end_bit, st));
        b.tmp.ensure(tb);
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(b.tmp.get(), tb, keys, b.srow.get(),
                                                   b.ids.get(), b.sids.get(), (int)op.nnz, 0,
                                                   end_bit, st));
    }
    const int64_t *rptr = op.rowptr.get(), *cst = b.cstart.get();
    if (words) {
        if (op.nnz > 0) {
            hipLaunchKernelGGL(k_csr_place_words, dim3(grid_rows(op.nnz)), dim3(kBlock), 0, st,
                               op.nnz, b.srow.get(), (const uint32_t *)b.sids.get(), cst, rptr,
                               op.ccol.get(), (uint16_t *)op.cval.get());
            SEV_LAUNCH_CHECK();
        }
    } else if (op.vtype == kValF64) {
        launch_csr_place<double, double>(op.nnz, n, colptr, b.srow.get(), b.sids.get(), cst,
                                         rptr, val, op.ccol.get(), op.cval.get(), st);
    } else {
        launch_csr_place<float, float>(op.nnz, n, colptr, b.srow.get(), b.sids.get(), cst, rptr,
                                       val, op.ccol.get(), op.cval.get(), st);
    }
    op.col12 = op.vtype == kValBF16 && n <= 4096;
    if (op.col12) pack12(op.ccol.get(), op.csr_nnz, op.ccol12, st);
    const double avg = (double)op.csr_nnz / (double)m;
    if (op.vtype == kValBF16)
        op.lanes_per_row = avg > 256 ? 16 : (avg > 64 ? 8 : 4);
    else
        op.lanes_per_row = avg > 96 ? 16 : 8;
    // A' w: lanes per column segment, enough for one pass over a segment of
    // the mean length (8 entries per lane and step in the BF16 form, 4
    // otherwise); a rule of the shape, like the lanes per row, since it
    // fixes how each segment is split into partial sums
    const double seg = (double)op.nnz / ((double)op.nslab * (double)n);
    const double per = op.vtype == kValBF16 ? 8.0 : 4.0;
    op.atw_lanes = 4;
    while (op.atw_lanes < kWave && op.atw_lanes * per < seg) op.atw_lanes <<= 1;
    // the build scratch is released when b goes out of scope: wait for the fill
    SEV_HIP(hipStreamSynchronize(st));
}

namespace {

template <class VT, class MT, int G>
void launch_av_g(const SlabCsc &op, int use_lds, const double *v, const double *rf,
                 const double *wprev, double *y, hipStream_t st) {
    const size_t smem = use_lds ? sizeof(double) * (size_t)op.n : 0;
    const unsigned grid = (unsigned)std::min<int64_t>(
        ceil_div(op.m, (int64_t)kWaves * (kWave / G)), (int64_t)device_cu_count() * 8);
    hipLaunchKernelGGL((k_av_csr<VT, MT, G>), dim3(grid), dim3(kBlock), smem, st, op.m, op.n,
                       op.rowptr.get(), op.ccol.get(), (const VT *)op.cval.get(),
                       (const MT *)op.mu, v, rf, wprev, y, use_lds);
    SEV_LAUNCH_CHECK();
}

template <class VT, class MT>
void launch_av_t(const SlabCsc &op, int use_lds, const double *v, const double *rf,
                 const double *wprev, double *y, hipStream_t st) {
    switch (op.lanes_per_row) {
        case 8: launch_av_g<VT, MT, 8>(op, use_lds, v, rf, wprev, y, st); break;
        case 16: launch_av_g<VT, MT, 16>(op, use_lds, v, rf, wprev, y, st); break;
        case 32: launch_av_g<VT, MT, 32>(op, use_lds, v, rf, wprev, y, st); break;
        default: launch_av_g<VT, MT, 64>(op, use_lds, v, rf, wprev, y, st); break;
    }
}

template <int G>
void launch_av_bf16_g(const SlabCsc &op, const double *v, const double *rf, const double *wprev,
                      double *y, hipStream_t st) {
    const unsigned grid = (unsigned)std::min<int64_t>(
        ceil_div(op.m, (int64_t)kWaves * (kWave / G)), (int64_t)device_cu_count() * 8);
    if (op.col12)
        hipLaunchKernelGGL((k_av_bf16<G, true>), dim3(grid), dim3(kBlock),
                           sizeof(float) * (size_t)op.n, st, op.m, op.n, op.rowptr.get(),
                           (const void *)op.ccol12.get(), (const uint16_t *)op.cval.get(),
                           (const float *)op.mu, v, rf, wprev, y);
    else
        hipLaunchKernelGGL((k_av_bf16<G, false>), dim3(grid), dim3(kBlock),
                           sizeof(float) * (size_t)op.n, st, op.m, op.n, op.rowptr.get(),
                           (const void *)op.ccol.get(), (const uint16_t *)op.cval.get(),
                           (const float *)op.mu, v, rf, wprev, y);
    SEV_LAUNCH_CHECK();
}

void launch_av_bf16(const SlabCsc &op, const double *v, const double *rf, const double *wprev,
                    double *y, hipStream_t st) {
    switch (op.lanes_per_row) {
        case 4: launch_av_bf16_g<4>(op, v, rf, wprev, y, st); break;
        case 8: launch_av_bf16_g<8>(op, v, rf, wprev, y, st); break;
        case 16: launch_av_bf16_g<16>(op, v, rf, wprev, y, st); break;
        default: launch_av_bf16_g<32>(op, v, rf, wprev, y, st); break;
    }
}

void launch_av(const SlabCsc &op, const double *v, const double *rf, const double *wprev,
               double *y, hipStream_t st) {
    const int use_lds = op.n <= kMaxFL ? 1 : 0;
    const bool bf = op.vtype == kValBF16 && op.n <= kMaxFL32;
    ProfScope ps("spmv_av", st,
                 (double)op.csr_nnz * (val_bytes(op.vtype) + (bf && op.col12 ? 1.5 : 2.0)) +
                     8.0 * op.m * (wprev ? 2 : 1) + 8.0 * (op.m + 1) + 8.0 * op.n);
    if (op.vtype == kValBF16 && op.n <= kMaxFL32)
        launch_av_bf16(op, v, rf, wprev, y, st);
    else if (op.vtype == kValBF16)
        launch_av_t<uint16_t, float>(op, use_lds, v, rf, wprev, y, st);
    else if (op.vtype == kValF64)
        launch_av_t<double, double>(op, use_lds, v, rf, wprev, y, st);
    else
        launch_av_t<float, float>(op, use_lds, v, rf, wprev, y, st);
}

int atw_groups_default(const SlabCsc &op) {
    const int64_t want = ceil_div((int64_t)device_cu_count() * 4, op.nslab);
    const int64_t cap = std::max<int64_t>(1, op.n / (4 * kWaves));
    return (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(want, 16), cap));
}

int atw_groups(const SlabCsc &op) { return op.atw_groups > 0 ? op.atw_groups : atw_groups_default(op); }

template <int G>
void launch_atw_g(const SlabCsc &op, const double *y, double *part, hipStream_t st) {
    const dim3 grid((unsigned)op.nslab, (unsigned)atw_groups(op));
    if (op.vtype == kValBF16)
        hipLaunchKernelGGL(k_atw_bf16<G>, grid, dim3(kBlock), 0, st, op.m, op.n, op.off.get(),
                           op.lrow12.get(), (const uint16_t *)op.vals, y, part);
    else if (op.vtype == kValF64)
        hipLaunchKernelGGL((k_atw<double, G>), grid, dim3(kBlock), 0, st, op.m, op.n,
                           op.off.get(), op.lrow.get(), (const double *)op.vals, y, part);
    else
        hipLaunchKernelGGL((k_atw<float, G>), grid, dim3(kBlock), 0, st, op.m, op.n,
                           op.off.get(), op.lrow.get(), (const float *)op.vals, y, part);
    SEV_LAUNCH_CHECK();
}

void launch_atw_only(const SlabCsc &op, const double *y, double *part, hipStream_t st) {
    ProfScope ps("spmv_atw", st,
                 (double)op.nnz * (val_bytes(op.vtype) + (op.vtype == kValBF16 ? 1.5 : 2.0)) +
                     8.0 * op.m + 8.0 * (op.nslab + 1) * op.n);
    switch (op.atw_lanes) {
        case 4: launch_atw_g<4>(op, y, part, st); break;
        case 8: launch_atw_g<8>(op, y, part, st); break;
        case 16: launch_atw_g<16>(op, y, part, st); break;
        case 32: launch_atw_g<32>(op, y, part, st); break;
        default: launch_atw_g<64>(op, y, part, st); break;
    }
}

// y' = y - W[:,0:jc] T into yout (plus norm/sum partials), then the slab
// partials of B' y'.
template <class TW>
void launch_upd_atw(const SlabCsc &op, const TW *W, int jc, const double *T,
                    const double *y, double *yout, double *part, hipStream_t st) {
    {
        ProfScope ps("reorth_apply", st, (double)op.m * (sizeof(TW) * jc + 16));
        hipLaunchKernelGGL(k_apply<TW>, dim3((unsigned)op.nslab), dim3(kApplyThreads), 0, st,
                           op.m, op.n, W, jc, T, y, yout, part);
        SEV_LAUNCH_CHECK();
    }
    launch_atw_only(op, yout, part, st);
}

void launch_right_step(const SlabCsc &op, int j, int work, const double *part2, double *V,
                       double *Fv, double *B, double *scal, int *flags, uint64_t seed,
                       double *rpart, hipStream_t st) {
    ProfScope ps("right_step", st, 16.0 * op.n * (j + 1));
    const int64_t n = op.n;
    const int nblk = (int)ceil_div(n, kRsBlock);
    double *npart = rpart + (size_t)nblk * kMaxWork;
    // the dots against V[:, 0:j] in column groups of 32, one block each
    const dim3 ga(nblk, (unsigned)((j + 32) / 32));
    if (op.mu64)
        hipLaunchKernelGGL(k_right_a<double>, ga, dim3(kRsBlock), 0, st, n, j, part2,
                           (const double *)op.mu, V, Fv, rpart);
    else
        hipLaunchKernelGGL(k_right_a<float>, ga, dim3(kRsBlock), 0, st, n, j, part2,
                           (const float *)op.mu, V, Fv, rpart);
    SEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_right_b, dim3(nblk), dim3(kRsBlock), 0, st, n, j, nblk, rpart, V, Fv,
                       npart);
    SEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_right_c<double>, dim3(1), dim3(1024), 0, st, n, j, work, nblk, part2,
                       npart, V, Fv, B, scal, flags, seed);
    SEV_LAUNCH_CHECK();
}

// Calls f(std::integral_constant<int, ns>) for ns = nks in 1..MAXNS.
template <int MAXNS, class F>
void switch_steps(int nks, F &&f) {
    static_assert(MAXNS <= 32, "instantiated up to 32 k-steps");
    if (nks < 1 || nks > MAXNS) throw std::runtime_error("restart product: work size out of range");
    switch (nks) {
    case 1: if constexpr (1 <= MAXNS) f(std::integral_constant<int, 1>()); break;
    case 2: if constexpr (2 <= MAXNS) f(std::integral_constant<int, 2>()); break;
    case 3: if constexpr (3 <= MAXNS) f(std::integral_constant<int, 3>()); break;
    case 4: if constexpr (4 <= MAXNS) f(std::integral_constant<int, 4>()); break;
    case 5: if constexpr (5 <= MAXNS) f(std::integral_constant<int, 5>()); break;
    case 6: if constexpr (6 <= MAXNS) f(std::integral_constant<int, 6>()); break;
    case 7: if constexpr (7 <= MAXNS) f(std::integral_constant<int, 7>()); break;
    case 8: if constexpr (8 <= MAXNS) f(std::integral_constant<int, 8>()); break;
    case 9: if constexpr (9 <= MAXNS) f(std::integral_constant<int, 9>()); break;
    case 10: if constexpr (10 <= MAXNS) f(std::integral_constant<int, 10>()); break;
    case 11: if constexpr (11 <= MAXNS) f(std::integral_constant<int, 11>()); break;
    case 12: if constexpr (12 <= MAXNS) f(std::integral_constant<int, 12>()); break;
    case 13: if constexpr (13 <= MAXNS) f(std::integral_constant<int, 13>()); break;
    case 14: if constexpr (14 <= MAXNS) f(std::integral_constant<int, 14>()); break;
    case 15: if constexpr (15 <= MAXNS) f(std::integral_constant<int, 15>()); break;
    case 16: if constexpr (16 <= MAXNS) f(std::integral_constant<int, 16>()); break;
    case 17: if constexpr (17 <= MAXNS) f(std::integral_constant<int, 17>()); break;
    case 18: if constexpr (18 <= MAXNS) f(std::integral_constant<int, 18>()); break;
    case 19: if constexpr (19 <= MAXNS) f(std::integral_constant<int, 19>()); break;
    case 20: if constexpr (20 <= MAXNS) f(std::integral_constant<int, 20>()); break;
    case 21: if constexpr (21 <= MAXNS) f(std::integral_constant<int, 21>()); break;
    case 22: if constexpr (22 <= MAXNS) f(std::integral_constant<int, 22>()); break;
    case 23: if constexpr (23 <= MAXNS) f(std::integral_constant<int, 23>()); break;
    case 24: if constexpr (24 <= MAXNS) f(std::integral_constant<int, 24>()); break;
    case 25: if constexpr (25 <= MAXNS) f(std::integral_constant<int, 25>()); break;
    case 26: if constexpr (26 <= MAXNS) f(std::integral_constant<int, 26>()); break;
    case 27: if constexpr (27 <= MAXNS) f(std::integral_constant<int, 27>()); break;
    case 28: if constexpr (28 <= MAXNS) f(std::integral_constant<int, 28>()); break;
    case 29: if constexpr (29 <= MAXNS) f(std::integral_constant<int, 29>()); break;
    case 30: if constexpr (30 <= MAXNS) f(std::integral_constant<int, 30>()); break;
    case 31: if constexpr (31 <= MAXNS) f(std::integral_constant<int, 31>()); break;
    case 32: if constexpr (32 <= MAXNS) f(std::integral_constant<int, 32>()); break;
    default: break;
    }
}

// Y[:, 0:kout] = X[:, 0:work] Cf[:, 0:kout] in launches of at most
// gemm_cols(NS) columns (extra, when given, follows the last column).
template <class TX, class TY>
void launch_gemm(const char *name, int64_t m, int work, const TX *X, const double *Cf,
                 int kout, TY *Y, const double *extra, const double *scal, hipStream_t st) {
    ProfScope ps(name, st, (double)m * (sizeof(TX) * work + sizeof(TY) * kout),
                 2.0 * m * work * kout);
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(ceil_div(m, 32), kWaves),
                                                      (int64_t)device_cu_count() * 8);
    const int nks = (work + 3) / 4;
    auto go = [&](auto ns) {
        constexpr int NS = decltype(ns)::value;
        constexpr int GC = gemm_cols(NS);
        for (int c0 = 0; c0 < kout; c0 += GC) {
            const int kc = kout - c0 < GC ? kout - c0 : GC;
            const bool last = c0 + GC >= kout;
            hipLaunchKernelGGL((k_gemm_mfma<NS, TX, TY>), dim3(grid), dim3(kBlock), 0, st, m,
                               work, X, Cf, c0, kc, Y, last ? extra : nullptr, scal);
        }
    };
    switch_steps<kMaxWork / 4>(nks, go);
    SEV_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Autotuner: for an operator shape not seen before (and large enough for the
// choice to matter) time every A v lane-group width and a ladder of A' w
// column-group counts on the operator itself, keep the fastest, and cache it
// per (rows, features, stored entries, value type).
// ---------------------------------------------------------------------------
struct TuneEntry {
    int lanes = 0, groups = 0;
};

std::map<std::tuple<int64_t, int64_t, int64_t, int>, TuneEntry> &tune_cache() {
    static std::map<std::tuple<int64_t, int64_t, int64_t, int>, TuneEntry> c;
    return c;
}

float time_launches(hipStream_t st, const std::function<void()> &f, int reps) {
    hipEvent_t a, b;
    SEV_HIP(hipEventCreate(&a));
    SEV_HIP(hipEventCreate(&b));
    f();  // warm-up
    SEV_HIP(hipEventRecord(a, st));
    for (int r = 0; r < reps; ++r) f();
    SEV_HIP(hipEventRecord(b, st));
    SEV_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SEV_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace

void autotune_operator(SlabCsc &op, hipStream_t st) {
    const auto key = std::make_tuple(op.m, op.n, op.csr_nnz, op.vtype);
    auto &cache = tune_cache();
    auto it = cache.find(key);
    if (it != cache.end()) {
        op.lanes_per_row = it->second.lanes;
        op.atw_groups = it->second.groups;
        return;
    }
    if (op.m < 50000) return;  // heuristics are as good as timing here
    ProfScope ps("autotune", st);
    DevBuf<double> v, y, part;
    v.ensure((size_t)op.n);
    y.ensure((size_t)op.m);
    part.ensure((size_t)(op.nslab * (op.n + 2)));
    SEV_HIP(hipMemsetAsync(v.get(), 0, sizeof(double) * op.n, st));
    SEV_HIP(hipMemsetAsync(y.get(), 0, sizeof(double) * op.m, st));
    Prof *saved = t_prof;
    t_prof = nullptr;  // keep the trial launches out of the kernel profile
    TuneEntry best;
    float best_ms = 1e30f;
    // The lanes per CSR row of A v keep the build's rule: they set how each
    // row's entries are split into partial sums, so a choice made by timing
    // would let the rounding, and so the results, differ between runs.  The
    // column groups of A' w are timed: every column segment is still summed
    // by one wavefront in the same order.
    best.lanes = op.lanes_per_row;
    const int cap = (int)std::max<int64_t>(1, op.n / (4 * kWaves));
    for (int g : {1, 2, 3, 4, 6, 8, 12, 16}) {
        if (g > cap) break;
        op.atw_groups = g;
        const float ms = time_launches(st, [&] { launch_atw_only(op, y.get(), part.get(), st); }, 3);
        if (ms < best_ms) {
            best_ms = ms;
            best.groups = g;
        }
    }
    op.atw_groups = best.groups;
    t_prof = saved;
    cache[key] = best;
    SEV_HIP(hipStreamSynchronize(st));
}

namespace {

}  // namespace

namespace {
template <class TW>
void irlba_run(const SlabCsc &op, const IrlbaParams &p, const double *init, IrlbaWork &w,
               DevBuf<TW> &Wb, DevBuf<TW> &Wb2, double *s, double *U, double *V, hipStream_t st,
               IrlbaOut *out, const HostOperator *hop) {
    // an FP32 basis keeps the current vector and the last normalised one in
    // FP64 buffers of their own
    constexpr bool b32 = std::is_same<TW, float>::value;
    const int64_t m = op.m, n = op.n;
    const int64_t nu = p.nu;
    int64_t workl = p.work;
    if (workl < nu) workl = nu + 1;
    if (workl > m) workl = m;
    if (workl > n) workl = n;
    const int work = (int)workl;
    SEV_REQUIRE(nu >= 1 && nu <= work, "irlba: need 1 <= nu <= min(m, n)");
    SEV_REQUIRE(work <= kMaxWork, "irlba: work (nu + 7) exceeds 128 Lanczos vectors, the most the device path keeps (nu <= 121)");
    const double eps = 2.220446049250313e-16;

    Wb.ensure((size_t)(m * work));
    Wb2.ensure((size_t)(m * work));
    if (b32) {
        w.ycur.ensure((size_t)m);
        w.wprev.ensure((size_t)m);
    }
    w.V.ensure((size_t)(n * work));
    w.V2.ensure((size_t)(n * work));
    w.B.ensure((size_t)(work * work));
    w.BU.ensure((size_t)(work * work));
    w.BV.ensure((size_t)(work * work));
    w.T.ensure(kMaxWork);
    w.part.ensure((size_t)(op.nslab * (n + 2)));
    w.part2.ensure((size_t)(kRedChunks * (n + 2)));
    w.red.ensure((size_t)op.nslab * kMaxWork);
    w.Fv.ensure((size_t)(n + m));  // F (n) followed by the updated y (m)
    w.rpart.ensure((size_t)ceil_div(n, kRsBlock) * (kMaxWork + 1));
    w.scal.ensure(4);
    w.flags.ensure(4);
    w.BS.ensure(kMaxWork);
    w.svr.ensure(kMaxWork);
    w.info.ensure(16);

    SEV_REQUIRE(!hop || (hop->matmul && !op.mu), "irlba: host operator needs matmul, no mu");
    const int64_t d0 = p.restart;
    SEV_REQUIRE(d0 >= 0 && d0 < work, "irlba: restart must be below the work size");
    SEV_REQUIRE(d0 == 0 || (p.rU && p.rs && p.rV), "irlba: restart needs the previous U, s, V");

    // starting vector v_d0 = init, orthogonalised against the d0 previous
    // right vectors (two classical Gram-Schmidt passes), normalised
    std::vector<double> v0(init, init + n);
    for (int pass = 0; pass < (d0 > 0 ? 2 : 0); ++pass) {
        std::vector<double> t((size_t)d0, 0.0);
        for (int64_t c = 0; c < d0; ++c)
            for (int64_t i = 0; i < n; ++i) t[c] += p.rV[c * n + i] * v0[i];
        for (int64_t c = 0; c < d0; ++c)
            for (int64_t i = 0; i < n; ++i) v0[i] -= t[c] * p.rV[c * n + i];
    }
    double nr = 0.0;
    for (int64_t i = 0; i < n; ++i) nr += v0[i] * v0[i];
    nr = std::sqrt(nr);
    if (nr < 2 * eps) fail(kErrConvergence, "irlba: starting vector is zero");
    for (int64_t i = 0; i < n; ++i) v0[i] /= nr;
    SEV_HIP(hipMemcpyAsync(w.V.get() + d0 * n, v0.data(), sizeof(double) * n,
                           hipMemcpyHostToDevice, st));
    SEV_HIP(hipMemsetAsync(w.B.get(), 0, sizeof(double) * work * work, st));
    SEV_HIP(hipMemsetAsync(w.flags.get(), 0, sizeof(int) * 4, st));
    SEV_HIP(hipMemsetAsync(w.scal.get(), 0, sizeof(double) * 4, st));
    if (d0 > 0) {
        // V[:,0:d0] = V_prev, W[:,0:d0] = U_prev, B = diag(s_prev); the
        // coupling column B[0:d0, d0] is taken from the first
        // reorthogonalisation (W[:,0:d0]' A v_d0)
        SEV_HIP(hipMemcpyAsync(w.V.get(), p.rV, sizeof(double) * n * d0, hipMemcpyHostToDevice,
                               st));
        SEV_REQUIRE(!b32, "irlba: a warm restart keeps the FP64 basis");
        SEV_HIP(hipMemcpyAsync(Wb.get(), p.rU, sizeof(TW) * m * d0, hipMemcpyHostToDevice, st));
        std::vector<double> Bd((size_t)work * work, 0.0);
        for (int64_t c = 0; c < d0; ++c) Bd[(size_t)c * work + c] = p.rs[c];
        SEV_HIP(hipMemcpyAsync(w.B.get(), Bd.data(), sizeof(double) * work * work,
                               hipMemcpyHostToDevice, st));
        SEV_HIP(hipStreamSynchronize(st));
    }

    // host operator staging: every product copies its input vector to the
    // host, calls matmul, and copies the result back
    std::vector<double> hx, hy, htemp;
    if (hop) {
        hx.resize((size_t)std::max(m, n));
        hy.resize((size_t)std::max(m, n));
        htemp.resize((size_t)std::max(m, n));
        // k_atw never runs: the column entries of the slab partials stay zero
        SEV_HIP(hipMemsetAsync(w.part.get(), 0, sizeof(double) * op.nslab * (n + 2), st));
    }
    auto host_mul = [&](char trans, const double *xd, int64_t nx, double *yd, int64_t ny) {
        ProfScope ps("host_matmul", st, 8.0 * (nx + ny));
        SEV_HIP(hipMemcpyAsync(hx.data(), xd, sizeof(double) * nx, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        hop->matmul(hy.data(), trans, hx.data(), htemp.data(), hop->data);
        SEV_HIP(hipMemcpyAsync(yd, hy.data(), sizeof(double) * ny, hipMemcpyHostToDevice, st));
        SEV_HIP(hipStreamSynchronize(st));
    };

    const int dot_blocks = (int)op.nslab;

    std::vector<double> Bh((size_t)work * work), BUh((size_t)work * work),
        BVh((size_t)work * work), BSh(work);
    int flags_h[4];
    int info_h[8] = {0, (int)nu, 0, 0, 0, 0, 0, 0};
    SEV_HIP(hipMemcpyAsync(w.info.get(), info_h, sizeof info_h, hipMemcpyHostToDevice, st));
    SEV_HIP(hipMemsetAsync(w.svr.get(), 0, sizeof(double) * kMaxWork, st));
    const size_t svd_lds = svd_lds_bytes(work);
    SEV_REQUIRE(svd_lds + 1024 <= (size_t)device_lds_per_block(),
                "irlba: projected matrix too large for the device SVD");
    int64_t iter = 0, nmult = 0, k = nu;
    bool converged = false;

    while (iter < p.maxit) {
        const int j0 = (iter == 0) ? (int)d0 : (int)k;
        for (int j = j0; j < work; ++j) {
            TW *Wd = Wb.get();
            double *Vd = w.V.get();
            // y = the product; with an FP64 basis it is built in column j
            double *y = b32 ? w.ycur.get() : (double *)(Wd + (int64_t)j * m);
            const double *wprev =
                j > j0 ? (b32 ? w.wprev.get() : (const double *)(Wd + (int64_t)(j - 1) * m))
                       : nullptr;
            // W[:,j] = A v_j - RF w_{j-1}
            if (hop) {
                host_mul('N', Vd + (int64_t)j * n, n, y, m);
                if (j > j0) {
                    hipLaunchKernelGGL(k_sub_rf, dim3(grid_rows(m)), dim3(kBlock), 0, st, y, m,
                                       w.scal.get() + 1, wprev);
                    SEV_LAUNCH_CHECK();
                }
            } else {
                launch_av(op, Vd + (int64_t)j * n, j > j0 ? w.scal.get() + 1 : nullptr, wprev, y,
                          st);
            }
            nmult++;
            // reorthogonalise against W[:,0:j]
            const int jc = j;
            if (jc > 0) {
                ProfScope ps("reorth_dot", st, (double)m * (sizeof(TW) * jc + 8));
                hipLaunchKernelGGL(k_dot_w<TW>, dim3(dot_blocks), dim3(kDotThreads), 0, st, m, Wd,
                                   jc, y, w.red.get());
                SEV_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_red_t, dim3(jc), dim3(kBlock), 0, st, dot_blocks,
                                   w.red.get(), w.T.get());
                SEV_LAUNCH_CHECK();
            }
            if (iter == 0 && j == j0 && d0 > 0)
                SEV_HIP(hipMemcpyAsync(w.B.get() + (int64_t)d0 * work, w.T.get(),
                                       sizeof(double) * d0, hipMemcpyDeviceToDevice, st));
            if (hop) {
                ProfScope ps("reorth_apply", st, (double)m * (sizeof(TW) * jc + 16));
                hipLaunchKernelGGL(k_apply<TW>, dim3((unsigned)op.nslab), dim3(kApplyThreads), 0,
                                   st, m, n, Wd, jc, w.T.get(), y, w.Fv.get() + n, w.part.get());
                SEV_LAUNCH_CHECK();
            } else {
                launch_upd_atw(op, Wd, jc, w.T.get(), y, w.Fv.get() + n, w.part.get(), st);
            }
            nmult++;
            {
                ProfScope ps("reduce_slabs", st, 8.0 * op.nslab * (n + 2));
                hipLaunchKernelGGL(k_red_slabs, dim3((unsigned)ceil_div(n + 2, kBlock), kRedChunks),
                                   dim3(kBlock), 0, st, op.nslab, n + 2, w.part.get(),
                                   w.part2.get());
                SEV_LAUNCH_CHECK();
            }
            // host operator: B' y' lands in the first reduction row (the
            // other rows hold zeros for the columns)
            if (hop) host_mul('T', w.Fv.get() + n, m, w.part2.get(), n);
            launch_right_step(op, j, work, w.part2.get(), Vd, w.Fv.get(), w.B.get(),
                              w.scal.get(), w.flags.get(), (uint64_t)(iter * work + j) + 17u,
                              w.rpart.get(), st);
            {
                if constexpr (b32) {
                    ProfScope ps("scale_col", st, 20.0 * m);
                    hipLaunchKernelGGL(k_scale_col32, dim3(grid_rows(m)), dim3(kBlock), 0, st,
                                       w.Fv.get() + n, Wd + (int64_t)j * m, w.wprev.get(), m,
                                       w.scal.get());
                } else {
                    ProfScope ps("scale_col", st, 16.0 * m);
                    hipLaunchKernelGGL(k_scale_col, dim3(grid_rows(m)), dim3(kBlock), 0, st,
                                       w.Fv.get() + n, y, m, w.scal.get());
                }
                SEV_LAUNCH_CHECK();
            }
        }
        // SVD of B and the restart decision on the device; the host only
        // reads the decision (converged, k) back
        {
            ProfScope ps("small_svd", st);
            launch_small_svd(work, w.B.get(), w.BU.get(), w.BS.get(), w.BV.get(),
                             w.info.get() + 8, st);
        }
        const int last = iter + 1 >= p.maxit ? 1 : 0;
        hipLaunchKernelGGL(k_restart_check, dim3(1), dim3(kBlock), 0, st, work, (int)nu, p.tol,
                           p.svtol, w.BU.get(), w.BS.get(), w.scal.get(), w.info.get() + 8,
                           w.svr.get(), w.B.get(), w.info.get(), last);
        SEV_LAUNCH_CHECK();
        SEV_HIP(hipMemcpyAsync(info_h, w.info.get(), sizeof(int) * 8, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipMemcpyAsync(flags_h, w.flags.get(), sizeof flags_h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        if (flags_h[0])
            fail(kErrConvergence, "irlba: invariant subspace encountered (rank-deficient operator)");
        if (info_h[4]) {
            // a singular value of B vanished: complete the basis on the host
            SEV_HIP(hipMemcpyAsync(Bh.data(), w.B.get(), sizeof(double) * work * work,
                                   hipMemcpyDeviceToHost, st));
            SEV_HIP(hipStreamSynchronize(st));
            small_svd(work, Bh.data(), BUh.data(), BSh.data(), BVh.data());
            SEV_HIP(hipMemcpyAsync(w.BU.get(), BUh.data(), sizeof(double) * work * work,
                                   hipMemcpyHostToDevice, st));
            SEV_HIP(hipMemcpyAsync(w.BV.get(), BVh.data(), sizeof(double) * work * work,
                                   hipMemcpyHostToDevice, st));
            SEV_HIP(hipMemcpyAsync(w.BS.get(), BSh.data(), sizeof(double) * work,
                                   hipMemcpyHostToDevice, st));
            SEV_HIP(hipMemsetAsync(w.info.get() + 8, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_restart_check, dim3(1), dim3(kBlock), 0, st, work, (int)nu, p.tol,
                               p.svtol, w.BU.get(), w.BS.get(), w.scal.get(), w.info.get() + 8,
                               w.svr.get(), w.B.get(), w.info.get(), last);
            SEV_LAUNCH_CHECK();
            SEV_HIP(hipMemcpyAsync(info_h, w.info.get(), sizeof(int) * 8, hipMemcpyDeviceToHost,
                                   st));
            SEV_HIP(hipStreamSynchronize(st));
        }
        iter++;
        if (info_h[0]) {
            converged = true;
            break;
        }
        if (iter >= p.maxit) break;
        k = info_h[1];
        if (info_h[2]) fail(kErrConvergence, "irlba: zero residual vector at restart");
        // restart: V <- [V BV[:,0:k], F/||F||], W <- W BU[:,0:k]; B was
        // rebuilt by the restart decision
        launch_gemm("restart_v", n, work, (const double *)w.V.get(), w.BV.get(), (int)k,
                    w.V2.get(), (const double *)w.Fv.get(), w.scal.get(), st);
        w.V.swap(w.V2);
        launch_gemm("restart_w", m, work, Wb.get(), w.BU.get(), (int)k, Wb2.get(),
                    (const double *)nullptr, w.scal.get(), st);
        Wb.swap(Wb2);
    }
    if (out) {
        out->iters = iter;
        out->nmult = nmult;
        out->svd_sweeps = info_h[5];
    }
    if (std::getenv("SEV_VERBOSE"))
        std::fprintf(stderr, "[severo_hip] irlba %lld x %lld nu=%lld work=%d: %lld restarts, %lld "
                     "products, %d Jacobi sweeps in the projected SVDs\n", (long long)m,
                     (long long)n, (long long)nu, work, (long long)iter, (long long)nmult,
                     info_h[5]);
    if (!converged) fail(kErrConvergence, "irlba: convergence failed (maxit reached)");
    SEV_HIP(hipMemcpyAsync(BSh.data(), w.BS.get(), sizeof(double) * work, hipMemcpyDeviceToHost,
                           st));
    SEV_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < nu; ++i) s[i] = BSh[i];
    launch_gemm("final_u", m, work, (const TW *)Wb.get(), w.BU.get(), (int)nu, U,
                (const double *)nullptr, w.scal.get(), st);
    launch_gemm("final_v", n, work, (const double *)w.V.get(), w.BV.get(), (int)nu, V,
                (const double *)nullptr, w.scal.get(), st);
    SEV_HIP(hipStreamSynchronize(st));
}
}  // namespace

void irlba_device(const SlabCsc &op, const IrlbaParams &p, const double *init, IrlbaWork &w,
                  double *s, double *U, double *V, hipStream_t st, IrlbaOut *out,
                  const HostOperator *hop) {
    // the FP32 left basis serves the device operator from a cold start
    if (p.basis32 && !hop && p.restart == 0) {
        w.W.release();
        w.W2.release();
        irlba_run<float>(op, p, init, w, w.Wf, w.Wf2, s, U, V, st, out, hop);
    } else {
        w.Wf.release();
        w.Wf2.release();
        w.ycur.release();
        w.wprev.release();
        irlba_run<double>(op, p, init, w, w.W, w.W2, s, U, V, st, out, hop);
    }
}

void op_apply(const SlabCsc &op, int trans, const double *x, double *y, IrlbaWork &w,
              hipStream_t st) {
    const int64_t n = op.n;
    if (!trans) {
        launch_av(op, x, nullptr, nullptr, y, st);
        return;
    }
    w.part.ensure((size_t)(op.nslab * (n + 2)));
    w.part2.ensure((size_t)(kRedChunks * (n + 2)));
    w.Fv.ensure((size_t)(n + op.m));
    launch_upd_atw(op, (const double *)nullptr, 0, nullptr, x, w.Fv.get() + n, w.part.get(), st);
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
}

void embedding_f32(const double *U, const double *s_dev, int64_t m, int64_t d, float *Z,
                   int64_t rs, int64_t cs, hipStream_t st) {
    hipLaunchKernelGGL(k_embed, dim3(grid_rows(m)), dim3(kBlock), 0, st, U, s_dev, m, d, Z, rs,
                       cs);
    SEV_LAUNCH_CHECK();
}

// ===========================================================================
// Batched IRLBA: S independent centered problems on the stacked rows of one
// operator, run in lockstep over the Lanczos steps.  Each sample keeps its
// own restart size k, convergence test and restart; at step j a sample takes
// part when it is not done and j >= its first step j0 (0 before the first
// restart, k after).  Per-sample state lives on the device, so a restart
// round is a fixed launch sequence for a given first step and buffer parity.
// ===========================================================================
namespace {

struct BDev {
    int64_t M, n;
    int work, S;
    const int32_t *row_sample;    // M
    const int64_t *slab_row0;     // nslab + 1
    const int32_t *slab_sample;   // nslab
    const int64_t *sample_slab0;  // S + 1
    const int32_t *blk_sample;    // nblk
    const int64_t *blk_r0;        // nblk + 1
    int64_t nslab;
    int *j0, *done, *iters, *fail;  // S each (state)
};

__device__ __forceinline__ bool b_active(const BDev &d, int s, int j) {
    return !d.done[s] && j >= d.j0[s];
}

template <int G, bool C12>
__global__ void __launch_bounds__(kBlock)
k_b_av_bf16(BDev d, const int64_t *rowptr, const void *col, const uint16_t *val,
            const float *mu_all, const double *Vall, int j, const double *scal,
            const double *wprevp, double *y) {
    extern __shared__ float vf[];
    __shared__ double dmu_s;
    const int s = d.blk_sample[blockIdx.x];
    if (!b_active(d, s, j)) return;
    const bool first = j == d.j0[s];
    const int64_t n = d.n;
    const double *v = Vall + ((int64_t)s * d.work + j) * n;
    const float *mu = mu_all + (int64_t)s * n;
    const int w = wave_index(), lane = lane_id();
    for (int64_t i = threadIdx.x; i < n; i += kBlock) vf[i] = (float)v[i];
    if (w == 0) {
        double dd = 0.0;
        for (int64_t i = lane; i < n; i += kWave) dd = fma((double)mu[i], v[i], dd);
        dd = wave_sum(dd);
        if (lane == 0) dmu_s = dd;
    }
    __syncthreads();
    const double dmu = dmu_s;
    const double rfv = first ? 0.0 : scal[4 * s + 1];
    const double *wprev = first ? nullptr : wprevp;
    constexpr int RPW = kWave / G;
    const int gl = lane % G, gi = lane / G;
    const int64_t rb = d.blk_r0[blockIdx.x], re = d.blk_r0[blockIdx.x + 1];
    for (int64_t r0 = rb + (int64_t)w * RPW; r0 < re; r0 += (int64_t)kWaves * RPW) {
        const int64_t r = r0 + gi;
        float a0 = 0.0f, a1 = 0.0f;
        if (r < re) {
            const int64_t b = rowptr[r], e = rowptr[r + 1];
            for (int64_t p = b + 8 * gl; p < e; p += 8 * G) {
                uint32_t c[8];
                load_idx8<C12>(col, p, c);
                const uint4 v8 = *(const uint4 *)(val + p);
                a0 = fmaf(bf_lo(v8.x), vf[c[0]], a0);
                a1 = fmaf(bf_hi(v8.x), vf[c[1]], a1);
                a0 = fmaf(bf_lo(v8.y), vf[c[2]], a0);
                a1 = fmaf(bf_hi(v8.y), vf[c[3]], a1);
                a0 = fmaf(bf_lo(v8.z), vf[c[4]], a0);
                a1 = fmaf(bf_hi(v8.z), vf[c[5]], a1);
                a0 = fmaf(bf_lo(v8.w), vf[c[6]], a0);
                a1 = fmaf(bf_hi(v8.w), vf[c[7]], a1);
            }
        }
        float acc = a0 + a1;
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        acc = __shfl(acc, gi * G, kWave);
        if (gl == 0 && r < re) {
            double t = (double)acc - dmu;
            if (wprev) t -= rfv * wprev[r];
            y[r] = t;
        }
    }
}

template <class TW>
__global__ void __launch_bounds__(kDotThreads)
k_b_dot_w(BDev d, const TW *W, int j, const double *y, double *part) {
    constexpr int NW = kDotThreads / kWave;
    __shared__ double red[NW][32];
    const int slab = blockIdx.x;
    const int s = d.slab_sample[slab];
    if (j == 0 || !b_active(d, s, j)) return;
    const int jc = j;
    const int64_t m = d.M;
    const int64_t rb = d.slab_row0[slab], re = d.slab_row0[slab + 1];
    const int w = threadIdx.x / kWave, lane = lane_id();
    for (int c0 = 0; c0 < jc; c0 += 32) {
        double acc[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc[i] = 0.0;
        for (int64_t r = rb + threadIdx.x; r < re; r += kDotThreads) {
            const double yr = y[r];
            // every load is issued (columns past jc read column jc - 1; their
            // sums are never stored), so the 32 loads of a row are in flight
            // together
            double wv[32];
#pragma unroll
            for (int i = 0; i < 32; ++i)
                wv[i] = (double)W[(int64_t)(c0 + i < jc ? c0 + i : jc - 1) * m + r];
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the FMAs
#pragma unroll
            for (int i = 0; i < 32; ++i) acc[i] = fma(wv[i], yr, acc[i]);
        }
        const double t = butterfly32(acc);
        if ((lane & 1) == 0) red[w][lane >> 1] = t;
        __syncthreads();
        if (threadIdx.x < 32 && c0 + (int)threadIdx.x < jc) {
            double u = 0.0;
            for (int q = 0; q < NW; ++q) u += red[q][threadIdx.x];
            part[(int64_t)(c0 + threadIdx.x) * d.nslab + slab] = u;
        }
        __syncthreads();
    }
}

// T_s[c] = sum over the sample's slabs (same order as the single problem)
__global__ void __launch_bounds__(kBlock)
k_b_red_t(BDev d, int j, const double *part, double *T) {
    __shared__ double scratch[kWaves];
    const int c = blockIdx.x, s = blockIdx.y;
    if (j == 0 || !b_active(d, s, j)) return;
    const int64_t sb = d.sample_slab0[s];
    const int nblk = (int)(d.sample_slab0[s + 1] - sb);
    const double *pc = part + (int64_t)c * d.nslab + sb;
    double t0 = 0.0, t1 = 0.0;
    int b = threadIdx.x;
    for (; b + kBlock < nblk; b += 2 * kBlock) {
        t0 += pc[b];
        t1 += pc[b + kBlock];
    }
    if (b < nblk) t0 += pc[b];
    const double t = block_sum(t0 + t1, scratch);
    if (threadIdx.x == 0) T[(int64_t)s * kMaxWork + c] = t;
}

template <class TW>
__global__ void __launch_bounds__(kApplyThreads)
k_b_apply(BDev d, const TW *W, int j, const double *T, const double *y, double *yout,
          double *part) {
    __shared__ double tl[kMaxWork];
    __shared__ double scratch[kApplyThreads / kWave];
    const int slab = blockIdx.x;
    const int s = d.slab_sample[slab];
    if (!b_active(d, s, j)) return;
    const int jc = j;
    const int64_t m = d.M, n = d.n;
    const int64_t r0 = d.slab_row0[slab];
    const int64_t nrows = d.slab_row0[slab + 1] - r0;
    if ((int)threadIdx.x < jc) tl[threadIdx.x] = T[(int64_t)s * kMaxWork + threadIdx.x];
    __syncthreads();
    double nrm = 0.0, sm = 0.0;
    for (int64_t lr = threadIdx.x; lr < nrows; lr += kApplyThreads) {
        const int64_t r = r0 + lr;
        double yv = y[r];
        if (jc > 0) {
            double dd = 0.0;
#pragma unroll 8
            for (int i = 0; i < jc; ++i) dd = fma((double)W[(int64_t)i * m + r], tl[i], dd);
            yv -= dd;
        }
        yout[r] = yv;
        nrm = fma(yv, yv, nrm);
        sm += yv;
    }
    nrm = block_sum(nrm, scratch);
    sm = block_sum(sm, scratch);
    if (threadIdx.x == 0) {
        part[(int64_t)slab * (n + 2) + n] = nrm;
        part[(int64_t)slab * (n + 2) + n + 1] = sm;
    }
}

__global__ void __launch_bounds__(kBlock)
k_b_atw_bf16(BDev d, int j, const int64_t *off, const uint32_t *lrow12, const uint16_t *val,
             const double *y, double *part) {
    __shared__ float wl[kSlabRows];
    const int64_t slab = blockIdx.x;
    const int s = d.slab_sample[slab];
    if (!b_active(d, s, j)) return;
    const int64_t n = d.n;
    const int g = blockIdx.y, ng = gridDim.y;
    const int64_t r0 = d.slab_row0[slab];
    const int64_t nrows = d.slab_row0[slab + 1] - r0;
    const int w = wave_index(), lane = lane_id();
    for (int64_t lr = threadIdx.x; lr < kSlabRows; lr += kBlock)
        wl[lr] = lr < nrows ? (float)y[r0 + lr] : 0.0f;
    __syncthreads();
    double *prow = part + slab * (n + 2);
    for (int64_t jj = (int64_t)g * kWaves + w; jj < n; jj += (int64_t)ng * kWaves) {
        const int64_t a = off[slab * n + jj], b = off[(slab + 1) * n + jj];
        const float acc = wave_sum(atw_seg_bf16(a, b, lrow12, val, wl, lane));
        if (lane == 0) prow[jj] = (double)acc;
    }
}

// part2_s[q][c] = sum over the slabs of chunk q of sample s
__global__ void __launch_bounds__(kBlock)
k_b_red_slabs(BDev d, int j, const double *part, double *part2) {
    const int s = blockIdx.z;
    if (!b_active(d, s, j)) return;
    const int64_t ncol = d.n + 2;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t q = blockIdx.y;
    if (c >= ncol) return;
    const int64_t sb = d.sample_slab0[s], ns = d.sample_slab0[s + 1] - sb;
    const int64_t per = (ns + kRedChunks - 1) / kRedChunks;
    const int64_t b0 = q * per, b1 = b0 + per < ns ? b0 + per : ns;
    double t = 0.0;
    for (int64_t b = b0; b < b1; ++b) t += part[(sb + b) * ncol + c];
    part2[((int64_t)s * kRedChunks + q) * ncol + c] = t;
}

__global__ void __launch_bounds__(kRsBlock)
k_b_right_a(BDev d, int j, const double *part2, const float *mu_all, const double *Vall,
            double *Fall, double *rpart, int nblk) {
    const int s = blockIdx.y;
    if (!b_active(d, s, j)) return;
    const int64_t n = d.n;
    right_a_body<float>(n, j, part2 + (int64_t)s * kRedChunks * (n + 2), mu_all + (int64_t)s * n,
                        Vall + (int64_t)s * d.work * n, Fall + (int64_t)s * n,
                        rpart + (int64_t)s * nblk * (kMaxWork + 1), blockIdx.x);
}

__global__ void __launch_bounds__(kRsBlock)
k_b_right_b(BDev d, int j, int nblk, const double *rpart, const double *Vall, double *Fall) {
    const int s = blockIdx.y;
    if (!b_active(d, s, j)) return;
    const int64_t n = d.n;
    const double *rp = rpart + (int64_t)s * nblk * (kMaxWork + 1);
    right_b_body(n, j, nblk, rp, Vall + (int64_t)s * d.work * n, Fall + (int64_t)s * n,
                 (double *)rp + (size_t)nblk * kMaxWork, blockIdx.x);
}

__global__ void __launch_bounds__(1024)
k_b_right_c(BDev d, int j, int nblk, const double *part2, const double *rpart, double *Vall,
            double *Fall, double *Ball, double *scal, int *flags) {
    const int s = blockIdx.x;
    if (!b_active(d, s, j)) return;
    const int64_t n = d.n;
    const int work = d.work;
    const double *rp = rpart + (int64_t)s * nblk * (kMaxWork + 1);
    right_c_body(n, j, work, nblk, part2 + (int64_t)s * kRedChunks * (n + 2),
                 rp + (size_t)nblk * kMaxWork, Vall + (int64_t)s * work * n,
                 Fall + (int64_t)s * n, Ball + (int64_t)s * work * work, scal + 4 * s,
                 flags + 4 * s, (uint64_t)(d.iters[s] * work + j) + 17u);
}

// W[:, j] = yout / S for the rows of active samples (and, for a basis in
// FP32, the FP64 value into w64 for the next product's recurrence term)
template <class TW>
__global__ void __launch_bounds__(kBlock)
k_b_scale_col(BDev d, int j, const double *yout, TW *W, double *w64, const double *scal) {
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < d.M;
         r += (int64_t)gridDim.x * kBlock) {
        const int s = d.row_sample[r];
        if (b_active(d, s, j)) {
            const double v = yout[r] / scal[4 * s];
            W[(int64_t)j * d.M + r] = (TW)v;
            if (w64) w64[r] = v;
        }
    }
}

__global__ void __launch_bounds__(kSvdThreads)
k_b_small_svd(BDev d, const double *Ball, double *BUall, double *BSall, double *BVall,
              int *info) {
    const int s = blockIdx.x;
    if (d.done[s]) return;
    const int work = d.work;
    const int64_t o = (int64_t)s * work * work;
    small_svd_body(work, Ball + o, BUall + o, BSall + (int64_t)s * kMaxWork, BVall + o,
                   info + 16 * s + 8);
}

__global__ void __launch_bounds__(kSvd8Threads)
k_b_small_svd8(BDev d, const double *Ball, double *BUall, double *BSall, double *BVall,
               int *info) {
    const int s = blockIdx.x;
    if (d.done[s]) return;
    const int work = d.work;
    const int64_t o = (int64_t)s * work * work;
    small_svd8_body(work, Ball + o, BUall + o, BSall + (int64_t)s * kMaxWork, BVall + o,
                    info + 16 * s + 8);
}

// The restart decision of each sample; then its state for the next round:
// j0 = k, iters + 1, done on convergence, failure or the last iteration.
__global__ void __launch_bounds__(kBlock)
k_b_restart_check(BDev d, int nu, int maxit, double tol, double svtol, const double *BUall,
                  const double *BSall, const double *scal, double *svr, double *Ball, int *info,
                  const int *flags) {
    const int s = blockIdx.x;
    if (d.done[s]) return;
    const int work = d.work;
    const int64_t o = (int64_t)s * work * work;
    int *inf = info + 16 * s;
    const int last = d.iters[s] + 1 >= maxit ? 1 : 0;
    const int inv = flags[4 * s];
    if (!inv)
        restart_check_body(work, nu, tol, svtol, BUall + o, BSall + (int64_t)s * kMaxWork,
                           scal + 4 * s, inf + 8, svr + (int64_t)s * kMaxWork, Ball + o, inf,
                           last);
    __syncthreads();
    if (threadIdx.x == 0) {
        d.iters[s] += 1;
        if (inv) {
            d.fail[s] = 1;  // invariant subspace
            d.done[s] = 1;
        } else if (inf[4]) {
            d.fail[s] = 4;  // needs the host SVD: reported as a failure here
            d.done[s] = 1;
        } else if (inf[0]) {
            d.done[s] = 1;  // converged
        } else if (last) {
            d.fail[s] = 2;  // maxit
            d.done[s] = 1;
        } else if (inf[2]) {
            d.fail[s] = 3;  // zero residual at restart
            d.done[s] = 1;
        } else {
            d.j0[s] = inf[1];
        }
    }
}

// Y_s[:, c] = X_s[:, 0:work] C_s[:, c] for c < kout_s: rows of the stacked
// W (ld M, sample from row_sample) or of the per-sample V blocks (n rows,
// ld n, sample = block).  The extra column kout_s = F_s / ||F_s|| for V.
// Samples that are done are skipped (restart) or selected by sel (final).
template <class TX>
__global__ void __launch_bounds__(kBlock)
k_b_gemm(BDev d, int vmode, const TX *X0, const TX *X1, const int *xsel, double *Y,
         int64_t ldy, const double *Call, int kfix, const double *Fall, const double *scal,
         bool restart) {
    const int64_t rows = vmode ? (int64_t)d.S * d.n : d.M;
    const int work = d.work;
    for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g < rows;
         g += (int64_t)gridDim.x * kBlock) {
        int s;
        int64_t r, ld;
        if (vmode) {
            s = (int)(g / d.n);
            r = g - (int64_t)s * d.n;
            ld = d.n;
        } else {
            s = d.row_sample[g];
            r = g;
            ld = d.M;
        }
        if (restart && d.done[s]) continue;
        const int kout = kfix > 0 ? kfix : d.j0[s];
        const TX *X = (xsel && xsel[s]) ? X1 : X0;
        const TX *Xs = vmode ? X + (int64_t)s * work * d.n : X;
        double *Ys = vmode ? Y + (int64_t)s * ldy * d.n : Y;
        const double *Cs = Call + (int64_t)s * work * work;
        for (int c0 = 0; c0 < kout; c0 += 16) {
            double acc[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) acc[c] = 0.0;
            for (int l = 0; l < work; ++l) {
                const double x = (double)Xs[(int64_t)l * ld + r];
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (c0 + c < kout) acc[c] = fma(x, Cs[(int64_t)(c0 + c) * work + l], acc[c]);
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c0 + c < kout) Ys[(int64_t)(c0 + c) * ld + r] = acc[c];
        }
        if (vmode && restart) Ys[(int64_t)kout * ld + r] = Fall[(int64_t)s * d.n + r] / scal[4 * s + 2];
    }
}

// The same on MFMA (the tiles of k_gemm_mfma): W side (vmode 0) one block
// per A v row block (rows of one sample, ld M); V side (vmode 1) block
// (x, s) takes rows 256x .. of sample s's n x work block (ld n).
template <int NS, class TX = double>
__global__ void __launch_bounds__(kBlock)
k_b_gemm_mfma(BDev d, int vmode, const TX *X0, const TX *X1, const int *xsel, TX *Y,
              int64_t ldy, const double *Call, int kfix, const double *Fall, const double *scal,
              bool restart) {
    constexpr int LD = kMaxWorkB + 1;
    __shared__ double cfl[kMaxWorkB * LD];
    const int work = d.work;
    int s;
    int64_t rb0, re0, ld;
    const TX *X;
    TX *Ys;
    if (vmode) {
        s = blockIdx.y;
        rb0 = (int64_t)blockIdx.x * 256;
        re0 = rb0 + 256 < d.n ? rb0 + 256 : d.n;
        ld = d.n;
    } else {
        s = d.blk_sample[blockIdx.x];
        rb0 = d.blk_r0[blockIdx.x];
        re0 = d.blk_r0[blockIdx.x + 1];
        ld = d.M;
    }
    if (restart && d.done[s]) return;
    X = (xsel && xsel[s]) ? X1 : X0;
    if (vmode) {
        X += (int64_t)s * work * d.n;
        Ys = Y + (int64_t)s * ldy * d.n;
    } else {
        Ys = Y;
    }
    const int kout = kfix > 0 ? kfix : d.j0[s];
    const double *Cf = Call + (int64_t)s * work * work;
    for (int i = threadIdx.x; i < kMaxWorkB * kMaxWorkB; i += kBlock) {
        const int c = i / kMaxWorkB, l = i % kMaxWorkB;
        cfl[c * LD + l] = (c < kout && l < work) ? Cf[c * work + l] : 0.0;
    }
    __syncthreads();
    const int w = wave_index(), lane = lane_id();
    const int nr = lane & 15, kq = lane >> 4;
    const int ncb = (kout + 15) / 16;
    for (int64_t t0 = rb0 + (int64_t)w * 16; t0 < re0; t0 += (int64_t)kWaves * 16) {
        const int64_t row = t0 + nr;
        const bool ok = row < re0;
        // all NS = ceil(work / 4) loads of the 16 x work slice first
        // (clamped indices, out-of-range values multiplied by zero), then
        // the MFMA chains
        const int64_t rc = ok ? row : re0 - 1;
        double bv[NS];
#pragma unroll
        for (int st = 0; st < NS; ++st) {
            const int k = 4 * st + kq;
            bv[st] = (double)X[(int64_t)(k < work ? k : work - 1) * ld + rc] *
                     ((k < work && ok) ? 1.0 : 0.0);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            if (cb >= ncb) break;
            v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int st = 0; st < NS; ++st) {
                const double av = cfl[(cb * 16 + nr) * LD + 4 * st + kq];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[st], acc, 0, 0, 0);
            }
            if (ok) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = cb * 16 + kq + 4 * i;
                    if (c < kout) Ys[(int64_t)c * ld + row] = (TX)acc[i];
                }
            }
        }
        if (ok) {
            if (vmode && restart && kq == 0)
                Ys[(int64_t)kout * ld + row] = (TX)(Fall[(int64_t)s * d.n + row] / scal[4 * s + 2]);
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_b_embed(const double *U, const double *s_all, const int32_t *row_sample, int64_t M, int64_t dd,
          float *Z, int64_t rs, int64_t cs) {
    for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < M;
         r += (int64_t)gridDim.x * kBlock) {
        const int s = row_sample[r];
        for (int64_t c = 0; c < dd; ++c)
            Z[r * rs + c * cs] = (float)(U[c * M + r] * s_all[(int64_t)s * dd + c]);
    }
}

__global__ void __launch_bounds__(kBlock)
k_var_slab_offsets(int64_t n, int64_t nslab, const int32_t *slab_of_row, const int64_t *slab_row0,
                   const int64_t *colptr, const int32_t *row, int64_t *off, uint16_t *lrow) {
    // column blockIdx.x; its entries are shared by the gridDim.y blocks of
    // that column (each entry's writes are its own)
    const int64_t j = blockIdx.x;
    const int64_t b = colptr[j], e = colptr[j + 1];
    if (b == e) {
        if (blockIdx.y == 0)
            for (int64_t s = threadIdx.x; s <= nslab; s += kBlock) off[s * n + j] = b;
        return;
    }
    for (int64_t p = b + (int64_t)blockIdx.y * kBlock + threadIdx.x; p < e;
         p += (int64_t)gridDim.y * kBlock) {
        const int64_t r = row[p];
        const int64_t sc = slab_of_row[r];
        const int64_t sp = (p == b) ? -1 : (int64_t)slab_of_row[row[p - 1]];
        for (int64_t s = sp + 1; s <= sc; ++s) off[s * n + j] = p;
        if (p == e - 1)
            for (int64_t s = sc + 1; s <= nslab; ++s) off[s * n + j] = e;
        lrow[p] = (uint16_t)(r - slab_row0[sc]);
    }
}

__global__ void __launch_bounds__(kBlock)
k_fill_slab_of_row(int64_t M, const int64_t *slab_row0, int64_t nslab, int32_t *out) {
    for (int64_t s = blockIdx.x; s < nslab; s += gridDim.x)
        for (int64_t r = slab_row0[s] + threadIdx.x; r < slab_row0[s + 1]; r += kBlock)
            out[r] = (int32_t)s;
}

__global__ void __launch_bounds__(kBlock)
k_fill_i32(int32_t *a, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        a[i] = v;
}

__global__ void __launch_bounds__(kBlock)
k_fill_f64(double *a, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        a[i] = v;
}

}  // namespace

void batch_layout(BatchLayout &L, const std::vector<int64_t> &cells, hipStream_t st) {
    L.S = (int64_t)cells.size();
    SEV_REQUIRE(L.S >= 1, "batch: no samples");
    L.row0.assign((size_t)L.S + 1, 0);
    for (int64_t s = 0; s < L.S; ++s) {
        SEV_REQUIRE(cells[(size_t)s] >= 2, "batch: every sample needs at least two cells");
        L.row0[(size_t)s + 1] = L.row0[(size_t)s] + cells[(size_t)s];
    }
    L.M = L.row0[(size_t)L.S];
    L.slab_row0.clear();
    L.slab_sample.clear();
    L.sample_slab0.assign((size_t)L.S + 1, 0);
    L.blk_sample.clear();
    L.blk_r0.clear();
    constexpr int64_t kBlkRows = 128;  // A v and restart row blocks
    for (int64_t s = 0; s < L.S; ++s) {
        L.sample_slab0[(size_t)s] = (int64_t)L.slab_sample.size();
        for (int64_t r = L.row0[(size_t)s]; r < L.row0[(size_t)s + 1]; r += kSlabRows) {
            L.slab_row0.push_back(r);
            L.slab_sample.push_back((int32_t)s);
        }
        for (int64_t r = L.row0[(size_t)s]; r < L.row0[(size_t)s + 1]; r += kBlkRows) {
            L.blk_r0.push_back(r);
            L.blk_sample.push_back((int32_t)s);
        }
    }
    L.sample_slab0[(size_t)L.S] = (int64_t)L.slab_sample.size();
    L.slab_row0.push_back(L.M);
    L.blk_r0.push_back(L.M);
    std::vector<int32_t> rs((size_t)L.M);
    for (int64_t s = 0; s < L.S; ++s)
        for (int64_t r = L.row0[(size_t)s]; r < L.row0[(size_t)s + 1]; ++r) rs[(size_t)r] = (int32_t)s;
    auto up64 = [&](DevBuf<int64_t> &b, const std::vector<int64_t> &h) {
        b.ensure(h.size());
        SEV_HIP(hipMemcpyAsync(b.get(), h.data(), sizeof(int64_t) * h.size(), hipMemcpyHostToDevice,
                               st));
    };
    auto up32 = [&](DevBuf<int32_t> &b, const std::vector<int32_t> &h) {
        b.ensure(h.size());
        SEV_HIP(hipMemcpyAsync(b.get(), h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice,
                               st));
    };
    up64(L.row0_d, L.row0);
    up64(L.slab_row0_d, L.slab_row0);
    up64(L.sample_slab0_d, L.sample_slab0);
    up64(L.blk_r0_d, L.blk_r0);
    up32(L.row_sample_d, rs);
    up32(L.slab_sample_d, L.slab_sample);
    up32(L.blk_sample_d, L.blk_sample);
    SEV_HIP(hipStreamSynchronize(st));
}

void batch_slabs(SlabCsc &op, const BatchLayout &L, const int64_t *colptr, const int32_t *row,
                 hipStream_t st) {
    SEV_REQUIRE(op.m == L.M, "batch: operator rows differ from the layout");
    const int64_t nslab = (int64_t)L.slab_sample.size();
    DevBuf<int32_t> slab_of_row;
    slab_of_row.ensure((size_t)L.M);
    hipLaunchKernelGGL(k_fill_slab_of_row, dim3((unsigned)std::min<int64_t>(nslab, 65535)),
                       dim3(kBlock), 0, st, L.M, L.slab_row0_d.get(), nslab, slab_of_row.get());
    SEV_LAUNCH_CHECK();
    op.nslab = nslab;
    op.off.ensure((size_t)((nslab + 1) * op.n));
    hipLaunchKernelGGL(k_var_slab_offsets, dim3((unsigned)op.n, 8), dim3(kBlock), 0, st, op.n, nslab,
                       slab_of_row.get(), L.slab_row0_d.get(), colptr, row, op.off.get(),
                       op.lrow.get());
    SEV_LAUNCH_CHECK();
    if (op.vtype == kValBF16) pack12(op.lrow.get(), op.nnz, op.lrow12, st);
    SEV_HIP(hipStreamSynchronize(st));
}

BatchIrlbaWork::~BatchIrlbaWork() {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
}

namespace {
template <class TW>
void irlba_batch_run(const SlabCsc &op, const BatchLayout &L, const float *mu_all,
                     const IrlbaParams &p, const double *init, BatchIrlbaWork &w,
                     DevBuf<TW> (&Wb)[2], double *s_out, double *U, double *V,
                     int64_t *iters_out, int64_t *nmult_out, bool graph, hipStream_t st) {
    // an FP32 basis keeps the current product and the last normalised vector
    // in FP64 buffers of their own
    constexpr bool b32 = std::is_same<TW, float>::value;
    SEV_REQUIRE(op.vtype == kValBF16 && op.n <= kMaxFL32,
                "batched irlba: BF16 operator with at most 14336 features");
    const int64_t S = L.S, M = L.M, n = op.n;
    const int nu = (int)p.nu;
    int64_t workl = p.work;
    int64_t minm = M;
    for (int64_t s = 0; s < S; ++s) minm = std::min<int64_t>(minm, L.row0[s + 1] - L.row0[s]);
    if (workl > minm) workl = minm;
    if (workl > n) workl = n;
    const int work = (int)workl;
    SEV_REQUIRE(nu >= 1 && nu <= work, "batched irlba: need 1 <= nu <= min(cells, features)");
    SEV_REQUIRE(work <= kMaxWorkB, "batched irlba: work exceeds 64");
    const int64_t nslab = (int64_t)L.slab_sample.size();
    const int nblk_rs = (int)ceil_div(n, kRsBlock);
    Wb[0].ensure((size_t)(M * work));
    Wb[1].ensure((size_t)(M * work));
    if (b32) {
        w.ycur.ensure((size_t)M);
        w.wprev.ensure((size_t)M);
    }
    w.V[0].ensure((size_t)(S * n * work));
    w.V[1].ensure((size_t)(S * n * work));
    w.B.ensure((size_t)(S * work * work));
    w.BU.ensure((size_t)(S * work * work));
    w.BV.ensure((size_t)(S * work * work));
    w.BS.ensure((size_t)(S * kMaxWork));
    w.svr.ensure((size_t)(S * kMaxWork));
    w.T.ensure((size_t)(S * kMaxWork));
    w.part.ensure((size_t)(nslab * (n + 2)));
    w.part2.ensure((size_t)(S * kRedChunks * (n + 2)));
    w.red.ensure((size_t)(nslab * kMaxWork));
    w.Fy.ensure((size_t)M);
    w.F.ensure((size_t)(S * n));
    w.scal.ensure((size_t)(4 * S));
    w.rpart.ensure((size_t)(S * nblk_rs * (kMaxWork + 1)));
    w.flags.ensure((size_t)(4 * S));
    w.info.ensure((size_t)(16 * S));
    w.state.ensure((size_t)(4 * S));
    const size_t svd_lds = svd_lds_bytes(work);
    SEV_REQUIRE(svd_lds + 1024 <= (size_t)device_lds_per_block(),
                "batched irlba: projected matrix too large for the device SVD");

    // starting vectors and state
    // column 0 of each sample's V block (stride work * n); the other
    // columns start at zero
    std::vector<double> v0((size_t)(S * n));
    for (int64_t s = 0; s < S; ++s) {
        const double *in = init + s * n;
        double nr = 0.0;
        for (int64_t i = 0; i < n; ++i) nr += in[i] * in[i];
        nr = std::sqrt(nr);
        if (nr < 4.5e-16) fail(kErrConvergence, "batched irlba: a starting vector is zero");
        for (int64_t i = 0; i < n; ++i) v0[(size_t)(s * n + i)] = in[i] / nr;
    }
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_rows(S * n * work)), dim3(kBlock), 0, st,
                       w.V[0].get(), S * n * work, 0.0);
    SEV_LAUNCH_CHECK();
    SEV_HIP(hipMemcpy2DAsync(w.V[0].get(), sizeof(double) * n * work, v0.data(),
                             sizeof(double) * n, sizeof(double) * n, (size_t)S,
                             hipMemcpyHostToDevice, st));
    std::vector<int> st_h((size_t)(4 * S), 0);  // j0 | done | iters | fail
    std::vector<int> info_h((size_t)(16 * S), 0);
    for (int64_t s = 0; s < S; ++s) info_h[(size_t)(16 * s + 1)] = nu;
    SEV_HIP(hipMemcpyAsync(w.state.get(), st_h.data(), sizeof(int) * st_h.size(),
                           hipMemcpyHostToDevice, st));
    SEV_HIP(hipMemcpyAsync(w.info.get(), info_h.data(), sizeof(int) * info_h.size(),
                           hipMemcpyHostToDevice, st));
    auto zero_f64 = [&](DevBuf<double> &b, int64_t cnt) {
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_rows(cnt)), dim3(kBlock), 0, st, b.get(), cnt,
                           0.0);
        SEV_LAUNCH_CHECK();
    };
    zero_f64(w.B, S * work * work);
    zero_f64(w.scal, 4 * S);
    zero_f64(w.svr, S * kMaxWork);
    zero_f64(w.part, nslab * (n + 2));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, st, w.flags.get(), 4 * S, 0);
    SEV_LAUNCH_CHECK();

    int *state = w.state.get();
    BDev d;
    d.M = M;
    d.n = n;
    d.work = work;
    d.S = (int)S;
    d.row_sample = L.row_sample_d.get();
    d.slab_row0 = L.slab_row0_d.get();
    d.slab_sample = L.slab_sample_d.get();
    d.sample_slab0 = L.sample_slab0_d.get();
    d.blk_sample = L.blk_sample_d.get();
    d.blk_r0 = L.blk_r0_d.get();
    d.nslab = nslab;
    d.j0 = state;
    d.done = state + S;
    d.iters = state + 2 * S;
    d.fail = state + 3 * S;
    const int64_t nblk_av = (int64_t)L.blk_sample.size();
    const int G = op.lanes_per_row;
    // A' w column groups: enough blocks per slab for the late rounds, when
    // only the slabs of a few samples are active
    const int ng = (int)std::max<int64_t>(1, std::min<int64_t>(16, n / (4 * kWaves)));
    const double tol = p.tol, svtol = p.svtol;
    const int maxit = (int)std::min<int64_t>(p.maxit, 1 << 30);
    // recorded graphs hold every pointer and size below: they are replayed
    // only while all of them are unchanged
    {
        const void *bufs[] = {Wb[0].get(), Wb[1].get(), w.ycur.get(), w.wprev.get(),
                              w.V[0].get(), w.V[1].get(),
                              w.B.get(), w.BU.get(), w.BV.get(), w.BS.get(), w.svr.get(),
                              w.T.get(), w.part.get(), w.part2.get(), w.red.get(), w.Fy.get(),
                              w.F.get(), w.scal.get(), w.rpart.get()};
        std::vector<uint64_t> sig;
        for (const void *b : bufs) sig.push_back((uint64_t)(uintptr_t)b);
        const void *ptrs[] = {w.flags.get(), w.info.get(), w.state.get(), mu_all,
                              op.rowptr.get(), op.ccol.get(), op.cval.get(), op.off.get(),
                              op.lrow.get(), op.vals, L.row_sample_d.get(), L.slab_row0_d.get(),
                              L.slab_sample_d.get(), L.sample_slab0_d.get(),
                              L.blk_sample_d.get(), L.blk_r0_d.get()};
        for (const void *q : ptrs) sig.push_back((uint64_t)(uintptr_t)q);
        uint64_t tolb, svb;
        std::memcpy(&tolb, &tol, 8);
        std::memcpy(&svb, &svtol, 8);
        const uint64_t sc[] = {(uint64_t)M, (uint64_t)n, (uint64_t)S, (uint64_t)work,
                               (uint64_t)nu, (uint64_t)maxit, tolb, svb, (uint64_t)G,
                               (uint64_t)ng, (uint64_t)nslab, (uint64_t)nblk_av};
        sig.insert(sig.end(), std::begin(sc), std::end(sc));
        if (sig != w.sig) {
            for (auto &kv : w.graphs) (void)hipGraphExecDestroy(kv.second);
            w.graphs.clear();
            w.sig = sig;
        }
    }

    auto round = [&](int jmin, int cur) {
        TW *Wc = Wb[cur].get(), *Wn = Wb[cur ^ 1].get();
        double *Vc = w.V[cur].get(), *Vn = w.V[cur ^ 1].get();
        for (int j = jmin; j < work; ++j) {
            // the product; with an FP64 basis it is built in column j
            double *y = b32 ? w.ycur.get() : (double *)(Wc + (int64_t)j * M);
            const double *wp =
                j > 0 ? (b32 ? w.wprev.get() : (const double *)(Wc + (int64_t)(j - 1) * M))
                      : nullptr;
            {
                const void *colp = op.col12 ? (const void *)op.ccol12.get()
                                            : (const void *)op.ccol.get();
                const size_t lds = sizeof(float) * n;
                const dim3 gb((unsigned)nblk_av);
#define SEV_B_AV(GG)                                                                       \
    do {                                                                                   \
        if (op.col12)                                                                      \
            hipLaunchKernelGGL((k_b_av_bf16<GG, true>), gb, dim3(kBlock), lds, st, d,      \
                               op.rowptr.get(), colp, (const uint16_t *)op.cval.get(),     \
                               mu_all, Vc, j, w.scal.get(), wp, y);                        \
        else                                                                               \
            hipLaunchKernelGGL((k_b_av_bf16<GG, false>), gb, dim3(kBlock), lds, st, d,     \
                               op.rowptr.get(), colp, (const uint16_t *)op.cval.get(),     \
                               mu_all, Vc, j, w.scal.get(), wp, y);                        \
    } while (0)
                switch (G) {
                    case 4: SEV_B_AV(4); break;
                    case 8: SEV_B_AV(8); break;
                    case 16: SEV_B_AV(16); break;
                    default: SEV_B_AV(32); break;
                }
#undef SEV_B_AV
            }
            SEV_LAUNCH_CHECK();
            if (j > 0) {
                hipLaunchKernelGGL(k_b_dot_w<TW>, dim3((unsigned)nslab), dim3(kDotThreads), 0, st,
                                   d, (const TW *)Wc, j, y, w.red.get());
                SEV_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_b_red_t, dim3(j, (unsigned)S), dim3(kBlock), 0, st, d, j,
                                   w.red.get(), w.T.get());
                SEV_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_b_apply<TW>, dim3((unsigned)nslab), dim3(kApplyThreads), 0, st, d,
                               (const TW *)Wc, j, w.T.get(), y, w.Fy.get(), w.part.get());
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_b_atw_bf16, dim3((unsigned)nslab, ng), dim3(kBlock), 0, st, d, j,
                               op.off.get(), op.lrow12.get(), (const uint16_t *)op.vals, w.Fy.get(),
                               w.part.get());
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_b_red_slabs,
                               dim3((unsigned)ceil_div(n + 2, kBlock), kRedChunks, (unsigned)S),
                               dim3(kBlock), 0, st, d, j, w.part.get(), w.part2.get());
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_b_right_a, dim3(nblk_rs, (unsigned)S), dim3(kRsBlock), 0, st, d,
                               j, w.part2.get(), mu_all, Vc, w.F.get(), w.rpart.get(), nblk_rs);
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_b_right_b, dim3(nblk_rs, (unsigned)S), dim3(kRsBlock), 0, st, d,
                               j, nblk_rs, w.rpart.get(), Vc, w.F.get());
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_b_right_c, dim3((unsigned)S), dim3(1024), 0, st, d, j, nblk_rs,
                               w.part2.get(), w.rpart.get(), Vc, w.F.get(), w.B.get(),
                               w.scal.get(), w.flags.get());
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_b_scale_col<TW>, dim3(grid_rows(M)), dim3(kBlock), 0, st, d, j,
                               w.Fy.get(), Wc, b32 ? w.wprev.get() : nullptr, w.scal.get());
            SEV_LAUNCH_CHECK();
        }
        if (work <= kSvd8MaxWork)
            hipLaunchKernelGGL(k_b_small_svd8, dim3((unsigned)S), dim3(kSvd8Threads), svd_lds, st,
                               d, w.B.get(), w.BU.get(), w.BS.get(), w.BV.get(), w.info.get());
        else
            hipLaunchKernelGGL(k_b_small_svd, dim3((unsigned)S), dim3(kSvdThreads), svd_lds, st,
                               d, w.B.get(), w.BU.get(), w.BS.get(), w.BV.get(), w.info.get());
        SEV_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_b_restart_check, dim3((unsigned)S), dim3(kBlock), 0, st, d, nu, maxit,
                           tol, svtol, w.BU.get(), w.BS.get(), w.scal.get(), w.svr.get(),
                           w.B.get(), w.info.get(), w.flags.get());
        SEV_LAUNCH_CHECK();
        // restart of the continuing samples into the other buffers
        switch_steps<kMaxWorkB / 4>((work + 3) / 4, [&](auto ns) {
            constexpr int NS = decltype(ns)::value;
            hipLaunchKernelGGL(k_b_gemm_mfma<NS>, dim3((unsigned)ceil_div(n, 256), (unsigned)S),
                               dim3(kBlock), 0, st, d, 1, (const double *)Vc, (const double *)Vc,
                               (const int *)nullptr, Vn,
                               (int64_t)work, w.BV.get(), 0, w.F.get(), w.scal.get(), true);
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_b_gemm_mfma<NS, TW>), dim3((unsigned)nblk_av), dim3(kBlock), 0,
                               st, d, 0, (const TW *)Wc, (const TW *)Wc, (const int *)nullptr, Wn,
                               M, w.BU.get(), 0, (const double *)nullptr, w.scal.get(), true);
            SEV_LAUNCH_CHECK();
        });
    };

    std::vector<int> fin_buf((size_t)S, -1);
    int cur = 0, jmin = 0, nround = 0;
    int64_t nmult = 0;
    std::vector<int> j0_prev((size_t)S, 0);
    Prof *saved = t_prof;
    while (true) {
        {
            ProfScope ps("batch_round", st);
            if (graph) {
                const auto key = std::make_pair(jmin, cur);
                auto it = w.graphs.find(key);
                if (it == w.graphs.end()) {
                    t_prof = nullptr;  // no profiler events inside the capture
                    hipGraph_t g;
                    SEV_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    round(jmin, cur);
                    SEV_HIP(hipStreamEndCapture(st, &g));
                    hipGraphExec_t ge;
                    SEV_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                    (void)hipGraphDestroy(g);
                    it = w.graphs.emplace(key, ge).first;
                    t_prof = saved;
                }
                SEV_HIP(hipGraphLaunch(it->second, st));
            } else {
                round(jmin, cur);
            }
        }
        nround++;
        SEV_HIP(hipMemcpyAsync(st_h.data(), state, sizeof(int) * st_h.size(),
                               hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        bool any = false;
        jmin = work;
        for (int64_t s = 0; s < S; ++s) {
            const int done = st_h[(size_t)(S + s)];
            if (fin_buf[(size_t)s] < 0) nmult += 2 * (work - j0_prev[(size_t)s]);
            if (done && fin_buf[(size_t)s] < 0) fin_buf[(size_t)s] = cur;
            if (!done) {
                any = true;
                jmin = std::min(jmin, st_h[(size_t)s]);
                j0_prev[(size_t)s] = st_h[(size_t)s];
            }
        }
        if (!any) break;
        cur ^= 1;
    }
    t_prof = saved;
    if (std::getenv("SEV_VERBOSE"))
        std::fprintf(stderr, "[severo_hip] batched irlba: %lld samples, %lld rows x %lld, nu=%d "
                     "work=%d: %d rounds, %lld products, %zu graphs\n", (long long)S,
                     (long long)M, (long long)n, nu, work, nround, (long long)nmult,
                     w.graphs.size());
    for (int64_t s = 0; s < S; ++s) {
        const int f = st_h[(size_t)(3 * S + s)];
        if (iters_out) iters_out[s] = st_h[(size_t)(2 * S + s)];
        if (f) {
            char msg[160];
            std::snprintf(msg, sizeof msg, "batched irlba: sample %lld %s", (long long)s,
                          f == 2 ? "did not converge (maxit reached)"
                                 : f == 1 ? "hit an invariant subspace (rank-deficient operator)"
                                          : f == 3 ? "had a zero residual at restart"
                                                   : "needs a singular-vector completion");
            fail(kErrConvergence, msg);
        }
    }
    if (nmult_out) *nmult_out = nmult;
    // final vectors: U = W BU[:, 0:nu] and V = V BV[:, 0:nu] from the buffer
    // that held each sample when it finished
    DevBuf<int> sel;
    sel.ensure((size_t)S);
    SEV_HIP(hipMemcpyAsync(sel.get(), fin_buf.data(), sizeof(int) * S, hipMemcpyHostToDevice, st));
    {
        ProfScope ps("batch_final", st);
        hipLaunchKernelGGL(k_b_gemm<TW>, dim3(grid_rows(M)), dim3(kBlock), 0, st, d, 0,
                           (const TW *)Wb[0].get(), (const TW *)Wb[1].get(),
                           (const int *)sel.get(), U, M, w.BU.get(), nu, (const double *)nullptr,
                           w.scal.get(), false);
        SEV_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_b_gemm<double>, dim3(grid_rows(S * n)), dim3(kBlock), 0, st, d, 1,
                           (const double *)w.V[0].get(), (const double *)w.V[1].get(),
                           (const int *)sel.get(), V, (int64_t)nu, w.BV.get(), nu,
                           (const double *)nullptr, w.scal.get(), false);
        SEV_LAUNCH_CHECK();
    }
    std::vector<double> bs((size_t)(S * kMaxWork));
    SEV_HIP(hipMemcpyAsync(bs.data(), w.BS.get(), sizeof(double) * bs.size(),
                           hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    for (int64_t s = 0; s < S; ++s)
        for (int i = 0; i < nu; ++i) s_out[s * nu + i] = bs[(size_t)(s * kMaxWork + i)];
}
}  // namespace

void irlba_batch(const SlabCsc &op, const BatchLayout &L, const float *mu_all,
                 const IrlbaParams &p, const double *init, BatchIrlbaWork &w, double *s_out,
                 double *U, double *V, int64_t *iters_out, int64_t *nmult_out, bool graph,
                 hipStream_t st) {
    if (p.basis32) {
        w.W[0].release();
        w.W[1].release();
        irlba_batch_run<float>(op, L, mu_all, p, init, w, w.Wf, s_out, U, V, iters_out,
                               nmult_out, graph, st);
    } else {
        w.Wf[0].release();
        w.Wf[1].release();
        w.ycur.release();
        w.wprev.release();
        irlba_batch_run<double>(op, L, mu_all, p, init, w, w.W, s_out, U, V, iters_out,
                                nmult_out, graph, st);
    }
}

void batch_embedding(const double *U, const double *s_dev, const BatchLayout &L, int64_t dd,
                     float *Z, int64_t rs, int64_t cs, hipStream_t st) {
    hipLaunchKernelGGL(k_b_embed, dim3(grid_rows(L.M)), dim3(kBlock), 0, st, U, s_dev,
                       L.row_sample_d.get(), L.M, dd, Z, rs, cs);
    SEV_LAUNCH_CHECK();
}

}  // namespace sev
