// Input-side kernels of the path: synthetic counts, normalize_cells,
// feature gather and scale_features.
//
// normalize_cells:  src/normalize.jl:17-38 (row_norm, log_norm)
// feature gather:   Y[:, hvf] (docs/src/pbmc.md:121, src/scaling.jl:350-352)
// scale_features:   src/scaling.jl:18-34 (mean_var), :199-217 (scale_data)
//
// Built with -ffp-contract=off: every FP operation rounds where the Julia
// source rounds (no fused multiply-add), and '/' and sqrt are the IEEE
// correctly rounded operations.
#include "sev_kernels.hpp"

namespace sev {

namespace {

constexpr int kBlock = 256;

// --------------------------------------------------------------------------
// Synthetic negative-binomial counts.  Entry (g, c): u = u01(seed, g, c);
// zero iff u < p0; otherwise inverse CDF with the recurrence
// P(x+1) = P(x) * q * (x + r) / (x + 1).
// --------------------------------------------------------------------------
constexpr uint32_t kClassStream = 0x7FFFFFFFu;
constexpr int32_t kNbXmax = 100000;

__device__ __forceinline__ int32_t nb_draw(double u, double p0, double q, double r) {
    if (u < p0) return 0;
    double p = p0, cdf = p0;
    int32_t x = 0;
    while (u >= cdf && x < kNbXmax) {
        p = ((p * q) * ((double)x + r)) / ((double)x + 1.0);
        x += 1;
        cdf = cdf + p;
    }
    return x;
}

__global__ void __launch_bounds__(kBlock)
k_synth_class(uint8_t *cls, int64_t C, uint64_t seedm, double frac) {
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < C;
         c += (int64_t)gridDim.x * kBlock)
        cls[c] = u01(seedm, kClassStream, (uint32_t)c) < frac ? 1 : 0;
}

// Block (chunk, gene): number of nonzeros among the chunk's cells.
__global__ void __launch_bounds__(kBlock)
k_synth_count(int64_t C, int64_t nchunk, int64_t chunk, uint64_t seedm, const double *p0,
              const uint8_t *cls, int64_t *cnt) {
    const int64_t g = blockIdx.x / nchunk, ch = blockIdx.x % nchunk;
    const int64_t c0 = ch * chunk;
    const int64_t c1 = c0 + chunk < C ? c0 + chunk : C;
    const double pa = p0[2 * g], pb = p0[2 * g + 1];
    int64_t n = 0;
    for (int64_t c = c0 + threadIdx.x; c < c1; c += kBlock) {
        double u = u01(seedm, (uint32_t)g, (uint32_t)c);
        n += (u >= (cls[c] ? pb : pa)) ? 1 : 0;
    }
    __shared__ int64_t red[kBlock / kWave];
    n = wave_sum_i64(n);
    if (lane_id() == 0) red[threadIdx.x / kWave] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w];
        cnt[blockIdx.x] = s;
    }
}

// Block (chunk, gene): write rows (ascending) and counts of the nonzeros.
__global__ void __launch_bounds__(kBlock)
k_synth_fill(int64_t C, int64_t nchunk, int64_t chunk, uint64_t seedm, double r,
             const double *p0, const double *q, const uint8_t *cls, const int64_t *off,
             int32_t *rowval, int32_t *nzval) {
    const int64_t g = blockIdx.x / nchunk, ch = blockIdx.x % nchunk;
    const int64_t c0 = ch * chunk;
    const int64_t c1 = c0 + chunk < C ? c0 + chunk : C;
    __shared__ int32_t wcnt[kBlock / kWave];
    int64_t pos = off[blockIdx.x];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    for (int64_t cb = c0; cb < c1; cb += kBlock) {
        const int64_t c = cb + threadIdx.x;
        int32_t x = 0;
        if (c < c1) {
            const int k = cls[c];
            double u = u01(seedm, (uint32_t)g, (uint32_t)c);
            x = nb_draw(u, p0[2 * g + k], q[2 * g + k], r);
        }
        const uint64_t m = __ballot(x != 0);
        if (lane == 0) wcnt[w] = __popcll(m);
        __syncthreads();
        int32_t before = 0, total = 0;
        for (int i = 0; i < kBlock / kWave; ++i) {
            before += i < w ? wcnt[i] : 0;
            total += wcnt[i];
        }
        if (x != 0) {
            const uint64_t lower = lane ? (m & ((~0ull) >> (kWave - lane))) : 0ull;
            const int64_t p = pos + before + __popcll(lower);
            rowval[p] = (int32_t)c;
            nzval[p] = x;
        }
        pos += total;
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// normalize_cells
// --------------------------------------------------------------------------

// s = sum(A, dims=2): integer row (cell) totals over the CSC nonzeros (or
// the number of stored entries per row when cnt is null).  Block (x, y)
// sums the entries of columns [x*cpb, ...) whose rows fall in
// [y*kTotRows, ...) into an LDS histogram (one lane per column finds its
// segment by binary search in the column's sorted rows; the wave then walks
// the 64 segments, four strides of loads in flight at a time), and adds the
// histogram to the totals with one global atomic per nonzero row.  The
// launch picks cpb (columns per block).
constexpr int kTotThreads = 1024;
constexpr int kTotWaves = kTotThreads / kWave;
constexpr int64_t kTotRows = 16384;  // rows per histogram (64 KiB of LDS)
constexpr int64_t kTotBlocks = 256;

__device__ __forceinline__ int64_t lower_row(const int32_t *row, int64_t lo, int64_t hi,
                                             int64_t key) {
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (row[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kTotThreads)
k_row_totals(int64_t m, int64_t n, int64_t cpb, const int64_t *colptr, const int32_t *row,
             const int32_t *cnt, uint32_t *tot) {
    __shared__ uint32_t h[kTotRows];
    const int64_t r0 = (int64_t)blockIdx.y * kTotRows;
    const int64_t r1 = r0 + kTotRows < m ? r0 + kTotRows : m;
    const int64_t c0 = (int64_t)blockIdx.x * cpb;
    const int64_t c1 = c0 + cpb < n ? c0 + cpb : n;
    for (int64_t i = threadIdx.x; i < r1 - r0; i += kTotThreads) h[i] = 0;
    __syncthreads();
    const int w = threadIdx.x / kWave, lane = lane_id();
    for (int64_t cb = c0 + (int64_t)w * kWave; cb < c1; cb += (int64_t)kTotWaves * kWave) {
        const int64_t c = cb + lane;
        int64_t a = 0, b = 0;
        if (c < c1) {
            const int64_t lo = colptr[c], hi = colptr[c + 1];
            a = lower_row(row, lo, hi, r0);
            b = lower_row(row, a, hi, r1);
        }
        const int nc = (int)(c1 - cb < kWave ? c1 - cb : kWave);
        for (int t = 0; t < nc; ++t) {
            const int64_t sa = __shfl(a, t, kWave), sb = __shfl(b, t, kWave);
            for (int64_t p = sa + lane; p < sb; p += 4 * kWave) {
                // clamped indices: every load is issued, extra ones add nothing
                int32_t rr[4], cc[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t q = p + u * kWave < sb ? p + u * kWave : sb - 1;
                    rr[u] = row[q];
                    cc[u] = cnt ? cnt[q] : 1;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint64_t lr = (uint64_t)((int64_t)rr[u] - r0);
                    if (p + u * kWave < sb && lr < (uint64_t)kTotRows)
                        atomicAdd(&h[lr], (uint32_t)cc[u]);
                }
            }
        }
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < r1 - r0; i += kTotThreads)
        if (h[i]) atomicAdd(&tot[r0 + i], h[i]);
}

// bad = 1 if the rows of some column are not strictly increasing (one
// wavefront per column)
__global__ void __launch_bounds__(kBlock)
k_check_sorted(int64_t n, const int64_t *colptr, const int32_t *row, int32_t *bad) {
    const int64_t j = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (j >= n) return;
    const int lane = lane_id();
    const int64_t b = colptr[j], e = colptr[j + 1];
    bool ok = true;
    for (int64_t p = b + 1 + lane; p < e; p += kWave) ok = ok && row[p] > row[p - 1];
    if (__ballot(!ok) && lane == 0) atomicOr(bad, 1);
}

// nzB = sf * nzA / s[row] in Float32, left to right; then log1p, defined as
// the double-precision log1p rounded to float.
__global__ void __launch_bounds__(kBlock)
k_lognorm_f32(const int32_t *row, const int32_t *cnt, const uint32_t *tot, int64_t nnz,
              float sf, int dolog, float *out) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * kBlock) {
        const float v = (sf * (float)cnt[p]) / (float)tot[row[p]];
        out[p] = dolog ? (float)log1p((double)v) : v;
    }
}

__global__ void __launch_bounds__(kBlock)
k_lognorm_f64(const int32_t *row, const int32_t *cnt, const uint32_t *tot, int64_t nnz,
              double sf, int dolog, double *out) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * kBlock) {
        const double v = (sf * (double)cnt[p]) / (double)tot[row[p]];
        out[p] = dolog ? log1p(v) : v;
    }
}

// --------------------------------------------------------------------------
// Feature gather
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_gather_len(const int64_t *feat, const int64_t *colptr, int64_t nfeat, int64_t *len) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j < nfeat) {
        const int64_t g = feat[j];
        len[j] = colptr[g + 1] - colptr[g];
    }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_gather_copy(const int64_t *feat, const int64_t *colptr, const int32_t *row, const T *val,
              const int64_t *fcolptr, int32_t *orow, T *oval) {
    // column blockIdx.x, shared by the gridDim.y blocks of that column, four
    // loads of a thread in flight at once
    const int64_t j = blockIdx.x;
    const int64_t g = feat[j];
    const int64_t b = colptr[g], len = colptr[g + 1] - b, ob = fcolptr[j];
    const int64_t step = (int64_t)gridDim.y * kBlock;
    int64_t i = (int64_t)blockIdx.y * kBlock + threadIdx.x;
    for (; i + 3 * step < len; i += 4 * step) {
        int32_t r[4];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            r[u] = row[b + i + u * step];
            v[u] = val[b + i + u * step];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            orow[ob + i + u * step] = r[u];
            oval[ob + i + u * step] = v[u];
        }
    }
    for (; i < len; i += step) {
        orow[ob + i] = row[b + i];
        oval[ob + i] = val[b + i];
    }
}

// --------------------------------------------------------------------------
// scale_data.  One wavefront per column runs the reference's sequential
// Welford recurrence (src/scaling.jl:21-30) in T, every lane computing the
// same values:
//   count = m - nnz; for v: count += 1; delta = v - mu;
//   mu += delta / T(count); s += delta * (v - mu)
// var = s / T(m - 1); sigma = sqrt(var); mu' = mu / sigma;
// clip = scale_max + mu'.
// Each chunk of 64 values is loaded with one coalesced read (one per lane)
// and consumed in order through readlane; the 64 reciprocals of the chunk's
// counts are computed in parallel by the lanes.
// --------------------------------------------------------------------------
__device__ __forceinline__ float lane_value(float x, int t) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), t));
}
__device__ __forceinline__ double lane_value(double x, int t) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, t);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), t);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// delta / T(count) rounded to T (the reference divides by count converted
// to T): the FP64 form divides.
__device__ __forceinline__ double welford_quot(double delta, double, double countf, bool) {
    return delta / countf;
}

// The FP32 quotient RN(delta/F), F = Float32(count), by Markstein's
// correction step (P. Markstein, IBM J. Res. Dev. 34(1), 1990): with
// y = RN(1/F) and q = RN(delta*y) within one ulp of delta/F, r = delta - q*F
// is exact (one fma) and RN(q + r*y) is RN(delta/F), the IEEE FP32
// division's result.  y = RN32(RN64(1/F)) is RN32(1/F): for an
// integer-valued F, 1/F is at least 2^-49 (relative) from every FP32
// rounding boundary, so the second rounding cannot move it.  Every operation
// of the recurrence's chain is FP32; deltas below 2^-90 in magnitude, whose
// products could leave the normal range, take the exact division instead.
__device__ __forceinline__ float welford_quot32(float delta, float F, float y) {
    const float q = delta * y;
    const float r = fmaf(-q, F, delta);
    return fmaf(r, y, q);
}
__device__ __forceinline__ bool welford_tiny(float delta) {
    return fabsf(delta) < 0x1p-90f && delta != 0.0f;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_scale_stats(int64_t m, int64_t ncol, const int64_t *colptr, const T *x, T scale_max, T *sdv,
              T *mupv, T *clipv, const int64_t *row0, int64_t S) {
    const int64_t col = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (col >= ncol) return;
    const int lane = lane_id();
    if (row0) {
        // column (j, s) of a stacked batch: the rows of sample s = col % S
        const int64_t s = col % S;
        m = row0[s + 1] - row0[s];
    }
    const int64_t b = colptr[col], e = colptr[col + 1];
    int64_t count = m - (e - b);
    const bool fast = sizeof(T) == 4;
    T mu = 0, s = 0;
    // the next chunk's values are loaded while the recurrence runs over the
    // current one (indices clamped into the column, so every load is issued
    // and only the current chunk's is waited for; lanes past the end read 0)
    T vraw = b < e ? x[b + lane < e ? b + lane : e - 1] : T(0);
    for (int64_t base = b; base < e; base += kWave) {
        const int64_t pn = base + kWave + lane;
        const T vnext = x[pn < e ? pn : e - 1];
        const T vl = base + lane < e ? vraw : T(0);
        vraw = vnext;
        // lane l: F = Float32(count + l + 1) and RN32(1/F) for element l of the chunk
        const float fl = fast ? (float)(count + lane + 1) : 0.0f;
        const float yl = fast ? (float)(1.0 / (double)fl) : 0.0f;
        const int nt = (int)((e - base) < kWave ? (e - base) : kWave);
        if constexpr (sizeof(T) == 4) {
            // every lane holds the same mu and delta, so the rare exact
            // division (tiny nonzero deltas) is a wave-uniform scalar branch,
            // and full chunks run without per-element bounds checks
            auto step = [&](int t) {
                const float v = lane_value(vl, t);
                const float delta = v - mu;
                float q = welford_quot32(delta, lane_value(fl, t), lane_value(yl, t));
                if (__builtin_expect(__builtin_amdgcn_readfirstlane((int)welford_tiny(delta)), 0))
                    q = delta / (float)(count + t + 1);
                mu = mu + q;
                s = s + delta * (v - mu);
            };
            if (nt == kWave) {
                // full chunk without the per-element branch: tiny deltas are
                // only flagged (off the recurrence chain), and a flagged
                // chunk is redone with the exact division
                const float mu0 = mu, s0 = s;
                bool tiny = false;
#pragma unroll
                for (int t = 0; t < kWave; ++t) {
                    const float v = lane_value(vl, t);
                    const float delta = v - mu;
                    tiny = tiny || welford_tiny(delta);
                    mu = mu + welford_quot32(delta, lane_value(fl, t), lane_value(yl, t));
                    s = s + delta * (v - mu);
                }
                if (__builtin_expect(__builtin_amdgcn_readfirstlane((int)tiny), 0)) {
                    mu = mu0;
                    s = s0;
                    for (int t = 0; t < kWave; ++t) step(t);
                }
            } else {
                for (int t = 0; t < nt; ++t) step(t);
            }
        } else {
#pragma unroll
            for (int t = 0; t < kWave; ++t) {
                if (t < nt) {
                    const T v = lane_value(vl, t);
                    const T delta = v - mu;
                    mu = mu + welford_quot(delta, 0.0, (T)(count + t + 1), false);
                    s = s + delta * (v - mu);
                }
            }
        }
        count += nt;
    }
    const T var = s / (T)(m - 1);
    const T sd = sqrt(var);
    const T mup = mu / sd;
    if (lane == 0) {
        sdv[col] = sd;
        mupv[col] = mup;
        clipv[col] = scale_max + mup;
    }
}

// The same statistics for many short columns (the stacked batch, column
// (j, s) holding feature j of sample s = col % S): one lane per column, each
// lane running its column's recurrence in the same operations and order as
// k_scale_stats (Markstein's FP32 quotient, exact division for the tiny
// deltas), so the results are the same bits.
__global__ void __launch_bounds__(kBlock)
k_scale_stats_lane(int64_t ncol, const int64_t *colptr, const float *x, float scale_max,
                   float *sdv, float *mupv, float *clipv, const int64_t *row0, int64_t S) {
    const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (col >= ncol) return;
    const int64_t smp = col % S;
    const int64_t m = row0[smp + 1] - row0[smp];
    const int64_t b = colptr[col], e = colptr[col + 1];
    int64_t count = m - (e - b);
    float mu = 0.0f, s = 0.0f;
    auto step = [&](float v) {
        ++count;
        const float F = (float)count;
        const float delta = v - mu;
        float q;
        if (welford_tiny(delta))
            q = delta / F;
        else
            q = welford_quot32(delta, F, (float)(1.0 / (double)F));
        mu = mu + q;
        s = s + delta * (v - mu);
    };
    // groups of 16 loads of the lane's column: the next group is in flight
    // while the recurrence runs over the current one
    constexpr int U = 16;
    int64_t p = b;
    float v[U], vn[U];
    if (p + U <= e) {
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = x[p + u];
    }
    for (; p + U <= e; p += U) {
        // the following group, indices clamped into the column
#pragma unroll
        for (int u = 0; u < U; ++u) vn[u] = x[p + U + u < e ? p + U + u : e - 1];
#pragma unroll
        for (int u = 0; u < U; ++u) step(v[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = vn[u];
    }
    for (; p < e; ++p) step(x[p]);
    const float var = s / (float)(m - 1);
    const float sd = sqrtf(var);
    const float mup = mu / sd;
    sdv[col] = sd;
    mupv[col] = mup;
    clipv[col] = scale_max + mup;
}

// mean_std(Float64, A) of the Int32 counts (src/scaling.jl:18-38,119-130):
// the same one-wavefront-per-column Welford recurrence in FP64 (IEEE
// division by the count), sd = sqrt(s / (m - 1)).
__global__ void __launch_bounds__(kBlock)
k_count_mean_std(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *x, double *muv,
                 double *sdv) {
    const int64_t col = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (col >= ncol) return;
    const int lane = lane_id();
    const int64_t b = colptr[col], e = colptr[col + 1];
    int64_t count = m - (e - b);
    double mu = 0.0, s = 0.0;
    for (int64_t base = b; base < e; base += kWave) {
        const int64_t p = base + lane;
        const double vl = p < e ? (double)x[p] : 0.0;
        const int nt = (int)((e - base) < kWave ? (e - base) : kWave);
        for (int t = 0; t < nt; ++t) {
            const double v = lane_value(vl, t);
            const double delta = v - mu;
            mu = mu + delta / (double)(count + t + 1);
            s = s + delta * (v - mu);
        }
        count += nt;
    }
    if (lane == 0) {
        muv[col] = mu;
        sdv[col] = sqrt(s / (double)(m - 1));
    }
}

// mean_var(Float64, row_norm(A, 1.0)) (src/scaling.jl:18-34,132-142 on
// src/normalize.jl:24-36): per column, FP64 Welford over the relative counts
// x / s_row in row order (1.0 * x is exact; IEEE division), the count
// starting at m - nnz; var = s / (m - 1).  One wavefront per column.
__global__ void __launch_bounds__(kBlock)
k_rel_mean_var(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *row,
               const int32_t *x, const uint32_t *tot, double *muv, double *varv) {
    const int64_t col = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (col >= ncol) return;
    const int lane = lane_id();
    const int64_t b = colptr[col], e = colptr[col + 1];
    int64_t count = m - (e - b);
    double mu = 0.0, s = 0.0;
    for (int64_t base = b; base < e; base += kWave) {
        const int64_t p = base + lane;
        const double vl = p < e ? (double)x[p] / (double)tot[row[p]] : 0.0;
        const int nt = (int)((e - base) < kWave ? (e - base) : kWave);
        for (int t = 0; t < nt; ++t) {
            const double v = lane_value(vl, t);
            const double delta = v - mu;
            mu = mu + delta / (double)(count + t + 1);
            s = s + delta * (v - mu);
        }
        count += nt;
    }
    if (lane == 0) {
        muv[col] = mu;
        varv[col] = s / (double)(m - 1);
    }
}

// b = min(x / sigma, clip) with only the upper clip (src/scaling.jl:210-213).
// Column blockIdx.x, shared by the gridDim.y blocks of that column (the
// densest columns hold most of the entries); four loads of a thread in
// flight at once.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_scale_map(const int64_t *colptr, const T *x, const T *sdv, const T *clipv, T *out) {
    const int64_t j = blockIdx.x;
    const int64_t b = colptr[j], e = colptr[j + 1];
    const T sd = sdv[j], clip = clipv[j];
    const int64_t step = (int64_t)gridDim.y * kBlock;
    int64_t p = b + (int64_t)blockIdx.y * kBlock + threadIdx.x;
    for (; p + 3 * step < e; p += 4 * step) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = x[p + u * step];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const T q = v[u] / sd;
            out[p + u * step] = (q > clip) ? clip : q;
        }
    }
    for (; p < e; p += step) {
        const T v = x[p] / sd;
        out[p] = (v > clip) ? clip : v;
    }
}

unsigned grid_for(int64_t n, int per_cu = 8) {
    int64_t want = ceil_div(n > 0 ? n : 1, kBlock);
    int64_t cap = (int64_t)device_cu_count() * per_cu;
    return (unsigned)(want < cap ? want : cap);
}

}  // namespace

void synth_classes(int64_t C, uint64_t seed, double boost_frac, uint8_t *cls_d, hipStream_t st) {
    hipLaunchKernelGGL(k_synth_class, dim3(grid_for(C)), dim3(kBlock), 0, st, cls_d, C,
                       mix64(seed), boost_frac);
    SEV_LAUNCH_CHECK();
}

int64_t synth_count(int64_t C, int64_t G, uint64_t seed, double r, const double *p0_d,
                    const double *q_d, double boost_frac, uint8_t *cls_d, int64_t *chunk_cnt_d,
                    int64_t *chunk_off_d, int64_t *colptr_d, int64_t *scratch, hipStream_t st,
                    SynthPlan *plan) {
    (void)q_d;
    (void)r;
    plan->cells = C;
    plan->genes = G;
    plan->chunk = 16384;
    plan->nchunk = C > 0 ? ceil_div(C, plan->chunk) : 1;
    SEV_REQUIRE(G * plan->nchunk < (int64_t)0x7FFFFFFF, "synthetic matrix too large");
    synth_classes(C, seed, boost_frac, cls_d, st);
    const int64_t nb = G * plan->nchunk;
    hipLaunchKernelGGL(k_synth_count, dim3((unsigned)nb), dim3(kBlock), 0, st, C, plan->nchunk,
                       plan->chunk, mix64(seed), p0_d, cls_d, chunk_cnt_d);
    SEV_LAUNCH_CHECK();
    scan_exclusive_i64(chunk_cnt_d, chunk_off_d, nb, scratch, st);
    // colptr[g] = chunk_off[g * nchunk]; colptr[G] = total
    int64_t total = 0;
    SEV_HIP(hipMemcpyAsync(&total, chunk_off_d + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SEV_HIP(hipMemcpy2DAsync(colptr_d, sizeof(int64_t), chunk_off_d,
                             sizeof(int64_t) * plan->nchunk, sizeof(int64_t), G,
                             hipMemcpyDeviceToDevice, st));
    SEV_HIP(hipMemcpyAsync(colptr_d + G, chunk_off_d + nb, sizeof(int64_t),
                           hipMemcpyDeviceToDevice, st));
    SEV_HIP(hipStreamSynchronize(st));
    return total;
}

void synth_fill(const SynthPlan &plan, uint64_t seed, double r, const double *p0_d,
                const double *q_d, const uint8_t *cls_d, const int64_t *chunk_off_d,
                int32_t *rowval_d, int32_t *nzval_d, hipStream_t st) {
    const int64_t nb = plan.genes * plan.nchunk;
    if (nb == 0) return;
    hipLaunchKernelGGL(k_synth_fill, dim3((unsigned)nb), dim3(kBlock), 0, st, plan.cells,
                       plan.nchunk, plan.chunk, mix64(seed), r, p0_d, q_d, cls_d, chunk_off_d,
                       rowval_d, nzval_d);
    SEV_LAUNCH_CHECK();
}

// row ranges x column chunks: a chunk is a multiple of 64 columns (one per
// lane), at most one per wavefront of the block, and small enough for about
// kTotBlocks blocks in all
static void launch_row_totals(int64_t m, int64_t n, const int64_t *colptr, const int32_t *row,
                              const int32_t *cnt, uint32_t *tot, hipStream_t st) {
    if (m == 0 || n == 0) return;
    const int64_t nrange = ceil_div(m, kTotRows);
    int64_t cpb = ceil_div(ceil_div(n * nrange, kTotBlocks), (int64_t)kWave) * kWave;
    cpb = std::max<int64_t>(kWave, std::min<int64_t>(cpb, (int64_t)kTotWaves * kWave));
    const dim3 grid((unsigned)ceil_div(n, cpb), (unsigned)nrange);
    hipLaunchKernelGGL(k_row_totals, grid, dim3(kTotThreads), 0, st, m, n, cpb, colptr, row, cnt,
                       tot);
    SEV_LAUNCH_CHECK();
}

void row_totals(const int64_t *colptr, const int32_t *row, const int32_t *cnt, int64_t nnz,
                int64_t ncells, int64_t ngenes, uint32_t *tot, hipStream_t st) {
    SEV_HIP(hipMemsetAsync(tot, 0, sizeof(uint32_t) * (size_t)(ncells > 0 ? ncells : 1), st));
    if (nnz == 0) return;
    ProfScope ps("row_totals", st, 8.0 * nnz + 4.0 * ncells);
    launch_row_totals(ncells, ngenes, colptr, row, cnt, tot, st);
}

__global__ void __launch_bounds__(kBlock) k_zero_u32(uint32_t *a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        a[i] = 0u;
}

void row_totals_fill(const int64_t *colptr, const int32_t *row, const int32_t *cnt, int64_t nnz,
                     int64_t ncells, int64_t ngenes, uint32_t *tot, hipStream_t st) {
    if (ncells > 0) {
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ncells, 4)), dim3(kBlock), 0, st, tot,
                           ncells);
        SEV_LAUNCH_CHECK();
    }
    if (nnz == 0) return;
    launch_row_totals(ncells, ngenes, colptr, row, cnt, tot, st);
}

void row_counts(const int64_t *colptr, const int32_t *row, int64_t nnz, int64_t m, int64_t n,
                int32_t *cnt, hipStream_t st) {
    SEV_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (size_t)(m > 0 ? m : 1), st));
    if (nnz == 0) return;
    launch_row_totals(m, n, colptr, row, nullptr, (uint32_t *)cnt, st);
}

void check_sorted(int64_t n, const int64_t *colptr, const int32_t *row, int32_t *bad,
                  hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_check_sorted, dim3((unsigned)ceil_div(n * kWave, kBlock)), dim3(kBlock),
                       0, st, n, colptr, row, bad);
    SEV_LAUNCH_CHECK();
}

void lognorm_f32(const int32_t *row, const int32_t *cnt, const uint32_t *tot, int64_t nnz,
                 float sf, int dolog, float *out, hipStream_t st) {
    if (nnz == 0) return;
    ProfScope ps("lognorm", st, 12.0 * nnz);
    hipLaunchKernelGGL(k_lognorm_f32, dim3(grid_for(nnz, 16)), dim3(kBlock), 0, st, row, cnt, tot,
                       nnz, sf, dolog, out);
    SEV_LAUNCH_CHECK();
}

void lognorm_f64(const int32_t *row, const int32_t *cnt, const uint32_t *tot, int64_t nnz,
                 double sf, int dolog, double *out, hipStream_t st) {
    if (nnz == 0) return;
    ProfScope ps("lognorm", st, 16.0 * nnz);
    hipLaunchKernelGGL(k_lognorm_f64, dim3(grid_for(nnz, 16)), dim3(kBlock), 0, st, row, cnt, tot,
                       nnz, sf, dolog, out);
    SEV_LAUNCH_CHECK();
}

void gather_lengths(const int64_t *feat, const int64_t *colptr, int64_t nfeat, int64_t *len,
                    hipStream_t st) {
    if (nfeat == 0) return;
    hipLaunchKernelGGL(k_gather_len, dim3((unsigned)ceil_div(nfeat, kBlock)), dim3(kBlock), 0, st,
                       feat, colptr, nfeat, len);
    SEV_LAUNCH_CHECK();
}

template <class T>
void gather_copy(const int64_t *feat, const int64_t *colptr, const int32_t *row, const T *val,
                 int64_t nfeat, const int64_t *fcolptr, int32_t *orow, T *oval, hipStream_t st) {
    if (nfeat == 0) return;
    ProfScope ps("gather", st);
    // a sample's selected columns are few and long: 8 blocks each
    hipLaunchKernelGGL(k_gather_copy<T>, dim3((unsigned)nfeat, 8), dim3(kBlock), 0, st, feat,
                       colptr, row, val, fcolptr, orow, oval);
    SEV_LAUNCH_CHECK();
}

template <class T>
void scale_stats(int64_t m, int64_t ncol, const int64_t *colptr, const T *x, T scale_max, T *sd,
                 T *mup, T *clip, hipStream_t st) {
    if (ncol == 0) return;
    const int64_t threads = ncol * kWave;
    ProfScope ps("scale_stats", st);
    hipLaunchKernelGGL(k_scale_stats<T>, dim3((unsigned)ceil_div(threads, kBlock)), dim3(kBlock),
                       0, st, m, ncol, colptr, x, scale_max, sd, mup, clip,
                       (const int64_t *)nullptr, (int64_t)1);
    SEV_LAUNCH_CHECK();
}

template <class T>
void scale_map(int64_t ncol, const int64_t *colptr, const T *x, const T *sd, const T *clip,
               T *out, hipStream_t st) {
    if (ncol == 0) return;
    ProfScope ps("scale_map", st);
    // a few long columns (one sample's features) take 8 blocks each; the
    // many short columns of a batch one
    const unsigned slices = ncol >= 8192 ? 1u : 8u;
    hipLaunchKernelGGL(k_scale_map<T>, dim3((unsigned)ncol, slices), dim3(kBlock), 0, st, colptr,
                       x, sd, clip, out);
    SEV_LAUNCH_CHECK();
}

template void gather_copy<float>(const int64_t *, const int64_t *, const int32_t *, const float *,
                                 int64_t, const int64_t *, int32_t *, float *, hipStream_t);
template void gather_copy<double>(const int64_t *, const int64_t *, const int32_t *,
                                  const double *, int64_t, const int64_t *, int32_t *, double *,
                                  hipStream_t);
void count_mean_std(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *x, double *mu,
                    double *sd, hipStream_t st) {
    if (ncol == 0) return;
    ProfScope ps("count_stats", st);
    hipLaunchKernelGGL(k_count_mean_std, dim3((unsigned)ceil_div(ncol * kWave, kBlock)),
                       dim3(kBlock), 0, st, m, ncol, colptr, x, mu, sd);
    SEV_LAUNCH_CHECK();
}

void rel_mean_var(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *row,
                  const int32_t *x, const uint32_t *tot, double *mu, double *var, hipStream_t st) {
    if (ncol == 0) return;
    ProfScope ps("rel_stats", st);
    hipLaunchKernelGGL(k_rel_mean_var, dim3((unsigned)ceil_div(ncol * kWave, kBlock)),
                       dim3(kBlock), 0, st, m, ncol, colptr, row, x, tot, mu, var);
    SEV_LAUNCH_CHECK();
}

template void scale_stats<float>(int64_t, int64_t, const int64_t *, const float *, float, float *,
                                 float *, float *, hipStream_t);
template void scale_stats<double>(int64_t, int64_t, const int64_t *, const double *, double,
                                  double *, double *, double *, hipStream_t);
template void scale_map<float>(int64_t, const int64_t *, const float *, const float *,
                               const float *, float *, hipStream_t);
template void scale_map<double>(int64_t, const int64_t *, const double *, const double *,
                                const double *, double *, hipStream_t);

// ---------------------------------------------------------------------------
// HBM copy probe: the calibration kernel for byte counters and the achievable
// copy rate of the device.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

double hbm_copy_probe() {
    const int64_t bytes = (int64_t)1 << 30;
    const int64_t n = bytes / (int64_t)sizeof(double2);
    DevBuf<double2> a, b;
    a.ensure((size_t)n);
    b.ensure((size_t)n);
    hipStream_t st;
    SEV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    SEV_HIP(hipMemsetAsync(a.get(), 0, (size_t)bytes, st));
    hipEvent_t e0, e1;
    SEV_HIP(hipEventCreate(&e0));
    SEV_HIP(hipEventCreate(&e1));
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n, 256), 256 * 32);
    float best = 1e30f;
    for (int r = 0; r < 6; ++r) {
        SEV_HIP(hipEventRecord(e0, st));
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, st, a.get(), b.get(), n);
        SEV_LAUNCH_CHECK();
        SEV_HIP(hipEventRecord(e1, st));
        SEV_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SEV_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0 && ms < best) best = ms;  // launch 0 warms up
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SEV_HIP(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
    return 2.0 * (double)bytes / (best * 1e-3) / 1e9;
}

namespace {
__global__ void __launch_bounds__(kBlock)
k_rows_to_device(const int64_t *in, int64_t len, int64_t m, int32_t *out, int32_t *bad) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < len;
         p += (int64_t)gridDim.x * kBlock) {
        const int64_t r = in[p] - 1;
        if (r < 0 || r >= m) {
            atomicOr(bad, 1);
            out[p] = 0;
        } else {
            out[p] = (int32_t)r;
        }
    }
}
}  // namespace

void rows_to_device(const int64_t *in, int64_t len, int64_t m, int32_t *out, int32_t *bad,
                    hipStream_t st) {
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div(len, kBlock), device_cu_count() * 16);
    hipLaunchKernelGGL(k_rows_to_device, dim3(g), dim3(kBlock), 0, st, in, len, m, out, bad);
    SEV_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Batches: S samples stacked by rows (sample s owns rows row0[s] ..
// row0[s+1]-1 of one cells x genes CSC).  Per-sample stages read the
// segment of a column that belongs to one sample.
// ---------------------------------------------------------------------------
namespace {

// soff[s*G + g] = first position of column g whose row belongs to sample
// >= s, for s = 0..S (rows are sorted within each column)
__global__ void __launch_bounds__(kBlock)
k_sample_offsets(int64_t G, int64_t S, const int32_t *row_sample, const int64_t *colptr,
                 const int32_t *row, int64_t *soff) {
    const int64_t g = blockIdx.x;
    const int64_t b = colptr[g], e = colptr[g + 1];
    if (b == e) {
        for (int64_t s = threadIdx.x; s <= S; s += kBlock) soff[s * G + g] = b;
        return;
    }
    for (int64_t p = b + threadIdx.x; p < e; p += kBlock) {
        const int64_t sc = row_sample[row[p]];
        const int64_t sp = p == b ? -1 : (int64_t)row_sample[row[p - 1]];
        for (int64_t s = sp + 1; s <= sc; ++s) soff[s * G + g] = p;
        if (p == e - 1)
            for (int64_t s = sc + 1; s <= S; ++s) soff[s * G + g] = e;
    }
}

// virtual column v = j*S + s: the segment of gene feat[s*F + j] in sample s
__global__ void __launch_bounds__(kBlock)
k_seg_len(int64_t F, int64_t S, int64_t G, const int64_t *feat, const int64_t *soff,
          int64_t *src_b, int64_t *len) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= F * S) return;
    const int64_t j = v / S, s = v % S;
    const int64_t g = feat[s * F + j];
    const int64_t b = soff[s * G + g], e = soff[(s + 1) * G + g];
    src_b[v] = b;
    len[v] = e - b;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_gather_seg(const int64_t *src_b, const int64_t *vcp, const int32_t *row, const T *val,
             int32_t *orow, T *oval) {
    const int64_t v = blockIdx.x;
    const int64_t b = src_b[v], ob = vcp[v], len = vcp[v + 1] - ob;
    for (int64_t i = threadIdx.x; i < len; i += kBlock) {
        orow[ob + i] = row[b + i];
        oval[ob + i] = val[b + i];
    }
}

// out[j] = in[j*S] for j = 0..n (column pointers of the n real columns from
// those of the n*S virtual ones)
__global__ void __launch_bounds__(kBlock)
k_stride_ptr(const int64_t *in, int64_t n, int64_t S, int64_t *out) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j <= n) out[j] = in[j * S];
}

// mu_all[s*F + j] = mup[j*S + s]
__global__ void __launch_bounds__(kBlock)
k_mu_by_sample(const float *mup, int64_t F, int64_t S, float *mu_all) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= F * S) return;
    const int64_t j = v / S, s = v % S;
    mu_all[s * F + j] = mup[v];
}

// len[g*S + s] = entries of column g in sample s's own CSC
__global__ void __launch_bounds__(kBlock)
k_stack_len(int64_t G, int64_t S, const int64_t *const *cps, int64_t *len) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= G * S) return;
    const int64_t g = i / S, s = i % S;
    len[i] = cps[s][g + 1] - cps[s][g];
}

__global__ void __launch_bounds__(kBlock)
k_stack_copy(int64_t S, int64_t s, const int64_t *cp, const int32_t *row, const int32_t *cnt,
             int64_t roff, const int64_t *starts, int32_t *orow, int32_t *ocnt) {
    const int64_t g = blockIdx.x;
    const int64_t b = cp[g], len = cp[g + 1] - b, ob = starts[g * S + s];
    for (int64_t i = threadIdx.x; i < len; i += kBlock) {
        orow[ob + i] = row[b + i] + (int32_t)roff;
        ocnt[ob + i] = cnt[b + i];
    }
}

}  // namespace

void sample_offsets(int64_t G, int64_t S, const int32_t *row_sample, const int64_t *colptr,
                    const int32_t *row, int64_t *soff, hipStream_t st) {
    hipLaunchKernelGGL(k_sample_offsets, dim3((unsigned)G), dim3(kBlock), 0, st, G, S, row_sample,
                       colptr, row, soff);
    SEV_LAUNCH_CHECK();
}

void batch_seg_lengths(int64_t F, int64_t S, int64_t G, const int64_t *feat, const int64_t *soff,
                       int64_t *src_b, int64_t *len, hipStream_t st) {
    hipLaunchKernelGGL(k_seg_len, dim3((unsigned)ceil_div(F * S, kBlock)), dim3(kBlock), 0, st, F,
                       S, G, feat, soff, src_b, len);
    SEV_LAUNCH_CHECK();
}

void batch_gather(int64_t nv, const int64_t *src_b, const int64_t *vcp, const int32_t *row,
                  const float *val, int32_t *orow, float *oval, hipStream_t st) {
    ProfScope ps("gather", st);
    hipLaunchKernelGGL(k_gather_seg<float>, dim3((unsigned)nv), dim3(kBlock), 0, st, src_b, vcp,
                       row, val, orow, oval);
    SEV_LAUNCH_CHECK();
}

void batch_scale_stats(int64_t ncol, const int64_t *colptr, const float *x, float scale_max,
                       float *sd, float *mup, float *clip, const int64_t *row0, int64_t S,
                       hipStream_t st) {
    ProfScope ps("scale_stats", st);
    hipLaunchKernelGGL(k_scale_stats_lane, dim3((unsigned)ceil_div(ncol, kBlock)), dim3(kBlock),
                       0, st, ncol, colptr, x, scale_max, sd, mup, clip, row0, S);
    SEV_LAUNCH_CHECK();
}

void stride_ptr(const int64_t *in, int64_t n, int64_t S, int64_t *out, hipStream_t st) {
    hipLaunchKernelGGL(k_stride_ptr, dim3((unsigned)ceil_div(n + 1, kBlock)), dim3(kBlock), 0, st,
                       in, n, S, out);
    SEV_LAUNCH_CHECK();
}

void mu_by_sample(const float *mup, int64_t F, int64_t S, float *mu_all, hipStream_t st) {
    hipLaunchKernelGGL(k_mu_by_sample, dim3((unsigned)ceil_div(F * S, kBlock)), dim3(kBlock), 0,
                       st, mup, F, S, mu_all);
    SEV_LAUNCH_CHECK();
}

void stack_lengths(int64_t G, int64_t S, const int64_t *const *cps_d, int64_t *len,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_stack_len, dim3((unsigned)ceil_div(G * S, kBlock)), dim3(kBlock), 0, st,
                       G, S, cps_d, len);
    SEV_LAUNCH_CHECK();
}

void stack_copy(int64_t G, int64_t S, int64_t s, const int64_t *cp, const int32_t *row,
                const int32_t *cnt, int64_t roff, const int64_t *starts, int32_t *orow,
                int32_t *ocnt, hipStream_t st) {
    hipLaunchKernelGGL(k_stack_copy, dim3((unsigned)G), dim3(kBlock), 0, st, S, s, cp, row, cnt,
                       roff, starts, orow, ocnt);
    SEV_LAUNCH_CHECK();
}

}  // namespace sev
