// Exact Euclidean k-nearest neighbours (brute force, bit-exact definition).
//
// Replaces the native FindNeighboursEuclidean32/64 called from
// src/neighbours.jl:32-63 (which the reference backs with an approximate
// multi-table search).  Here the neighbours are exact under a stated order of
// operations, so every implementation of the definition agrees bit for bit:
//   D(i,j) = sum_{t=0}^{d-1} (x_it - x_jt)^2, each difference and square
//            rounded in T, accumulated in t order, no FMA;
//   ranking by (D, j) lexicographically (lower index wins ties, the order of
//   the reference test's partialsortperm, test/test_nn.jl:33);
//   j == i skipped only when include_self == 0; distance = sqrt(D).
//
// One block handles kQB queries; every thread owns one candidate of a
// 256-candidate tile and computes its distance to all kQB queries (query
// coordinates broadcast from LDS, candidate coordinates read coalesced from
// the column-major X).  Candidates below a query's current k-th distance are
// appended to that query's LDS pending list; after each tile one wavefront
// per query merges the pending entries into a sorted list held one entry per
// lane (insertion by ballot + shift).
#include "sev_kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>

namespace sev {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

template <class T>
__device__ __forceinline__ T inf_val();
template <>
__device__ __forceinline__ float inf_val<float>() {
    return __builtin_huge_valf();
}
template <>
__device__ __forceinline__ double inf_val<double>() {
    return __builtin_huge_val();
}

template <class T, int kQB>
__global__ void __launch_bounds__(kBlock)
k_knn_exact(const T *X, int64_t n, int64_t d, int64_t ld, int k, int include_self,
            const int64_t *qidx, int64_t nq, int scatter, int32_t *oidx, T *odist,
            const int64_t *seg, const int32_t *row_sample) {
    extern __shared__ unsigned char smem_raw[];
    T *qx = (T *)smem_raw;  // kQB * d
    __shared__ T pendD[kQB][kBlock];
    __shared__ int32_t pendJ[kQB][kBlock];
    __shared__ int pendN[kQB];
    __shared__ T listD[kQB][kMaxK];
    __shared__ int32_t listJ[kQB][kMaxK];
    __shared__ int listN[kQB];
    __shared__ T thr[kQB];
    __shared__ int64_t qrow[kQB];

    const int64_t q0 = (int64_t)blockIdx.x * kQB;
    const int nqb = (int)((nq - q0) < kQB ? (nq - q0) : kQB);
    if (threadIdx.x < kQB) {
        const int q = threadIdx.x;
        qrow[q] = q < nqb ? (qidx ? qidx[q0 + q] : q0 + q) : -1;
        pendN[q] = 0;
        listN[q] = 0;
        thr[q] = inf_val<T>();
    }
    __syncthreads();
    for (int64_t e = threadIdx.x; e < (int64_t)kQB * d; e += kBlock) {
        const int q = (int)(e / d);
        const int64_t t = e % d;
        qx[e] = q < nqb ? X[t * ld + qrow[q]] : T(0);
    }
    __syncthreads();

    // candidates: all n points, or (batches) the rows of the sample of the
    // block's first query -- every query of a block is from that sample
    int64_t cb = 0, ce = n;
    if (seg) {
        const int32_t s = row_sample[qrow[0]];
        cb = seg[s];
        ce = seg[s + 1];
    }
    const int lane = lane_id(), w = threadIdx.x / kWave;
    for (int64_t tile = cb; tile < ce; tile += kBlock) {
        const int64_t j = tile + threadIdx.x;
        if (j < ce) {
            T D[kQB];
#pragma unroll
            for (int q = 0; q < kQB; ++q) D[q] = T(0);
            for (int64_t t = 0; t < d; ++t) {
                const T xj = X[t * ld + j];
#pragma unroll
                for (int q = 0; q < kQB; ++q) {
                    const T diff = qx[q * d + t] - xj;
                    D[q] = D[q] + diff * diff;
                }
            }
#pragma unroll
            for (int q = 0; q < kQB; ++q) {
                if (q < nqb && D[q] < thr[q] && (include_self || j != qrow[q])) {
                    const int pos = atomicAdd(&pendN[q], 1);
                    pendD[q][pos] = D[q];
                    pendJ[q][pos] = (int32_t)j;
                }
            }
        }
        __syncthreads();
        for (int q = w; q < nqb; q += kWaves) {
            const int P = pendN[q];
            if (P == 0) continue;
            int cnt = listN[q];
            T ld_ = lane < cnt ? listD[q][lane] : inf_val<T>();
            int32_t lj = lane < cnt ? listJ[q][lane] : INT_MAX;
            for (int e = 0; e < P; ++e) {
                const T Dn = pendD[q][e];
                const int32_t jn = pendJ[q][e];
                const bool less = lane < cnt && (ld_ < Dn || (ld_ == Dn && lj < jn));
                const int pos = __popcll(__ballot(less));
                const T upD = __shfl_up(ld_, 1, kWave);
                const int32_t upJ = __shfl_up(lj, 1, kWave);
                if (pos < k) {
                    if (lane == pos) {
                        ld_ = Dn;
                        lj = jn;
                    } else if (lane > pos) {
                        ld_ = upD;
                        lj = upJ;
                    }
                    cnt = cnt + 1 < k ? cnt + 1 : k;
                }
            }
            if (lane < kMaxK) {
                listD[q][lane] = ld_;
                listJ[q][lane] = lj;
            }
            const T kth = __shfl(ld_, k - 1, kWave);
            if (lane == 0) {
                listN[q] = cnt;
                thr[q] = cnt == k ? kth : inf_val<T>();
                pendN[q] = 0;
            }
        }
        __syncthreads();
    }
    for (int q = w; q < nqb; q += kWaves) {
        const int cnt = listN[q];
        if (lane < k) {
            const int64_t r = scatter ? qrow[q] : q0 + q;
            oidx[r * k + lane] = lane < cnt ? listJ[q][lane] : -1;
            odist[r * k + lane] = lane < cnt ? sqrt(listD[q][lane]) : inf_val<T>();
        }
    }
}

}  // namespace

template <class T>
void knn_exact_impl(const T *X, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                    const int64_t *qidx, int64_t nq, int scatter, int32_t *idx, T *dist,
                    hipStream_t st, const int64_t *seg = nullptr,
                    const int32_t *row_sample = nullptr) {
    SEV_REQUIRE(k >= 1 && k <= kMaxK, "knn: k must be in [1, 64]");
    SEV_REQUIRE(d >= 1, "knn: d must be >= 1");
    // 16 queries per block in FP32, 8 in FP64 (keeps LDS under 64 KB)
    constexpr int QB = sizeof(T) == 4 ? 16 : 8;
    const size_t smem = sizeof(T) * (size_t)QB * (size_t)d;
    SEV_REQUIRE(smem <= 16384, "knn: too many dimensions for the exact kernel");
    if (nq == 0) return;
    ProfScope ps("knn_exact", st, 0.0, 2.0 * (double)nq * (double)n * (double)d);
    hipLaunchKernelGGL((k_knn_exact<T, QB>), dim3((unsigned)ceil_div(nq, QB)), dim3(kBlock), smem,
                       st, X, n, d, ld, (int)k, include_self, qidx, nq, scatter, idx, dist, seg,
                       row_sample);
    SEV_LAUNCH_CHECK();
}

template <class T>
void knn_exact(const T *X, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
               const int64_t *qidx, int64_t nq, int32_t *idx, T *dist, hipStream_t st) {
    knn_exact_impl<T>(X, n, d, ld, k, include_self, qidx, nq, 0, idx, dist, st);
}

template <class T>
void knn_exact_rows(const T *X, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                    const int64_t *qidx, int64_t nq, int32_t *idx, T *dist, hipStream_t st) {
    SEV_REQUIRE(qidx != nullptr, "knn_exact_rows needs row indices");
    knn_exact_impl<T>(X, n, d, ld, k, include_self, qidx, nq, 1, idx, dist, st);
}

void knn_exact_rows_batch(const float *X, int64_t n, int64_t d, int64_t ld, int64_t k,
                          int include_self, const int64_t *qidx, int64_t nq, const int64_t *seg,
                          const int32_t *row_sample, int32_t *idx, float *dist, hipStream_t st) {
    SEV_REQUIRE(qidx != nullptr && seg != nullptr, "knn_exact_rows_batch needs rows and samples");
    knn_exact_impl<float>(X, n, d, ld, k, include_self, qidx, nq, 1, idx, dist, st, seg,
                          row_sample);
}

// --------------------------------------------------------------------------
// Cosine neighbours (the CosineDist form of src/neighbours.jl:41-50 and
// :63-72).  Each row is scaled to unit length in T (sum of squares in t
// order, IEEE sqrt and division; an all-zero row stays zero); the exact
// Euclidean search then runs on the unit rows, whose D ranks pairs as
// 1 - cos does in exact arithmetic; the reported distance is D / 2, with D
// recomputed for the found pairs by the same definition.
// --------------------------------------------------------------------------
namespace {

template <class T>
__global__ void __launch_bounds__(kBlock)
k_unit_rows(T *X, int64_t n, int64_t d) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        T s = 0;
        for (int64_t t = 0; t < d; ++t) {
            const T v = X[t * n + i];
            s = s + v * v;
        }
        if (s > T(0)) {
            const T nrm = sqrt(s);
            for (int64_t t = 0; t < d; ++t) X[t * n + i] = X[t * n + i] / nrm;
        }
    }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_half_pair_dist(const T *X, int64_t n, int64_t d, const int32_t *idx, int64_t k, T *dist) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n * k;
         p += (int64_t)gridDim.x * kBlock) {
        const int64_t i = p / k;
        const int32_t j = idx[p];
        if (j < 0 || j >= n) continue;
        T D = 0;
        for (int64_t t = 0; t < d; ++t) {
            const T diff = X[t * n + i] - X[t * n + j];
            D = D + diff * diff;
        }
        dist[p] = T(0.5) * D;
    }
}

unsigned rows_grid(int64_t n) {
    const int64_t g = ceil_div(n > 0 ? n : 1, (int64_t)kBlock);
    return (unsigned)std::min<int64_t>(g, (int64_t)device_cu_count() * 16);
}

}  // namespace

template <class T>
void knn_unit_rows(T *X, int64_t n, int64_t d, hipStream_t st) {
    hipLaunchKernelGGL(k_unit_rows<T>, dim3(rows_grid(n)), dim3(kBlock), 0, st, X, n, d);
    SEV_LAUNCH_CHECK();
}

template <class T>
void knn_half_pair_dist(const T *X, int64_t n, int64_t d, const int32_t *idx, int64_t k, T *dist,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_half_pair_dist<T>, dim3(rows_grid(n * k)), dim3(kBlock), 0, st, X, n, d,
                       idx, k, dist);
    SEV_LAUNCH_CHECK();
}

template void knn_unit_rows<float>(float *, int64_t, int64_t, hipStream_t);
template void knn_unit_rows<double>(double *, int64_t, int64_t, hipStream_t);
template void knn_half_pair_dist<float>(const float *, int64_t, int64_t, const int32_t *, int64_t,
                                        float *, hipStream_t);
template void knn_half_pair_dist<double>(const double *, int64_t, int64_t, const int32_t *,
                                         int64_t, double *, hipStream_t);

template void knn_exact<float>(const float *, int64_t, int64_t, int64_t, int64_t, int,
                               const int64_t *, int64_t, int32_t *, float *, hipStream_t);
template void knn_exact<double>(const double *, int64_t, int64_t, int64_t, int64_t, int,
                                const int64_t *, int64_t, int32_t *, double *, hipStream_t);
template void knn_exact_rows<float>(const float *, int64_t, int64_t, int64_t, int64_t, int,
                                    const int64_t *, int64_t, int32_t *, float *, hipStream_t);
template void knn_exact_rows<double>(const double *, int64_t, int64_t, int64_t, int64_t, int,
                                     const int64_t *, int64_t, int32_t *, double *, hipStream_t);

}  // namespace sev
