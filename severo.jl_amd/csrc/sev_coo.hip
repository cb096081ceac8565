// Coordinate list -> CSC of the transpose on the device (the sparse(I, J, V)
// and copy(X') of _read_10X, src/input.jl:115 and :207).  The keys
// gene << 32 | cell are radix-sorted with their values, runs of equal keys
// are summed (duplicates combine with +, stored zeros stay), and the column
// pointers are the lower bounds of each gene in the sorted distinct keys, so
// the cells of every gene column come out ascending.
#include "sev_kernels.hpp"

#include <hipcub/hipcub.hpp>

#include <climits>

namespace sev {

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock)
k_coo_split(const uint64_t *ukeys, const int64_t *sums, const int64_t *nruns, int32_t *row,
            int32_t *val, int32_t *bad) {
    const int64_t nu = *nruns;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nu;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t s = sums[e];
        if (s < INT_MIN || s > INT_MAX) *bad = 1;
        row[e] = (int32_t)(ukeys[e] & 0xFFFFFFFFull);
        val[e] = (int32_t)s;
    }
}

// colptr[g] = number of distinct keys with gene < g
__global__ void __launch_bounds__(kBlock)
k_coo_colptr(const uint64_t *ukeys, const int64_t *nruns, int64_t ngenes, int64_t *colptr) {
    const int64_t nu = *nruns;
    for (int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x; g <= ngenes;
         g += (int64_t)gridDim.x * kBlock) {
        const uint64_t key = (uint64_t)g << 32;
        int64_t lo = 0, hi = nu;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (ukeys[mid] < key)
                lo = mid + 1;
            else
                hi = mid;
        }
        colptr[g] = lo;
    }
}

unsigned grid_for(int64_t n) {
    const int64_t want = (n + kBlock - 1) / kBlock;
    const int64_t cap = (int64_t)device_cu_count() * 8;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, cap));
}

}  // namespace

void coo_to_csc(int64_t ngenes, uint64_t *keys, int64_t *vals, int64_t n, DevBuf<int64_t> &cp,
                DevBuf<int32_t> &row, DevBuf<int32_t> &val, int64_t &nnz, hipStream_t st) {
    SEV_REQUIRE(n >= 0 && n < (int64_t)INT_MAX, "coordinate list too long (2^31 - 1 entries)");
    SEV_REQUIRE(ngenes >= 0 && ngenes < ((int64_t)1 << 31), "too many genes");
    cp.ensure((size_t)ngenes + 1);
    DevBuf<int64_t> nruns;
    nruns.ensure(1);
    SEV_HIP(hipMemsetAsync(nruns.get(), 0, sizeof(int64_t), st));
    DevBuf<uint64_t> k2, uk;
    DevBuf<int64_t> v2, sums;
    DevBuf<unsigned char> tmp;
    if (n > 0) {
        k2.ensure((size_t)n);
        v2.ensure((size_t)n);
        int end_bit = 32;
        while (end_bit < 64 && ((uint64_t)ngenes >> (end_bit - 32)) != 0) ++end_bit;
        size_t tb = 0;
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, k2.get(), vals, v2.get(),
                                                   (int)n, 0, end_bit, st));
        tmp.ensure(tb);
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tb, keys, k2.get(), vals, v2.get(),
                                                   (int)n, 0, end_bit, st));
        // the inputs are free now: distinct keys and sums go there
        uint64_t *ukeys = keys;
        int64_t *usums = vals;
        size_t tb2 = 0;
        SEV_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, tb2, k2.get(), ukeys, v2.get(), usums,
                                                  nruns.get(), hipcub::Sum(), (int)n, st));
        tmp.ensure(tb2);
        SEV_HIP(hipcub::DeviceReduce::ReduceByKey(tmp.get(), tb2, k2.get(), ukeys, v2.get(),
                                                  usums, nruns.get(), hipcub::Sum(), (int)n, st));
    }
    int64_t nu = 0;
    SEV_HIP(hipMemcpyAsync(&nu, nruns.get(), sizeof nu, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    row.ensure((size_t)nu);
    val.ensure((size_t)nu);
    DevBuf<int32_t> bad;
    bad.ensure(1);
    SEV_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
    if (nu > 0) {
        hipLaunchKernelGGL(k_coo_split, dim3(grid_for(nu)), dim3(kBlock), 0, st, keys, vals,
                           nruns.get(), row.get(), val.get(), bad.get());
        SEV_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_coo_colptr, dim3(grid_for(ngenes + 1)), dim3(kBlock), 0, st, keys,
                       nruns.get(), ngenes, cp.get());
    SEV_LAUNCH_CHECK();
    int32_t bad_h = 0;
    SEV_HIP(hipMemcpyAsync(&bad_h, bad.get(), sizeof bad_h, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    SEV_REQUIRE(bad_h == 0, "a count does not fit Int32 (the device counts are Int32)");
    nnz = nu;
}

}  // namespace sev
