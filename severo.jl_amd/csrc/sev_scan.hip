// Device-wide exclusive prefix sums (three launches: tile sums, top scan,
// tile fix-up).  Deterministic integer arithmetic.
#include "sev_kernels.hpp"

namespace sev {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int64_t kScanTile = kScanBlock * kScanItems;

// Exclusive scan of one value per thread over the block; returns the
// exclusive prefix and writes the block total to *total.
__device__ int64_t block_exclusive_scan(int64_t v, int64_t *total) {
    __shared__ int64_t wsum[kScanBlock / kWave];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int64_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int64_t s = lane < kScanBlock / kWave ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < kScanBlock / kWave; o <<= 1) {
            int64_t y = __shfl_up(s, o, kWave);
            if (lane >= o) s += y;
        }
        if (lane < kScanBlock / kWave) wsum[lane] = s;
    }
    __syncthreads();
    int64_t base = w > 0 ? wsum[w - 1] : 0;
    *total = wsum[kScanBlock / kWave - 1];
    __syncthreads();
    return base + x - v;
}

template <class In>
__global__ void __launch_bounds__(kScanBlock)
k_scan_tiles(const In *in, int64_t *out, int64_t n, int64_t *tile_sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + i;
        v[i] = idx < n ? (int64_t)in[idx] : 0;
        s += v[i];
    }
    int64_t total;
    int64_t pre = block_exclusive_scan(s, &total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t idx = base + i;
        if (idx < n) out[idx] = pre;
        pre += v[i];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place; tile_sums[ntiles]
// receives the grand total.
__global__ void __launch_bounds__(kScanBlock) k_scan_top(int64_t *tile_sums, int64_t ntiles) {
    const int64_t per = (ntiles + kScanBlock - 1) / kScanBlock;
    const int64_t b = (int64_t)threadIdx.x * per;
    const int64_t e = b + per < ntiles ? b + per : ntiles;
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += tile_sums[i];
    int64_t total;
    int64_t pre = block_exclusive_scan(s, &total);
    for (int64_t i = b; i < e; ++i) {
        int64_t t = tile_sums[i];
        tile_sums[i] = pre;
        pre += t;
    }
    if (threadIdx.x == 0) tile_sums[ntiles] = total;
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_fix(int64_t *out, int64_t n, const int64_t *tile_off, int64_t ntiles) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    const int64_t add = tile_off[blockIdx.x];
    for (int64_t i = threadIdx.x; i < kScanTile; i += kScanBlock) {
        int64_t idx = base + i;
        if (idx < n) out[idx] += add;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = tile_off[ntiles];
}

template <class In>
void scan_impl(const In *in, int64_t *out, int64_t n, int64_t *scratch, hipStream_t st) {
    const int64_t ntiles = n > 0 ? ceil_div(n, kScanTile) : 1;
    hipLaunchKernelGGL(k_scan_tiles<In>, dim3((unsigned)ntiles), dim3(kScanBlock), 0, st, in, out,
                       n, scratch);
    SEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, st, scratch, ntiles);
    SEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_fix, dim3((unsigned)ntiles), dim3(kScanBlock), 0, st, out, n,
                       scratch, ntiles);
    SEV_LAUNCH_CHECK();
}

}  // namespace

size_t scan_scratch_elems(int64_t n) { return (size_t)(n > 0 ? ceil_div(n, kScanTile) : 1) + 1; }

void scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *scratch,
                        hipStream_t st) {
    scan_impl<int64_t>(in, out, n, scratch, st);
}

void scan_exclusive_i32(const int32_t *in, int64_t *out, int64_t n, int64_t *scratch,
                        hipStream_t st) {
    scan_impl<int32_t>(in, out, n, scratch, st);
}

}  // namespace sev
