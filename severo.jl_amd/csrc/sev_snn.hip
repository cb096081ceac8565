// Jaccard shared-nearest-neighbour graph.
//
// Replaces src/neighbours.jl:79 (the k-NN indicator matrix nn, column j
// holding N(j)) and :88-93 (_jaccard_index: S = nn'nn, w = x/(k + (k - x)),
// droptol!(S, prune)).  S[i,j] = |N(i) n N(j)| is computed per output column
// j without forming nn'nn: with the reverse lists R(l) = {i : l in N(i)},
//   S[i,j] = #{ l in N(j) : i in R(l) },
// so column j is the multiset union of R(l) over l in N(j), counted.
//
// One wavefront per column counts the union in a private LDS open-addressing
// table (64-bit slots: key+1 in the high word, count in the low word,
// inserted with compare-and-swap / add), keeps entries with count >= xmin
// (w > prune, decided exactly on the host from the k+1 possible weights),
// sorts the kept entries by row with a bitonic sort and writes them.
// Columns whose union exceeds the wave table go to a block kernel with an
// LDS table, or a global-memory table for the largest hubs.  Counts are
// integers and the output is sorted, so the result is deterministic.
#include "sev_kernels.hpp"

#include <algorithm>
#include <vector>

namespace sev {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kWaveSlots = 2048;   // per-wave table
constexpr int kWaveMaxP = 1024;    // candidates handled by the wave path
constexpr int kBlockSlots = 4096;  // per-block LDS table
constexpr int kBlockMaxP = 2048;
constexpr int kBlockKept = 2048;   // LDS sort buffer of kept entries

// Entries outside [0, n) (an incomplete neighbour list) are ignored.
__global__ void __launch_bounds__(kBlock)
k_indeg(const int32_t *nn, int64_t total, int64_t n, int32_t *deg) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < total;
         p += (int64_t)gridDim.x * kBlock) {
        const int32_t l = nn[p];
        if (l >= 0 && l < n) atomicAdd(&deg[l], 1);
    }
}

__global__ void __launch_bounds__(kBlock)
k_rev_fill(const int32_t *nn, int64_t total, int64_t n, int64_t k, const int64_t *rptr,
           int32_t *cur, int32_t *rev) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < total;
         p += (int64_t)gridDim.x * kBlock) {
        const int32_t l = nn[p];
        if (l < 0 || l >= n) continue;
        const int32_t pos = atomicAdd(&cur[l], 1);
        rev[rptr[l] + pos] = (int32_t)(p / k);
    }
}

__device__ __forceinline__ uint32_t slot_hash(uint32_t key) {
    uint32_t h = key * 0x9E3779B1u;
    return h ^ (h >> 15);
}

// Insert key into an open-addressing table of nslots (power of two).
__device__ __forceinline__ void tab_insert(uint64_t *tab, uint32_t nslots, uint32_t key) {
    const uint64_t tag = (uint64_t)(key + 1u) << 32;
    uint32_t h = slot_hash(key) & (nslots - 1);
    while (true) {
        uint64_t cur = tab[h];
        if (cur == 0) {
            const uint64_t old = atomicCAS((unsigned long long *)&tab[h], 0ull,
                                           (unsigned long long)(tag | 1u));
            if (old == 0) return;
            cur = old;
        }
        if ((cur >> 32) == (uint64_t)(key + 1u)) {
            atomicAdd((unsigned long long *)&tab[h], 1ull);
            return;
        }
        h = (h + 1) & (nslots - 1);
    }
}

// Per-column neighbour setup held in lanes 0..k-1: deduplicated l, the length
// of R(l) and the exclusive prefix of the lengths.  Returns the total P.
__device__ int64_t column_setup(const int32_t *nn, int64_t n, int64_t j, int k,
                                const int64_t *rptr, int64_t &mybeg, int64_t &mypre,
                                int64_t &mylen) {
    const int lane = lane_id();
    int32_t l = lane < k ? nn[j * k + lane] : -1;
    // drop duplicates (a neighbour listed twice counts once, as sparse() ORs)
    bool dup = false;
    for (int t = 0; t < k; ++t) {
        const int32_t o = __shfl(l, t, kWave);
        if (t < lane && o == l) dup = true;
    }
    int64_t len = 0;
    mybeg = 0;
    if (lane < k && !dup && l >= 0 && l < n) {
        mybeg = rptr[l];
        len = rptr[l + 1] - mybeg;
    }
    int64_t pre = len;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int64_t y = __shfl_up(pre, o, kWave);
        if (lane >= o) pre += y;
    }
    const int64_t total = __shfl(pre, kWave - 1, kWave);
    mypre = pre - len;
    mylen = len;
    return total;
}

// Map a flat candidate index e (< P) to (owner lane) by binary search over
// the lanes' exclusive prefixes (which are nondecreasing in lane order).
__device__ __forceinline__ int owner_lane(int64_t e, int64_t mypre) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int64_t p = __shfl(mypre, lo + step, kWave);
        if (lo + step < kWave && p <= e) lo += step;
    }
    return lo;
}

// Bitonic sort of n (power of two) 64-bit values in ascending order by one
// wavefront (n <= slots of the table region).
__device__ void wave_bitonic(uint64_t *a, int n) {
    const int lane = lane_id();
    for (int size = 2; size <= n; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1)
            for (int i = lane; i < n / 2; i += kWave) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const uint64_t x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
}

__device__ void block_bitonic(uint64_t *a, int n) {
    for (int size = 2; size <= n; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n / 2; i += blockDim.x) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const uint64_t x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
}

// Wave path.  count mode (rows == nullptr): cnt[j] = kept entries, or -1 and
// j appended to the heavy list when P exceeds the wave table.  fill mode:
// writes the sorted kept entries of the light columns at colptr[j].
template <class T>
__global__ void __launch_bounds__(kBlock)
k_snn_wave(const int32_t *nn, int64_t n, int k, int xmin, const int64_t *rptr,
           const int32_t *rev, int32_t *cnt, int32_t *heavy, int64_t *heavyP, int32_t *nheavy,
           const T *wtab, const int64_t *colptr, int32_t *rows, T *vals) {
    __shared__ uint64_t tabs[kWaves][kWaveSlots];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    uint64_t *tab = tabs[w];
    for (int i = lane; i < kWaveSlots; i += kWave) tab[i] = 0;
    const int64_t stride = (int64_t)gridDim.x * kWaves;
    for (int64_t j = (int64_t)blockIdx.x * kWaves + w; j < n; j += stride) {
        int64_t mybeg, mypre, mylen;
        const int64_t P = column_setup(nn, n, j, k, rptr, mybeg, mypre, mylen);
        if (P > kWaveMaxP) {
            if (!rows && lane == 0) {
                cnt[j] = -1;
                const int h = atomicAdd(nheavy, 1);
                heavy[h] = (int32_t)j;
                heavyP[h] = P;
            }
            continue;
        }
        // every lane runs the shuffles each round; lanes past P insert nothing
        for (int64_t base = 0; base < P; base += kWave) {
            const int64_t e = base + lane;
            const int o = owner_lane(e < P ? e : P - 1, mypre);
            const int64_t ob = __shfl(mybeg, o, kWave);
            const int64_t op = __shfl(mypre, o, kWave);
            if (e < P) tab_insert(tab, kWaveSlots, (uint32_t)rev[ob + (e - op)]);
        }
        // compact kept entries to the front of the table
        uint64_t kept[kWaveSlots / kWave];
        int nk = 0;
#pragma unroll
        for (int i = 0; i < kWaveSlots / kWave; ++i) {
            const uint64_t v = tab[lane + i * kWave];
            kept[i] = v;
            if (v != 0 && (int)(v & 0xFFFFFFFFu) >= xmin) nk++;
        }
        int pre = nk;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(pre, o, kWave);
            if (lane >= o) pre += y;
        }
        const int L = __shfl(pre, kWave - 1, kWave);
        if (!rows) {
            if (lane == 0) cnt[j] = L;
            for (int i = lane; i < kWaveSlots; i += kWave) tab[i] = 0;
            continue;
        }
        int pos = pre - nk;
        for (int i = lane; i < kWaveSlots; i += kWave) tab[i] = 0;
#pragma unroll
        for (int i = 0; i < kWaveSlots / kWave; ++i) {
            const uint64_t v = kept[i];
            if (v != 0 && (int)(v & 0xFFFFFFFFu) >= xmin) tab[pos++] = v;
        }
        int Lp = 1;
        while (Lp < L) Lp <<= 1;
        for (int i = L + lane; i < Lp; i += kWave) tab[i] = ~0ull;
        wave_bitonic(tab, Lp);
        const int64_t ob = colptr[j];
        for (int i = lane; i < L; i += kWave) {
            const uint64_t v = tab[i];
            rows[ob + i] = (int32_t)((v >> 32) - 1);
            vals[ob + i] = wtab[v & 0xFFFFFFFFu];
        }
        for (int i = lane; i < Lp; i += kWave) tab[i] = 0;
    }
}

// Heavy columns: one block per column, table in LDS (P <= kBlockMaxP) or in
// global memory at gtab + gtab_off[h] with gtab_size[h] slots.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_snn_block(const int32_t *nn, int64_t n, int k, int xmin, const int64_t *rptr,
            const int32_t *rev,
            const int32_t *heavy, int64_t nheavy, int32_t *cnt, uint64_t *gtab,
            const int64_t *gtab_off, const T *wtab, const int64_t *colptr, int32_t *rows,
            T *vals) {
    __shared__ uint64_t ltab[kBlockSlots];
    __shared__ uint64_t lkept[kBlockKept];
    __shared__ int64_t sbeg[kMaxK], spre[kMaxK + 1];
    __shared__ int wk[kWaves];
    __shared__ int nkept;
    const int64_t h = blockIdx.x;
    if (h >= nheavy) return;
    const int64_t j = heavy[h];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    if (w == 0) {
        int64_t mybeg, mypre, mylen;
        const int64_t P = column_setup(nn, n, j, k, rptr, mybeg, mypre, mylen);
        if (lane < kMaxK) {
            sbeg[lane] = mybeg;
            spre[lane] = mypre;
        }
        if (lane == 0) spre[kMaxK] = P;
    }
    __syncthreads();
    const int64_t P = spre[kMaxK];
    uint64_t *tab, *kept;
    uint32_t nslots;
    if (P <= kBlockMaxP) {
        tab = ltab;
        kept = lkept;
        nslots = kBlockSlots;
    } else {
        // global region: nslots table slots followed by nslots/2 sort slots
        tab = gtab + gtab_off[h];
        nslots = (uint32_t)((gtab_off[h + 1] - gtab_off[h]) / 3 * 2);
        kept = tab + nslots;
    }
    for (uint32_t i = threadIdx.x; i < nslots; i += kBlock) tab[i] = 0;
    __syncthreads();
    for (int64_t e = threadIdx.x; e < P; e += kBlock) {
        // owner = last neighbour slot whose prefix is <= e (deduplicated
        // slots have zero length and share the next slot's prefix)
        int lo = 0;
        for (int t = 1; t < k; ++t)
            if (spre[t] <= e) lo = t;
        tab_insert(tab, nslots, (uint32_t)rev[sbeg[lo] + (e - spre[lo])]);
    }
    __syncthreads();
    // count kept entries
    int nk = 0;
    for (uint32_t i = threadIdx.x; i < nslots; i += kBlock) {
        const uint64_t v = tab[i];
        if (v != 0 && (int)(v & 0xFFFFFFFFu) >= xmin) nk++;
    }
    int wsum = nk;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(wsum, o, kWave);
        if (lane >= o) wsum += y;
    }
    if (lane == kWave - 1) wk[w] = wsum;
    __syncthreads();
    int L = 0;
    for (int q = 0; q < kWaves; ++q) L += wk[q];
    if (!rows) {
        if (threadIdx.x == 0) cnt[j] = L;
        return;
    }
    // gather the L kept entries (any order), pad to a power of two with the
    // largest key and sort ascending: rows come out in order
    if (threadIdx.x == 0) nkept = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nslots; i += kBlock) {
        const uint64_t v = tab[i];
        if (v != 0 && (int)(v & 0xFFFFFFFFu) >= xmin) kept[atomicAdd(&nkept, 1)] = v;
    }
    __syncthreads();
    int Lp = 1;
    while (Lp < L) Lp <<= 1;
    for (int i = L + threadIdx.x; i < Lp; i += kBlock) kept[i] = ~0ull;
    __syncthreads();
    block_bitonic(kept, Lp);
    const int64_t ob = colptr[j];
    for (int64_t i = threadIdx.x; i < L; i += kBlock) {
        const uint64_t v = kept[i];
        rows[ob + i] = (int32_t)((v >> 32) - 1);
        vals[ob + i] = wtab[v & 0xFFFFFFFFu];
    }
}

unsigned grid_for(int64_t n, int per_cu = 8) {
    int64_t want = ceil_div(n > 0 ? n : 1, kBlock);
    int64_t cap = (int64_t)device_cu_count() * per_cu;
    return (unsigned)(want < cap ? want : cap);
}

}  // namespace

template <class T>
int64_t snn_count(const int32_t *nn, int64_t n, int64_t k, int xmin, SnnWork &w,
                  int64_t *colptr, hipStream_t st) {
    SEV_REQUIRE(k >= 1 && k <= kMaxK, "snn: k must be in [1, 64]");
    const int64_t total = n * k;
    w.deg.ensure((size_t)n);
    w.cur.ensure((size_t)n);
    w.rev.ensure((size_t)total);
    w.cnt.ensure((size_t)n);
    w.heavy.ensure((size_t)n);
    w.heavyP.ensure((size_t)n);
    w.nheavy.ensure(1);
    w.rptr.ensure((size_t)n + 1);
    w.scratch.ensure(scan_scratch_elems(n + 1));
    SEV_HIP(hipMemsetAsync(w.deg.get(), 0, sizeof(int32_t) * n, st));
    SEV_HIP(hipMemsetAsync(w.cur.get(), 0, sizeof(int32_t) * n, st));
    SEV_HIP(hipMemsetAsync(w.nheavy.get(), 0, sizeof(int32_t), st));
    hipLaunchKernelGGL(k_indeg, dim3(grid_for(total)), dim3(kBlock), 0, st, nn, total, n,
                       w.deg.get());
    SEV_LAUNCH_CHECK();
    scan_exclusive_i32(w.deg.get(), w.rptr.get(), n, w.scratch.get(), st);
    hipLaunchKernelGGL(k_rev_fill, dim3(grid_for(total)), dim3(kBlock), 0, st, nn, total, n, k,
                       w.rptr.get(), w.cur.get(), w.rev.get());
    SEV_LAUNCH_CHECK();
    const unsigned wg = (unsigned)std::min<int64_t>(ceil_div(n, kWaves), device_cu_count() * 8);
    ProfScope ps("snn_count", st);
    hipLaunchKernelGGL(k_snn_wave<T>, dim3(wg), dim3(kBlock), 0, st, nn, n, (int)k, xmin,
                       w.rptr.get(), w.rev.get(), w.cnt.get(), w.heavy.get(), w.heavyP.get(),
                       w.nheavy.get(), (const T *)nullptr, (const int64_t *)nullptr, (int32_t *)nullptr,
                       (T *)nullptr);
    SEV_LAUNCH_CHECK();
    int32_t nh = 0;
    SEV_HIP(hipMemcpyAsync(&nh, w.nheavy.get(), sizeof nh, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    w.nheavy_h = nh;
    if (nh > 0) {
        // columns above the LDS capacity get a global table of
        // 2 * next_pow2(P) slots
        std::vector<int64_t> hp(nh);
        SEV_HIP(hipMemcpy(hp.data(), w.heavyP.get(), sizeof(int64_t) * nh, hipMemcpyDeviceToHost));
        std::vector<int64_t> offs(nh + 1, 0);
        for (int32_t h = 0; h < nh; ++h) {
            int64_t slots = 0;
            if (hp[h] > kBlockMaxP) {
                slots = 1;
                while (slots < 2 * hp[h]) slots <<= 1;
                slots += slots / 2;  // table + sort buffer (>= next_pow2(P))
            }
            offs[h + 1] = offs[h] + slots;
        }
        w.htab_off.ensure((size_t)nh + 1);
        SEV_HIP(hipMemcpy(w.htab_off.get(), offs.data(), sizeof(int64_t) * (nh + 1),
                          hipMemcpyHostToDevice));
        w.htab.ensure((size_t)(offs[nh] > 0 ? offs[nh] : 1));
        hipLaunchKernelGGL(k_snn_block<T>, dim3((unsigned)nh), dim3(kBlock), 0, st, nn, n,
                           (int)k, xmin, w.rptr.get(), w.rev.get(), w.heavy.get(), (int64_t)nh,
                           w.cnt.get(), w.htab.get(), w.htab_off.get(), (const T *)nullptr,
                           (const int64_t *)nullptr, (int32_t *)nullptr, (T *)nullptr);
        SEV_LAUNCH_CHECK();
    }
    scan_exclusive_i32(w.cnt.get(), colptr, n, w.scratch.get(), st);
    int64_t nnz = 0;
    SEV_HIP(hipMemcpyAsync(&nnz, colptr + n, sizeof nnz, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    return nnz;
}

template <class T>
void snn_fill(const int32_t *nn, int64_t n, int64_t k, int xmin, const T *wtab, SnnWork &w,
              const int64_t *colptr, int32_t *rows, T *vals, hipStream_t st) {
    const unsigned wg = (unsigned)std::min<int64_t>(ceil_div(n, kWaves), device_cu_count() * 8);
    ProfScope ps("snn_fill", st);
    hipLaunchKernelGGL(k_snn_wave<T>, dim3(wg), dim3(kBlock), 0, st, nn, n, (int)k, xmin,
                       w.rptr.get(), w.rev.get(), w.cnt.get(), w.heavy.get(), w.heavyP.get(),
                       w.nheavy.get(), wtab, colptr, rows, vals);
    SEV_LAUNCH_CHECK();
    if (w.nheavy_h > 0) {
        hipLaunchKernelGGL(k_snn_block<T>, dim3((unsigned)w.nheavy_h), dim3(kBlock), 0, st, nn,
                           n, (int)k, xmin, w.rptr.get(), w.rev.get(), w.heavy.get(), w.nheavy_h,
                           w.cnt.get(), w.htab.get(), w.htab_off.get(), wtab, colptr, rows, vals);
        SEV_LAUNCH_CHECK();
    }
}

template int64_t snn_count<float>(const int32_t *, int64_t, int64_t, int, SnnWork &, int64_t *,
                                  hipStream_t);
template int64_t snn_count<double>(const int32_t *, int64_t, int64_t, int, SnnWork &, int64_t *,
                                   hipStream_t);
template void snn_fill<float>(const int32_t *, int64_t, int64_t, int, const float *, SnnWork &,
                              const int64_t *, int32_t *, float *, hipStream_t);
template void snn_fill<double>(const int32_t *, int64_t, int64_t, int, const double *, SnnWork &,
                               const int64_t *, int32_t *, double *, hipStream_t);

}  // namespace sev
