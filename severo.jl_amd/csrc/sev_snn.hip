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
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace sev {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kWaveKept = 256;    // kept entries a wave sorts in LDS
constexpr int kBlockKept = 2048;  // kept entries a block sorts in its buffer
constexpr int kTableBytes = 16384;    // per-wave hash table
constexpr int kBlockTableBytes = 65536;  // per-block hash table

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

// Hash-table slots: key + 1 in the high bits, the count in the low bits, so
// that sorting slots sorts by row.  32-bit slots (rows < 2^25, counts < 128)
// halve the table size; 64-bit slots serve larger matrices.
template <class S>
struct Slot;
template <>
struct Slot<uint32_t> {
    static constexpr int kCountBits = 7;
    static constexpr uint32_t kCountMask = 127u;
};
template <>
struct Slot<uint64_t> {
    static constexpr int kCountBits = 32;
    static constexpr uint64_t kCountMask = 0xFFFFFFFFull;
};

__device__ __forceinline__ uint32_t slot_hash(uint32_t key) {
    uint32_t h = key * 0x9E3779B1u;
    return h ^ (h >> 15);
}

template <class S>
__device__ __forceinline__ S slot_cas(S *a, S cmp, S val) {
    return atomicCAS(a, cmp, val);
}
template <>
__device__ __forceinline__ uint64_t slot_cas<uint64_t>(uint64_t *a, uint64_t cmp, uint64_t val) {
    return (uint64_t)atomicCAS((unsigned long long *)a, (unsigned long long)cmp,
                               (unsigned long long)val);
}
__device__ __forceinline__ void slot_inc(uint32_t *a) { atomicAdd(a, 1u); }
__device__ __forceinline__ void slot_inc(uint64_t *a) {
    atomicAdd((unsigned long long *)a, 1ull);
}

// Insert key into an open-addressing table of nslots (a power of two).
template <class S>
__device__ __forceinline__ void tab_insert(S *tab, uint32_t nslots, uint32_t key) {
    const S tag = (S)(key + 1u) << Slot<S>::kCountBits;
    uint32_t h = slot_hash(key) & (nslots - 1);
    while (true) {
        S cur = tab[h];
        if (cur == 0) {
            const S old = slot_cas<S>(&tab[h], (S)0, (S)(tag | 1u));
            if (old == 0) return;
            cur = old;
        }
        if ((cur >> Slot<S>::kCountBits) == (S)(key + 1u)) {
            slot_inc(&tab[h]);
            return;
        }
        h = (h + 1) & (nslots - 1);
    }
}

template <class S>
__device__ __forceinline__ bool slot_kept(S v, int xmin) {
    return v != 0 && (int)(v & Slot<S>::kCountMask) >= xmin;
}
template <class S>
__device__ __forceinline__ int32_t slot_row(S v) {
    return (int32_t)((v >> Slot<S>::kCountBits) - 1);
}
template <class S>
__device__ __forceinline__ int slot_count(S v) {
    return (int)(v & Slot<S>::kCountMask);
}

// Table size for P candidates: the next power of two >= 2P (at least 64).
__device__ __forceinline__ uint32_t table_size(int64_t P) {
    uint32_t t = 64;
    while ((int64_t)t < 2 * P) t <<= 1;
    return t;
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

// Insert the candidates e = first, first + stride, ... < P of a column into
// tab.  Entry e belongs to the neighbour list t with pre[t] <= e < pre[t+1]
// (found by binary search over the 64 prefixes; lists dropped as duplicates
// have zero length and are skipped by taking the last such t) and is the row
// rev[beg[t] + e - pre[t]].  Eight entries per thread are looked up and their
// rows loaded before any is inserted, so the loads overlap.
template <class S>
__device__ __forceinline__ void insert_entries(S *tab, uint32_t nslots, int64_t P, int64_t first,
                                               int64_t stride, const int64_t *beg,
                                               const int64_t *pre, const int32_t *rev) {
    constexpr int R = 8;
    for (int64_t base = first; base < P; base += R * stride) {
        uint32_t key[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t e = base + r * stride;
            int t = 0;
#pragma unroll
            for (int step = kMaxK / 2; step > 0; step >>= 1)
                if (pre[t + step] <= e) t += step;
            key[r] = e < P ? (uint32_t)rev[beg[t] + (e - pre[t])] : 0u;
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (base + r * stride < P) tab_insert<S>(tab, nslots, key[r]);
    }
}

// Bitonic sort of n (power of two) values ascending by one wavefront.
template <class S>
__device__ void wave_bitonic(S *a, int n) {
    const int lane = lane_id();
    for (int size = 2; size <= n; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1)
            for (int i = lane; i < n / 2; i += kWave) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const S x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
}

template <class S>
__device__ void block_bitonic(S *a, int64_t n) {
    for (int64_t size = 2; size <= n; size <<= 1)
        for (int64_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (int64_t i = threadIdx.x; i < n / 2; i += blockDim.x) {
                const int64_t lo = 2 * i - (i & (stride - 1));
                const int64_t hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const S x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
}

// Wave path: one wavefront per column with a private LDS table, for columns
// with P <= table/2 candidates and at most kWaveKept kept entries; other
// columns are listed as heavy (count mode) or skipped (fill mode) by the
// same rule.  count mode (rows == nullptr): cnt[j] = kept entries, and with
// stage given the kept slots sorted by row at stage[j * kWaveKept] (staged[j]
// = 1), so that the fill only copies them.  fill mode: the kept entries
// sorted by row at colptr[j].
template <class T, class S>
__global__ void __launch_bounds__(kBlock)
k_snn_wave(const int32_t *nn, int64_t n, int k, int xmin, const int64_t *rptr,
           const int32_t *rev, int32_t *cnt, int32_t *heavy, int64_t *heavyP, int32_t *nheavy,
           const T *wtab, const int64_t *colptr, int32_t *rows, T *vals, S *stage,
           int32_t *staged) {
    constexpr int NS = kTableBytes / (int)sizeof(S);
    __shared__ S tabs[kWaves][NS];
    __shared__ S kbufs[kWaves][kWaveKept];
    __shared__ int64_t wbeg[kWaves][kMaxK], wpre[kWaves][kMaxK];
    static_assert(kMaxK == kWave, "one neighbour list per lane");
    const int lane = lane_id(), w = threadIdx.x / kWave;
    S *tab = tabs[w];
    S *kbuf = kbufs[w];
    for (int i = lane; i < NS; i += kWave) tab[i] = 0;
    const int64_t stride = (int64_t)gridDim.x * kWaves;
    for (int64_t j = (int64_t)blockIdx.x * kWaves + w; j < n; j += stride) {
        int64_t mybeg, mypre, mylen;
        const int64_t P = column_setup(nn, n, j, k, rptr, mybeg, mypre, mylen);
        if (P > NS / 2) {
            if (!rows && lane == 0) {
                const int h = atomicAdd(nheavy, 1);
                heavy[h] = (int32_t)j;
                heavyP[h] = P;
            }
            continue;
        }
        const uint32_t T_ = table_size(P);
        // lanes past k hold prefix P, so the search never selects them
        wbeg[w][lane] = mybeg;
        wpre[w][lane] = mypre;
        insert_entries<S>(tab, T_, P, lane, kWave, wbeg[w], wpre[w], rev);
        // one pass over the table: the kept slots go to kbuf in slot order
        // (the first kWaveKept of them) and every slot is cleared
        const uint64_t ltmask = (1ull << lane) - 1;
        int L = 0;
        for (uint32_t i0 = 0; i0 < T_; i0 += kWave) {
            const S v = tab[i0 + lane];
            const bool kept = slot_kept<S>(v, xmin);
            const uint64_t km = __ballot(kept);
            const int pos = L + __popcll(km & ltmask);
            if (kept && pos < kWaveKept) kbuf[pos] = v;
            L += __popcll(km);
            tab[i0 + lane] = 0;
        }
        if (L > kWaveKept || (!rows && !stage)) {
            if (!rows && lane == 0) {
                if (L > kWaveKept) {
                    const int h = atomicAdd(nheavy, 1);
                    heavy[h] = (int32_t)j;
                    heavyP[h] = P;
                } else {
                    cnt[j] = L;
                }
            }
            continue;
        }
        int Lp = 1;
        while (Lp < L) Lp <<= 1;
        for (int i = L + lane; i < Lp; i += kWave) kbuf[i] = ~(S)0;
        wave_bitonic<S>(kbuf, Lp);
        if (!rows) {
            // count mode with staging
            S *dst = stage + j * kWaveKept;
            for (int i = lane; i < L; i += kWave) dst[i] = kbuf[i];
            if (lane == 0) {
                cnt[j] = L;
                staged[j] = 1;
            }
            continue;
        }
        const int64_t ob = colptr[j];
        for (int i = lane; i < L; i += kWave) {
            const S v = kbuf[i];
            rows[ob + i] = slot_row<S>(v);
            vals[ob + i] = wtab[slot_count<S>(v)];
        }
    }
}

// Fill of the staged columns: copy the sorted kept slots to colptr[j].
template <class T, class S>
__global__ void __launch_bounds__(kBlock)
k_snn_stage_copy(int64_t n, const S *stage, const int32_t *staged, const T *wtab,
                 const int64_t *colptr, int32_t *rows, T *vals) {
    const int lane = lane_id();
    const int64_t stride = (int64_t)gridDim.x * kWaves;
    for (int64_t j = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave; j < n; j += stride) {
        if (!staged[j]) continue;
        const int64_t ob = colptr[j];
        const int L = (int)(colptr[j + 1] - ob);
        const S *src = stage + j * kWaveKept;
        for (int i = lane; i < L; i += kWave) {
            const S v = src[i];
            rows[ob + i] = slot_row<S>(v);
            vals[ob + i] = wtab[slot_count<S>(v)];
        }
    }
}

// Count the candidates of one column into tab (nslots slots, zeroed here)
// with the whole block, and return the number of kept entries.
template <class S, int BS>
__device__ int block_count(S *tab, uint32_t nslots, int64_t P, int k, const int32_t *rev,
                           const int64_t *sbeg, const int64_t *spre, int xmin, int *wk) {
    const int lane = lane_id(), w = threadIdx.x / kWave;
    for (uint32_t i = threadIdx.x; i < nslots; i += BS) tab[i] = 0;
    __syncthreads();
    insert_entries<S>(tab, nslots, P, threadIdx.x, BS, sbeg, spre, rev);
    __syncthreads();
    int nk = 0;
    for (uint32_t i = threadIdx.x; i < nslots; i += BS) nk += slot_kept<S>(tab[i], xmin);
    int wsum = nk;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(wsum, o, kWave);
        if (lane >= o) wsum += y;
    }
    if (lane == kWave - 1) wk[w] = wsum;
    __syncthreads();
    int L = 0;
    for (int q = 0; q < BS / kWave; ++q) L += wk[q];
    __syncthreads();
    return L;
}

// Write the kept entries of tab sorted by row: gathered into kept (when
// given and large enough) and sorted there, or sorted in place in the table
// after the other slots are set to the largest key.
template <class S, class T, int BS>
__device__ void block_emit(S *tab, uint32_t nslots, S *kept, int64_t kcap, int L, int xmin,
                           int *nkept, const T *wtab, int32_t *rows, T *vals) {
    S *src;
    int64_t Lp = 1;
    while (Lp < L) Lp <<= 1;
    if (kept && L <= kcap) {
        if (threadIdx.x == 0) *nkept = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nslots; i += BS) {
            const S v = tab[i];
            if (slot_kept<S>(v, xmin)) kept[atomicAdd(nkept, 1)] = v;
        }
        __syncthreads();
        for (int64_t i = L + threadIdx.x; i < Lp; i += BS) kept[i] = ~(S)0;
        __syncthreads();
        block_bitonic<S>(kept, Lp);
        src = kept;
    } else {
        for (uint32_t i = threadIdx.x; i < nslots; i += BS)
            if (!slot_kept<S>(tab[i], xmin)) tab[i] = ~(S)0;
        __syncthreads();
        block_bitonic<S>(tab, nslots);
        src = tab;
    }
    for (int64_t i = threadIdx.x; i < L; i += BS) {
        const S v = src[i];
        rows[i] = slot_row<S>(v);
        vals[i] = wtab[slot_count<S>(v)];
    }
}

// Heavy columns: persistent blocks of BS threads take one column at a time
// from their list, with a hash table of nslots_max slots in dynamic LDS for
// columns with P <= lds_maxP (table size table_size(P), at most nslots_max,
// so at most 3/4 full), else in global memory at gtab + gtab_off[h] (64-bit
// slots: 2/3 of the region is the table, 1/3 the sort buffer).  With stage
// given, the count pass writes the sorted kept slots of LDS-table columns
// with at most kWaveKept of them to stage (staged[j] = 1), as the wave path
// does, and the fill pass skips those; it recounts the other LDS-table
// columns and reuses the global tables filled by the count pass.
template <class T, class S, int BS>
__global__ void __launch_bounds__(BS)
k_snn_block(const int32_t *nn, int64_t n, int k, int xmin, const int64_t *rptr,
            const int32_t *rev, const int32_t *heavy, int64_t nheavy, uint32_t nslots_max,
            int64_t lds_maxP, int32_t *cnt, uint64_t *gtab,
            const int64_t *gtab_off, const T *wtab, const int64_t *colptr, int32_t *rows,
            T *vals, S *stage, int32_t *staged) {
    extern __shared__ unsigned char snn_lds[];
    S *ltab = (S *)snn_lds;
    __shared__ S lkept[kBlockKept];
    __shared__ int64_t sbeg[kMaxK], spre[kMaxK + 1];
    __shared__ int wk[BS / kWave];
    __shared__ int nkept;
    const int lane = lane_id(), w = threadIdx.x / kWave;
    // the LDS table is clear at the start of every column: cleared here
    // once, and after each column by the pass that reads it
    for (uint32_t i = threadIdx.x; i < nslots_max; i += BS) ltab[i] = 0;
    for (int64_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const int64_t j = heavy[h];
        if (rows && staged[j]) continue;  // emitted by the count pass
        __syncthreads();  // the previous column is done with the shared state
        if (w == 0) {
            int64_t mybeg, mypre, mylen;
            const int64_t P = column_setup(nn, n, j, k, rptr, mybeg, mypre, mylen);
            if (lane < kMaxK) {
                sbeg[lane] = mybeg;
                spre[lane] = mypre;
            }
            if (lane == 0) {
                spre[kMaxK] = P;
                nkept = 0;
            }
        }
        __syncthreads();
        const int64_t P = spre[kMaxK];
        if (P <= lds_maxP) {
            const uint32_t T_ = table_size(P) < nslots_max ? table_size(P) : nslots_max;
            if (!rows) {
                insert_entries<S>(ltab, T_, P, threadIdx.x, BS, sbeg, spre, rev);
                __syncthreads();
                // one pass: count the kept slots, keep the first kBlockKept
                // of them, clear the table
                for (uint32_t i = threadIdx.x; i < T_; i += BS) {
                    const S v = ltab[i];
                    if (slot_kept<S>(v, xmin)) {
                        const int q = atomicAdd(&nkept, 1);
                        if (q < kBlockKept) lkept[q] = v;
                    }
                    ltab[i] = 0;
                }
                __syncthreads();
                const int L = nkept;
                if (threadIdx.x == 0) cnt[j] = L;
                if (stage && L <= kWaveKept && w == 0) {
                    // sort the kept slots by row (one wavefront) and stage them
                    int Lp = 1;
                    while (Lp < L) Lp <<= 1;
                    for (int i = L + lane; i < Lp; i += kWave) lkept[i] = ~(S)0;
                    wave_bitonic<S>(lkept, Lp);
                    S *dst = stage + j * kWaveKept;
                    for (int i = lane; i < L; i += kWave) dst[i] = lkept[i];
                    if (lane == 0) staged[j] = 1;
                }
                continue;
            }
            const int L = block_count<S, BS>(ltab, T_, P, k, rev, sbeg, spre, xmin, wk);
            block_emit<S, T, BS>(ltab, T_, lkept, kBlockKept, L, xmin, &nkept, wtab,
                                 rows + colptr[j], vals + colptr[j]);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < T_; i += BS) ltab[i] = 0;
        } else {
            // the global table of this column keeps its counts from the count
            // pass (its region is private to the column): the fill only emits
            uint64_t *tab = gtab + gtab_off[h];
            const uint32_t T_ = (uint32_t)((gtab_off[h + 1] - gtab_off[h]) / 3 * 2);
            if (!rows) {
                const int L =
                    block_count<uint64_t, BS>(tab, T_, P, k, rev, sbeg, spre, xmin, wk);
                if (threadIdx.x == 0) cnt[j] = L;
                continue;
            }
            const int L = (int)(colptr[j + 1] - colptr[j]);
            block_emit<uint64_t, T, BS>(tab, T_, tab + T_, T_ / 2, L, xmin, &nkept, wtab,
                                        rows + colptr[j], vals + colptr[j]);
        }
    }
}

unsigned grid_for(int64_t n, int per_cu = 8) {
    int64_t want = ceil_div(n > 0 ? n : 1, kBlock);
    int64_t cap = (int64_t)device_cu_count() * per_cu;
    return (unsigned)(want < cap ? want : cap);
}

// rows below 2^25 (and counts below 128) fit 32-bit slots
bool small_slots(int64_t n) { return n < ((int64_t)1 << 25) - 1; }

// The heavy columns in three classes: an LDS table of kBlockTableBytes in
// blocks of kBlock threads (two per CU) for P up to half its slots, one of
// kBigTableBytes in blocks of 1024 threads (one per CU) for P up to 3/4 of
// its slots, and global-memory tables above that.
constexpr int64_t kBigTableBytes = 131072;
constexpr int kBigBlock = 1024;

int64_t slot_bytes(int64_t n) { return small_slots(n) ? 4 : 8; }
uint32_t small_table_slots(int64_t n) { return (uint32_t)(kBlockTableBytes / slot_bytes(n)); }
uint32_t big_table_slots(int64_t n) { return (uint32_t)(kBigTableBytes / slot_bytes(n)); }
int64_t small_maxP(int64_t n) { return small_table_slots(n) / 2; }
int64_t big_maxP(int64_t n) { return (int64_t)big_table_slots(n) * 3 / 4; }

// Staging of the wave columns' kept entries (n x kWaveKept slots) when it
// fits this budget; otherwise the fill recounts those columns.
constexpr int64_t kStageBudget = (int64_t)8 << 30;

bool use_stage(int64_t n) {
    return n * kWaveKept * (small_slots(n) ? 4 : 8) <= kStageBudget;
}

template <class T>
void launch_wave(bool count, const int32_t *nn, int64_t n, int k, int xmin, SnnWork &w,
                 const T *wtab, const int64_t *colptr, int32_t *rows, T *vals, hipStream_t st) {
    const unsigned wg = (unsigned)std::min<int64_t>(ceil_div(n, kWaves), device_cu_count() * 8);
    const bool stage = count && use_stage(n);
    if (small_slots(n))
        hipLaunchKernelGGL((k_snn_wave<T, uint32_t>), dim3(wg), dim3(kBlock), 0, st, nn, n, k,
                           xmin, w.rptr.get(), w.rev.get(), w.cnt.get(), w.heavy.get(),
                           w.heavyP.get(), w.nheavy.get(), wtab, colptr, count ? nullptr : rows,
                           vals, stage ? (uint32_t *)w.stage.get() : nullptr, w.staged.get());
    else
        hipLaunchKernelGGL((k_snn_wave<T, uint64_t>), dim3(wg), dim3(kBlock), 0, st, nn, n, k,
                           xmin, w.rptr.get(), w.rev.get(), w.cnt.get(), w.heavy.get(),
                           w.heavyP.get(), w.nheavy.get(), wtab, colptr, count ? nullptr : rows,
                           vals, stage ? (uint64_t *)w.stage.get() : nullptr, w.staged.get());
    SEV_LAUNCH_CHECK();
}

template <class T>
void launch_stage_copy(int64_t n, SnnWork &w, const T *wtab, const int64_t *colptr,
                       int32_t *rows, T *vals, hipStream_t st) {
    const unsigned wg = (unsigned)std::min<int64_t>(ceil_div(n, kWaves), device_cu_count() * 8);
    if (small_slots(n))
        hipLaunchKernelGGL((k_snn_stage_copy<T, uint32_t>), dim3(wg), dim3(kBlock), 0, st, n,
                           (const uint32_t *)w.stage.get(), w.staged.get(), wtab, colptr, rows,
                           vals);
    else
        hipLaunchKernelGGL((k_snn_stage_copy<T, uint64_t>), dim3(wg), dim3(kBlock), 0, st, n,
                           (const uint64_t *)w.stage.get(), w.staged.get(), wtab, colptr, rows,
                           vals);
    SEV_LAUNCH_CHECK();
}

template <class T, class S, int BS>
void launch_block_class(bool count, const int32_t *nn, int64_t n, int k, int xmin, SnnWork &w,
                        int64_t first, int64_t cnt_cols, uint32_t slots, int64_t maxP,
                        int per_cu, const T *wtab, const int64_t *colptr, int32_t *rows,
                        T *vals, hipStream_t st) {
    if (cnt_cols <= 0) return;
    const size_t lds = (size_t)slots * sizeof(S);
    if (lds > 65536) {
        static bool attr = false;
        if (!attr) {
            SEV_HIP(hipFuncSetAttribute((const void *)k_snn_block<T, S, BS>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            attr = true;
        }
    }
    const bool stage = use_stage(n);
    const unsigned grid =
        (unsigned)std::min<int64_t>(cnt_cols, (int64_t)per_cu * device_cu_count());
    hipLaunchKernelGGL((k_snn_block<T, S, BS>), dim3(grid), dim3(BS), lds, st, nn, n, k, xmin,
                       w.rptr.get(), w.rev.get(), w.heavy.get() + first, cnt_cols, slots, maxP,
                       w.cnt.get(), w.htab.get(), w.htab_off.get(), wtab, colptr,
                       count ? nullptr : rows, vals, stage ? (S *)w.stage.get() : nullptr,
                       w.staged.get());
    SEV_LAUNCH_CHECK();
}

template <class T, class S>
void launch_block_s(bool count, const int32_t *nn, int64_t n, int k, int xmin, SnnWork &w,
                    const T *wtab, const int64_t *colptr, int32_t *rows, T *vals,
                    hipStream_t st) {
    // persistent blocks: two per CU fit the small table, one the big one
    launch_block_class<T, S, kBlock>(count, nn, n, k, xmin, w, 0, w.nheavy_cls[0],
                                     small_table_slots(n), small_maxP(n), 2, wtab, colptr, rows,
                                     vals, st);
    launch_block_class<T, S, kBigBlock>(count, nn, n, k, xmin, w, w.nheavy_cls[0],
                                        w.nheavy_cls[1], big_table_slots(n), big_maxP(n), 1,
                                        wtab, colptr, rows, vals, st);
    launch_block_class<T, S, kBlock>(count, nn, n, k, xmin, w,
                                     w.nheavy_cls[0] + w.nheavy_cls[1], w.nheavy_cls[2], 0, -1, 2,
                                     wtab, colptr, rows, vals, st);
}

template <class T>
void launch_block(bool count, const int32_t *nn, int64_t n, int k, int xmin, SnnWork &w,
                  const T *wtab, const int64_t *colptr, int32_t *rows, T *vals, hipStream_t st) {
    if (w.nheavy_h <= 0) return;
    if (small_slots(n))
        launch_block_s<T, uint32_t>(count, nn, n, k, xmin, w, wtab, colptr, rows, vals, st);
    else
        launch_block_s<T, uint64_t>(count, nn, n, k, xmin, w, wtab, colptr, rows, vals, st);
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
    w.staged.ensure((size_t)n);
    SEV_HIP(hipMemsetAsync(w.staged.get(), 0, sizeof(int32_t) * n, st));
    if (use_stage(n)) w.stage.ensure((size_t)(n * kWaveKept * (small_slots(n) ? 4 : 8)));
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
    ProfScope ps("snn_count", st);
    launch_wave<T>(true, nn, n, (int)k, xmin, w, nullptr, nullptr, nullptr, nullptr, st);
    int32_t nh = 0;
    SEV_HIP(hipMemcpyAsync(&nh, w.nheavy.get(), sizeof nh, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    w.nheavy_h = nh;
    if (nh > 0) {
        // the heavy columns ordered by class (small LDS table, big LDS table,
        // global table); the global ones get a table of 2 * next_pow2(P)
        // slots plus a sort buffer of half that
        // pinned host copies on the stream (the pageable path's staging cost
        // milliseconds here)
        int64_t *hp = w.hP.ensure((size_t)nh);
        int32_t *hj = w.hcol.ensure((size_t)nh);
        int32_t *order = w.horder.ensure((size_t)nh);
        SEV_HIP(hipMemcpyAsync(hp, w.heavyP.get(), sizeof(int64_t) * nh, hipMemcpyDeviceToHost,
                               st));
        SEV_HIP(hipMemcpyAsync(hj, w.heavy.get(), sizeof(int32_t) * nh, hipMemcpyDeviceToHost,
                               st));
        SEV_HIP(hipStreamSynchronize(st));
        const int64_t lim[2] = {small_maxP(n), big_maxP(n)};
        int64_t no = 0;
        std::vector<int64_t> gp;
        for (int c = 0; c < 3; ++c) {
            const int64_t before = no;
            for (int32_t h = 0; h < nh; ++h) {
                const int cls = hp[h] <= lim[0] ? 0 : (hp[h] <= lim[1] ? 1 : 2);
                if (cls != c) continue;
                order[no++] = hj[h];
                if (c == 2) gp.push_back(hp[h]);
            }
            w.nheavy_cls[c] = no - before;
        }
        // columns are independent, so the order within a class is free: kept
        // as found for determinism of the layout (every result is sorted)
        SEV_HIP(hipMemcpyAsync(w.heavy.get(), order, sizeof(int32_t) * nh,
                               hipMemcpyHostToDevice, st));
        const int64_t ng = w.nheavy_cls[2];
        int64_t *offs = w.hoff.ensure((size_t)ng + 1);
        offs[0] = 0;
        int64_t mx = 0;
        for (int64_t h = 0; h < ng; ++h) {
            int64_t slots = 1;
            while (slots < 2 * gp[(size_t)h]) slots <<= 1;
            slots += slots / 2;
            mx = std::max(mx, gp[(size_t)h]);
            offs[(size_t)h + 1] = offs[(size_t)h] + slots;
        }
        if (std::getenv("SEV_VERBOSE"))
            std::fprintf(stderr, "[severo_hip] snn n=%lld k=%lld: %d block columns (%lld with "
                         "P <= %lld, %lld with P <= %lld, %lld with a global table, max P "
                         "%lld)\n", (long long)n, (long long)k, nh,
                         (long long)w.nheavy_cls[0], (long long)lim[0],
                         (long long)w.nheavy_cls[1], (long long)lim[1], (long long)ng,
                         (long long)mx);
        w.htab_off.ensure((size_t)ng + 1);
        SEV_HIP(hipMemcpyAsync(w.htab_off.get(), offs, sizeof(int64_t) * (ng + 1),
                               hipMemcpyHostToDevice, st));
        w.htab.ensure((size_t)(offs[(size_t)ng] > 0 ? offs[(size_t)ng] : 1));
        launch_block<T>(true, nn, n, (int)k, xmin, w, nullptr, nullptr, nullptr, nullptr, st);
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
    ProfScope ps("snn_fill", st);
    if (use_stage(n))
        launch_stage_copy<T>(n, w, wtab, colptr, rows, vals, st);
    else
        launch_wave<T>(false, nn, n, (int)k, xmin, w, wtab, colptr, rows, vals, st);
    launch_block<T>(false, nn, n, (int)k, xmin, w, wtab, colptr, rows, vals, st);
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
