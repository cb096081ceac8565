// Louvain clustering of a neighbour graph on the device: cluster(SNN;
// algorithm=:louvain, resolution, nrandomstarts, niterations,
// group_singletons) of src/clustering.jl:59-104 with the modularity search of
// src/modularity.jl:165-380.
//
// The reference moves one node at a time in a shuffled order.  Here, in each
// sweep, a keyed pseudo-random half of the nodes of a level choose their best
// neighbouring cluster at once against the same cluster totals (moving half
// the nodes keeps simultaneous moves from undoing each other); a sweep that
// does not raise the modularity is discarded and retried with fewer nodes
// (down to 1/64).  The labels
// are permuted at random before each local-moving phase, which is what
// differs between random starts.  The level structure (local moving,
// renumbering, the reduced network, merging, and repetition while the gain
// reaches min_modularity) and the choice of the best start follow the
// reference.
//
// Edge weights are held in fixed point (w * 2^S rounded, S <= 40 chosen so
// that nnz * max(w) * 2^S stays below 2^61), so every accumulation is an exact integer
// sum: results do not depend on the order of atomic updates, and repeated
// runs give identical labels.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "sev_common.hpp"
#include "sev_kernels.hpp"

namespace sev {

namespace {

using u64 = unsigned long long;

constexpr int kLvThreads = 256;
constexpr int kLvWaves = kLvThreads / kWave;
constexpr int kLightDeg = 256;   // nodes with at most this many edges hash in LDS
constexpr int kSmallDeg = 128;   // ... in 256-slot tables (12 KB a block, 8 waves a SIMD)
constexpr int kMaxSweeps = 200;  // per local-moving phase
constexpr int kSqBlocks = 256;   // fixed grid of the sum of squared totals
constexpr int64_t kSinglesPerLaunch = 4096;

inline unsigned grid_of(int64_t n, int threads = kLvThreads) {
    return (unsigned)std::max<int64_t>(1, ceil_div(n, threads));
}

// Open addressing with linear probing; cap exceeds twice the number of
// distinct keys, so a free slot is always reached.  Power-of-two tables
// (shift = 32 - log2(cap)) take the high bits of the multiplicative hash.
__device__ __forceinline__ void hash_add(int *keys, u64 *vals, uint32_t cap, int key, u64 w) {
    uint32_t h = ((uint32_t)key * 2654435761u) % cap;
    while (true) {
        const int prev = atomicCAS(&keys[h], -1, key);
        if (prev == -1 || prev == key) {
            atomicAdd(&vals[h], w);
            return;
        }
        h = h + 1 == cap ? 0 : h + 1;
    }
}

__device__ __forceinline__ void hash_add_pow2(int *keys, u64 *vals, uint32_t cap, int shift,
                                              int key, u64 w) {
    uint32_t h = ((uint32_t)key * 2654435761u) >> shift;
    while (true) {
        const int prev = atomicCAS(&keys[h], -1, key);
        if (prev == -1 || prev == key) {
            atomicAdd(&vals[h], w);
            return;
        }
        h = (h + 1) & (cap - 1);
    }
}

__device__ __forceinline__ void best_merge(double &s, int &c, double os, int oc) {
    if (os > s || (os == s && oc < c)) {
        s = os;
        c = oc;
    }
}

__device__ __forceinline__ void wave_best(double &s, int &c) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double os = __shfl_xor(s, o, kWave);
        const int oc = __shfl_xor(c, o, kWave);
        best_merge(s, c, os, oc);
    }
}

__device__ __forceinline__ u64 wave_max_u64(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const u64 ov = (u64)__shfl_xor((long long)v, o, kWave);
        v = ov > v ? ov : v;
    }
    return v;
}

__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (u64)__shfl_xor((long long)v, o, kWave);
    return v;
}

// Whether node v may move in a sweep: a keyed pseudo-random subset of about
// athresh / 2^24 of the nodes (all of them at athresh = 2^24).
constexpr uint32_t kActAll = 1u << 24;
__device__ __forceinline__ bool active(int64_t v, uint64_t akey, uint32_t athresh) {
    if (athresh >= kActAll) return true;
    uint64_t z = (uint64_t)v + akey + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 40) < athresh;
}

// Score of joining cluster C (its total w_tot excluding the node):
// kin - res * (w_tot * k) / totw, best_local_move (src/modularity.jl:243-253).
// f = res * k / totw is formed once per node.
struct MoveRule {
    int cv;
    double f;
    __device__ __forceinline__ void consider(int C, u64 kin, const u64 *wtot, double &bs, int &bc,
                                             u64 &kcv) const {
        if (C == cv) {
            kcv = kin;
            return;
        }
        const double sc = (double)kin - (double)wtot[C] * f;
        best_merge(bs, bc, sc, C);
    }
    // move when the best allowed cluster beats staying (gain > 0)
    __device__ __forceinline__ int decide(double bs, int bc, u64 kcv, u64 wcv, u64 kvi) const {
        const double init = (double)kcv - (double)(wcv - kvi) * f;
        return (bc != INT_MAX && bs > init) ? bc : cv;
    }
};

// One wave per node of at most kCap / 2 edges: all nodes (list == nullptr;
// nodes with more edges are left to the listed launches) or the nodes of a
// list.
template <int kCap>
__global__ void __launch_bounds__(kLvThreads)
k_move_light(int64_t n, const int32_t *list, const int64_t *ptr, const int32_t *adj,
             const u64 *w, const u64 *k, const int32_t *c, const u64 *wtot, int32_t *cnew,
             double totw, double res, uint64_t akey, uint32_t athresh, int *moved) {
    __shared__ int skey[kLvWaves][kCap];
    __shared__ u64 sval[kLvWaves][kCap];
    const int lane = lane_id(), wv = (int)(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * kLvWaves + wv;
    if (t >= n) return;
    const int64_t v = list ? (int64_t)list[t] : t;
    const int64_t p0 = ptr[v];
    const int64_t deg = ptr[v + 1] - p0;
    if (deg > kCap / 2) return;
    const int cv = c[v];
    if (deg == 0 || !active(v, akey, athresh)) {
        if (lane == 0) cnew[v] = cv;
        return;
    }
    // each lane's first edge is gathered while the table is cleared
    int c0 = -1;
    u64 w0 = 0;
    if (lane < deg) {
        c0 = c[adj[p0 + lane]];
        w0 = w[p0 + lane];
    }
    int *keys = skey[wv];
    u64 *vals = sval[wv];
    // table of the smallest power of two >= 2 deg, at least one slot a lane
    uint32_t cap = kWave;
    int shift = 26;
    while (cap < 2 * (uint32_t)deg) {
        cap <<= 1;
        --shift;
    }
    for (uint32_t s = lane; s < cap; s += kWave) {
        keys[s] = -1;
        vals[s] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (lane < deg) hash_add_pow2(keys, vals, cap, shift, c0, w0);
    for (int64_t e = lane + kWave; e < deg; e += kWave)
        hash_add_pow2(keys, vals, cap, shift, c[adj[p0 + e]], w[p0 + e]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const u64 kvi = k[v];
    const MoveRule r{cv, res * (double)kvi / totw};
    double bs = -INFINITY;
    int bc = INT_MAX;
    u64 kcv = 0;
    for (uint32_t s = lane; s < cap; s += kWave) {
        const int C = keys[s];
        if (C >= 0) r.consider(C, vals[s], wtot, bs, bc, kcv);
    }
    wave_best(bs, bc);
    kcv = wave_max_u64(kcv);
    if (lane == 0) {
        const int to = r.decide(bs, bc, kcv, wtot[cv], kvi);
        cnew[v] = to;
        if (to != cv) atomicAdd(moved, 1);
    }
}

// Nodes with more than kLightDeg edges: one block each, hashing into a
// global table of 2 * deg slots at offset 2 * ptr[v].
__global__ void __launch_bounds__(kLvThreads)
k_move_heavy(const int32_t *heavy, const int64_t *ptr, const int32_t *adj, const u64 *w,
             const u64 *k, const int32_t *c, const u64 *wtot, int32_t *cnew, int *hkeys,
             u64 *hvals, double totw, double res, uint64_t akey, uint32_t athresh, int *moved) {
    __shared__ double rs[kLvWaves];
    __shared__ int rc[kLvWaves];
    __shared__ u64 rk[kLvWaves];
    const int64_t v = heavy[blockIdx.x];
    if (!active(v, akey, athresh)) {
        if (threadIdx.x == 0) cnew[v] = c[v];
        return;
    }
    const int64_t p0 = ptr[v];
    const int64_t deg = ptr[v + 1] - p0;
    const uint32_t cap = (uint32_t)(2 * deg);
    int *keys = hkeys + 2 * p0;
    u64 *vals = hvals + 2 * p0;
    for (uint32_t s = threadIdx.x; s < cap; s += kLvThreads) {
        keys[s] = -1;
        vals[s] = 0;
    }
    __threadfence();
    __syncthreads();
    for (int64_t e = threadIdx.x; e < deg; e += kLvThreads)
        hash_add(keys, vals, cap, c[adj[p0 + e]], w[p0 + e]);
    __threadfence();
    __syncthreads();
    const int cv = c[v];
    const u64 kvi = k[v];
    const MoveRule r{cv, res * (double)kvi / totw};
    double bs = -INFINITY;
    int bc = INT_MAX;
    u64 kcv = 0;
    for (uint32_t s = threadIdx.x; s < cap; s += kLvThreads) {
        const int C = __hip_atomic_load(&keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (C >= 0)
            r.consider(C, __hip_atomic_load(&vals[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       wtot, bs, bc, kcv);
    }
    wave_best(bs, bc);
    kcv = wave_max_u64(kcv);
    const int wv = (int)(threadIdx.x >> 6);
    if (lane_id() == 0) {
        rs[wv] = bs;
        rc[wv] = bc;
        rk[wv] = kcv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kLvWaves; ++i) {
            best_merge(bs, bc, rs[i], rc[i]);
            kcv = rk[i] > kcv ? rk[i] : kcv;
        }
        const int to = r.decide(bs, bc, kcv, wtot[cv], kvi);
        cnew[v] = to;
        if (to != cv) atomicAdd(moved, 1);
    }
}

// nodes of kSmallDeg < deg <= kLightDeg (medium) and deg > kLightDeg (heavy)
__global__ void k_deg_lists(int64_t n, const int64_t *ptr, int32_t *medium, int32_t *heavy,
                            int *count) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int64_t deg = ptr[v + 1] - ptr[v];
    if (deg > kLightDeg) heavy[atomicAdd(&count[1], 1)] = (int32_t)v;
    else if (deg > kSmallDeg) medium[atomicAdd(&count[0], 1)] = (int32_t)v;
}

__global__ void k_fill_u64(u64 *a, int64_t n, u64 val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = val;
}

__global__ void k_fill_i32(int32_t *a, int64_t n, int32_t val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = val;
}

__global__ void k_iota(int32_t *a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

// w_tot[C] = sum of the node weights of C
__global__ void k_wtot(int64_t n, const int32_t *c, const u64 *k, u64 *wtot) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicAdd(&wtot[c[v]], k[v]);
}

// totals after a sweep from those before it (wtot_new holds a copy): only
// the nodes that moved change them (exact in two's-complement integers)
__global__ void k_wtot_moves(int64_t n, const int32_t *c, const int32_t *cnew, const u64 *k,
                             u64 *wtot_new) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int32_t a = c[v], b = cnew[v];
    if (a != b) {
        atomicAdd(&wtot_new[a], (u64)0 - k[v]);
        atomicAdd(&wtot_new[b], k[v]);
    }
}

// weight of the edges inside clusters (sum_w_in without tot_self), one
// edge per thread over a fixed grid
constexpr int kInBlocks = 2048;
__global__ void __launch_bounds__(kLvThreads)
k_mod_in(int64_t m, const int32_t *src, const int32_t *adj, const u64 *w, const int32_t *c,
         u64 *acc) {
    __shared__ u64 part[kLvWaves];
    u64 s = 0;
    for (int64_t e = (int64_t)blockIdx.x * kLvThreads + threadIdx.x; e < m;
         e += (int64_t)kInBlocks * kLvThreads)
        if (c[src[e]] == c[adj[e]]) s += w[e];
    s = wave_sum_u64(s);
    if (lane_id() == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        u64 t = 0;
        for (int i = 0; i < kLvWaves; ++i) t += part[i];
        if (t) atomicAdd(acc, t);
    }
}

// sum over clusters of (w_tot / totw)^2 in a fixed order: a fixed grid of
// kSqBlocks strided partial sums, a fixed tree per block, then one block
__global__ void __launch_bounds__(kLvThreads)
k_mod_sq(int64_t L, const u64 *wtot, double totw, double *part) {
    __shared__ double sh[kLvThreads];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kLvThreads + threadIdx.x; i < L;
         i += (int64_t)kSqBlocks * kLvThreads) {
        const double x = (double)wtot[i] / totw;
        s += x * x;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kLvThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kLvThreads) k_mod_sq_final(const double *part, double *out) {
    __shared__ double sh[kSqBlocks];
    sh[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int o = kSqBlocks / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// renumbering by first occurrence (renumber!, src/modularity.jl:289-315)
__global__ void k_first(int64_t n, const int32_t *c, int32_t *first) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicMin(&first[c[v]], (int32_t)v);
}

__global__ void k_is_first(int64_t n, const int32_t *c, const int32_t *first, int32_t *flag) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = first[c[v]] == (int32_t)v ? 1 : 0;
}

__global__ void k_relabel(int64_t n, int32_t *c, const int32_t *first, const int64_t *pos) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) c[v] = (int32_t)pos[first[c[v]]];
}

// A keyed bijection of [0, L): invertible mixing on `bits` bits with cycle
// walking back into range.
__device__ __forceinline__ uint32_t perm_label(uint32_t x, uint32_t L, int bits, uint64_t key) {
    const uint32_t mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1u);
    const uint32_t a1 = (uint32_t)key | 1u, k1 = (uint32_t)(key >> 32);
    const uint64_t key2 = key * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    const uint32_t a2 = (uint32_t)key2 | 1u, k2 = (uint32_t)(key2 >> 32);
    const int r = bits / 2 + 1;
    do {
        x = (x * a1 + k1) & mask;
        x ^= x >> r;
        x = (x * a2 + k2) & mask;
        x ^= x >> r;
    } while (x >= L);
    return x;
}

__global__ void k_permute(int64_t n, int32_t *c, uint32_t L, int bits, uint64_t key) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) c[v] = (int32_t)perm_label((uint32_t)c[v], L, bits, key);
}

__global__ void k_merge(int64_t n, int32_t *c, const int32_t *cr) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) c[v] = cr[c[v]];
}

// reduced_network (src/modularity.jl:196-241): edge (c[u], c[v], w) keys
__global__ void k_edge_keys(int64_t n, const int64_t *ptr, const int32_t *adj, const u64 *w,
                            const int32_t *c, int b, u64 *key, u64 *val) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const u64 hi = (u64)c[v] << b;
    for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
        key[e] = hi | (u64)c[adj[e]];
        val[e] = w[e];
    }
}

// flag the first entry of each run of equal keys between distinct clusters
__global__ void k_run_heads(int64_t m, const u64 *key, int b, int32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const u64 kk = key[i];
    const u64 mask = (1ull << b) - 1ull;
    const bool head = i == 0 || key[i - 1] != kk;
    flag[i] = (head && (kk >> b) != (kk & mask)) ? 1 : 0;
}

__global__ void __launch_bounds__(kLvThreads)
k_run_emit(int64_t m, const u64 *key, const u64 *val, int b, const int32_t *flag,
           const int64_t *pos, int32_t *adj_r, int32_t *src_r, u64 *w_r, int32_t *rowcnt,
           u64 *self_acc) {
    const int64_t i = (int64_t)blockIdx.x * kLvThreads + threadIdx.x;
    u64 self = 0;
    if (i < m) {
        const u64 kk = key[i];
        const u64 mask = (1ull << b) - 1ull;
        const u64 row = kk >> b, col = kk & mask;
        if (row == col) {
            self = val[i];
        } else {
            const int64_t g = pos[i] + flag[i] - 1;
            if (flag[i]) {
                adj_r[g] = (int32_t)col;
                src_r[g] = (int32_t)row;
                atomicAdd(&rowcnt[row], 1);
            }
            atomicAdd(&w_r[g], val[i]);
        }
    }
    self = wave_sum_u64(self);
    if (lane_id() == 0 && self) atomicAdd(self_acc, self);
}

__global__ void k_set_i64(int64_t *a, int64_t val) { a[0] = val; }

// Network(snn) (src/modularity.jl:165-194): the self-link of each column is
// skipped; weights go to fixed point.
__global__ void k_count_nonself(int64_t n, const int64_t *colptr, const int32_t *row,
                                int32_t *cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t t = 0;
    for (int64_t p = colptr[i]; p < colptr[i + 1]; ++p) t += row[p] != (int32_t)i ? 1 : 0;
    cnt[i] = t;
}

__global__ void k_build_net(int64_t n, const int64_t *colptr, const int32_t *row,
                            const double *val, double scale, const int64_t *ptr, int32_t *adj,
                            int32_t *src, u64 *w, u64 *k) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t e = ptr[i];
    u64 s = 0;
    for (int64_t p = colptr[i]; p < colptr[i + 1]; ++p) {
        if (row[p] == (int32_t)i) continue;
        const u64 q = (u64)llrint(val[p] * scale);
        adj[e] = row[p];
        src[e] = (int32_t)i;
        w[e] = q;
        s += q;
        ++e;
    }
    k[i] = s;
}

// largest weight as its bit pattern (ordered like the value for w >= 0)
template <class T>
__global__ void __launch_bounds__(kLvThreads) k_max_bits(int64_t m, const T *val, u64 *acc) {
    const int64_t i = (int64_t)blockIdx.x * kLvThreads + threadIdx.x;
    const double x = i < m ? (double)val[i] : 0.0;
    const u64 b = wave_max_u64((u64)__double_as_longlong(x));
    if (lane_id() == 0 && b) atomicMax(acc, b);
}

template <class T>
__global__ void k_to_f64(int64_t m, const T *in, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = (double)in[i];
}

__global__ void __launch_bounds__(kLvThreads) k_sum_u64(int64_t n, const u64 *a, u64 *acc) {
    const int64_t i = (int64_t)blockIdx.x * kLvThreads + threadIdx.x;
    const u64 s = wave_sum_u64(i < n ? a[i] : 0);
    if (lane_id() == 0 && s) atomicAdd(acc, s);
}

// group_singletons! (src/clustering.jl:79-98)
__global__ void k_count_labels(int64_t n, const int32_t *lab, int32_t *cnt) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicAdd(&cnt[lab[v]], 1);
}

__global__ void k_flag_multi(int64_t L, const int32_t *cnt, int32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L) flag[i] = cnt[i] > 1 ? 1 : 0;
}

__global__ void k_flag_single(int64_t n, const int32_t *lab, const int32_t *cnt, int32_t *flag) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = cnt[lab[v]] <= 1 ? 1 : 0;
}

__global__ void k_compact(int64_t L, const int32_t *flag, const int64_t *pos, int32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L && flag[i]) out[pos[i]] = (int32_t)i;
}

// Singles joined to another single (in either direction) keep the reference's
// sequential order; the others read no label that moves, so they are placed
// in parallel.
__global__ void k_single_dep(int64_t ns, const int32_t *singles, const int64_t *colptr,
                             const int32_t *row, const int32_t *lab, const int32_t *cnt,
                             int32_t *dep) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const int32_t j = singles[t];
    for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) {
        const int32_t i = row[p];
        if (i != j && cnt[lab[i]] <= 1) {
            dep[j] = 1;
            dep[i] = 1;
        }
    }
}

// One wave per independent single; the choice goes to choice[t] (-1 for
// dependent ones) and is applied afterwards.
__global__ void __launch_bounds__(kLvThreads)
k_group_indep(int64_t ns, const int32_t *singles, const int32_t *dep, const int64_t *colptr,
              const int32_t *row, const double *val, const int32_t *others, int64_t nothers,
              const int32_t *lab, int32_t *choice) {
    const int lane = lane_id();
    const int64_t t = (int64_t)blockIdx.x * kLvWaves + (threadIdx.x >> 6);
    if (t >= ns) return;
    const int32_t j = singles[t];
    if (dep[j]) {
        if (lane == 0) choice[t] = -1;
        return;
    }
    const int64_t p0 = colptr[j], p1 = colptr[j + 1];
    int64_t found = INT64_MAX;
    for (int64_t base = 0; base < nothers && found == INT64_MAX; base += kWave) {
        const int64_t o = base + lane;
        int64_t mine = INT64_MAX;
        if (o < nothers) {
            const int cl = others[o];
            bool none = true;
            for (int64_t p = p0; p < p1 && none; ++p) none = lab[row[p]] != cl;
            if (none) mine = o;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int64_t other = __shfl_xor(mine, off, kWave);
            mine = other < mine ? other : mine;
        }
        found = mine;
    }
    int closest;
    if (found != INT64_MAX) {
        closest = others[found];
    } else {
        double bs = -INFINITY;
        int bo = INT_MAX;
        for (int64_t o = lane; o < nothers; o += kWave) {
            const int cl = others[o];
            double sum = 0.0;
            int64_t cnt = 0;
            for (int64_t p = p0; p < p1; ++p)
                if (lab[row[p]] == cl) {
                    sum += val[p];
                    ++cnt;
                }
            best_merge(bs, bo, sum / (double)cnt, (int)o);
        }
        wave_best(bs, bo);
        closest = others[bo];
    }
    if (lane == 0) choice[t] = closest;
}

__global__ void k_apply_choice(int64_t ns, const int32_t *singles, const int32_t *choice,
                               int32_t *lab) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ns && choice[t] >= 0) lab[singles[t]] = choice[t];
}

__global__ void k_flag_dep(int64_t ns, const int32_t *singles, const int32_t *dep,
                           int32_t *flag) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ns) flag[t] = dep[singles[t]];
}

__global__ void k_compact_vals(int64_t L, const int32_t *flag, const int64_t *pos,
                               const int32_t *vals, int32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L && flag[i]) out[pos[i]] = vals[i];
}

// One block moves singles[s0..s1) in order: the first cluster of size > 1
// (ascending label) that the node has no entry into wins (the reference's
// mean over no entries is NaN, which findmax returns at once); otherwise the
// cluster with the largest mean entry weight, the first on ties.  Labels are
// read and written at device scope so that each node sees the moves before
// it.
__global__ void __launch_bounds__(kLvThreads)
k_group_singletons(int64_t s0, int64_t s1, const int32_t *singles, const int64_t *colptr,
                   const int32_t *row, const double *val, const int32_t *others, int64_t nothers,
                   int32_t *lab) {
    __shared__ int64_t sh_first[kLvWaves];
    __shared__ double sh_s[kLvWaves];
    __shared__ int sh_c[kLvWaves];
    const int wv = (int)(threadIdx.x >> 6), lane = lane_id();
    for (int64_t t = s0; t < s1; ++t) {
        const int64_t j = singles[t];
        const int64_t p0 = colptr[j], p1 = colptr[j + 1];
        int64_t found = INT64_MAX;
        for (int64_t base = 0; base < nothers && found == INT64_MAX; base += kLvThreads) {
            const int64_t o = base + threadIdx.x;
            int64_t mine = INT64_MAX;
            if (o < nothers) {
                const int cl = others[o];
                bool none = true;
                for (int64_t p = p0; p < p1 && none; ++p)
                    none = __hip_atomic_load(&lab[row[p]], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) != cl;
                if (none) mine = o;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const int64_t other = __shfl_xor(mine, off, kWave);
                mine = other < mine ? other : mine;
            }
            if (lane == 0) sh_first[wv] = mine;
            __syncthreads();
            for (int i = 0; i < kLvWaves; ++i) found = sh_first[i] < found ? sh_first[i] : found;
            __syncthreads();
        }
        int closest;
        if (found != INT64_MAX) {
            closest = others[found];
        } else {
            double bs = -INFINITY;
            int bo = INT_MAX;
            for (int64_t o = threadIdx.x; o < nothers; o += kLvThreads) {
                const int cl = others[o];
                double s = 0.0;
                int64_t cnt = 0;
                for (int64_t p = p0; p < p1; ++p)
                    if (__hip_atomic_load(&lab[row[p]], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) == cl) {
                        s += val[p];
                        ++cnt;
                    }
                best_merge(bs, bo, s / (double)cnt, (int)o);
            }
            wave_best(bs, bo);
            if (lane == 0) {
                sh_s[wv] = bs;
                sh_c[wv] = bo;
            }
            __syncthreads();
            for (int i = 0; i < kLvWaves; ++i) best_merge(bs, bo, sh_s[i], sh_c[i]);
            closest = others[bo];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            __hip_atomic_store(&lab[j], closest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        __syncthreads();
    }
}

// -------------------------------------------------------------------------
// Host side
// -------------------------------------------------------------------------

struct LvGraph {
    int64_t n = 0, nnz = 0;
    u64 tot_self = 0;
    DevBuf<int64_t> ptr;
    DevBuf<int32_t> adj, src;  // target and source node of each edge
    DevBuf<u64> w, k;
    DevBuf<int32_t> medium, heavy;
    int64_t nmedium = 0, nheavy = 0;
};

struct LvLevel {
    LvGraph g;           // reduced network of this depth (unused at depth 0)
    DevBuf<int32_t> c;   // clustering of g's nodes
};

struct Louvain {
    hipStream_t st;
    double res, min_mod;
    u64 totw = 0;
    double totw_d = 0.0;
    uint64_t seed = 0;
    uint64_t calls = 0;
    std::vector<std::unique_ptr<LvLevel>> levels;  // stable addresses across depths
    // scratch shared by all depths (used within one phase at a time)
    DevBuf<int32_t> cnew, first, flag;
    DevBuf<u64> wtot, wtot_new, key, val, key2, val2;
    DevBuf<int64_t> pos, scan_scratch;
    DevBuf<int> hkeys, counter;
    DevBuf<u64> hvals, acc;
    DevBuf<double> part, sq;
    DevBuf<char> sort_tmp;

    int64_t scan_i32(const int32_t *in, int64_t n) {
        pos.ensure((size_t)n + 1);
        scan_scratch.ensure(scan_scratch_elems(n));
        scan_exclusive_i32(in, pos.get(), n, scan_scratch.get(), st);
        int64_t last = 0;
        int32_t lf = 0;
        SEV_HIP(hipMemcpyAsync(&last, pos.get() + n - 1, sizeof last, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipMemcpyAsync(&lf, in + n - 1, sizeof lf, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        return last + lf;
    }

    void heavy_list(LvGraph &g) {
        g.medium.ensure((size_t)g.n);
        g.heavy.ensure((size_t)g.n);
        counter.ensure(2);
        SEV_HIP(hipMemsetAsync(counter.get(), 0, 2 * sizeof(int), st));
        k_deg_lists<<<grid_of(g.n), kLvThreads, 0, st>>>(g.n, g.ptr.get(), g.medium.get(),
                                                         g.heavy.get(), counter.get());
        SEV_LAUNCH_CHECK();
        int h[2] = {0, 0};
        SEV_HIP(hipMemcpyAsync(h, counter.get(), sizeof h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        g.nmedium = h[0];
        g.nheavy = h[1];
    }

    void compute_wtot(const LvGraph &g, const int32_t *c, u64 *wt) {
        k_fill_u64<<<grid_of(g.n), kLvThreads, 0, st>>>(wt, g.n, 0);
        k_wtot<<<grid_of(g.n), kLvThreads, 0, st>>>(g.n, c, g.k.get(), wt);
        SEV_LAUNCH_CHECK();
    }

    // modularity (src/modularity.jl:125-148) of clustering c with totals wt
    double modularity(const LvGraph &g, const int32_t *c, const u64 *wt) {
        acc.ensure(1);
        part.ensure(kSqBlocks);
        sq.ensure(1);
        k_fill_u64<<<1, 1, 0, st>>>(acc.get(), 1, 0);
        if (g.nnz > 0)
            k_mod_in<<<(unsigned)std::min<int64_t>(kInBlocks, grid_of(g.nnz)), kLvThreads, 0, st>>>(
                g.nnz, g.src.get(), g.adj.get(), g.w.get(), c, acc.get());
        k_mod_sq<<<kSqBlocks, kLvThreads, 0, st>>>(g.n, wt, totw_d, part.get());
        k_mod_sq_final<<<1, kSqBlocks, 0, st>>>(part.get(), sq.get());
        SEV_LAUNCH_CHECK();
        u64 in = 0;
        double s = 0.0;
        SEV_HIP(hipMemcpyAsync(&in, acc.get(), sizeof in, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipMemcpyAsync(&s, sq.get(), sizeof s, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        return (double)(g.tot_self + in) / totw_d - res * s;
    }

    // local_move! (src/modularity.jl:255-287) in parallel sweeps; returns the
    // modularity gained
    double local_move(LvGraph &g, int32_t *c) {
        const int64_t n = g.n;
        if (g.heavy.get() == nullptr) heavy_list(g);
        {
            int bits = 0;
            while (bits < 31 && (1ll << bits) < n) ++bits;
            const uint64_t key = seed ^ (0x9E3779B97F4A7C15ull * (++calls));
            k_permute<<<grid_of(n), kLvThreads, 0, st>>>(n, c, (uint32_t)n, bits, key);
            SEV_LAUNCH_CHECK();
        }
        cnew.ensure((size_t)n);
        wtot.ensure((size_t)n);
        wtot_new.ensure((size_t)n);
        counter.ensure(1);
        if (g.nheavy > 0) {
            hkeys.ensure((size_t)(2 * g.nnz));
            hvals.ensure((size_t)(2 * g.nnz));
        }
        compute_wtot(g, c, wtot.get());
        const double q0 = modularity(g, c, wtot.get());
        double q = q0, round_gain = 0.0;
        uint32_t athresh = kActAll >> 1;
        bool rejected = false;
        // the active subsets differ between sweeps, and between the phases
        // (levels and repeats) of one start
        const uint64_t pkey = (seed ^ (0xC2B2AE3D27D4EB4Full * (uint64_t)calls)) *
                              0xD1B54A32D192ED03ull;
        for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
            const uint64_t akey = pkey + (uint64_t)sweep;
            SEV_HIP(hipMemsetAsync(counter.get(), 0, sizeof(int), st));
            k_move_light<2 * kSmallDeg><<<grid_of(n, kLvWaves), kLvThreads, 0, st>>>(
                n, nullptr, g.ptr.get(), g.adj.get(), g.w.get(), g.k.get(), c, wtot.get(),
                cnew.get(), totw_d, res, akey, athresh, counter.get());
            if (g.nmedium > 0)
                k_move_light<2 * kLightDeg><<<grid_of(g.nmedium, kLvWaves), kLvThreads, 0, st>>>(
                    g.nmedium, g.medium.get(), g.ptr.get(), g.adj.get(), g.w.get(), g.k.get(), c,
                    wtot.get(), cnew.get(), totw_d, res, akey, athresh, counter.get());
            if (g.nheavy > 0)
                k_move_heavy<<<(unsigned)g.nheavy, kLvThreads, 0, st>>>(
                    g.heavy.get(), g.ptr.get(), g.adj.get(), g.w.get(), g.k.get(), c, wtot.get(),
                    cnew.get(), hkeys.get(), hvals.get(), totw_d, res, akey, athresh,
                    counter.get());
            SEV_LAUNCH_CHECK();
            int moved = 0;
            SEV_HIP(hipMemcpyAsync(&moved, counter.get(), sizeof moved, hipMemcpyDeviceToHost, st));
            SEV_HIP(hipStreamSynchronize(st));
            // a sweep in which no chosen node wants to move gains nothing; the
            // pair test below then ends the phase
            if (moved != 0) {
                SEV_HIP(hipMemcpyAsync(wtot_new.get(), wtot.get(), sizeof(u64) * (size_t)n,
                                       hipMemcpyDeviceToDevice, st));
                k_wtot_moves<<<grid_of(n), kLvThreads, 0, st>>>(n, c, cnew.get(), g.k.get(),
                                                                wtot_new.get());
                SEV_LAUNCH_CHECK();
                const double qn = modularity(g, cnew.get(), wtot_new.get());
                if (qn > q) {
                    SEV_HIP(hipMemcpyAsync(c, cnew.get(), sizeof(int32_t) * (size_t)n,
                                           hipMemcpyDeviceToDevice, st));
                    wtot.swap(wtot_new);
                    round_gain += qn - q;
                    q = qn;
                    athresh = kActAll >> 1;
                } else {
                    // simultaneous moves interfered: retry with half as many
                    // nodes, down to 1/64 of them
                    rejected = true;
                    athresh >>= 1;
                    if (athresh < (kActAll >> 6)) break;
                }
            }
            if (sweep & 1) {
                if (!rejected && round_gain < min_mod) break;
                round_gain = 0.0;
                rejected = false;
            }
        }
        return q - q0;
    }

    int64_t renumber(int64_t n, int32_t *c) {
        first.ensure((size_t)n);
        flag.ensure((size_t)n);
        k_fill_i32<<<grid_of(n), kLvThreads, 0, st>>>(first.get(), n, INT_MAX);
        k_first<<<grid_of(n), kLvThreads, 0, st>>>(n, c, first.get());
        k_is_first<<<grid_of(n), kLvThreads, 0, st>>>(n, c, first.get(), flag.get());
        SEV_LAUNCH_CHECK();
        const int64_t nc = scan_i32(flag.get(), n);
        k_relabel<<<grid_of(n), kLvThreads, 0, st>>>(n, c, first.get(), pos.get());
        SEV_LAUNCH_CHECK();
        return nc;
    }

    void reduce(const LvGraph &g, const int32_t *c, int64_t nc, LvGraph &r) {
        int b = 1;
        while ((1ll << b) < nc) ++b;
        const int64_t m = g.nnz;
        SEV_REQUIRE(m < INT_MAX, "too many edges for one radix sort");
        r.n = nc;
        r.tot_self = g.tot_self;
        r.heavy.release();
        r.medium.release();
        r.nheavy = 0;
        r.nmedium = 0;
        r.ptr.ensure((size_t)nc + 1);
        r.k.ensure((size_t)nc);
        k_fill_u64<<<grid_of(nc), kLvThreads, 0, st>>>(r.k.get(), nc, 0);
        k_wtot<<<grid_of(g.n), kLvThreads, 0, st>>>(g.n, c, g.k.get(), r.k.get());
        SEV_LAUNCH_CHECK();
        if (m == 0) {
            SEV_HIP(hipMemsetAsync(r.ptr.get(), 0, sizeof(int64_t) * (size_t)(nc + 1), st));
            r.nnz = 0;
            r.adj.ensure(1);
            r.src.ensure(1);
            r.w.ensure(1);
            return;
        }
        key.ensure((size_t)m);
        val.ensure((size_t)m);
        key2.ensure((size_t)m);
        val2.ensure((size_t)m);
        k_edge_keys<<<grid_of(g.n), kLvThreads, 0, st>>>(g.n, g.ptr.get(), g.adj.get(), g.w.get(),
                                                         c, b, key.get(), val.get());
        SEV_LAUNCH_CHECK();
        size_t tb = 0;
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.get(), key2.get(), val.get(),
                                                   val2.get(), (int)m, 0, 2 * b, st));
        sort_tmp.ensure(tb);
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp.get(), tb, key.get(), key2.get(),
                                                   val.get(), val2.get(), (int)m, 0, 2 * b, st));
        flag.ensure((size_t)m);
        k_run_heads<<<grid_of(m), kLvThreads, 0, st>>>(m, key2.get(), b, flag.get());
        SEV_LAUNCH_CHECK();
        const int64_t ne = scan_i32(flag.get(), m);
        r.nnz = ne;
        r.adj.ensure((size_t)std::max<int64_t>(ne, 1));
        r.src.ensure((size_t)std::max<int64_t>(ne, 1));
        r.w.ensure((size_t)std::max<int64_t>(ne, 1));
        DevBuf<int32_t> rowcnt;
        rowcnt.ensure((size_t)nc);
        acc.ensure(1);
        k_fill_u64<<<grid_of(ne), kLvThreads, 0, st>>>(r.w.get(), ne, 0);
        k_fill_i32<<<grid_of(nc), kLvThreads, 0, st>>>(rowcnt.get(), nc, 0);
        k_fill_u64<<<1, 1, 0, st>>>(acc.get(), 1, 0);
        k_run_emit<<<grid_of(m), kLvThreads, 0, st>>>(m, key2.get(), val2.get(), b, flag.get(),
                                                      pos.get(), r.adj.get(), r.src.get(),
                                                      r.w.get(), rowcnt.get(), acc.get());
        SEV_LAUNCH_CHECK();
        u64 self = 0;
        SEV_HIP(hipMemcpyAsync(&self, acc.get(), sizeof self, hipMemcpyDeviceToHost, st));
        scan_scratch.ensure(scan_scratch_elems(nc));
        scan_exclusive_i32(rowcnt.get(), r.ptr.get(), nc, scan_scratch.get(), st);
        k_set_i64<<<1, 1, 0, st>>>(r.ptr.get() + nc, ne);
        SEV_LAUNCH_CHECK();
        SEV_HIP(hipStreamSynchronize(st));
        r.tot_self += self;
    }

    // louvain! (src/modularity.jl:317-331) at depth d on graph g
    double level(LvGraph &g, int32_t *c, size_t d) {
        if (g.n <= 1) return 0.0;
        double gain = local_move(g, c);
        const int64_t nc = renumber(g.n, c);
        if (nc < g.n) {
            while (levels.size() <= d + 1) levels.push_back(std::make_unique<LvLevel>());
            LvLevel &L = *levels[d + 1];
            reduce(g, c, nc, L.g);
            L.c.ensure((size_t)nc);
            k_iota<<<grid_of(nc), kLvThreads, 0, st>>>(L.c.get(), nc);
            SEV_LAUNCH_CHECK();
            gain += level(L.g, L.c.get(), d + 1);
            k_merge<<<grid_of(g.n), kLvThreads, 0, st>>>(g.n, c, L.c.get());
            SEV_LAUNCH_CHECK();
        }
        return gain;
    }
};

}  // namespace

int64_t louvain_cluster(int64_t n, const int64_t *colptr_d, const int32_t *row_d,
                        const void *val_in, bool f32, int64_t nnz, double res, int64_t nstarts,
                        int64_t niter, double min_mod, uint64_t seed, bool group,
                        int32_t *labels_d, double *q_out, hipStream_t st) {
    SEV_REQUIRE(n >= 1 && n < INT_MAX, "number of nodes out of range");
    SEV_REQUIRE(nstarts >= 1 && niter >= 0, "nrandomstarts must be >= 1, niterations >= 0");
    SEV_REQUIRE(std::isfinite(res), "resolution must be finite");
    DevBuf<double> val64;
    const double *val_d = (const double *)val_in;
    u64 maxbits = 0;
    {
        DevBuf<u64> mx;
        mx.ensure(1);
        k_fill_u64<<<1, 1, 0, st>>>(mx.get(), 1, 0);
        if (f32) {
            val64.ensure((size_t)std::max<int64_t>(nnz, 1));
            k_to_f64<<<grid_of(nnz), kLvThreads, 0, st>>>(nnz, (const float *)val_in,
                                                          val64.get());
            val_d = val64.get();
        }
        k_max_bits<<<grid_of(nnz), kLvThreads, 0, st>>>(nnz, val_d, mx.get());
        SEV_LAUNCH_CHECK();
        SEV_HIP(hipMemcpyAsync(&maxbits, mx.get(), sizeof maxbits, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
    }
    double wmax = 0.0;
    std::memcpy(&wmax, &maxbits, sizeof wmax);
    const double wbound = wmax * (double)nnz;
    SEV_REQUIRE(std::isfinite(wbound), "graph weights must be finite");
    int e2 = 0;
    if (wbound > 0.0) (void)std::frexp(wbound, &e2);
    const int S = std::min(40, 61 - e2);
    SEV_REQUIRE(S >= 0, "total edge weight too large for the fixed-point sums");
    Louvain lv;
    lv.st = st;
    lv.res = res;
    lv.min_mod = min_mod;
    lv.seed = seed;
    lv.levels.push_back(std::make_unique<LvLevel>());
    LvGraph &g = lv.levels[0]->g;
    g.n = n;
    g.ptr.ensure((size_t)n + 1);
    g.k.ensure((size_t)n);
    {
        DevBuf<int32_t> cnt;
        cnt.ensure((size_t)n);
        k_count_nonself<<<grid_of(n), kLvThreads, 0, st>>>(n, colptr_d, row_d, cnt.get());
        SEV_LAUNCH_CHECK();
        lv.scan_scratch.ensure(scan_scratch_elems(n));
        scan_exclusive_i32(cnt.get(), g.ptr.get(), n, lv.scan_scratch.get(), st);
        int64_t last = 0;
        int32_t lc = 0;
        SEV_HIP(hipMemcpyAsync(&last, g.ptr.get() + n - 1, sizeof last, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipMemcpyAsync(&lc, cnt.get() + n - 1, sizeof lc, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        g.nnz = last + lc;
        k_set_i64<<<1, 1, 0, st>>>(g.ptr.get() + n, g.nnz);
        g.adj.ensure((size_t)std::max<int64_t>(g.nnz, 1));
        g.src.ensure((size_t)std::max<int64_t>(g.nnz, 1));
        g.w.ensure((size_t)std::max<int64_t>(g.nnz, 1));
        k_build_net<<<grid_of(n), kLvThreads, 0, st>>>(n, colptr_d, row_d, val_d,
                                                       std::ldexp(1.0, S), g.ptr.get(),
                                                       g.adj.get(), g.src.get(), g.w.get(),
                                                       g.k.get());
        lv.acc.ensure(1);
        k_fill_u64<<<1, 1, 0, st>>>(lv.acc.get(), 1, 0);
        k_sum_u64<<<grid_of(n), kLvThreads, 0, st>>>(n, g.k.get(), lv.acc.get());
        SEV_LAUNCH_CHECK();
        SEV_HIP(hipMemcpyAsync(&lv.totw, lv.acc.get(), sizeof(u64), hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        lv.totw_d = (double)lv.totw;
    }
    double qbest = NAN;
    if (lv.totw == 0) {
        // no edges between distinct nodes: every node stays alone
        k_iota<<<grid_of(n), kLvThreads, 0, st>>>(labels_d, n);
        SEV_LAUNCH_CHECK();
    } else {
        DevBuf<int32_t> cur;
        cur.ensure((size_t)n);
        for (int64_t s = 0; s < nstarts; ++s) {
            lv.seed = seed * 0x9E3779B97F4A7C15ull + (uint64_t)s * 0xD1B54A32D192ED03ull + 1;
            k_iota<<<grid_of(n), kLvThreads, 0, st>>>(cur.get(), n);
            SEV_LAUNCH_CHECK();
            double gain = lv.level(g, cur.get(), 0);
            for (int64_t it = 1; gain >= min_mod && it <= niter; ++it)
                gain = lv.level(g, cur.get(), 0);
            lv.compute_wtot(g, cur.get(), lv.wtot.ensure((size_t)n));
            const double q = lv.modularity(g, cur.get(), lv.wtot.get());
            if (s == 0 || qbest < q) {
                qbest = q;
                SEV_HIP(hipMemcpyAsync(labels_d, cur.get(), sizeof(int32_t) * (size_t)n,
                                       hipMemcpyDeviceToDevice, st));
            }
        }
    }
    int64_t nc = lv.renumber(n, labels_d);
    if (group) {
        DevBuf<int32_t> cnt, others, singles;
        cnt.ensure((size_t)nc);
        k_fill_i32<<<grid_of(nc), kLvThreads, 0, st>>>(cnt.get(), nc, 0);
        k_count_labels<<<grid_of(n), kLvThreads, 0, st>>>(n, labels_d, cnt.get());
        lv.flag.ensure((size_t)std::max(n, nc));
        k_flag_single<<<grid_of(n), kLvThreads, 0, st>>>(n, labels_d, cnt.get(), lv.flag.get());
        SEV_LAUNCH_CHECK();
        const int64_t nsingle = lv.scan_i32(lv.flag.get(), n);
        if (nsingle > 0) {
            singles.ensure((size_t)nsingle);
            k_compact<<<grid_of(n), kLvThreads, 0, st>>>(n, lv.flag.get(), lv.pos.get(),
                                                         singles.get());
            k_flag_multi<<<grid_of(nc), kLvThreads, 0, st>>>(nc, cnt.get(), lv.flag.get());
            SEV_LAUNCH_CHECK();
            const int64_t nothers = lv.scan_i32(lv.flag.get(), nc);
            SEV_REQUIRE(nothers > 0,
                        "group_singletons: no cluster has more than one member (every node is "
                        "alone; the reference's findmax over no clusters fails)");
            others.ensure((size_t)nothers);
            k_compact<<<grid_of(nc), kLvThreads, 0, st>>>(nc, lv.flag.get(), lv.pos.get(),
                                                          others.get());
            SEV_LAUNCH_CHECK();
            DevBuf<int32_t> dep, choice, seq;
            dep.ensure((size_t)n);
            choice.ensure((size_t)nsingle);
            k_fill_i32<<<grid_of(n), kLvThreads, 0, st>>>(dep.get(), n, 0);
            k_single_dep<<<grid_of(nsingle), kLvThreads, 0, st>>>(
                nsingle, singles.get(), colptr_d, row_d, labels_d, cnt.get(), dep.get());
            k_group_indep<<<grid_of(nsingle, kLvWaves), kLvThreads, 0, st>>>(
                nsingle, singles.get(), dep.get(), colptr_d, row_d, val_d, others.get(), nothers,
                labels_d, choice.get());
            k_apply_choice<<<grid_of(nsingle), kLvThreads, 0, st>>>(nsingle, singles.get(),
                                                                    choice.get(), labels_d);
            k_flag_dep<<<grid_of(nsingle), kLvThreads, 0, st>>>(nsingle, singles.get(), dep.get(),
                                                                lv.flag.get());
            SEV_LAUNCH_CHECK();
            const int64_t ndep = lv.scan_i32(lv.flag.get(), nsingle);
            if (ndep > 0) {
                seq.ensure((size_t)ndep);
                k_compact_vals<<<grid_of(nsingle), kLvThreads, 0, st>>>(
                    nsingle, lv.flag.get(), lv.pos.get(), singles.get(), seq.get());
                SEV_LAUNCH_CHECK();
            }
            for (int64_t s0 = 0; s0 < ndep; s0 += kSinglesPerLaunch) {
                const int64_t s1 = std::min(ndep, s0 + kSinglesPerLaunch);
                k_group_singletons<<<1, kLvThreads, 0, st>>>(s0, s1, seq.get(), colptr_d, row_d,
                                                             val_d, others.get(), nothers,
                                                             labels_d);
                SEV_LAUNCH_CHECK();
            }
            nc = lv.renumber(n, labels_d);
        }
    }
    SEV_HIP(hipStreamSynchronize(st));
    if (q_out) *q_out = qbest;
    return nc;
}

}  // namespace sev
