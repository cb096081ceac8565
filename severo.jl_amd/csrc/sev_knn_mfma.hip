// Exact kNN by MFMA candidate search + exact re-rank + certificate.
//
// The result is identical to the exact definition of knn_exact (sev_knn.hip):
// candidates are proposed by a matrix-core distance GEMM, then every
// candidate is re-ranked with the exact FP32 distance, and a rigorous error
// bound certifies that no candidate outside the list could enter the top k.
// Rows that fail the certificate are recomputed by the exhaustive exact
// kernel.
//
// Distance GEMM.  With q' = [q, 1, (1)] and c' = [c, -|c|^2/2 split over one
// (FP32) or two (BF16 hi + lo) extra columns], dot(q', c') = q.c - |c|^2/2,
// so |q - c|^2 = |q|^2 - 2 dot(q', c'): per query the ranking is by dot'
// descending and the epilogue is one compare per pair.
//   FP32 form: v_mfma_f32_16x16x4_f32, operands the FP32 embedding.
//   BF16 form: v_mfma_f32_16x16x32_bf16, operands rounded to BF16 (per point
//              rounding error eps = |bf16(z) - z|_2 is kept), FP32 accumulate.
//
//   FP8 form:  v_mfma_f32_16x16x32_fp8_fp8 on e4m3 operands, each point
//              scaled by the power of two sigma of its 128-row block (block
//              max |z| / sigma <= 240); the score is sigma_q sigma_c acc -
//              |c^|^2/2 with the norm term kept in FP32 (no norm columns).
//
// Certificate.  Every candidate outside a query's final list has
// dot' <= thr (the list's K'-th value), i.e. d~ >= d~_K' = |q|^2 - 2 thr.
// The computed d~ differs from |q^ - c^|^2 by at most
// eta (|q^|^2 + max|c^|^2) + 2^-23 d~ (FP32 accumulation over KP terms and
// the split of |c|^2/2), |q - c| >= |q^ - c^| - eps_q - max eps, and the
// exact sequential FP32 distance is >= |q - c|^2 (1 - 2^-24)^(d+3).  If that
// lower bound exceeds the k-th exact distance among the list, the top k is
// exact.
#include "sev_kernels.hpp"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <type_traits>
#include <cstdlib>
#include <cmath>
#include <algorithm>
#include <vector>

namespace sev {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kBlock = 256;
constexpr int kNW = kBlock / kWave;  // waves per block
constexpr int kCT = 64;              // candidates per tile

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float ninf() { return -__builtin_huge_valf(); }

// ---------------------------------------------------------------------------
// Operand preparation (thread per point).  Z column-major (x_it at Z[t*ld+i]).
// FP32: Zq[i] = [z, 1, 0..], Zc[i] = [z, -n/2, 0..], n = |z|^2.
// BF16: Zq[i] = [zb, 1, 1, 0..], Zc[i] = [zb, hi, lo, 0..], hi + lo = -|zb|^2/2.
// Also the row-major FP32 copy Zr used by the exact re-rank, n^ (FP32,
// rounded up) and eps (FP32, rounded up), and the global maxima of both.
// ---------------------------------------------------------------------------
// Block maxima for the FP8 form: bmax[i >> 7] = max_t |z_it| (float bits).
__global__ void __launch_bounds__(kBlock)
k_knn_blockmax(const float *Z, int64_t n, int64_t d, int64_t ld, unsigned *bmax) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        float mx = 0.0f;
        for (int64_t t = 0; t < d; ++t) mx = fmaxf(mx, fabsf(Z[t * ld + i]));
        atomicMax(&bmax[i >> 7], __float_as_uint(mx));
    }
}

// FP8 (e4m3) of x (|x| <= 240) and back, by the hardware conversions.
__device__ __forceinline__ uint32_t f2fp8x2(float a, float b) {
    return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false) & 0xFFFFu;
}

template <int MODE, int KP>
__global__ void __launch_bounds__(kBlock)
k_knn_prep(const float *Z, int64_t n, int64_t d, int64_t ld, void *Zq, void *Zc, float *Zr,
           float *nhat, float *eps, unsigned *gmax, const unsigned *bmax, float *scf,
           float *hn) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        double nn = 0.0, ee = 0.0;
        if (MODE == 2) {
            // sigma = the power of two with block max / sigma in (120, 240]
            const float bm = __uint_as_float(bmax[i >> 7]);
            float sig = 1.0f;
            if (bm > 0.0f) {
                int e;
                (void)frexpf(bm / 240.0f, &e);
                sig = ldexpf(1.0f, e);
            }
            uint8_t *q = (uint8_t *)Zq + i * KP;
            for (int64_t t = 0; t < d; t += 2) {
                const float z0 = Z[t * ld + i];
                const float z1 = t + 1 < d ? Z[(t + 1) * ld + i] : 0.0f;
                const int pk = (int)f2fp8x2(z0 / sig, z1 / sig);
                const float c0 = __builtin_amdgcn_cvt_f32_fp8(pk, 0) * sig;
                const float c1 = __builtin_amdgcn_cvt_f32_fp8(pk, 1) * sig;
                q[t] = (uint8_t)(pk & 0xFF);
                Zr[i * d + t] = z0;
                nn += (double)c0 * (double)c0;
                ee += ((double)c0 - (double)z0) * ((double)c0 - (double)z0);
                if (t + 1 < d) {
                    q[t + 1] = (uint8_t)(pk >> 8);
                    Zr[i * d + t + 1] = z1;
                    nn += (double)c1 * (double)c1;
                    ee += ((double)c1 - (double)z1) * ((double)c1 - (double)z1);
                }
            }
            for (int64_t t = d; t < KP; ++t) q[t] = 0;
            scf[i] = sig;
            hn[i] = (float)(-0.5 * nn);
        } else if (MODE == 0) {
            float *q = (float *)Zq + i * KP;
            float *c = (float *)Zc + i * KP;
            for (int64_t t = 0; t < d; ++t) {
                const float z = Z[t * ld + i];
                q[t] = z;
                c[t] = z;
                Zr[i * d + t] = z;
                nn += (double)z * (double)z;
            }
            q[d] = 1.0f;
            c[d] = (float)(-0.5 * nn);
            for (int64_t t = d + 1; t < KP; ++t) {
                q[t] = 0.0f;
                c[t] = 0.0f;
            }
        } else {
            // the row is assembled in registers (pairs of BF16 per word) and
            // written as 16-byte pieces: 2-byte stores strided by the row
            // length left most of each written line to be merged
            uint32_t wz[KP / 2];
#pragma unroll
            for (int u = 0; u < KP / 2; ++u) wz[u] = 0u;
            float *zr = Zr + i * d;
#pragma unroll
            for (int t = 0; t < KP; ++t) {
                if (t < d) {
                    const float z = Z[(int64_t)t * ld + i];
                    const uint16_t b = f2bf(z);
                    const float zb = bf2f(b);
                    wz[t / 2] |= (uint32_t)b << (16 * (t & 1));
                    zr[t] = z;
                    nn += (double)zb * (double)zb;
                    const double e = (double)zb - (double)z;
                    ee += e * e;
                }
            }
            const float half = (float)(-0.5 * nn);
            const uint16_t hi = f2bf(half);
            const uint16_t lo = f2bf((float)((double)half - (double)bf2f(hi)));
            uint32_t wq[KP / 2], wc[KP / 2];
#pragma unroll
            for (int u = 0; u < KP / 2; ++u) {
                uint32_t q2 = wz[u], c2 = wz[u];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int t = 2 * u + h;
                    if (t == d || t == d + 1) {
                        q2 |= 0x3F80u << (16 * h);  // 1.0
                        c2 |= (uint32_t)(t == d ? hi : lo) << (16 * h);
                    }
                }
                wq[u] = q2;
                wc[u] = c2;
            }
            if constexpr (KP % 8 == 0) {  // the BF16 row lengths (64, 128)
                uint4 *q4 = (uint4 *)((uint16_t *)Zq + i * KP);
                uint4 *c4 = (uint4 *)((uint16_t *)Zc + i * KP);
#pragma unroll
                for (int u = 0; u < KP / 8; ++u) {
                    q4[u] = make_uint4(wq[4 * u], wq[4 * u + 1], wq[4 * u + 2], wq[4 * u + 3]);
                    c4[u] = make_uint4(wc[4 * u], wc[4 * u + 1], wc[4 * u + 2], wc[4 * u + 3]);
                }
            } else {
                uint32_t *q1 = (uint32_t *)((uint16_t *)Zq + i * KP);
                uint32_t *c1 = (uint32_t *)((uint16_t *)Zc + i * KP);
#pragma unroll
                for (int u = 0; u < KP / 2; ++u) {
                    q1[u] = wq[u];
                    c1[u] = wc[u];
                }
            }
        }
        const float nf = (float)(nn * (1.0 + 1e-6)) + 1e-30f;
        const float ef = (float)(sqrt(ee) * (1.0 + 1e-6));
        nhat[i] = nf;
        eps[i] = ef;
        atomicMax(&gmax[0], __float_as_uint(nf));
        atomicMax(&gmax[1], __float_as_uint(ef));
    }
}

// ---------------------------------------------------------------------------
// Candidate search.  A block owns QB = kNW*QW*16 queries (wave w owns
// QW row-blocks of 16); candidate tiles of kCT are staged in LDS and
// multiplied with MFMA; each query keeps its KL largest dot' values (ties by
// lower index) in LDS, updated by one wavefront-wide insertion per passing
// pair.
// ---------------------------------------------------------------------------
template <int MODE>
struct Elem;
template <>
struct Elem<0> {
    typedef float T;
};
template <>
struct Elem<1> {
    typedef uint16_t T;
};
template <>
struct Elem<2> {
    typedef uint8_t T;
};

// Queries: rows q0..q0+QB-1 of the query list (qidx[i], or i itself when
// qidx is null; nq entries; qidx[i] < 0 marks an unused slot).  Candidates:
// chunk blockIdx.y of gridDim.y equal chunks of [0, n) (multiples of the
// tile), or, with seg, of the rows seg[s] .. seg[s+1]-1 of the sample s of
// the block's first query (batches: every query of a block is from s).
// Outputs at [chunk][list index]: cidx (KL ids), cthr, ccnt, and the list
// values in cval when it is given.  cthr bounds every candidate left out of
// the list: dot' <= cthr (-inf: every candidate is in the list).
//
// Candidate order.  With perm, row p of Zc is the point perm[p] (the points
// by ascending |c|^2, so the candidates with the largest dot' for a typical
// query come first) and the lists hold perm[p]; positions past the chunk
// carry id -1.  The lists are the KL best under (dot' desc, id asc) in any
// order; the order only changes how many insertions it takes to get there.
//
// Seed.  With nseed > 0 every query first sweeps the head [0, nseed) of the
// candidates without lists, keeping the largest dot' of each of the 64
// residues of the tile position, and starts its threshold at the 32nd
// largest of those 64 maxima: a value reached by 32 real candidates, so
// admitting only candidates above it loses none of the final list whenever
// the list fills, and the list's threshold stays a bound for the candidates
// left out when it does not.
constexpr int kSeedRank = 32;
template <int MODE, int KP, int QW, int KL>
__global__ void __launch_bounds__(kBlock)
k_knn_cand(const void *Zq_, const void *Zc_, int64_t n, int include_self, const int64_t *qidx,
           int64_t nq, int32_t *cidx, float *cthr, int32_t *ccnt, float *cval,
           const int64_t *seg, const int32_t *row_sample, const float *scf, const float *hn,
           const int32_t *perm, int64_t nseed) {
    typedef typename Elem<MODE>::T E;
    static_assert(KL % 16 == 0 && KL >= 32 && KL <= kWave, "lists of 32, 48 or 64");
    constexpr int QB = kNW * QW * 16;
    // LDS row stride (elements): odd for FP32; BF16 rows unpadded, their
    // 16-byte chunks stored at chunk ^ (row mod chunks per row), so the 8
    // lanes of a 16-byte read phase meet 8 different bank groups; FP8 rows
    // padded by 16 bytes
    constexpr int ROW = MODE == 0 ? KP + 1 : (MODE == 1 ? KP : KP + 16);
    constexpr int KS = MODE == 0 ? KP / 4 : KP / 32;  // MFMA steps per block
    constexpr int NCB = kCT / 16;                     // candidate blocks per tile
    __shared__ __attribute__((aligned(16))) E tiles[2][kCT * ROW];
    // FP8: per-candidate scale and -|c^|^2/2 of each tile
    __shared__ float tilef[2][MODE == 2 ? kCT : 1], tileh[2][MODE == 2 ? kCT : 1];
    __shared__ float listV[QB][KL];
    __shared__ int32_t listI[QB][KL];
    __shared__ int listN[QB];
    __shared__ int32_t qg[QB];  // query point of each list row (-1: none)
    __shared__ float seedv[QB];  // starting threshold of each list row
    __shared__ int32_t tileid[2][kCT];  // with perm: the point of each tile row

    const E *Zq = (const E *)Zq_;
    const E *Zc = (const E *)Zc_;
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int64_t q0 = (int64_t)blockIdx.x * QB;
    const int col = lane & 15, grp = lane >> 4;

    for (int e = threadIdx.x; e < QB * KL; e += kBlock) {
        (&listV[0][0])[e] = ninf();
        (&listI[0][0])[e] = INT_MAX;
    }
    for (int i = threadIdx.x; i < QB; i += kBlock) {
        listN[i] = 0;
        seedv[i] = ninf();
        const int64_t qi = q0 + i;
        qg[i] = qi < nq ? (int32_t)(qidx ? qidx[qi] : qi) : -1;
    }
    // query operands and the running thresholds of this lane's output rows
    // (rows 4*grp + i of each row block) in registers; rows past n never
    // accept a candidate
    float af[QW][MODE == 0 ? KS : 1];
    v8bf ab[QW][MODE == 1 ? KS : 1];
    long a8[QW][MODE == 2 ? KS : 1];
    float sq[QW][MODE == 2 ? 4 : 1];  // FP8: scale of the lane's output rows
    float thr[QW][4];
    __syncthreads();
    // candidate chunk
    int64_t rb = 0, re = n;
    if (seg) {
        const int32_t q = qg[0];
        const int32_t smp = q >= 0 ? row_sample[q] : 0;
        rb = q >= 0 ? seg[smp] : 0;
        re = q >= 0 ? seg[smp + 1] : 0;
    }
    const int64_t csz = ((re - rb + gridDim.y - 1) / gridDim.y + kCT - 1) / kCT * kCT;
    const int64_t cbeg = rb + (int64_t)blockIdx.y * csz;
    const int64_t cend = cbeg + csz < re ? cbeg + csz : re;
    // the seed head: [0, nseed), or with seg the lowest-norm 1/64 of the
    // block's sample (nseed > 0 then only enables it)
    const int64_t sh0 = seg ? rb : 0;
    if (seg && nseed > 0) nseed = ((re - rb) / 64) / kCT * kCT;
#pragma unroll
    for (int rb = 0; rb < QW; ++rb) {
        const int32_t qpt = qg[w * QW * 16 + rb * 16 + col];
        const bool ok = qpt >= 0;
        const int64_t qrow = ok ? qpt : 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int32_t qo = qg[w * QW * 16 + rb * 16 + 4 * grp + i];
            thr[rb][i] = qo >= 0 ? ninf() : __builtin_huge_valf();
            if constexpr (MODE == 2) sq[rb][i] = qo >= 0 ? scf[qo] : 1.0f;
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if constexpr (MODE == 2) {
                a8[rb][s] = ok ? *(const long *)((const uint8_t *)Zq + qrow * KP + 32 * s + 8 * grp)
                               : 0L;
            } else if (MODE == 0) {
                af[rb][s] = ok ? ((const float *)Zq)[qrow * KP + 4 * s + grp] : 0.0f;
            } else {
                v8bf v;
                if (ok) {
                    v = *(const v8bf *)((const uint16_t *)Zq + qrow * KP + 32 * s + 8 * grp);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.0f;
                }
                ab[rb][s] = v;
            }
        }
    }

    // Candidate tiles are double-buffered: the loads of tile t+1 are issued
    // into registers before tile t is multiplied and land in the other LDS
    // buffer afterwards, so one barrier per tile suffices.
    // 16-byte loads: 4 floats (FP32), 8 bf16 or 16 fp8 values per load
    constexpr int EPL = MODE == 0 ? 4 : (MODE == 1 ? 8 : 16);  // elements per load
    constexpr int CH = KP / EPL;                               // loads per candidate row
    constexpr int NLD = (kCT * CH + kBlock - 1) / kBlock;      // loads per thread
    static_assert(KP % EPL == 0, "row length must be a multiple of 16 bytes");
    static_assert(MODE != 1 || (CH & (CH - 1)) == 0, "BF16 rows of a power-of-two chunk count");
    // two register stage sets: the main loop keeps the loads of two tiles
    // in flight (the candidate rows mostly come from the MALL)
    uint4 stage[2][NLD];
    float fstage[2] = {1.0f, 1.0f}, hstage[2] = {0.0f, 0.0f};
    int32_t idstage[2] = {-1, -1};
    auto fetch = [&](int64_t t0, int64_t lim, int k) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < NLD; ++u) {
            const int e = threadIdx.x + u * kBlock;
            const int r = e / CH, part = e % CH;
            const int64_t c = t0 + r;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (e < kCT * CH && c < lim)
                v = *(const uint4 *)((const E *)Zc + c * KP + part * EPL);
            stage[k][u] = v;
        }
        if (perm && threadIdx.x < kCT) {
            const int64_t c = t0 + threadIdx.x;
            idstage[k] = c < lim ? perm[c] : -1;
        }
        if constexpr (MODE == 2) {
            const int64_t c = t0 + threadIdx.x;
            if (threadIdx.x < kCT) {
                fstage[k] = c < lim ? scf[c] : 1.0f;
                hstage[k] = c < lim ? hn[c] : 0.0f;
            }
        }
    };
    auto deposit = [&](E *tile, int b, int k) __attribute__((always_inline)) {
        if (perm && threadIdx.x < kCT) tileid[b][threadIdx.x] = idstage[k];
        if constexpr (MODE == 2) {
            if (threadIdx.x < kCT) {
                tilef[b][threadIdx.x] = fstage[k];
                tileh[b][threadIdx.x] = hstage[k];
            }
        }
#pragma unroll
        for (int u = 0; u < NLD; ++u) {
            const int e = threadIdx.x + u * kBlock;
            if (e >= kCT * CH) break;
            const int r = e / CH, part = e % CH;
            if (MODE == 0) {
                float *d = (float *)tile + r * ROW + part * 4;  // odd row stride: scalar stores
                d[0] = __uint_as_float(stage[k][u].x);
                d[1] = __uint_as_float(stage[k][u].y);
                d[2] = __uint_as_float(stage[k][u].z);
                d[3] = __uint_as_float(stage[k][u].w);
            } else {
                const int pc = MODE == 1 ? (part ^ (r & (CH - 1))) : part;
                *(uint4 *)(&tile[r * ROW + pc * EPL]) = stage[k][u];
            }
        }
    };
    // all products of a tile: acc[cb][rb] = 16 queries x 16 candidates
    auto products = [&](const E *tile, int b, v4f (&acc)[NCB][QW]) __attribute__((always_inline)) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            float bfv[MODE == 0 ? KS : 1];
            v8bf bbv[MODE == 1 ? KS : 1];
            long b8[MODE == 2 ? KS : 1];
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if constexpr (MODE == 2)
                    b8[s] = *(const long *)((const uint8_t *)tile + (cb * 16 + col) * ROW +
                                            32 * s + 8 * grp);
                else if (MODE == 0)
                    bfv[s] = ((const float *)tile)[(cb * 16 + col) * ROW + 4 * s + grp];
                else
                    bbv[s] = *(const v8bf *)((const uint16_t *)tile + (cb * 16 + col) * ROW +
                                             ((4 * s + grp) ^ ((cb * 16 + col) & (CH - 1))) *
                                                 EPL);
            }
#pragma unroll
            for (int rb = 0; rb < QW; ++rb) {
                v4f a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    if constexpr (MODE == 2)
                        a = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a8[rb][s], b8[s], a, 0, 0,
                                                                       0);
                    else if (MODE == 0)
                        a = __builtin_amdgcn_mfma_f32_16x16x4f32(af[rb][s], bfv[s], a, 0, 0, 0);
                    else
                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab[rb][s], bbv[s], a, 0, 0, 0);
                }
                if constexpr (MODE == 2) {
                    // score = sigma_q sigma_c (q^'.c^') - |c^|^2/2 (the scales are
                    // powers of two: their product is exact)
                    const float fc = tilef[b][cb * 16 + col], hc = tileh[b][cb * 16 + col];
#pragma unroll
                    for (int i = 0; i < 4; ++i) a[i] = fmaf(a[i], sq[rb][i] * fc, hc);
                }
                acc[cb][rb] = a;
            }
        }
    };

    // seed: the head of the candidates without lists (see above)
    if (nseed > 0) {
        // the 64 residues: column cb*16 + col of tile parity h
        constexpr int NRES = kWave / kCT;
        static_assert(NRES * kCT == kWave, "tiles of 16, 32 or 64 candidates");
        float mx[NRES][NCB][QW][4];
#pragma unroll
        for (int h = 0; h < NRES; ++h)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int rb = 0; rb < QW; ++rb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) mx[h][cb][rb][i] = ninf();
        fetch(sh0, sh0 + nseed, 0);
        deposit(tiles[0], 0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        int sb = 0;
        for (int64_t t0 = sh0; t0 < sh0 + nseed; t0 += kCT, sb ^= 1) {
            const bool more = t0 + kCT < sh0 + nseed;
            if (more) fetch(t0 + kCT, sh0 + nseed, 0);
            v4f acc[NCB][QW];
            products(tiles[sb], sb, acc);
            const int par = (int)(((t0 - sh0) / kCT) % NRES);
#pragma unroll
            for (int h = 0; h < NRES; ++h) {
                if (h != par) continue;
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                    for (int rb = 0; rb < QW; ++rb)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            mx[h][cb][rb][i] = fmaxf(mx[h][cb][rb][i], acc[cb][rb][i]);
            }
            if (more) deposit(tiles[sb ^ 1], sb ^ 1, 0);
            __syncthreads();
        }
        // the 64 maxima of each row into its (still empty) list storage:
        // residue (h*NCB + cb)*16 + col at that slot of listV, continuing in
        // listI for lists of 32
#pragma unroll
        for (int rb = 0; rb < QW; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = w * QW * 16 + rb * 16 + 4 * grp + i;
#pragma unroll
                for (int h = 0; h < NRES; ++h)
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) {
                        const int slot = (h * NCB + cb) * 16 + col;
                        if (slot < KL)
                            listV[row][slot] = mx[h][cb][rb][i];
                        else
                            listI[row][slot - KL] = __float_as_int(mx[h][cb][rb][i]);
                    }
            }
        __syncthreads();
        // each wavefront: the kSeedRank-th largest of its rows' maxima
        for (int r = 0; r < QW * 16; ++r) {
            const int row = w * QW * 16 + r;
            float v = lane < KL ? listV[row][lane] : __int_as_float(listI[row][lane - KL]);
            for (int size = 2; size <= kWave; size <<= 1) {
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    const float ov = __shfl_xor(v, stride, kWave);
                    const bool up = (lane & size) == 0;
                    const bool lower = (lane & stride) == 0;
                    // descending: the lower position of an ascending pair
                    // keeps the larger value
                    if ((lower == up) ? ov > v : v > ov) v = ov;
                }
            }
            const float sd = __shfl(v, kSeedRank - 1, kWave);
            if (lane == 0) seedv[row] = qg[row] >= 0 ? sd : ninf();
        }
        __syncthreads();
        for (int e = threadIdx.x; e < QB * KL; e += kBlock) {
            (&listV[0][0])[e] = ninf();
            (&listI[0][0])[e] = INT_MAX;
        }
#pragma unroll
        for (int rb = 0; rb < QW; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = w * QW * 16 + rb * 16 + 4 * grp + i;
                if (qg[row] >= 0) thr[rb][i] = seedv[row];
            }
        __syncthreads();
    }
    // tile t is in LDS buffer t % 2 and its rows were loaded into stage set
    // t % 2 two tiles earlier
    fetch(cbeg, cend, 0);
    if (cbeg + kCT < cend) fetch(cbeg + kCT, cend, 1);
    deposit(tiles[0], 0, 0);
    // every global load issued so far (query operands, first two tiles) is
    // complete here, so the loop only waits for its own prefetches
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // one tile: buffer and stage set buf; the next tile's rows (stage set
    // buf ^ 1) go to the other buffer at the end
    auto tile_step = [&](int64_t t0, int buf) __attribute__((always_inline)) {
        const E *tile = tiles[buf];
        const bool more = t0 + kCT < cend;
        if (t0 + 2 * kCT < cend) fetch(t0 + 2 * kCT, cend, buf);
        // the tile holds padding columns, or (self pairs excluded) may hold
        // this block's own queries: such pairs are rejected in the insertion
        // path (their products may pass the threshold test)
        const bool edge = t0 + kCT > cend ||
                          (!include_self && (qidx || (t0 < q0 + QB && q0 < t0 + kCT)));
        const int t32 = (int)t0, n32 = (int)cend;
        v4f acc[NCB][QW];
        products(tile, buf, acc);
#pragma unroll
        for (int rb = 0; rb < QW; ++rb) {
            // one test per output row for the whole tile: the largest of its
            // 4 x 16 products against the row's threshold (IEEE maximum, so
            // the MFMA results are not canonicalised first; the tests are
            // combined as wave masks)
            uint64_t any = 0, anyi[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float m = __builtin_elementwise_maximum(
                    __builtin_elementwise_maximum(acc[0][rb][i], acc[1][rb][i]),
                    __builtin_elementwise_maximum(acc[2][rb][i], acc[3][rb][i]));
                anyi[i] = __ballot(m > thr[rb][i]);
                any |= anyi[i];
            }
            if (!any) continue;
            // the candidate id of this lane's column in each candidate block
            int32_t cids[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
                cids[cb] = perm ? tileid[buf][cb * 16 + col] : t32 + cb * 16 + col;
            // only the output rows i whose tile maximum passed can hold a
            // passing pair (thresholds only rise); their pairs are visited
            // block by block (the lists do not depend on the order)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!anyi[i]) continue;
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    const int32_t cid = cids[cb];
                    // the product as an integer for readlane (on vector
                    // elements it was seen to read element 0 for every element)
                    const int ai = __float_as_int(acc[cb][rb][i]);
                    uint64_t mask = __ballot(acc[cb][rb][i] > thr[rb][i]);
                    while (mask) {
                        const int src = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const int g = src >> 4;
                        const float v = __int_as_float(__builtin_amdgcn_readlane(ai, src));
                        // the row's threshold lives in the lanes of group g
                        const float ts =
                            __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr[rb][i]),
                                                                     16 * g));
                        if (!(v > ts)) continue;
                        const int qlo = w * QW * 16 + rb * 16 + 4 * g + i;
                        const int32_t cand = __builtin_amdgcn_readlane(cid, src);
                        if (perm) {
                            if (cand < 0 || (!include_self && cand == qg[qlo])) continue;
                        } else if (edge && (cand >= n32 || (!include_self && cand == qg[qlo]))) {
                            continue;
                        }
                        // insert (v, cand) into the descending list of qlo:
                        // entries from pos on move one slot up by LDS stores
                        const int cnt = listN[qlo];
                        const float lv = lane < KL ? listV[qlo][lane] : ninf();
                        const int32_t li = lane < KL ? listI[qlo][lane] : INT_MAX;
                        const bool better = lane < cnt && (lv > v || (lv == v && li < cand));
                        const int pos = __popcll(__ballot(better));
                        const int ncnt = cnt + 1 < KL ? cnt + 1 : KL;
                        if (lane >= pos && lane + 1 < ncnt) {
                            listV[qlo][lane + 1] = lv;
                            listI[qlo][lane + 1] = li;
                        }
                        if (lane == pos) {
                            listV[qlo][pos] = v;
                            listI[qlo][pos] = cand;
                        }
                        const float last =
                            pos == KL - 1 ? v
                                          : __int_as_float(__builtin_amdgcn_readlane(
                                                __float_as_int(lv), KL - 2));
                        // a list that is not full keeps its starting threshold
                        const float nthr = ncnt == KL ? last : ts;
                        if (lane == 0) listN[qlo] = ncnt;
                        if (grp == g) thr[rb][i] = nthr;
                    }
                }
            }
        }
        if (more) deposit(tiles[buf ^ 1], buf ^ 1, buf ^ 1);
        __syncthreads();
    };
    for (int64_t t0 = cbeg; t0 < cend;) {
        tile_step(t0, 0);
        t0 += kCT;
        if (t0 >= cend) break;
        tile_step(t0, 1);
        t0 += kCT;
    }
    for (int ql = w; ql < QB; ql += kNW) {
        const int64_t qi = q0 + ql;
        if (qi >= nq) continue;
        const int64_t o = (int64_t)blockIdx.y * nq + qi;
        const int cnt = listN[ql];
        if (lane < KL) {
            cidx[o * KL + lane] = lane < cnt ? listI[ql][lane] : -1;
            if (cval) cval[o * KL + lane] = listV[ql][lane];
        }
        if (lane == 0) {
            cthr[o] = cnt == KL ? listV[ql][KL - 1] : seedv[ql];
            ccnt[o] = cnt;
        }
    }
}

// ---------------------------------------------------------------------------
// Exact re-rank and certificate (one wavefront per query).
// ---------------------------------------------------------------------------
// (v, id) ranks before (ov, oid): larger value, ties by lower id.
__device__ __forceinline__ bool ranks_before(float v, int32_t id, float ov, int32_t oid) {
    return v > ov || (v == ov && id < oid);
}

// Bitonic sort of the 128 pairs held two per lane (position e*64 + lane)
// into rank order (ranks_before first).
__device__ __forceinline__ void wave_rank_sort2(float (&v)[2], int32_t (&id)[2]) {
    const int lane = lane_id();
    for (int size = 2; size <= 2 * kWave; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride == kWave) {
                // partner in the other register of the lane (size 128: the
                // lower position keeps the better pair)
                if (ranks_before(v[1], id[1], v[0], id[0])) {
                    const float tv = v[0];
                    const int32_t ti = id[0];
                    v[0] = v[1];
                    id[0] = id[1];
                    v[1] = tv;
                    id[1] = ti;
                }
                continue;
            }
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int pos = e * kWave + lane;
                const float ov = __shfl_xor(v[e], stride, kWave);
                const int32_t oi = __shfl_xor(id[e], stride, kWave);
                const bool up = (pos & size) == 0;
                const bool lower = (pos & stride) == 0;
                const bool take = (lower == up) ? ranks_before(ov, oi, v[e], id[e])
                                                : ranks_before(v[e], id[e], ov, oi);
                if (take) {
                    v[e] = ov;
                    id[e] = oi;
                }
            }
        }
    }
}

// Merge the per-chunk candidate lists of each query (one wavefront per list
// index): the running top KL (lanes 0..KL-1) and the next chunk's list
// (lanes KL..2KL-1) are sorted together by (value desc, id asc) and the first
// KL kept.  The merged threshold is the KL-th value when KL candidates were
// kept over all chunks.  Every candidate outside the merged list lies below
// its chunk's threshold, which is at most the merged one when KL were kept;
// otherwise the merged threshold is the largest chunk threshold (a chunk's
// list that is not full has its starting threshold).  For KL > 32 the
// running list is one register of the lanes and the chunk's list another
// (the running register keeps up to 64 entries; its first KL are the list).
template <int KL>
__global__ void __launch_bounds__(kBlock)
k_knn_merge(int64_t nq, int nchunk, const int32_t *pidx, const float *pval, const int32_t *pcnt,
            const float *pthr, int32_t *cidx, float *cthr, int32_t *ccnt) {
    static_assert(KL <= kWave, "a list must fit a wavefront");
    const int lane = lane_id();
    if constexpr (2 * KL > kWave) {
        const int64_t qi = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
        if (qi >= nq) return;
        float v[2] = {ninf(), ninf()};
        int32_t id[2] = {INT_MAX, INT_MAX};
        int tot = 0;
        float floor = ninf();
        for (int c = 0; c < nchunk; ++c) {
            const int64_t o = (int64_t)c * nq + qi;
            const int cnt = pcnt[o];
            tot += cnt;
            floor = fmaxf(floor, pthr[o]);
            v[1] = lane < cnt ? pval[o * KL + lane] : ninf();
            id[1] = lane < cnt ? pidx[o * KL + lane] : INT_MAX;
            wave_rank_sort2(v, id);
        }
        const int cnt = tot < KL ? tot : KL;
        if (lane < KL) cidx[qi * KL + lane] = lane < cnt ? id[0] : -1;
        const float last = __shfl(v[0], KL - 1, kWave);
        if (lane == 0) {
            ccnt[qi] = cnt;
            cthr[qi] = tot >= KL ? fmaxf(last, floor) : floor;
        }
        return;
    }
    const int64_t qi = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (qi >= nq) return;
    float v = ninf();
    int32_t id = INT_MAX;
    int tot = 0;
    float floor = ninf();
    for (int c = 0; c < nchunk; ++c) {
        const int64_t o = (int64_t)c * nq + qi;
        const int cnt = pcnt[o];
        tot += cnt;
        floor = fmaxf(floor, pthr[o]);
        if (lane >= KL && lane < 2 * KL) {
            const int l = lane - KL;
            v = l < cnt ? pval[o * KL + l] : ninf();
            id = l < cnt ? pidx[o * KL + l] : INT_MAX;
        } else if (lane >= 2 * KL) {
            v = ninf();
            id = INT_MAX;
        }
        // bitonic sort of the 64 lanes, descending by value, ascending by id
        for (int size = 2; size <= kWave; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const float ov = __shfl_xor(v, stride, kWave);
                const int32_t oi = __shfl_xor(id, stride, kWave);
                const bool other_first = ov > v || (ov == v && oi < id);
                const bool mine_first = v > ov || (v == ov && id < oi);
                const bool up = (lane & size) == 0;
                const bool lower = (lane & stride) == 0;
                if ((lower == up) ? other_first : mine_first) {
                    v = ov;
                    id = oi;
                }
            }
        }
    }
    const int cnt = tot < KL ? tot : KL;
    if (lane < KL) cidx[qi * KL + lane] = lane < cnt ? id : -1;
    const float last = __shfl(v, KL - 1, kWave);
    if (lane == 0) {
        ccnt[qi] = cnt;
        cthr[qi] = tot >= KL ? fmaxf(last, floor) : floor;
    }
}

// Queries: list index i maps to the point qidx[i] (or i when qidx is null);
// the candidate list of i is row i of cidx/cthr/ccnt.
template <int KL>
__global__ void __launch_bounds__(kBlock)
k_knn_rerank(const float *Zr, int64_t nq, const int64_t *qidx, int d, int k, const int32_t *cidx,
             const float *cthr, const int32_t *ccnt, const float *nhat, const float *eps,
             const unsigned *gmax, double eta, double spec_factor, int32_t *oidx, float *odist,
             int32_t *nfail, int64_t *fail) {
    const int lane = lane_id();
    const int64_t qi = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (qi >= nq) return;
    const int64_t q = qidx ? qidx[qi] : qi;
    if (q < 0) return;  // unused slot of a padded query list
    const int cnt = ccnt[qi];
    const int32_t c = lane < cnt && lane < KL ? cidx[qi * KL + lane] : -1;
    float D = __builtin_huge_valf();
    if (c >= 0) {
        const float *zq = Zr + q * d;
        const float *zc = Zr + (int64_t)c * d;
        float acc = 0.0f;
        if ((d & 1) == 0) {
            // even d: rows start on 8-byte boundaries (Zr is a device
            // allocation), so the coordinates are read in pairs; the sum
            // still runs over t in order
            const float2 *q2 = (const float2 *)zq, *c2 = (const float2 *)zc;
            for (int t = 0; t < d / 2; ++t) {
                const float2 a = q2[t], b = c2[t];
                const float d0 = a.x - b.x;
                acc = acc + d0 * d0;
                const float d1 = a.y - b.y;
                acc = acc + d1 * d1;
            }
        } else {
            for (int t = 0; t < d; ++t) {
                const float diff = zq[t] - zc[t];
                acc = acc + diff * diff;
            }
        }
        D = acc;
    }
    int32_t ci = c >= 0 ? c : INT_MAX;
    // bitonic sort of (D, ci) ascending across the 64 lanes
    for (int size = 2; size <= kWave; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const float oD = __shfl_xor(D, stride, kWave);
            const int32_t oc = __shfl_xor(ci, stride, kWave);
            const bool up = (lane & size) == 0;
            const bool lower = (lane & stride) == 0;
            const bool o_less = oD < D || (oD == D && oc < ci);
            const bool o_greater = oD > D || (oD == D && oc > ci);
            const bool take = (lower == up) ? o_less : o_greater;
            if (take) {
                D = oD;
                ci = oc;
            }
        }
    }
    const float Dk = __shfl(D, k - 1, kWave);
    const float th = cthr[qi];
    bool ok;
    if (th == ninf()) {
        ok = cnt >= k;  // every admissible candidate was seen and re-ranked
    } else {
        const double ncmax = (double)__uint_as_float(gmax[0]);
        const double emax = (double)__uint_as_float(gmax[1]);
        const double nq_ = (double)nhat[q];
        const double dt = nq_ - 2.0 * (double)th;
        const double lb2 = dt * (1.0 - 0x1p-22) - eta * (nq_ + ncmax) - 1e-30;
        const double lb = lb2 > 0.0 ? sqrt(lb2) - (double)eps[q] - emax : -1.0;
        ok = lb > 0.0 && lb * lb * spec_factor > (double)Dk;
    }
    ok = ok && !(Dk != Dk);
    if (ok) {
        if (lane < k) {
            oidx[q * k + lane] = ci;
            odist[q * k + lane] = sqrtf(D);
        }
    } else if (lane == 0) {
        fail[atomicAdd(nfail, 1)] = q;
    }
}

// ---------------------------------------------------------------------------
// Self-test of the MFMA operand layouts assumed above: computes a 16x16 tile
// with MFMA and with scalar code, returns the largest difference.
// ---------------------------------------------------------------------------
__global__ void k_selftest_mfma(int mode, float *out) {
    const int lane = lane_id();
    const int col = lane & 15, grp = lane >> 4;
    auto A = [](int m, int kk) { return (float)((m * 7 + kk * 3) % 11) - 5.0f; };
    auto B = [](int kk, int nn) { return (float)((kk * 5 + nn * 9) % 13) - 6.0f; };
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    int K;
    if (mode == 3) {
        // FP8 e4m3: small integers are exact; lanes hold 8 consecutive k
        K = 32;
        int lo = 0, hi = 0, blo = 0, bhi = 0;
        for (int e = 0; e < 8; e += 2) {
            const int pa = (int)f2fp8x2(A(col, 8 * grp + e), A(col, 8 * grp + e + 1));
            const int pb = (int)f2fp8x2(B(8 * grp + e, col), B(8 * grp + e + 1, col));
            if (e < 4) {
                lo |= pa << (8 * e);
                blo |= pb << (8 * e);
            } else {
                hi |= pa << (8 * (e - 4));
                bhi |= pb << (8 * (e - 4));
            }
        }
        const long a8 = (long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
        const long b8 = (long)(((uint64_t)(uint32_t)bhi << 32) | (uint32_t)blo);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a8, b8, acc, 0, 0, 0);
    } else if (mode == 2) {
        typedef double v4d __attribute__((ext_vector_type(4)));
        v4d a64 = {0.0, 0.0, 0.0, 0.0};
        a64 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)A(col, grp), (double)B(grp, col), a64,
                                                    0, 0, 0);
        for (int i = 0; i < 4; ++i) acc[i] = (float)a64[i];
        K = 4;
    } else if (mode == 0) {
        K = 4;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A(col, grp), B(grp, col), acc, 0, 0, 0);
    } else {
        K = 32;
        v8bf a, b;
        for (int e = 0; e < 8; ++e) {
            a[e] = (__bf16)A(col, 8 * grp + e);
            b[e] = (__bf16)B(8 * grp + e, col);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    float err = 0.0f;
    for (int i = 0; i < 4; ++i) {
        // FP32/BF16 16x16: lane holds rows 4*(l/16) + i; FP64 16x16x4: (l/16) + 4i
        const int m = mode == 2 ? grp + 4 * i : 4 * grp + i, nn = col;
        float ref = 0.0f;
        for (int kk = 0; kk < K; ++kk) ref += A(m, kk) * B(kk, nn);
        err = fmaxf(err, fabsf(ref - acc[i]));
    }
    for (int o = 32; o > 0; o >>= 1) err = fmaxf(err, __shfl_xor(err, o, kWave));
    if (lane == 0) out[0] = err;
}

}  // namespace

double mfma_selftest(int mode) {
    DevBuf<float> o;
    o.ensure(1);
    hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(kWave), 0, 0, mode, o.get());
    SEV_LAUNCH_CHECK();
    float h = -1.0f;
    SEV_HIP(hipMemcpy(&h, o.get(), sizeof h, hipMemcpyDeviceToHost));
    return h;
}

namespace {

__global__ void __launch_bounds__(kBlock)
k_norm_keys(const float *nhat, int64_t n, unsigned *key, int32_t *id) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        key[i] = __float_as_uint(nhat[i]);  // positive: the bits order as the values
        id[i] = (int32_t)i;
    }
}

__global__ void __launch_bounds__(kBlock)
k_norm_keys_seg(const float *nhat, const int32_t *row_sample, int64_t n, uint64_t *key,
                int32_t *id) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        key[i] = ((uint64_t)(uint32_t)row_sample[i] << 32) | __float_as_uint(nhat[i]);
        id[i] = (int32_t)i;
    }
}

// dst row p = src row perm[p]; rows of ch 16-byte chunks
__global__ void __launch_bounds__(kBlock)
k_gather_rows(const uint4 *src, const int32_t *perm, int64_t n, int ch, uint4 *dst) {
    const int64_t tot = n * ch;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t p = e / ch;
        const int part = (int)(e - p * ch);
        dst[e] = src[(int64_t)perm[p] * ch + part];
    }
}

// Candidate order for the search: w.perm = the points by ascending |c|^2
// (the stable radix sort keeps ties in index order) and w.zc's rows put in
// that order.
void norm_order(int64_t n, size_t row_bytes, KnnWork &w, hipStream_t st,
                const int32_t *row_sample = nullptr, int64_t nsamp = 0) {
    SEV_REQUIRE(row_bytes % 16 == 0, "knn: operand rows must be 16-byte multiples");
    ProfScope ps("knn_order", st, 2.0 * n * row_bytes + 24.0 * n);
    w.sval.ensure((size_t)n);
    w.perm.ensure((size_t)n);
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div(n, kBlock), device_cu_count() * 8);
    size_t tb = 0;
    if (row_sample) {
        // a batch: by (sample, |c|^2), so every sample's rows stay in its
        // own row range (the samples are stacked in order)
        w.skey64.ensure((size_t)n);
        w.skey64b.ensure((size_t)n);
        hipLaunchKernelGGL(k_norm_keys_seg, dim3(g), dim3(kBlock), 0, st, w.nhat.get(),
                           row_sample, n, w.skey64.get(), w.sval.get());
        SEV_LAUNCH_CHECK();
        int end_bit = 33;
        while (end_bit < 64 && (nsamp >> (end_bit - 32)) != 0) ++end_bit;
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, w.skey64.get(), w.skey64b.get(),
                                                   w.sval.get(), w.perm.get(), (int)n, 0,
                                                   end_bit, st));
        w.stmp.ensure(tb);
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(w.stmp.get(), tb, w.skey64.get(),
                                                   w.skey64b.get(), w.sval.get(), w.perm.get(),
                                                   (int)n, 0, end_bit, st));
    } else {
        w.skey.ensure((size_t)n);
        w.skey2.ensure((size_t)n);
        hipLaunchKernelGGL(k_norm_keys, dim3(g), dim3(kBlock), 0, st, w.nhat.get(), n,
                           w.skey.get(), w.sval.get());
        SEV_LAUNCH_CHECK();
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, w.skey.get(), w.skey2.get(),
                                                   w.sval.get(), w.perm.get(), (int)n, 0, 32, st));
        w.stmp.ensure(tb);
        SEV_HIP(hipcub::DeviceRadixSort::SortPairs(w.stmp.get(), tb, w.skey.get(), w.skey2.get(),
                                                   w.sval.get(), w.perm.get(), (int)n, 0, 32, st));
    }
    const int ch = (int)(row_bytes / 16);
    w.zcp.ensure((size_t)n * row_bytes);
    const unsigned g2 =
        (unsigned)std::min<int64_t>(ceil_div(n * ch, kBlock), device_cu_count() * 16);
    hipLaunchKernelGGL(k_gather_rows, dim3(g2), dim3(kBlock), 0, st,
                       (const uint4 *)w.zc.get(), w.perm.get(), n, ch, (uint4 *)w.zcp.get());
    SEV_LAUNCH_CHECK();
    w.zc.swap(w.zcp);
}

// Head of the ordered candidates swept for the seed thresholds: 1/64 of the
// points in whole tiles (none below one tile).
int64_t seed_head(int64_t n) { return (n / 64) / kCT * kCT; }

// Candidate search + merge (chunked) + re-rank for a query list.  Rows that
// fail the certificate are appended to w.fail (count in w.nfail).  order:
// 0 = candidates in index order; 1 = candidates by ascending norm (made
// here, prep required) with seed thresholds; 2 = the same with the norm
// order the last order-1 call made on the same operands.  With seg (a
// batch of nsamp samples) the order is by norm within each sample.
template <int MODE, int KP, int QW, int KL>
void cand_rerank(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                 const int64_t *qidx, int64_t nq, KnnWork &w, int32_t *idx, float *dist,
                 int32_t *nfail, int64_t *fail, hipStream_t st, const int64_t *seg = nullptr,
                 const int32_t *row_sample = nullptr, bool prep = true, int order = 0,
                 int64_t nsamp = 0) {
    const size_t es = MODE == 0 ? 4 : (MODE == 1 ? 2 : 1);
    w.zq.ensure((size_t)n * KP * es);
    if (MODE != 2) w.zc.ensure((size_t)n * KP * es);  // FP8: one operand array for both
    w.zr.ensure((size_t)(n * d));
    w.nhat.ensure((size_t)n);
    w.eps.ensure((size_t)n);
    w.gmax.ensure(2);
    if (MODE == 2) {
        w.bmax.ensure((size_t)ceil_div(n, 128));
        w.scf.ensure((size_t)n);
        w.hn.ensure((size_t)n);
    }
    if (prep) {
        SEV_HIP(hipMemsetAsync(w.gmax.get(), 0, 2 * sizeof(unsigned), st));
        ProfScope ps("knn_prep", st, 4.0 * n * d * 2 + 2.0 * n * KP * es);
        const unsigned g = (unsigned)std::min<int64_t>(ceil_div(n, kBlock), device_cu_count() * 8);
        if (MODE == 2) {
            SEV_HIP(hipMemsetAsync(w.bmax.get(), 0, sizeof(unsigned) * ceil_div(n, 128), st));
            hipLaunchKernelGGL(k_knn_blockmax, dim3(g), dim3(kBlock), 0, st, Z, n, d, ld,
                               w.bmax.get());
            SEV_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL((k_knn_prep<MODE, KP>), dim3(g), dim3(kBlock), 0, st, Z, n, d, ld,
                           (void *)w.zq.get(), (void *)(MODE == 2 ? w.zq.get() : w.zc.get()),
                           w.zr.get(), w.nhat.get(), w.eps.get(), w.gmax.get(),
                           (const unsigned *)w.bmax.get(), w.scf.get(), w.hn.get());
        SEV_LAUNCH_CHECK();
    }
    const int32_t *perm = nullptr;
    int64_t nseed = 0;
    if (order != 0) {
        SEV_REQUIRE(MODE != 2, "knn: norm order needs separate operands");
        SEV_REQUIRE(order == 2 || prep, "knn: norm order is made with the operands");
        if (order == 1) norm_order(n, (size_t)KP * es, w, st, seg ? row_sample : nullptr, nsamp);
        perm = w.perm.get();
        nseed = seg ? 1 : seed_head(n);  // with seg the kernel sizes each sample's head
    }
    const void *zc_op = MODE == 2 ? (const void *)w.zq.get() : (const void *)w.zc.get();
    constexpr int QB = kNW * QW * 16;
    const int64_t qblocks = ceil_div(nq, QB);
    // split the candidates when the query blocks alone cannot fill the device
    int nchunk = 1;
    const int64_t want = (int64_t)device_cu_count() * 8;
    if (KL <= kWave && qblocks < want)
        nchunk = (int)std::min<int64_t>(64, std::max<int64_t>(1, want / qblocks));
    nchunk = (int)std::min<int64_t>(nchunk, std::max<int64_t>(1, n / (4 * kCT)));
    w.cidx.ensure((size_t)nq * KL);
    w.cthr.ensure((size_t)nq);
    w.ccnt.ensure((size_t)nq);
    {
        ProfScope ps(MODE == 0 ? "knn_cand_fp32"
                               : (MODE == 1 ? (KL == 64 ? "knn_cand_bf16_kl64"
                                                        : (KL == 48 ? "knn_cand_bf16_kl48"
                                                                    : "knn_cand_bf16"))
                                            : "knn_cand_fp8"),
                     st, 0.0,
                     2.0 * (double)nq * (double)n * (double)d);
        if (nchunk == 1) {
            hipLaunchKernelGGL((k_knn_cand<MODE, KP, QW, KL>), dim3((unsigned)qblocks),
                               dim3(kBlock), 0, st, (const void *)w.zq.get(), zc_op, n,
                               include_self, qidx, nq, w.cidx.get(),
                               w.cthr.get(), w.ccnt.get(), (float *)nullptr, seg, row_sample,
                               w.scf.get(), w.hn.get(), perm, nseed);
            SEV_LAUNCH_CHECK();
        } else if constexpr (KL <= kWave) {
            w.pidx.ensure((size_t)nchunk * nq * KL);
            w.pval.ensure((size_t)nchunk * nq * KL);
            w.pthr.ensure((size_t)nchunk * nq);
            w.pcnt.ensure((size_t)nchunk * nq);
            hipLaunchKernelGGL((k_knn_cand<MODE, KP, QW, KL>), dim3((unsigned)qblocks, nchunk),
                               dim3(kBlock), 0, st, (const void *)w.zq.get(), zc_op, n,
                               include_self, qidx, nq, w.pidx.get(),
                               w.pthr.get(), w.pcnt.get(), w.pval.get(), seg, row_sample,
                               w.scf.get(), w.hn.get(), perm, nseed);
            SEV_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_knn_merge<KL>), dim3((unsigned)ceil_div(nq * kWave, kBlock)),
                               dim3(kBlock), 0, st, nq, nchunk, w.pidx.get(), w.pval.get(),
                               w.pcnt.get(), w.pthr.get(), w.cidx.get(), w.cthr.get(),
                               w.ccnt.get());
            SEV_LAUNCH_CHECK();
        }
    }
    {
        // FP32: accumulation + the FP32 norm column; BF16: + the hi/lo split of
        // the norm; FP8: accumulation of exact products, the exact power-of-two
        // descale, the FMA and the FP32 norm term
        const double eta = 4.0 * (KP + 2) * 0x1p-24 + (MODE == 1 ? 0x1p-14 : 0x1p-22);
        const double spec = pow(1.0 - 0x1p-24, (double)(d + 3));
        ProfScope ps("knn_rerank", st, (double)nq * KL * (4.0 + 4.0 * d));
        hipLaunchKernelGGL((k_knn_rerank<KL>), dim3((unsigned)ceil_div(nq * kWave, kBlock)),
                           dim3(kBlock), 0, st, w.zr.get(), nq, qidx, (int)d, (int)k,
                           w.cidx.get(), w.cthr.get(), w.ccnt.get(), w.nhat.get(), w.eps.get(),
                           w.gmax.get(), eta, spec, idx, dist, nfail, fail);
        SEV_LAUNCH_CHECK();
    }
}

template <int MODE, int KP, int QW, int KL, int KP32>
void run_mfma(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
              KnnWork &w, int32_t *idx, float *dist, hipStream_t st, int64_t *n_fallback) {
    w.nfail.ensure(3);
    w.fail.ensure((size_t)n);
    SEV_HIP(hipMemsetAsync(w.nfail.get(), 0, 3 * sizeof(int32_t), st));
    cand_rerank<MODE, KP, QW, KL>(Z, n, d, ld, k, include_self, nullptr, n, w, idx, dist,
                                  w.nfail.get(), w.fail.get(), st, nullptr, nullptr, true, 1);
    int32_t nf = 0;
    SEV_HIP(hipMemcpyAsync(&nf, w.nfail.get(), sizeof nf, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    const int64_t *left = w.fail.get();
    if constexpr (MODE == 1 && KL < 64 && KP32 > 0) {
        if (nf > 0) {
            // BF16 rows without a certificate: BF16 candidates again with a
            // list of 64 (the operands of the first pass), whose larger gap
            // to the k-th neighbour certifies most of them
            w.fail2.ensure((size_t)nf);
            SEV_HIP(hipMemcpyAsync(w.fail2.get(), w.fail.get(), sizeof(int64_t) * nf,
                                   hipMemcpyDeviceToDevice, st));
            cand_rerank<1, KP, 1, 64>(Z, n, d, ld, k, include_self, w.fail2.get(), nf, w, idx,
                                      dist, w.nfail.get() + 2, w.fail.get(), st, nullptr,
                                      nullptr, false, 2);
            SEV_HIP(hipMemcpyAsync(&nf, w.nfail.get() + 2, sizeof nf, hipMemcpyDeviceToHost, st));
            SEV_HIP(hipStreamSynchronize(st));
        }
    }
    if constexpr (MODE == 1 && KP32 > 0) {
        if (nf > 0) {
            // rows still without a certificate: FP32 candidates for those rows
            w.fail2.ensure((size_t)nf);
            SEV_HIP(hipMemcpyAsync(w.fail2.get(), w.fail.get(), sizeof(int64_t) * nf,
                                   hipMemcpyDeviceToDevice, st));
            cand_rerank<0, KP32, (KL > 32 ? 1 : 2), KL>(Z, n, d, ld, k, include_self,
                                                        w.fail2.get(), nf, w, idx,
                                        dist, w.nfail.get() + 1, w.fail.get(), st);
            SEV_HIP(hipMemcpyAsync(&nf, w.nfail.get() + 1, sizeof nf, hipMemcpyDeviceToHost, st));
            SEV_HIP(hipStreamSynchronize(st));
        }
    }
    if (nf > 0)
        knn_exact_rows<float>(Z, n, d, ld, k, include_self, left, nf, idx, dist, st);
    if (n_fallback) *n_fallback = nf;
}

// FP8 candidates, then for the rows without a certificate BF16 candidates,
// then FP32 candidates, then the exhaustive kernel.
template <int KP8, int QW, int KL, int KP16, int KP32>
void run_mfma_fp8(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                  KnnWork &w, int32_t *idx, float *dist, hipStream_t st, int64_t *n_fallback,
                  int64_t *passes) {
    w.nfail.ensure(4);
    w.fail.ensure((size_t)n);
    w.fail2.ensure((size_t)n);
    SEV_HIP(hipMemsetAsync(w.nfail.get(), 0, 4 * sizeof(int32_t), st));
    cand_rerank<2, KP8, QW, KL>(Z, n, d, ld, k, include_self, nullptr, n, w, idx, dist,
                                w.nfail.get(), w.fail.get(), st);
    auto count = [&](int c) {
        int32_t nf = 0;
        SEV_HIP(hipMemcpyAsync(&nf, w.nfail.get() + c, sizeof nf, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        return nf;
    };
    int32_t nf = count(0);
    if (passes) passes[0] = nf;
    if (nf > 0) {
        SEV_HIP(hipMemcpyAsync(w.fail2.get(), w.fail.get(), sizeof(int64_t) * nf,
                               hipMemcpyDeviceToDevice, st));
        cand_rerank<1, KP16, QW, KL>(Z, n, d, ld, k, include_self, w.fail2.get(), nf, w, idx,
                                     dist, w.nfail.get() + 1, w.fail.get(), st);
        nf = count(1);
    }
    if (passes) passes[1] = nf;
    if constexpr (KP32 > 0) {
        if (nf > 0) {
            SEV_HIP(hipMemcpyAsync(w.fail2.get(), w.fail.get(), sizeof(int64_t) * nf,
                                   hipMemcpyDeviceToDevice, st));
            cand_rerank<0, KP32, QW, KL>(Z, n, d, ld, k, include_self, w.fail2.get(), nf, w, idx,
                                         dist, w.nfail.get() + 2, w.fail.get(), st);
            nf = count(2);
        }
    }
    if (passes) passes[2] = nf;
    if (nf > 0)
        knn_exact_rows<float>(Z, n, d, ld, k, include_self, w.fail.get(), nf, idx, dist, st);
    if (n_fallback) *n_fallback = nf;
}

}  // namespace

void knn_mfma_fp8(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                  KnnWork &w, int32_t *idx, float *dist, hipStream_t st, int64_t *n_fallback,
                  int64_t *passes) {
    SEV_REQUIRE(k >= 1 && k <= kMaxK, "knn: k must be in [1, 64]");
    SEV_REQUIRE(d >= 1 && d + 2 <= 64, "knn: the FP8 form supports d <= 62");
    SEV_REQUIRE(n < ((int64_t)1 << 31) - 256, "knn: at most 2^31 - 256 points");
    if (k <= 24) {
        if (d <= 32 && d + 1 <= 52)
            run_mfma_fp8<32, 2, 32, 64, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                            n_fallback, passes);
        else if (d + 1 <= 52)
            run_mfma_fp8<64, 2, 32, 64, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                            n_fallback, passes);
        else
            run_mfma_fp8<64, 2, 32, 64, 0>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                           n_fallback, passes);
    } else {
        if (d <= 32 && d + 1 <= 52)
            run_mfma_fp8<32, 1, 64, 64, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                            n_fallback, passes);
        else if (d + 1 <= 52)
            run_mfma_fp8<64, 1, 64, 64, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                            n_fallback, passes);
        else
            run_mfma_fp8<64, 1, 64, 64, 0>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                           n_fallback, passes);
    }
}

void knn_mfma(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
              int bf16, KnnWork &w, int32_t *idx, float *dist, hipStream_t st,
              int64_t *n_fallback) {
    SEV_REQUIRE(k >= 1 && k <= kMaxK, "knn: k must be in [1, 64]");
    SEV_REQUIRE(d >= 1, "knn: d must be >= 1");
    SEV_REQUIRE(n < ((int64_t)1 << 31) - 256, "knn: at most 2^31 - 256 points");
    if (bf16) {
        SEV_REQUIRE(d + 2 <= 128, "knn: BF16 form supports d <= 126");
        if (k <= 24 && d + 2 <= 64 && d + 1 <= 52)
            run_mfma<1, 64, 2, 32, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st, n_fallback);
        else if (k <= 40 && d + 2 <= 64 && d + 1 <= 52)
            run_mfma<1, 64, 2, 48, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st, n_fallback);
        else if (k <= 24 && d + 2 <= 64)
            run_mfma<1, 64, 2, 32, 128>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                        n_fallback);
        else if (k <= 24)
            run_mfma<1, 128, 2, 32, 128>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                         n_fallback);
        else if (d + 2 <= 64 && d + 1 <= 52)
            run_mfma<1, 64, 1, 64, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                       n_fallback);
        else if (d + 2 <= 64)
            run_mfma<1, 64, 1, 64, 0>(Z, n, d, ld, k, include_self, w, idx, dist, st, n_fallback);
        else
            run_mfma<1, 128, 1, 64, 0>(Z, n, d, ld, k, include_self, w, idx, dist, st, n_fallback);
    } else {
        SEV_REQUIRE(d + 1 <= 128, "knn: FP32 MFMA form supports d <= 127");
        if (d + 1 <= 52) {
            if (k <= 24)
                run_mfma<0, 52, 2, 32, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                           n_fallback);
            else if (k <= 40)
                run_mfma<0, 52, 2, 48, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                           n_fallback);
            else
                run_mfma<0, 52, 1, 64, 52>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                           n_fallback);
        } else {
            if (k <= 24) {
                run_mfma<0, 128, 1, 32, 128>(Z, n, d, ld, k, include_self, w, idx, dist, st,
                                             n_fallback);
            } else {
                // wide rows with long lists exceed the LDS budget: exhaustive
                knn_exact<float>(Z, n, d, ld, k, include_self, nullptr, n, idx, dist, st);
                if (n_fallback) *n_fallback = n;
            }
        }
    }
}

namespace {

// Query list of a batch: the given rows (all rows when rows is null; else
// sorted ascending) grouped by sample, each sample's group padded to a
// multiple of QB -- with -1 (unused slots) or with copies of its last row.
std::vector<int64_t> batch_queries(const BatchLayout &L, const int64_t *rows, int64_t nrows,
                                   int QB, bool pad_copy) {
    std::vector<int64_t> q;
    auto pad = [&](int64_t last) {
        while (q.size() % (size_t)QB) q.push_back(pad_copy ? last : -1);
    };
    if (!rows) {
        q.reserve((size_t)(L.M + L.S * QB));
        for (int64_t s = 0; s < L.S; ++s) {
            for (int64_t r = L.row0[(size_t)s]; r < L.row0[(size_t)s + 1]; ++r) q.push_back(r);
            pad(L.row0[(size_t)s + 1] - 1);
        }
        return q;
    }
    int64_t s = 0;
    for (int64_t i = 0; i < nrows; ++i) {
        const int64_t r = rows[i];
        if (r >= L.row0[(size_t)s + 1]) {
            if (!q.empty()) pad(q.back());
            while (r >= L.row0[(size_t)s + 1]) ++s;
        }
        q.push_back(r);
    }
    if (!q.empty()) pad(q.back());
    return q;
}

void upload_sync(DevBuf<int64_t> &b, const std::vector<int64_t> &h, hipStream_t st) {
    b.ensure(h.size() ? h.size() : 1);
    if (!h.empty())
        SEV_HIP(hipMemcpyAsync(b.get(), h.data(), sizeof(int64_t) * h.size(),
                               hipMemcpyHostToDevice, st));
    SEV_HIP(hipStreamSynchronize(st));
}

std::vector<int64_t> read_fails(const int64_t *fail, int32_t nf, hipStream_t st) {
    std::vector<int64_t> f((size_t)nf);
    if (nf > 0)
        SEV_HIP(hipMemcpyAsync(f.data(), fail, sizeof(int64_t) * (size_t)nf,
                               hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    std::sort(f.begin(), f.end());
    return f;
}

template <int MODE, int KP, int QW, int KL, int KP32>
void run_mfma_batch(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
                    const BatchLayout &L, KnnWork &w, int32_t *idx, float *dist, hipStream_t st,
                    int64_t *n_fallback) {
    const int64_t *seg = L.row0_d.get();
    const int32_t *rs = L.row_sample_d.get();
    w.nfail.ensure(2);
    w.fail.ensure((size_t)n);
    SEV_HIP(hipMemsetAsync(w.nfail.get(), 0, 2 * sizeof(int32_t), st));
    const std::vector<int64_t> q = batch_queries(L, nullptr, 0, kNW * QW * 16, false);
    upload_sync(w.qpad, q, st);
    cand_rerank<MODE, KP, QW, KL>(Z, n, d, ld, k, include_self, w.qpad.get(), (int64_t)q.size(),
                                  w, idx, dist, w.nfail.get(), w.fail.get(), st, seg, rs, true,
                                  1, L.S);
    int32_t nf = 0;
    SEV_HIP(hipMemcpyAsync(&nf, w.nfail.get(), sizeof nf, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    std::vector<int64_t> f = read_fails(w.fail.get(), nf, st);
    if constexpr (MODE == 1 && KP32 > 0) {
        if (nf > 0) {
            // BF16 rows without a certificate: FP32 candidates for those rows
            constexpr int QW32 = KL > 32 ? 1 : 2;
            const std::vector<int64_t> q2 = batch_queries(L, f.data(), nf, kNW * QW32 * 16, false);
            upload_sync(w.fail2, q2, st);
            cand_rerank<0, KP32, QW32, KL>(Z, n, d, ld, k, include_self, w.fail2.get(),
                                        (int64_t)q2.size(), w, idx, dist, w.nfail.get() + 1,
                                        w.fail.get(), st, seg, rs);
            SEV_HIP(hipMemcpyAsync(&nf, w.nfail.get() + 1, sizeof nf, hipMemcpyDeviceToHost, st));
            SEV_HIP(hipStreamSynchronize(st));
            f = read_fails(w.fail.get(), nf, st);
        }
    }
    if (nf > 0) {
        const std::vector<int64_t> q3 = batch_queries(L, f.data(), nf, 16, true);
        upload_sync(w.qpad, q3, st);
        knn_exact_rows_batch(Z, n, d, ld, k, include_self, w.qpad.get(), (int64_t)q3.size(), seg,
                             rs, idx, dist, st);
    }
    if (n_fallback) *n_fallback = nf;
}

}  // namespace

void knn_mfma_batch(const float *Z, int64_t n, int64_t d, int64_t ld, int64_t k,
                    int include_self, int bf16, const BatchLayout &L, KnnWork &w, int32_t *idx,
                    float *dist, hipStream_t st, int64_t *n_fallback) {
    SEV_REQUIRE(k >= 1 && k <= kMaxK, "knn: k must be in [1, 64]");
    SEV_REQUIRE(d >= 1, "knn: d must be >= 1");
    SEV_REQUIRE(n == L.M, "knn: embedding rows differ from the batch");
    SEV_REQUIRE(n < ((int64_t)1 << 31) - 256, "knn: at most 2^31 - 256 points");
    for (int64_t s = 0; s < L.S; ++s)
        SEV_REQUIRE(L.row0[(size_t)s + 1] - L.row0[(size_t)s] >= k + (include_self ? 0 : 1),
                    "knn: k larger than the cells of a sample");
    if (bf16) {
        SEV_REQUIRE(d + 2 <= 128, "knn: BF16 form supports d <= 126");
        if (k <= 24 && d + 2 <= 64 && d + 1 <= 52)
            run_mfma_batch<1, 64, 2, 32, 52>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                             n_fallback);
        else if (k <= 24 && d + 2 <= 64)
            run_mfma_batch<1, 64, 2, 32, 128>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                              n_fallback);
        else if (k <= 24)
            run_mfma_batch<1, 128, 2, 32, 128>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                               n_fallback);
        else if (d + 2 <= 64 && d + 1 <= 52)
            run_mfma_batch<1, 64, 1, 64, 52>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                             n_fallback);
        else if (d + 2 <= 64)
            run_mfma_batch<1, 64, 1, 64, 0>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                            n_fallback);
        else
            run_mfma_batch<1, 128, 1, 64, 0>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                             n_fallback);
    } else {
        SEV_REQUIRE(d + 1 <= 128, "knn: FP32 MFMA form supports d <= 127");
        if (d + 1 <= 52) {
            if (k <= 24)
                run_mfma_batch<0, 52, 2, 32, 52>(Z, n, d, ld, k, include_self, L, w, idx, dist,
                                                 st, n_fallback);
            else
                run_mfma_batch<0, 52, 1, 64, 52>(Z, n, d, ld, k, include_self, L, w, idx, dist,
                                                 st, n_fallback);
        } else {
            SEV_REQUIRE(k <= 24, "knn: batches with d > 51 need k <= 24 in the FP32 form");
            run_mfma_batch<0, 128, 1, 32, 128>(Z, n, d, ld, k, include_self, L, w, idx, dist, st,
                                               n_fallback);
        }
    }
}

}  // namespace sev
