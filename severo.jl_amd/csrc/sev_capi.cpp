// C ABI of libsevero_hip.so (declared in include/severo_hip.h).
//
// Host entry points convert Severo's Julia layout (1-based Int64 CSC,
// column-major dense, 1-based neighbour indices) to the device layout
// (0-based int64 colptr, int32 rows, row-major neighbour lists), run the
// device kernels and convert back.  The sev_pipeline handle keeps every
// stage resident on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/severo_hip.h"
#include "sev_kernels.hpp"

#include <atomic>

using namespace sev;

thread_local Prof *sev::t_prof = nullptr;

namespace {

thread_local std::string g_err;

// Installs the handle's profiler for the duration of an entry point.
struct ProfInstall {
    Prof *prev;
    explicit ProfInstall(Prof *p) : prev(t_prof) { t_prof = p; }
    ~ProfInstall() { t_prof = prev; }
};

template <class F>
int guarded(F &&f) {
    try {
        f();
        return SEV_OK;
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return SEV_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SEV_ERR_INTERNAL;
    } catch (...) {
        g_err = "unknown error";
        return SEV_ERR_INTERNAL;
    }
}

// The library's code objects are built for gfx950 only; refuse anything else
// (and a machine without a device) instead of falling back.
void ensure_device() {
    thread_local bool ok = false;
    if (ok) return;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) fail(kErrNoDevice, "no HIP device available");
    int dev = 0;
    SEV_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SEV_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        fail(kErrNoDevice, std::string("libsevero_hip is built for gfx950, device is ") +
                               prop.gcnArchName);
    ok = true;
}

}  // namespace

int sev::guarded_call(void (*fn)(void *), void *ctx) {
    return guarded([&] { fn(ctx); });
}

void sev::require_device() { ensure_device(); }

namespace {

// Host-to-device transfer of large arrays through two reusable pinned
// buffers: host threads fill one (converting as they copy) while the DMA of
// the other runs.  The buffers are allocated at the first large transfer and
// kept until the process ends; transfers are serialised by a mutex.
struct PinnedStage {
    static constexpr size_t kBytes = (size_t)256 << 20;  // per buffer
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    std::mutex mu;
    void ensure() {
        for (int u = 0; u < 2; ++u) {
            if (!buf[u]) SEV_HIP(hipHostMalloc(&buf[u], kBytes, hipHostMallocDefault));
            if (!done[u]) SEV_HIP(hipEventCreateWithFlags(&done[u], hipEventDisableTiming));
        }
    }
};

PinnedStage &pinned_stage() {
    static PinnedStage *ps = new PinnedStage();  // lives until the process ends
    return *ps;
}

constexpr size_t kStageMinBytes = (size_t)64 << 20;  // smaller arrays go directly

// f(first, count, thread) over [0, n) split across host threads
template <class F>
void host_parallel(int64_t n, F &&f) {
    const int64_t nt = std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)std::thread::hardware_concurrency(), 16,
                              n / ((int64_t)1 << 20) + 1}));
    std::vector<std::thread> th;
    const int64_t per = (n + nt - 1) / nt;
    for (int64_t t = 1; t < nt; ++t) {
        const int64_t a = t * per, b = std::min<int64_t>(n, a + per);
        if (a < b) th.emplace_back([&f, a, b, t] { f(a, b - a, (int)t); });
    }
    f(0, std::min<int64_t>(n, per), 0);
    for (auto &x : th) x.join();
}

// dst[0..n) = convert(src[0..n)) through the pinned buffers; convert(in,
// out, count, thread) may flag errors through its own state.
template <class S, class D, class C>
void staged_upload(D *dst, const S *src, int64_t n, C &&convert, hipStream_t st) {
    PinnedStage &ps = pinned_stage();
    std::lock_guard<std::mutex> lock(ps.mu);
    ps.ensure();
    const int64_t per = (int64_t)(PinnedStage::kBytes / sizeof(D));
    int u = 0;
    for (int64_t p0 = 0; p0 < n; p0 += per, u ^= 1) {
        const int64_t len = std::min<int64_t>(per, n - p0);
        if (ps.pending[u]) SEV_HIP(hipEventSynchronize(ps.done[u]));
        D *h = (D *)ps.buf[u];
        host_parallel(len, [&](int64_t a, int64_t c, int t) { convert(src + p0 + a, h + a, c, t); });
        SEV_HIP(hipMemcpyAsync(dst + p0, h, sizeof(D) * (size_t)len, hipMemcpyHostToDevice, st));
        SEV_HIP(hipEventRecord(ps.done[u], st));
        ps.pending[u] = true;
    }
    for (int v = 0; v < 2; ++v)
        if (ps.pending[v]) {
            SEV_HIP(hipEventSynchronize(ps.done[v]));
            ps.pending[v] = false;
        }
}

// Int32 counts to the device (staged when large)
void upload_counts(DevBuf<int32_t> &b, const int32_t *h, size_t n, hipStream_t st) {
    b.ensure(n);
    if (n == 0) return;
    if (n * sizeof(int32_t) < kStageMinBytes) {
        SEV_HIP(hipMemcpyAsync(b.get(), h, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
        return;
    }
    staged_upload<int32_t, int32_t>(
        b.get(), h, (int64_t)n,
        [](const int32_t *in, int32_t *out, int64_t c, int) {
            std::memcpy(out, in, sizeof(int32_t) * (size_t)c);
        },
        st);
}

// Julia 1-based Int64 CSC structure -> device 0-based int64 colptr and int32
// rows (validated).
void csc_to_device(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                   DevBuf<int64_t> &dcp, DevBuf<int32_t> &drow, int64_t &nnz, hipStream_t st) {
    SEV_REQUIRE(m >= 0 && n >= 0, "negative dimension");
    SEV_REQUIRE(m < (int64_t)std::numeric_limits<int32_t>::max(), "too many rows for int32 rows");
    SEV_REQUIRE(colptr != nullptr, "colptr is NULL");
    SEV_REQUIRE(colptr[0] == 1, "colptr must be 1-based (colptr[0] == 1)");
    std::vector<int64_t> cp((size_t)n + 1);
    for (int64_t j = 0; j <= n; ++j) {
        cp[(size_t)j] = colptr[j] - 1;
        if (j > 0) SEV_REQUIRE(cp[(size_t)j] >= cp[(size_t)j - 1], "colptr not nondecreasing");
    }
    nnz = cp[(size_t)n];
    SEV_REQUIRE(nnz == 0 || rowval != nullptr, "rowval is NULL");
    dcp.ensure((size_t)n + 1);
    drow.ensure((size_t)nnz);
    SEV_HIP(hipMemcpyAsync(dcp.get(), cp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice,
                           st));
    if (nnz && (size_t)nnz * sizeof(int64_t) >= kStageMinBytes) {
        // large: converted to 0-based int32 (and range-checked) by host
        // threads into the pinned buffers, half the bytes over the link
        std::vector<char> badt(16, 0);
        staged_upload<int64_t, int32_t>(
            drow.get(), rowval, nnz,
            [&](const int64_t *in, int32_t *out, int64_t c, int t) {
                bool bad = false;
                for (int64_t i = 0; i < c; ++i) {
                    const int64_t r = in[i] - 1;
                    bad |= (r < 0) | (r >= m);
                    out[i] = (int32_t)r;
                }
                if (bad) badt[(size_t)t] = 1;
            },
            st);
        SEV_REQUIRE(std::find(badt.begin(), badt.end(), 1) == badt.end(),
                    "row index out of range");
        DevBuf<int32_t> bad;
        bad.ensure(1);
        SEV_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
        check_sorted(n, dcp.get(), drow.get(), bad.get(), st);
        int32_t bad_h = 0;
        SEV_HIP(hipMemcpyAsync(&bad_h, bad.get(), sizeof bad_h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        SEV_REQUIRE(bad_h == 0, "rows of a column not strictly increasing");
    } else if (nnz) {
        // the 1-based Int64 rows go up in chunks through two device staging
        // buffers and are converted (and range-checked) on the device
        const int64_t chunk = std::min<int64_t>(nnz, (int64_t)1 << 25);  // 256 MB of Int64
        DevBuf<int64_t> stage[2];
        DevBuf<int32_t> bad;
        stage[0].ensure((size_t)chunk);
        stage[1].ensure((size_t)std::min<int64_t>(chunk, std::max<int64_t>(nnz - chunk, 1)));
        bad.ensure(1);
        SEV_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int32_t), st));
        int u = 0;
        for (int64_t p0 = 0; p0 < nnz; p0 += chunk, u ^= 1) {
            const int64_t len = std::min<int64_t>(chunk, nnz - p0);
            SEV_HIP(hipMemcpyAsync(stage[u].get(), rowval + p0, sizeof(int64_t) * len,
                                   hipMemcpyHostToDevice, st));
            rows_to_device(stage[u].get(), len, m, drow.get() + p0, bad.get(), st);
        }
        int32_t bad_h = 0;
        SEV_HIP(hipMemcpyAsync(&bad_h, bad.get(), sizeof bad_h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        SEV_REQUIRE(bad_h == 0, "row index out of range");
        // a SparseMatrixCSC keeps the rows of each column strictly increasing;
        // the row-range searches of the later stages rely on it
        check_sorted(n, dcp.get(), drow.get(), bad.get(), st);
        SEV_HIP(hipMemcpyAsync(&bad_h, bad.get(), sizeof bad_h, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        SEV_REQUIRE(bad_h == 0, "rows of a column not strictly increasing");
    }
    SEV_HIP(hipStreamSynchronize(st));
}

struct Stream {
    hipStream_t st = nullptr;
    Stream() { SEV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); }
    ~Stream() {
        if (st) (void)hipStreamDestroy(st);
    }
};

template <class T>
void upload(DevBuf<T> &b, const T *h, size_t n, hipStream_t st) {
    b.ensure(n);
    if (n) SEV_HIP(hipMemcpyAsync(b.get(), h, sizeof(T) * n, hipMemcpyHostToDevice, st));
}

template <class T>
void download(T *h, const T *d, size_t n, hipStream_t st) {
    if (n) SEV_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, st));
}

// xmin and weight table for the Jaccard graph: w[x] = x/(k + (k - x)) in T,
// keep iff |w| > T(prune) (droptol!).
template <class T>
int snn_table(int64_t k, double prune, std::vector<T> &w) {
    w.resize((size_t)k + 1);
    const T kk = (T)k, pr = (T)prune;
    int xmin = (int)k + 1;
    for (int64_t x = 0; x <= k; ++x) {
        const T xf = (T)x;
        w[(size_t)x] = xf / (kk + (kk - xf));
        if (std::fabs(w[(size_t)x]) > pr && xmin > (int)x) xmin = (int)x;
    }
    // w is increasing in x, so keep iff x >= xmin; a count of 0 never occurs
    if (xmin < 1) xmin = 1;
    return xmin;
}

}  // namespace

// ---------------------------------------------------------------------------
// Pipeline handle
// ---------------------------------------------------------------------------
struct sev_pipeline {
    Stream stream;
    Prof prof;
    // counts (cells x genes CSC)
    int64_t m = 0, g = 0, nnz = 0;
    DevBuf<int64_t> cp;
    DevBuf<int32_t> row, cnt;
    bool have_counts = false;
    // normalised values
    DevBuf<uint32_t> tot;
    DevBuf<float> norm;
    bool have_norm = false;
    // selection + scaling
    int64_t F = 0, fnnz = 0;
    DevBuf<int64_t> feat, fcp, flen;
    DevBuf<int32_t> frow;
    DevBuf<float> fval, bval, mu, sd, clip;
    bool have_scaled = false;
    // PCA
    SlabCsc op;
    IrlbaWork iw;
    int64_t npcs = 0;
    std::vector<double> s;
    DevBuf<double> U, V, s_d;
    bool have_pca = false;
    // embedding (column-major m x zdim Float32)
    int64_t zdim = 0;
    DevBuf<float> Z;
    bool have_embedding = false;
    // neighbours (row-major m x k, 0-based)
    int64_t k = 0;
    DevBuf<int32_t> nn;
    DevBuf<float> dist;
    bool have_nn = false;
    // SNN
    KnnWork kw;
    int64_t knn_passes[3] = {0, 0, 0};  // FP8 form: rows left after each candidate pass
    SnnWork sw;
    int64_t snn_nnz = 0;
    int snn_dtype = SEV_FLOAT32;
    DevBuf<int64_t> scp;
    DevBuf<int32_t> srow;
    DevBuf<float> sval32, wtab32;
    DevBuf<double> sval64, wtab64;
    bool have_snn = false;
    // scratch
    DevBuf<int64_t> scan_scratch, chunk_cnt, chunk_off;
    DevBuf<uint8_t> cls;
    DevBuf<double> p0, q;
    // batch: S samples stacked by rows (cells), every per-sample stage runs
    // on its own rows
    BatchLayout bl;
    bool batched = false;
    DevBuf<int64_t> soff, vsrc, vcp;
    DevBuf<float> mu_all;
    BatchIrlbaWork biw;

    hipStream_t st() const { return stream.st; }
    int64_t nsamples() const { return batched ? bl.S : 1; }

    void invalidate_from_counts() {
        have_norm = have_scaled = have_pca = have_embedding = have_nn = have_snn = false;
    }
};

namespace {
template <class T>
void scale_host(int64_t m, int64_t n, const DevBuf<int64_t> &cp, int64_t nnz, const T *nzval,
                T scale_max, T *out, T *mu, hipStream_t st) {
    DevBuf<T> x, b, dmu, dsd, dclip;
    upload(x, nzval, (size_t)nnz, st);
    b.ensure((size_t)nnz);
    dmu.ensure((size_t)n);
    dsd.ensure((size_t)n);
    dclip.ensure((size_t)n);
    scale_stats<T>(m, n, cp.get(), x.get(), scale_max, dsd.get(), dmu.get(), dclip.get(), st);
    scale_map<T>(n, cp.get(), x.get(), dsd.get(), dclip.get(), b.get(), st);
    download(out, b.get(), (size_t)nnz, st);
    download(mu, dmu.get(), (size_t)n, st);
    SEV_HIP(hipStreamSynchronize(st));
}
}  // namespace

namespace {
template <class T>
void knn_host(const T *X, int64_t n, int64_t d, int64_t ld, int64_t k, int include_self,
              int precision, int32_t *nn_index, T *distances, int64_t *n_fallback,
              bool cosine = false) {
    SEV_REQUIRE(n >= 1 && d >= 1 && ld >= n, "bad X shape");
    SEV_REQUIRE(k >= 1 && k <= (include_self ? n : n - 1), "k larger than the number of points");
    Stream s;
    DevBuf<T> dx;
    dx.ensure((size_t)(n * d));
    SEV_HIP(hipMemcpy2DAsync(dx.get(), sizeof(T) * n, X, sizeof(T) * ld, sizeof(T) * n, d,
                             hipMemcpyHostToDevice, s.st));
    if (cosine) knn_unit_rows<T>(dx.get(), n, d, s.st);
    DevBuf<int32_t> idx;
    DevBuf<T> dist;
    idx.ensure((size_t)(n * k));
    dist.ensure((size_t)(n * k));
    int64_t nfb = n;
    if (sizeof(T) == 4 && precision == SEV_PREC_FP8) {
        KnnWork w;
        knn_mfma_fp8((const float *)dx.get(), n, d, n, k, include_self, w, idx.get(),
                     (float *)dist.get(), s.st, &nfb, nullptr);
    } else if (sizeof(T) == 4 && precision != SEV_PREC_EXACT) {
        KnnWork w;
        knn_mfma((const float *)dx.get(), n, d, n, k, include_self, precision == SEV_PREC_BF16, w,
                 idx.get(), (float *)dist.get(), s.st, &nfb);
    } else {
        knn_exact<T>(dx.get(), n, d, n, k, include_self, nullptr, n, idx.get(), dist.get(), s.st);
    }
    if (cosine) knn_half_pair_dist<T>(dx.get(), n, d, idx.get(), k, dist.get(), s.st);
    if (n_fallback) *n_fallback = nfb;
    std::vector<int32_t> hi((size_t)(n * k));
    std::vector<T> hd((size_t)(n * k));
    download(hi.data(), idx.get(), hi.size(), s.st);
    download(hd.data(), dist.get(), hd.size(), s.st);
    SEV_HIP(hipStreamSynchronize(s.st));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t t = 0; t < k; ++t) {
            nn_index[i + t * n] = hi[(size_t)(i * k + t)] + 1;
            distances[i + t * n] = hd[(size_t)(i * k + t)];
        }
}
}  // namespace

extern "C" {

const char *sev_last_error(void) { return g_err.c_str(); }

int sev_version(void) { return 1; }

int sev_selftest(int what, double *result) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(what >= 0 && what <= 4, "unknown self-test");
        SEV_REQUIRE(result != nullptr, "result is NULL");
        *result = what == 3 ? hbm_copy_probe() : mfma_selftest(what == 4 ? 3 : what);
    });
}

int sev_device_init(int device) {
    return guarded([&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            fail(kErrNoDevice, "no HIP device available");
        SEV_REQUIRE(device >= 0 && device < count, "device index out of range");
        SEV_HIP(hipSetDevice(device));
        ensure_device();
        noalloc_setup();
    });
}

int sev_pipeline_create(sev_pipeline **out) {
    return guarded([&] {
        SEV_REQUIRE(out != nullptr, "out is NULL");
        ensure_device();
        *out = new sev_pipeline();
    });
}

void sev_pipeline_destroy(sev_pipeline *p) { delete p; }

int sev_pipeline_synchronize(sev_pipeline *p) {
    return guarded([&] { SEV_HIP(hipStreamSynchronize(p->st())); });
}

int sev_pipeline_load_counts(sev_pipeline *p, int64_t m, int64_t n, const int64_t *colptr,
                             const int64_t *rowval, const int32_t *nzval) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        hipStream_t st = p->st();
        csc_to_device(m, n, colptr, rowval, p->cp, p->row, p->nnz, st);
        upload_counts(p->cnt, nzval, (size_t)p->nnz, st);
        SEV_HIP(hipStreamSynchronize(st));
        p->m = m;
        p->g = n;
        p->have_counts = true;
        p->batched = false;
        p->invalidate_from_counts();
    });
}

int sev_pipeline_load_mtx(sev_pipeline *p, const char *path, int64_t *out_m, int64_t *out_n,
                          int64_t *out_nnz) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        hipStream_t st = p->st();
        MtxReader rd(path);
        const int64_t genes = rd.rows(), cells = rd.cols(), E = rd.entries();
        SEV_REQUIRE(cells < (int64_t)0x7FFFFFFF && genes < (int64_t)0x7FFFFFFF,
                    "matrix too large for int32 cell and gene indices");
        SEV_REQUIRE(E < (int64_t)0x7FFFFFFF, "more than 2^31 - 1 entries");
        // (gene << 32 | cell, value) on the host, then to the device
        std::vector<uint64_t> hk((size_t)E);
        std::vector<int64_t> hv((size_t)E);
        std::atomic<int64_t> bad_index(-1), inexact(-1);
        const bool real = rd.is_real();
        rd.parse(
            [&](int64_t i, int64_t r, int64_t c, int64_t vi, double vf) {
                if (r < 1 || r > genes || c < 1 || c > cells) {
                    int64_t none = -1;
                    bad_index.compare_exchange_strong(none, i);
                    return;
                }
                int64_t v = vi;
                if (real) {
                    // convert_counts turns a real matrix into Int32 counts
                    if (!(vf == std::floor(vf)) || !(std::fabs(vf) < 9.0e18)) {
                        int64_t none = -1;
                        inexact.compare_exchange_strong(none, i);
                        return;
                    }
                    v = (int64_t)vf;
                }
                hk[(size_t)i] = ((uint64_t)(r - 1) << 32) | (uint64_t)(c - 1);
                hv[(size_t)i] = v;
            },
            nullptr);
        SEV_REQUIRE(bad_index.load() < 0,
                    "matrix entry " + std::to_string(bad_index.load() + 1) +
                        ": row or column index out of range");
        SEV_REQUIRE(inexact.load() < 0, "matrix entry " + std::to_string(inexact.load() + 1) +
                                            ": a real count that is not an integer");
        DevBuf<uint64_t> dk;
        DevBuf<int64_t> dv;
        dk.ensure((size_t)E);
        dv.ensure((size_t)E);
        if (E > 0 && (size_t)E * sizeof(uint64_t) >= kStageMinBytes) {
            staged_upload<uint64_t, uint64_t>(
                dk.get(), hk.data(), E,
                [](const uint64_t *in, uint64_t *out, int64_t c, int) {
                    std::memcpy(out, in, sizeof(uint64_t) * (size_t)c);
                },
                st);
            staged_upload<int64_t, int64_t>(
                dv.get(), hv.data(), E,
                [](const int64_t *in, int64_t *out, int64_t c, int) {
                    std::memcpy(out, in, sizeof(int64_t) * (size_t)c);
                },
                st);
        } else if (E > 0) {
            SEV_HIP(hipMemcpyAsync(dk.get(), hk.data(), sizeof(uint64_t) * (size_t)E,
                                   hipMemcpyHostToDevice, st));
            SEV_HIP(hipMemcpyAsync(dv.get(), hv.data(), sizeof(int64_t) * (size_t)E,
                                   hipMemcpyHostToDevice, st));
            SEV_HIP(hipStreamSynchronize(st));
        }
        int64_t nnz = 0;
        coo_to_csc(genes, dk.get(), dv.get(), E, p->cp, p->row, p->cnt, nnz, st);
        p->m = cells;
        p->g = genes;
        p->nnz = nnz;
        p->have_counts = true;
        p->batched = false;
        p->invalidate_from_counts();
        if (out_m) *out_m = cells;
        if (out_n) *out_n = genes;
        if (out_nnz) *out_nnz = nnz;
    });
}

int sev_pipeline_synth_counts(sev_pipeline *p, int64_t m, int64_t n, uint64_t seed, double r,
                              const double *p0, const double *q, double boost_frac) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(m > 0 && n > 0 && m < (int64_t)0x7FFFFFFF && n < (int64_t)0x7FFFFFFF,
                    "bad synthetic shape");
        hipStream_t st = p->st();
        upload(p->p0, p0, (size_t)(2 * n), st);
        upload(p->q, q, (size_t)(2 * n), st);
        p->cls.ensure((size_t)m);
        SynthPlan plan;
        const int64_t nchunk = ceil_div(m, 16384);
        p->chunk_cnt.ensure((size_t)(n * nchunk));
        p->chunk_off.ensure((size_t)(n * nchunk + 1));
        p->scan_scratch.ensure(scan_scratch_elems(n * nchunk));
        p->cp.ensure((size_t)n + 1);
        const int64_t nnz = synth_count(m, n, seed, r, p->p0.get(), p->q.get(), boost_frac,
                                        p->cls.get(), p->chunk_cnt.get(), p->chunk_off.get(),
                                        p->cp.get(), p->scan_scratch.get(), st, &plan);
        p->row.ensure((size_t)nnz);
        p->cnt.ensure((size_t)nnz);
        synth_fill(plan, seed, r, p->p0.get(), p->q.get(), p->cls.get(), p->chunk_off.get(),
                   p->row.get(), p->cnt.get(), st);
        SEV_HIP(hipStreamSynchronize(st));
        p->m = m;
        p->g = n;
        p->nnz = nnz;
        p->have_counts = true;
        p->batched = false;
        p->invalidate_from_counts();
    });
}

int sev_pipeline_synth_batch(sev_pipeline *p, int64_t S, const int64_t *cells, int64_t n,
                             const uint64_t *seeds, const double *r, const double *p0,
                             const double *q, const double *boost_frac) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(S >= 1 && n > 0 && n < (int64_t)0x7FFFFFFF, "bad batch shape");
        hipStream_t st = p->st();
        std::vector<int64_t> cv(cells, cells + S);
        int64_t M = 0;
        for (int64_t s = 0; s < S; ++s) {
            SEV_REQUIRE(cv[(size_t)s] >= 2, "every sample needs at least two cells");
            M += cv[(size_t)s];
        }
        SEV_REQUIRE(M < (int64_t)0x7FFFFFFF, "batch too large");
        // each sample into its own CSC, then stacked by rows
        std::vector<DevBuf<int64_t>> cps((size_t)S);
        std::vector<DevBuf<int32_t>> rows((size_t)S), cnts((size_t)S);
        std::vector<const int64_t *> cph((size_t)S);
        for (int64_t s = 0; s < S; ++s) {
            const int64_t m = cv[(size_t)s];
            upload(p->p0, p0 + (size_t)(2 * n * s), (size_t)(2 * n), st);
            upload(p->q, q + (size_t)(2 * n * s), (size_t)(2 * n), st);
            p->cls.ensure((size_t)m);
            SynthPlan plan;
            const int64_t nchunk = ceil_div(m, 16384);
            p->chunk_cnt.ensure((size_t)(n * nchunk));
            p->chunk_off.ensure((size_t)(n * nchunk + 1));
            p->scan_scratch.ensure(scan_scratch_elems(n * nchunk));
            cps[(size_t)s].ensure((size_t)n + 1);
            const int64_t nnz = synth_count(m, n, seeds[s], r[s], p->p0.get(), p->q.get(),
                                            boost_frac[s], p->cls.get(), p->chunk_cnt.get(),
                                            p->chunk_off.get(), cps[(size_t)s].get(),
                                            p->scan_scratch.get(), st, &plan);
            rows[(size_t)s].ensure((size_t)nnz + 1);
            cnts[(size_t)s].ensure((size_t)nnz + 1);
            synth_fill(plan, seeds[s], r[s], p->p0.get(), p->q.get(), p->cls.get(),
                       p->chunk_off.get(), rows[(size_t)s].get(), cnts[(size_t)s].get(), st);
            SEV_HIP(hipStreamSynchronize(st));
            cph[(size_t)s] = cps[(size_t)s].get();
        }
        DevBuf<const int64_t *> cps_d;
        upload(cps_d, cph.data(), (size_t)S, st);
        DevBuf<int64_t> starts;
        starts.ensure((size_t)(n * S + 1));
        p->scan_scratch.ensure(scan_scratch_elems(n * S));
        stack_lengths(n, S, cps_d.get(), starts.get(), st);
        scan_exclusive_i64(starts.get(), starts.get(), n * S, p->scan_scratch.get(), st);
        int64_t nnz = 0;
        SEV_HIP(hipMemcpyAsync(&nnz, starts.get() + n * S, sizeof nnz, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        p->cp.ensure((size_t)n + 1);
        p->row.ensure((size_t)nnz);
        p->cnt.ensure((size_t)nnz);
        stride_ptr(starts.get(), n, S, p->cp.get(), st);
        int64_t roff = 0;
        for (int64_t s = 0; s < S; ++s) {
            stack_copy(n, S, s, cps[(size_t)s].get(), rows[(size_t)s].get(), cnts[(size_t)s].get(),
                       roff, starts.get(), p->row.get(), p->cnt.get(), st);
            roff += cv[(size_t)s];
        }
        SEV_HIP(hipStreamSynchronize(st));
        p->m = M;
        p->g = n;
        p->nnz = nnz;
        p->have_counts = true;
        p->invalidate_from_counts();
        batch_layout(p->bl, cv, st);
        p->batched = true;
    });
}

int sev_pipeline_set_batch(sev_pipeline *p, int64_t S, const int64_t *cells) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_counts, "no counts loaded");
        if (S == 0) {
            p->batched = false;
        } else {
            SEV_REQUIRE(S >= 1 && cells, "bad batch");
            std::vector<int64_t> cv(cells, cells + S);
            int64_t M = 0;
            for (int64_t c : cv) M += c;
            SEV_REQUIRE(M == p->m, "sample cells do not add up to the loaded cells");
            batch_layout(p->bl, cv, p->st());
            p->batched = true;
        }
        p->have_scaled = p->have_pca = p->have_embedding = p->have_nn = p->have_snn = false;
    });
}

int sev_pipeline_counts_info(sev_pipeline *p, int64_t *m, int64_t *n, int64_t *nnz) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_counts, "no counts loaded");
        if (m) *m = p->m;
        if (n) *n = p->g;
        if (nnz) *nnz = p->nnz;
    });
}

int sev_pipeline_get_counts(sev_pipeline *p, int64_t *colptr, int64_t *rowval, int32_t *nzval) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_counts, "no counts loaded");
        hipStream_t st = p->st();
        std::vector<int32_t> rv((size_t)p->nnz);
        if (colptr) download(colptr, p->cp.get(), (size_t)p->g + 1, st);
        if (rowval) download(rv.data(), p->row.get(), (size_t)p->nnz, st);
        if (nzval) download(nzval, p->cnt.get(), (size_t)p->nnz, st);
        SEV_HIP(hipStreamSynchronize(st));
        if (colptr)
            for (int64_t j = 0; j <= p->g; ++j) colptr[j] += 1;
        if (rowval)
            for (int64_t i = 0; i < p->nnz; ++i) rowval[i] = (int64_t)rv[(size_t)i] + 1;
    });
}

int sev_pipeline_normalize(sev_pipeline *p, int method, double scale_factor) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_counts, "no counts loaded");
        SEV_REQUIRE(method == SEV_LOGNORMALIZE || method == SEV_RELATIVECOUNTS,
                    "unknown normalization method");
        hipStream_t st = p->st();
        p->tot.ensure((size_t)p->m);
        p->norm.ensure((size_t)p->nnz);
        row_totals(p->cp.get(), p->row.get(), p->cnt.get(), p->nnz, p->m, p->g, p->tot.get(), st);
        lognorm_f32(p->row.get(), p->cnt.get(), p->tot.get(), p->nnz, (float)scale_factor,
                    method == SEV_LOGNORMALIZE, p->norm.get(), st);
        p->have_norm = true;
        p->have_scaled = p->have_pca = p->have_embedding = p->have_nn = p->have_snn = false;
    });
}

int sev_pipeline_get_normalized(sev_pipeline *p, float *nzval) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_norm, "normalize has not run");
        download(nzval, p->norm.get(), (size_t)p->nnz, p->st());
        SEV_HIP(hipStreamSynchronize(p->st()));
    });
}

int sev_pipeline_select_scale(sev_pipeline *p, int64_t nfeat, const int64_t *features,
                              double scale_max) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_norm, "normalize has not run");
        SEV_REQUIRE(nfeat >= 1, "need at least one feature");
        hipStream_t st = p->st();
        std::vector<int64_t> f((size_t)nfeat);
        for (int64_t j = 0; j < nfeat; ++j) {
            SEV_REQUIRE(features[j] >= 1 && features[j] <= p->g, "feature index out of range");
            f[(size_t)j] = features[j] - 1;
        }
        upload(p->feat, f.data(), (size_t)nfeat, st);
        p->flen.ensure((size_t)nfeat + 1);
        p->fcp.ensure((size_t)nfeat + 1);
        p->scan_scratch.ensure(scan_scratch_elems(nfeat));
        gather_lengths(p->feat.get(), p->cp.get(), nfeat, p->flen.get(), st);
        scan_exclusive_i64(p->flen.get(), p->fcp.get(), nfeat, p->scan_scratch.get(), st);
        int64_t fnnz = 0;
        SEV_HIP(hipMemcpyAsync(&fnnz, p->fcp.get() + nfeat, sizeof fnnz, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        p->frow.ensure((size_t)fnnz);
        p->fval.ensure((size_t)fnnz);
        p->bval.ensure((size_t)fnnz + kValuePad);
        p->mu.ensure((size_t)nfeat);
        p->sd.ensure((size_t)nfeat);
        p->clip.ensure((size_t)nfeat);
        gather_copy<float>(p->feat.get(), p->cp.get(), p->row.get(), p->norm.get(), nfeat,
                           p->fcp.get(), p->frow.get(), p->fval.get(), st);
        scale_stats<float>(p->m, nfeat, p->fcp.get(), p->fval.get(), (float)scale_max, p->sd.get(),
                           p->mu.get(), p->clip.get(), st);
        scale_map<float>(nfeat, p->fcp.get(), p->fval.get(), p->sd.get(), p->clip.get(),
                         p->bval.get(), st);
        p->F = nfeat;
        p->fnnz = fnnz;
        p->have_scaled = true;
        p->have_pca = p->have_embedding = p->have_nn = p->have_snn = false;
    });
}

int sev_pipeline_select_scale_batch(sev_pipeline *p, int64_t nfeat, const int64_t *features,
                                    double scale_max) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_norm, "normalize has not run");
        SEV_REQUIRE(p->batched, "no batch declared");
        SEV_REQUIRE(nfeat >= 1, "need at least one feature");
        hipStream_t st = p->st();
        const int64_t S = p->bl.S, G = p->g, F = nfeat, NV = F * S;
        std::vector<int64_t> f((size_t)NV);
        for (int64_t i = 0; i < NV; ++i) {
            SEV_REQUIRE(features[i] >= 1 && features[i] <= G, "feature index out of range");
            f[(size_t)i] = features[i] - 1;
        }
        upload(p->feat, f.data(), (size_t)NV, st);
        // virtual columns v = j*S + s: feature j of sample s on its own rows
        p->soff.ensure((size_t)((S + 1) * G));
        sample_offsets(G, S, p->bl.row_sample_d.get(), p->cp.get(), p->row.get(), p->soff.get(),
                       st);
        p->vsrc.ensure((size_t)NV);
        p->vcp.ensure((size_t)NV + 1);
        p->scan_scratch.ensure(scan_scratch_elems(NV));
        batch_seg_lengths(F, S, G, p->feat.get(), p->soff.get(), p->vsrc.get(), p->vcp.get(), st);
        scan_exclusive_i64(p->vcp.get(), p->vcp.get(), NV, p->scan_scratch.get(), st);
        int64_t fnnz = 0;
        SEV_HIP(hipMemcpyAsync(&fnnz, p->vcp.get() + NV, sizeof fnnz, hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        p->frow.ensure((size_t)fnnz + 1);
        p->fval.ensure((size_t)fnnz + 1);
        p->bval.ensure((size_t)fnnz + kValuePad);
        p->mu.ensure((size_t)NV);
        p->sd.ensure((size_t)NV);
        p->clip.ensure((size_t)NV);
        p->mu_all.ensure((size_t)NV);
        p->fcp.ensure((size_t)F + 1);
        batch_gather(NV, p->vsrc.get(), p->vcp.get(), p->row.get(), p->norm.get(), p->frow.get(),
                     p->fval.get(), st);
        batch_scale_stats(NV, p->vcp.get(), p->fval.get(), (float)scale_max, p->sd.get(),
                          p->mu.get(), p->clip.get(), p->bl.row0_d.get(), S, st);
        scale_map<float>(NV, p->vcp.get(), p->fval.get(), p->sd.get(), p->clip.get(),
                         p->bval.get(), st);
        stride_ptr(p->vcp.get(), F, S, p->fcp.get(), st);
        mu_by_sample(p->mu.get(), F, S, p->mu_all.get(), st);
        SEV_HIP(hipStreamSynchronize(st));
        p->F = F;
        p->fnnz = fnnz;
        p->have_scaled = true;
        p->have_pca = p->have_embedding = p->have_nn = p->have_snn = false;
    });
}

int sev_pipeline_scaled_nnz(sev_pipeline *p, int64_t *nnz) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_scaled, "scale has not run");
        *nnz = p->fnnz;
    });
}

int sev_pipeline_get_scaled(sev_pipeline *p, int64_t *colptr, int64_t *rowval, float *nzval,
                            float *mu) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_scaled, "scale has not run");
        hipStream_t st = p->st();
        std::vector<int32_t> rv((size_t)p->fnnz);
        if (colptr) download(colptr, p->fcp.get(), (size_t)p->F + 1, st);
        if (rowval) download(rv.data(), p->frow.get(), (size_t)p->fnnz, st);
        if (nzval) download(nzval, p->bval.get(), (size_t)p->fnnz, st);
        if (mu) download(mu, p->batched ? p->mu_all.get() : p->mu.get(),
                         (size_t)(p->F * p->nsamples()), st);
        SEV_HIP(hipStreamSynchronize(st));
        if (colptr)
            for (int64_t j = 0; j <= p->F; ++j) colptr[j] += 1;
        if (rowval)
            for (int64_t i = 0; i < p->fnnz; ++i) rowval[i] = (int64_t)rv[(size_t)i] + 1;
    });
}

int sev_pipeline_pca(sev_pipeline *p, int64_t npcs, int64_t work, int64_t maxit, double tol,
                     const double *init, int precision, int64_t *iters, int64_t *nmult) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_scaled, "scale has not run");
        SEV_REQUIRE(!p->batched, "a batch runs sev_pipeline_pca_batch");
        SEV_REQUIRE(precision == SEV_PREC_FP32 || precision == SEV_PREC_BF16, "bad precision");
        hipStream_t st = p->st();
        const int64_t m = p->m, n = p->F;
        const int64_t nu = std::min(std::min(m, n), npcs);
        SEV_REQUIRE(nu >= 1, "npcs must be >= 1");
        SEV_REQUIRE(init != nullptr, "init is NULL");
        build_slab_csc(p->op, m, n, p->fcp.get(), p->frow.get(), p->bval.get(), false,
                       p->mu.get(), precision == SEV_PREC_BF16, st);
        autotune_operator(p->op, st);
        SEV_REQUIRE(work == 0 || (work >= nu && work <= 128),
                    "work must be 0 (npcs + 7) or in [npcs, 128]");
        IrlbaParams prm;
        prm.nu = nu;
        prm.work = work > 0 ? work : nu + 7;
        prm.maxit = maxit;
        prm.tol = tol;
        prm.svtol = tol;
        // the BF16 form keeps its left Lanczos basis in FP32
        prm.basis32 = precision == SEV_PREC_BF16;
        p->s.assign((size_t)nu, 0.0);
        p->U.ensure((size_t)(m * nu));
        p->V.ensure((size_t)(n * nu));
        IrlbaOut out;
        irlba_device(p->op, prm, init, p->iw, p->s.data(), p->U.get(), p->V.get(), st, &out);
        if (iters) *iters = out.iters;
        if (nmult) *nmult = out.nmult;
        p->npcs = nu;
        upload(p->s_d, p->s.data(), (size_t)nu, st);
        p->Z.ensure((size_t)(m * nu));
        embedding_f32(p->U.get(), p->s_d.get(), m, nu, p->Z.get(), 1, m, st);
        p->zdim = nu;
        p->have_pca = p->have_embedding = true;
        p->have_nn = p->have_snn = false;
    });
}

int sev_pipeline_pca_batch(sev_pipeline *p, int64_t npcs, int64_t work, int64_t maxit, double tol,
                           const double *init, int use_graph, int64_t *iters, int64_t *nmult) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_scaled && p->batched, "no scaled batch");
        SEV_REQUIRE(init != nullptr, "init is NULL");
        hipStream_t st = p->st();
        const int64_t M = p->m, n = p->F, S = p->bl.S;
        int64_t minm = M;
        for (int64_t s = 0; s < S; ++s)
            minm = std::min<int64_t>(minm, p->bl.row0[(size_t)s + 1] - p->bl.row0[(size_t)s]);
        const int64_t nu = std::min(std::min(minm, n), npcs);
        SEV_REQUIRE(nu >= 1, "npcs must be >= 1");
        SEV_REQUIRE(n <= 14336, "batches need at most 14336 features");
        // the stacked BF16 operator; lane widths tuned on it, then slabs cut
        // at the sample boundaries
        build_slab_csc(p->op, M, n, p->fcp.get(), p->frow.get(), p->bval.get(), false, nullptr,
                       true, st);
        autotune_operator(p->op, st);
        batch_slabs(p->op, p->bl, p->fcp.get(), p->frow.get(), st);
        IrlbaParams prm;
        prm.nu = nu;
        SEV_REQUIRE(work == 0 || (work >= nu + 1 && work <= 64),
                    "batch work must be 0 (npcs + 7) or in [npcs + 1, 64]");
        prm.work = work == 0 ? nu + 7 : work;
        prm.maxit = maxit;
        prm.tol = tol;
        prm.svtol = tol;
        prm.basis32 = 1;  // the BF16 form: left bases in FP32
        p->s.assign((size_t)(S * nu), 0.0);
        p->U.ensure((size_t)(M * nu));
        p->V.ensure((size_t)(S * n * nu));
        std::vector<int64_t> it((size_t)S, 0);
        irlba_batch(p->op, p->bl, p->mu_all.get(), prm, init, p->biw, p->s.data(), p->U.get(),
                    p->V.get(), it.data(), nmult, use_graph != 0, st);
        if (iters) std::memcpy(iters, it.data(), sizeof(int64_t) * (size_t)S);
        p->npcs = nu;
        upload(p->s_d, p->s.data(), (size_t)(S * nu), st);
        p->Z.ensure((size_t)(M * nu));
        batch_embedding(p->U.get(), p->s_d.get(), p->bl, nu, p->Z.get(), 1, M, st);
        SEV_HIP(hipStreamSynchronize(st));
        p->zdim = nu;
        p->have_pca = p->have_embedding = true;
        p->have_nn = p->have_snn = false;
    });
}

int sev_pipeline_get_pca(sev_pipeline *p, double *s, double *U, double *V, float *Z) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_pca, "pca has not run");
        hipStream_t st = p->st();
        const int64_t S = p->nsamples();
        if (s) std::memcpy(s, p->s.data(), sizeof(double) * (size_t)(S * p->npcs));
        if (U) download(U, p->U.get(), (size_t)(p->m * p->npcs), st);
        if (V) download(V, p->V.get(), (size_t)(S * p->F * p->npcs), st);
        if (Z) download(Z, p->Z.get(), (size_t)(p->m * p->npcs), st);
        SEV_HIP(hipStreamSynchronize(st));
    });
}

int sev_pipeline_set_embedding(sev_pipeline *p, int64_t m, int64_t d, const float *Z, int64_t ld) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(m >= 1 && d >= 1 && ld >= m, "bad embedding shape");
        hipStream_t st = p->st();
        p->Z.ensure((size_t)(m * d));
        SEV_HIP(hipMemcpy2DAsync(p->Z.get(), sizeof(float) * m, Z, sizeof(float) * ld,
                                 sizeof(float) * m, d, hipMemcpyHostToDevice, st));
        SEV_HIP(hipStreamSynchronize(st));
        if (p->batched && m != p->bl.M) p->batched = false;
        p->m = m;
        p->zdim = d;
        p->have_embedding = true;
        p->have_nn = p->have_snn = false;
    });
}

// The pipeline's kNN search on its embedding; cosine: on a copy of the
// embedding's first d columns scaled to unit rows, then D / 2 of the found
// pairs (the CosineDist form of the kNN entry points).
static int pipeline_neighbours(sev_pipeline *p, int64_t k, int include_self, int64_t ndims,
                               int precision, int64_t *n_fallback, bool cosine) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_embedding, "no embedding");
        SEV_REQUIRE(precision == SEV_PREC_FP32 || precision == SEV_PREC_BF16 ||
                        precision == SEV_PREC_EXACT || precision == SEV_PREC_FP8,
                    "bad precision");
        SEV_REQUIRE(!(p->batched && precision == SEV_PREC_FP8),
                    "the FP8 candidate form is not provided for batches");
        const int64_t m = p->m;
        const int64_t d = (ndims <= 0 || ndims > p->zdim) ? p->zdim : ndims;
        SEV_REQUIRE(k >= 1 && k <= (include_self ? m : m - 1), "k larger than the number of cells");
        hipStream_t st = p->st();
        p->nn.ensure((size_t)(m * k));
        p->dist.ensure((size_t)(m * k));
        DevBuf<float> zc;
        const float *Zs = p->Z.get();
        if (cosine) {
            zc.ensure((size_t)(m * d));
            SEV_HIP(hipMemcpyAsync(zc.get(), p->Z.get(), sizeof(float) * (size_t)(m * d),
                                   hipMemcpyDeviceToDevice, st));
            knn_unit_rows<float>(zc.get(), m, d, st);
            Zs = zc.get();
        }
        if (p->batched) {
            for (int64_t s = 0; s < p->bl.S; ++s)
                SEV_REQUIRE(p->bl.row0[(size_t)s + 1] - p->bl.row0[(size_t)s] >=
                                k + (include_self ? 0 : 1),
                            "k larger than the cells of a sample");
            if (precision == SEV_PREC_EXACT) {
                std::vector<int64_t> q;
                q.reserve((size_t)(m + 16 * p->bl.S));
                for (int64_t s = 0; s < p->bl.S; ++s) {
                    for (int64_t r = p->bl.row0[(size_t)s]; r < p->bl.row0[(size_t)s + 1]; ++r)
                        q.push_back(r);
                    while (q.size() % 16) q.push_back(p->bl.row0[(size_t)s + 1] - 1);
                }
                DevBuf<int64_t> qd;
                upload(qd, q.data(), q.size(), st);
                knn_exact_rows_batch(Zs, m, d, m, k, include_self, qd.get(),
                                     (int64_t)q.size(), p->bl.row0_d.get(),
                                     p->bl.row_sample_d.get(), p->nn.get(), p->dist.get(), st);
                SEV_HIP(hipStreamSynchronize(st));
                if (n_fallback) *n_fallback = m;
            } else {
                knn_mfma_batch(Zs, m, d, m, k, include_self, precision == SEV_PREC_BF16,
                               p->bl, p->kw, p->nn.get(), p->dist.get(), st, n_fallback);
            }
        } else if (precision == SEV_PREC_EXACT) {
            knn_exact<float>(Zs, m, d, m, k, include_self, nullptr, m, p->nn.get(),
                             p->dist.get(), st);
            if (n_fallback) *n_fallback = m;
        } else if (precision == SEV_PREC_FP8) {
            knn_mfma_fp8(Zs, m, d, m, k, include_self, p->kw, p->nn.get(), p->dist.get(),
                         st, n_fallback, p->knn_passes);
        } else {
            knn_mfma(Zs, m, d, m, k, include_self, precision == SEV_PREC_BF16, p->kw,
                     p->nn.get(), p->dist.get(), st, n_fallback);
        }
        if (cosine) knn_half_pair_dist<float>(Zs, m, d, p->nn.get(), k, p->dist.get(), st);
        p->k = k;
        p->have_nn = true;
        p->have_snn = false;
    });
}

int sev_pipeline_neighbours(sev_pipeline *p, int64_t k, int include_self, int64_t ndims,
                            int precision, int64_t *n_fallback) {
    return pipeline_neighbours(p, k, include_self, ndims, precision, n_fallback, false);
}

int sev_pipeline_neighbours_cosine(sev_pipeline *p, int64_t k, int include_self, int64_t ndims,
                                   int precision, int64_t *n_fallback) {
    return pipeline_neighbours(p, k, include_self, ndims, precision, n_fallback, true);
}

int sev_pipeline_find_variable_features(sev_pipeline *p, int64_t nfeatures, double loess_span,
                                        int64_t *selected, double *metric) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_counts, "no counts loaded");
        SEV_REQUIRE(selected != nullptr, "selected is NULL");
        find_variable_features(p->m, p->g, p->cp.get(), p->cnt.get(), nfeatures, loess_span,
                               selected, metric, p->st());
        for (int64_t i = 0; i < nfeatures; ++i) selected[i] += 1;
    });
}

int sev_pipeline_find_variable_features_method(sev_pipeline *p, int method, int64_t nfeatures,
                                               int64_t num_bins, int binning, double alpha_thresh,
                                               int64_t *selected, double *metric) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_counts, "no counts loaded");
        SEV_REQUIRE(p->m >= 2 && selected != nullptr, "bad arguments");
        SEV_REQUIRE(method == SEV_HVG_DISPERSION || method == SEV_HVG_MEANVARPLOT ||
                        method == SEV_HVG_SAUNDERS,
                    "unknown selection method (the vst method is "
                    "sev_pipeline_find_variable_features)");
        SEV_REQUIRE(binning == SEV_BIN_WIDTH,
                    "only width binning: frequency binning puts the smallest mean in bin 0 "
                    "(right-closed cut at the first quantile), outside the bin statistics");
        find_variable_features_method(p->m, p->g, p->cp.get(), p->row.get(), p->cnt.get(), method,
                                      nfeatures, num_bins, alpha_thresh, selected, metric,
                                      p->st());
        for (int64_t i = 0; i < nfeatures; ++i) selected[i] += 1;
    });
}

int sev_pipeline_knn_passes(sev_pipeline *p, int64_t *passes) {
    return guarded([&] {
        SEV_REQUIRE(passes != nullptr, "passes is NULL");
        for (int i = 0; i < 3; ++i) passes[i] = p->knn_passes[i];
    });
}

int sev_pipeline_get_neighbours(sev_pipeline *p, int32_t *nn_index, float *distances) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_nn, "neighbours have not run");
        hipStream_t st = p->st();
        const int64_t m = p->m, k = p->k;
        std::vector<int32_t> idx((size_t)(m * k));
        std::vector<float> dd((size_t)(m * k));
        download(idx.data(), p->nn.get(), idx.size(), st);
        download(dd.data(), p->dist.get(), dd.size(), st);
        SEV_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < m; ++i)
            for (int64_t t = 0; t < k; ++t) {
                if (nn_index) nn_index[i + t * m] = idx[(size_t)(i * k + t)] + 1;
                if (distances) distances[i + t * m] = dd[(size_t)(i * k + t)];
            }
    });
}

int sev_pipeline_snn(sev_pipeline *p, double prune, int dtype, int64_t *nnz) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_nn, "neighbours have not run");
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        hipStream_t st = p->st();
        const int64_t m = p->m, k = p->k;
        p->scp.ensure((size_t)m + 1);
        int64_t total;
        if (dtype == SEV_FLOAT32) {
            std::vector<float> w;
            const int xmin = snn_table<float>(k, prune, w);
            upload(p->wtab32, w.data(), w.size(), st);
            total = snn_count<float>(p->nn.get(), m, k, xmin, p->sw, p->scp.get(), st);
            p->srow.ensure((size_t)total);
            p->sval32.ensure((size_t)total);
            snn_fill<float>(p->nn.get(), m, k, xmin, p->wtab32.get(), p->sw, p->scp.get(),
                            p->srow.get(), p->sval32.get(), st);
        } else {
            std::vector<double> w;
            const int xmin = snn_table<double>(k, prune, w);
            upload(p->wtab64, w.data(), w.size(), st);
            total = snn_count<double>(p->nn.get(), m, k, xmin, p->sw, p->scp.get(), st);
            p->srow.ensure((size_t)total);
            p->sval64.ensure((size_t)total);
            snn_fill<double>(p->nn.get(), m, k, xmin, p->wtab64.get(), p->sw, p->scp.get(),
                             p->srow.get(), p->sval64.get(), st);
        }
        p->snn_nnz = total;
        p->snn_dtype = dtype;
        p->have_snn = true;
        if (nnz) *nnz = total;
    });
}

int sev_pipeline_cluster(sev_pipeline *p, double resolution, int64_t nrandomstarts,
                         int64_t niterations, uint64_t seed, int group_singletons,
                         int64_t *labels, double *modularity, int64_t *nclusters) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_snn, "the SNN graph has not been built");
        SEV_REQUIRE(labels != nullptr, "labels is NULL");
        SEV_REQUIRE(nrandomstarts >= 1 && niterations >= 0,
                    "nrandomstarts must be >= 1 and niterations >= 0");
        hipStream_t st = p->st();
        const int64_t m = p->m;
        const bool f32 = p->snn_dtype == SEV_FLOAT32;
        const void *val = f32 ? (const void *)p->sval32.get() : (const void *)p->sval64.get();
        DevBuf<int32_t> dlab;
        dlab.ensure((size_t)m);
        double q = 0.0;
        const int64_t nc = louvain_cluster(m, p->scp.get(), p->srow.get(), val, f32, p->snn_nnz,
                                           resolution, nrandomstarts, niterations, 1e-4, seed,
                                           group_singletons != 0, dlab.get(), &q, st);
        std::vector<int32_t> lab((size_t)m);
        SEV_HIP(hipMemcpyAsync(lab.data(), dlab.get(), sizeof(int32_t) * (size_t)m,
                               hipMemcpyDeviceToHost, st));
        SEV_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < m; ++i) labels[i] = (int64_t)lab[(size_t)i] + 1;
        if (modularity) *modularity = q;
        if (nclusters) *nclusters = nc;
    });
}

int sev_pipeline_profile(sev_pipeline *p, int enable) {
    return guarded([&] {
        SEV_HIP(hipStreamSynchronize(p->st()));
        p->prof.reset();
        p->prof.on = enable != 0;
        p->prof.only.clear();
    });
}

int sev_pipeline_profile_only(sev_pipeline *p, const char *name) {
    return guarded([&] {
        SEV_REQUIRE(name && name[0], "profile_only: empty region name");
        SEV_HIP(hipStreamSynchronize(p->st()));
        p->prof.reset();
        p->prof.on = true;
        p->prof.only = name;
    });
}

int sev_pipeline_profile_read(sev_pipeline *p, int max, char *names, int64_t *launches,
                              double *ms, double *bytes, double *flops, int *count) {
    return guarded([&] {
        SEV_HIP(hipStreamSynchronize(p->st()));
        auto agg = p->prof.aggregate();
        const int n = (int)std::min<size_t>(agg.size(), (size_t)(max > 0 ? max : 0));
        for (int i = 0; i < n; ++i) {
            if (names) {
                std::memset(names + 32 * i, 0, 32);
                std::strncpy(names + 32 * i, agg[(size_t)i].name.c_str(), 31);
            }
            if (launches) launches[i] = agg[(size_t)i].launches;
            if (ms) ms[i] = agg[(size_t)i].ms;
            if (bytes) bytes[i] = agg[(size_t)i].bytes;
            if (flops) flops[i] = agg[(size_t)i].flops;
        }
        if (count) *count = (int)agg.size();
    });
}

int sev_pipeline_get_snn(sev_pipeline *p, int64_t *colptr, int64_t *rowval, void *nzval) {
    return guarded([&] {
        ProfInstall pi_(&p->prof);
        SEV_REQUIRE(p->have_snn, "snn has not run");
        hipStream_t st = p->st();
        const int64_t m = p->m, nnz = p->snn_nnz;
        std::vector<int32_t> rv((size_t)nnz);
        if (colptr) download(colptr, p->scp.get(), (size_t)m + 1, st);
        if (rowval) download(rv.data(), p->srow.get(), (size_t)nnz, st);
        if (nzval) {
            if (p->snn_dtype == SEV_FLOAT32)
                download((float *)nzval, p->sval32.get(), (size_t)nnz, st);
            else
                download((double *)nzval, p->sval64.get(), (size_t)nnz, st);
        }
        SEV_HIP(hipStreamSynchronize(st));
        if (colptr)
            for (int64_t j = 0; j <= m; ++j) colptr[j] += 1;
        if (rowval)
            for (int64_t i = 0; i < nnz; ++i) rowval[i] = (int64_t)rv[(size_t)i] + 1;
    });
}

// ---------------------------------------------------------------------------
// Drop-ins over host buffers
// ---------------------------------------------------------------------------

int sev_normalize_cells(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                        const void *nzval, int count_type, int method, double scale_factor,
                        int dtype, void *out_nzval) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(method == SEV_LOGNORMALIZE || method == SEV_RELATIVECOUNTS,
                    "unknown normalization method");
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        SEV_REQUIRE(count_type == SEV_INT32 || count_type == SEV_INT64, "bad count type");
        Stream s;
        DevBuf<int64_t> cp;
        DevBuf<int32_t> row, cnt;
        DevBuf<uint32_t> tot;
        int64_t nnz = 0;
        csc_to_device(m, n, colptr, rowval, cp, row, nnz, s.st);
        std::vector<int32_t> c((size_t)nnz);
        for (int64_t i = 0; i < nnz; ++i) {
            const int64_t v = count_type == SEV_INT32 ? ((const int32_t *)nzval)[i]
                                                      : ((const int64_t *)nzval)[i];
            SEV_REQUIRE(v >= 0 && v <= (int64_t)std::numeric_limits<int32_t>::max(),
                        "counts must be non-negative and fit in Int32");
            c[(size_t)i] = (int32_t)v;
        }
        upload_counts(cnt, c.data(), (size_t)nnz, s.st);
        tot.ensure((size_t)m);
        row_totals(cp.get(), row.get(), cnt.get(), nnz, m, n, tot.get(), s.st);
        const int dolog = method == SEV_LOGNORMALIZE;
        if (dtype == SEV_FLOAT32) {
            DevBuf<float> out;
            out.ensure((size_t)nnz);
            lognorm_f32(row.get(), cnt.get(), tot.get(), nnz, (float)scale_factor, dolog, out.get(),
                        s.st);
            download((float *)out_nzval, out.get(), (size_t)nnz, s.st);
            SEV_HIP(hipStreamSynchronize(s.st));
        } else {
            DevBuf<double> out;
            out.ensure((size_t)nnz);
            lognorm_f64(row.get(), cnt.get(), tot.get(), nnz, scale_factor, dolog, out.get(),
                        s.st);
            download((double *)out_nzval, out.get(), (size_t)nnz, s.st);
            SEV_HIP(hipStreamSynchronize(s.st));
        }
    });
}


int sev_scale_features(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                       const void *nzval, int dtype, double scale_max, void *out_nzval,
                       void *out_mu) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        Stream s;
        DevBuf<int64_t> cp;
        DevBuf<int32_t> row;
        int64_t nnz = 0;
        csc_to_device(m, n, colptr, rowval, cp, row, nnz, s.st);
        if (dtype == SEV_FLOAT32)
            scale_host<float>(m, n, cp, nnz, (const float *)nzval, (float)scale_max,
                              (float *)out_nzval, (float *)out_mu, s.st);
        else
            scale_host<double>(m, n, cp, nnz, (const double *)nzval, scale_max,
                               (double *)out_nzval, (double *)out_mu, s.st);
    });
}

int sev_centered_mul(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                     const void *nzval, const void *mu, int dtype, char trans, const double *x,
                     double *y, double alpha, double beta) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        SEV_REQUIRE(trans == 'N' || trans == 'T', "trans must be 'N' or 'T'");
        Stream s;
        DevBuf<int64_t> cp;
        DevBuf<int32_t> row;
        int64_t nnz = 0;
        csc_to_device(m, n, colptr, rowval, cp, row, nnz, s.st);
        const size_t es = dtype == SEV_FLOAT32 ? 4 : 8;
        DevBuf<unsigned char> vals, dmu;
        vals.ensure((size_t)(nnz + kValuePad) * es);
        SEV_HIP(hipMemsetAsync(vals.get(), 0, (size_t)(nnz + kValuePad) * es, s.st));
        if (nnz)
            SEV_HIP(hipMemcpyAsync(vals.get(), nzval, (size_t)nnz * es, hipMemcpyHostToDevice, s.st));
        if (mu) {
            dmu.ensure((size_t)n * es);
            SEV_HIP(hipMemcpyAsync(dmu.get(), mu, (size_t)n * es, hipMemcpyHostToDevice, s.st));
        }
        SlabCsc op;
        IrlbaWork w;
        build_slab_csc(op, m, n, cp.get(), row.get(), vals.get(), dtype == SEV_FLOAT64,
                       mu ? (const void *)dmu.get() : nullptr, false, s.st);
        const int64_t lx = trans == 'N' ? n : m, ly = trans == 'N' ? m : n;
        DevBuf<double> dx, dy;
        upload(dx, x, (size_t)lx, s.st);
        dy.ensure((size_t)ly);
        op_apply(op, trans == 'T', dx.get(), dy.get(), w, s.st);
        std::vector<double> t((size_t)ly);
        download(t.data(), dy.get(), (size_t)ly, s.st);
        SEV_HIP(hipStreamSynchronize(s.st));
        for (int64_t i = 0; i < ly; ++i)
            y[i] = (beta == 0.0) ? alpha * t[(size_t)i] : alpha * t[(size_t)i] + beta * y[i];
    });
}

// A CSC matrix of the caller (Julia layout) and its mu on the device.
struct HeldCsc {
    DevBuf<int64_t> cp;
    DevBuf<int32_t> row;
    DevBuf<unsigned char> vals;
    DevBuf<double> mu;
    DevCsc view;
};

void hold_csc(HeldCsc &h, int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
              const void *nzval, const void *mu, int dtype, hipStream_t st) {
    SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
    int64_t nnz = 0;
    csc_to_device(m, n, colptr, rowval, h.cp, h.row, nnz, st);
    const size_t es = dtype == SEV_FLOAT32 ? 4 : 8;
    SEV_REQUIRE(nnz == 0 || nzval != nullptr, "nzval is NULL");
    h.vals.ensure((size_t)nnz * es);
    if (nnz)
        SEV_HIP(hipMemcpyAsync(h.vals.get(), nzval, (size_t)nnz * es, hipMemcpyHostToDevice, st));
    if (mu) {
        std::vector<double> m64((size_t)n);
        for (int64_t j = 0; j < n; ++j)
            m64[(size_t)j] = dtype == SEV_FLOAT32 ? (double)((const float *)mu)[j]
                                                  : ((const double *)mu)[j];
        upload(h.mu, m64.data(), (size_t)n, st);
        SEV_HIP(hipStreamSynchronize(st));
    }
    h.view.m = m;
    h.view.n = n;
    h.view.nnz = nnz;
    h.view.cp = h.cp.get();
    h.view.row = h.row.get();
    h.view.val = h.vals.get();
    h.view.f64 = dtype == SEV_FLOAT64;
    h.view.mu = mu ? h.mu.get() : nullptr;
}

int sev_centered_mul_mat(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                         const void *nzval, const void *mu, int dtype, char trans, int64_t p,
                         const double *X, int64_t ldx, double *Y, int64_t ldy, double alpha,
                         double beta) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(trans == 'N' || trans == 'T', "trans must be 'N' or 'T'");
        Stream s;
        HeldCsc A;
        hold_csc(A, m, n, colptr, rowval, nzval, mu, dtype, s.st);
        centered_mul_mat(A.view, trans == 'T', p, X, ldx, Y, ldy, alpha, beta, s.st);
    });
}

int sev_centered_gram(int64_t m, int64_t nl, const int64_t *colptr_l, const int64_t *rowval_l,
                      const void *nzval_l, const void *mu_l, int64_t nr,
                      const int64_t *colptr_r, const int64_t *rowval_r, const void *nzval_r,
                      const void *mu_r, int dtype, double *C, int64_t ldc, double alpha,
                      double beta) {
    return guarded([&] {
        ensure_device();
        Stream s;
        HeldCsc L, R;
        hold_csc(L, m, nl, colptr_l, rowval_l, nzval_l, mu_l, dtype, s.st);
        if (colptr_r) hold_csc(R, m, nr, colptr_r, rowval_r, nzval_r, mu_r, dtype, s.st);
        centered_gram(L.view, colptr_r ? &R.view : nullptr, C, ldc, alpha, beta, s.st);
    });
}

int sev_irlba_centered_csc(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                           const void *B, const void *mu, int dtype, int64_t nu, int64_t work,
                           int64_t maxit, double tol, const double *init, double *s, double *U,
                           double *V, int precision, int64_t *iters, int64_t *nmult) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(precision == SEV_PREC_FP32 || precision == SEV_PREC_BF16, "bad precision");
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        SEV_REQUIRE(init && s && U && V, "NULL output");
        Stream st;
        DevBuf<int64_t> cp;
        DevBuf<int32_t> row;
        DevBuf<unsigned char> vals, dmu;
        int64_t nnz = 0;
        csc_to_device(m, n, colptr, rowval, cp, row, nnz, st.st);
        const size_t es = dtype == SEV_FLOAT32 ? 4 : 8;
        vals.ensure((size_t)(nnz + kValuePad) * es);
        SEV_HIP(hipMemsetAsync(vals.get(), 0, (size_t)(nnz + kValuePad) * es, st.st));
        if (nnz)
            SEV_HIP(hipMemcpyAsync(vals.get(), B, (size_t)nnz * es, hipMemcpyHostToDevice, st.st));
        if (mu) {
            dmu.ensure((size_t)n * es);
            SEV_HIP(hipMemcpyAsync(dmu.get(), mu, (size_t)n * es, hipMemcpyHostToDevice, st.st));
        }
        SlabCsc op;
        build_slab_csc(op, m, n, cp.get(), row.get(), vals.get(), dtype == SEV_FLOAT64,
                       mu ? (const void *)dmu.get() : nullptr, precision == SEV_PREC_BF16, st.st);
        IrlbaParams prm;
        prm.nu = nu;
        prm.work = work;
        prm.maxit = maxit;
        prm.tol = tol;
        prm.svtol = tol;
        IrlbaWork w;
        DevBuf<double> dU, dV;
        dU.ensure((size_t)(m * nu));
        dV.ensure((size_t)(n * nu));
        IrlbaOut out;
        try {
            irlba_device(op, prm, init, w, s, dU.get(), dV.get(), st.st, &out);
        } catch (...) {
            if (iters) *iters = out.iters;
            if (nmult) *nmult = out.nmult;
            throw;
        }
        download(U, dU.get(), (size_t)(m * nu), st.st);
        download(V, dV.get(), (size_t)(n * nu), st.st);
        SEV_HIP(hipStreamSynchronize(st.st));
        if (iters) *iters = out.iters;
        if (nmult) *nmult = out.nmult;
    });
}

int irlba(int64_t m, int64_t n, int64_t nu, int64_t work, int64_t maxit, int64_t restart,
          double tol, double *init, double *s, double *U, double *V, sev_irlba_randv_fn randv,
          sev_irlba_matmul_fn matmul, void *data) {
    // src/irlba.jl:73 turns any nonzero status into error("convergence
    // failed"); the actual cause goes to stderr
    const int rc = guarded([&] {
        ensure_device();
        SEV_REQUIRE(m >= 1 && n >= 1 && nu >= 1, "irlba: empty operator or nu < 1");
        SEV_REQUIRE(matmul && s && U && V, "irlba: NULL matmul or output");
        SEV_REQUIRE(init || randv, "irlba: need init or randv");
        Stream st;
        SlabCsc op;  // shape only: the products are the caller's
        op.m = m;
        op.n = n;
        op.nslab = ceil_div(m, kSlabRows);
        std::vector<double> v0((size_t)n);
        if (init)
            std::copy(init, init + n, v0.begin());
        else
            randv(v0.data(), (int32_t)n, data);
        HostOperator hop;
        hop.matmul = matmul;
        hop.randv = randv;
        hop.data = data;
        IrlbaParams prm;
        prm.nu = nu;
        prm.work = work;
        prm.maxit = maxit;
        prm.tol = tol;
        prm.svtol = tol;
        // previous triplets are read before the outputs are written
        std::vector<double> rU, rs, rV;
        if (restart > 0) {
            SEV_REQUIRE(restart <= nu, "irlba: restart exceeds nu");
            rU.assign(U, U + m * restart);
            rs.assign(s, s + restart);
            rV.assign(V, V + n * restart);
            prm.restart = restart;
            prm.rU = rU.data();
            prm.rs = rs.data();
            prm.rV = rV.data();
        }
        IrlbaWork w;
        DevBuf<double> dU, dV;
        dU.ensure((size_t)(m * nu));
        dV.ensure((size_t)(n * nu));
        irlba_device(op, prm, v0.data(), w, s, dU.get(), dV.get(), st.st, nullptr, &hop);
        download(U, dU.get(), (size_t)(m * nu), st.st);
        download(V, dV.get(), (size_t)(n * nu), st.st);
        SEV_HIP(hipStreamSynchronize(st.st));
    });
    if (rc != SEV_OK) std::fprintf(stderr, "[severo_hip] irlba: %s\n", sev_last_error());
    return rc;
}

int sev_find_variable_features(int64_t m, int64_t n, const int64_t *colptr, const int64_t *rowval,
                               const int32_t *nzval, int64_t nfeatures, double loess_span,
                               int64_t *selected, double *metric) {
    (void)rowval;  // the statistics only need each column's values in row order
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(m >= 2 && n >= 1 && colptr && nzval && selected, "bad arguments");
        SEV_REQUIRE(colptr[0] == 1, "colptr must be 1-based");
        std::vector<int64_t> cp((size_t)n + 1);
        for (int64_t j = 0; j <= n; ++j) {
            cp[(size_t)j] = colptr[j] - 1;
            SEV_REQUIRE(j == 0 || cp[(size_t)j] >= cp[(size_t)j - 1], "colptr not monotone");
        }
        const int64_t nnz = cp[(size_t)n];
        Stream s;
        DevBuf<int64_t> dcp;
        DevBuf<int32_t> dcnt;
        upload(dcp, cp.data(), cp.size(), s.st);
        upload_counts(dcnt, nzval, (size_t)std::max<int64_t>(nnz, 1), s.st);
        find_variable_features(m, n, dcp.get(), dcnt.get(), nfeatures, loess_span, selected,
                               metric, s.st);
        for (int64_t i = 0; i < nfeatures; ++i) selected[i] += 1;
    });
}

int sev_find_variable_features_method(int64_t m, int64_t n, const int64_t *colptr,
                                      const int64_t *rowval, const int32_t *nzval, int method,
                                      int64_t nfeatures, int64_t num_bins, int binning,
                                      double alpha_thresh, int64_t *selected, double *metric) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(m >= 2 && n >= 1 && colptr && nzval && selected, "bad arguments");
        SEV_REQUIRE(method == SEV_HVG_DISPERSION || method == SEV_HVG_MEANVARPLOT ||
                        method == SEV_HVG_SAUNDERS,
                    "unknown selection method (the vst method is sev_find_variable_features)");
        SEV_REQUIRE(binning == SEV_BIN_WIDTH,
                    "only width binning: frequency binning puts the smallest mean in bin 0 "
                    "(right-closed cut at the first quantile), outside the bin statistics");
        Stream s;
        DevBuf<int64_t> dcp;
        DevBuf<int32_t> drow, dcnt;
        int64_t nnz = 0;
        csc_to_device(m, n, colptr, rowval, dcp, drow, nnz, s.st);
        for (int64_t p = 0; p < nnz; ++p)
            SEV_REQUIRE(nzval[p] >= 0, "counts must be non-negative");
        upload_counts(dcnt, nzval, (size_t)std::max<int64_t>(nnz, 1), s.st);
        find_variable_features_method(m, n, dcp.get(), drow.get(), dcnt.get(), method, nfeatures,
                                      num_bins, alpha_thresh, selected, metric, s.st);
        for (int64_t i = 0; i < nfeatures; ++i) selected[i] += 1;
    });
}

int sev_knn(const void *X, int64_t n, int64_t d, int64_t ld, int dtype, int64_t k,
            int include_self, int precision, int32_t *nn_index, void *distances,
            int64_t *n_fallback) {
    return guarded([&] {
        SEV_REQUIRE(k >= 1 && k <= 64,
                    "knn: k must be in [1, 64] (the longest neighbour list the search keeps)");
        ensure_device();
        SEV_REQUIRE(precision == SEV_PREC_FP32 || precision == SEV_PREC_BF16 ||
                        precision == SEV_PREC_EXACT || precision == SEV_PREC_FP8,
                    "bad precision");
        if (dtype == SEV_FLOAT32)
            knn_host<float>((const float *)X, n, d, ld, k, include_self, precision, nn_index,
                            (float *)distances, n_fallback);
        else if (dtype == SEV_FLOAT64)
            knn_host<double>((const double *)X, n, d, ld, k, include_self, precision, nn_index,
                             (double *)distances, n_fallback);
        else
            fail(kErrArg, "bad dtype");
    });
}

int sev_knn_cosine(const void *X, int64_t n, int64_t d, int64_t ld, int dtype, int64_t k,
                   int include_self, int precision, int32_t *nn_index, void *distances,
                   int64_t *n_fallback) {
    return guarded([&] {
        SEV_REQUIRE(k >= 1 && k <= 64,
                    "knn: k must be in [1, 64] (the longest neighbour list the search keeps)");
        ensure_device();
        SEV_REQUIRE(precision == SEV_PREC_FP32 || precision == SEV_PREC_BF16 ||
                        precision == SEV_PREC_EXACT || precision == SEV_PREC_FP8,
                    "bad precision");
        if (dtype == SEV_FLOAT32)
            knn_host<float>((const float *)X, n, d, ld, k, include_self, precision, nn_index,
                            (float *)distances, n_fallback, true);
        else if (dtype == SEV_FLOAT64)
            knn_host<double>((const double *)X, n, d, ld, k, include_self, precision, nn_index,
                             (double *)distances, n_fallback, true);
        else
            fail(kErrArg, "bad dtype");
    });
}

// The libcell drop-ins have no status (src/neighbours.jl:32-74 reads the
// outputs directly).  On failure the message goes to stderr, every index is
// set to 0 (outside Julia's 1-based range, so the sparse() of
// src/neighbours.jl:79 raises instead of consuming undefined memory) and the
// distances to NaN.
extern "C++" template <class T>
static void dropin_knn_status(int rc, const char *name, int n, int k, int32_t *nn_index,
                              T *distances) {
    if (rc == SEV_OK) return;
    std::fprintf(stderr, "[severo_hip] %s: %s\n", name, sev_last_error());
    if (n > 0 && k > 0) {
        const size_t cnt = (size_t)n * (size_t)k;
        if (nn_index) std::fill(nn_index, nn_index + cnt, 0);
        if (distances) std::fill(distances, distances + cnt, std::numeric_limits<T>::quiet_NaN());
    }
}

void FindNeighboursEuclidean32(const float *X, int n, int d, int ld, int k, int ntables,
                               int include_self, int seed, int32_t *nn_index, float *distances) {
    (void)ntables;
    (void)seed;
    dropin_knn_status(sev_knn(X, n, d, ld, SEV_FLOAT32, k, include_self, SEV_PREC_FP32, nn_index,
                              distances, nullptr),
                      "FindNeighboursEuclidean32", n, k, nn_index, distances);
}

void FindNeighboursEuclidean64(const double *X, int n, int d, int ld, int k, int ntables,
                               int include_self, int seed, int32_t *nn_index, double *distances) {
    (void)ntables;
    (void)seed;
    dropin_knn_status(sev_knn(X, n, d, ld, SEV_FLOAT64, k, include_self, SEV_PREC_FP32, nn_index,
                              distances, nullptr),
                      "FindNeighboursEuclidean64", n, k, nn_index, distances);
}

void FindNeighboursCosine32(const float *X, int n, int d, int ld, int k, int ntables,
                            int include_self, int seed, int32_t *nn_index, float *distances) {
    (void)ntables;
    (void)seed;
    dropin_knn_status(sev_knn_cosine(X, n, d, ld, SEV_FLOAT32, k, include_self, SEV_PREC_FP32,
                                     nn_index, distances, nullptr),
                      "FindNeighboursCosine32", n, k, nn_index, distances);
}

void FindNeighboursCosine64(const double *X, int n, int d, int ld, int k, int ntables,
                            int include_self, int seed, int32_t *nn_index, double *distances) {
    (void)ntables;
    (void)seed;
    dropin_knn_status(sev_knn_cosine(X, n, d, ld, SEV_FLOAT64, k, include_self, SEV_PREC_FP32,
                                     nn_index, distances, nullptr),
                      "FindNeighboursCosine64", n, k, nn_index, distances);
}

int sev_cluster_louvain(int64_t n, const int64_t *colptr, const int64_t *rowval,
                        const void *nzval, int dtype, double resolution, int64_t nrandomstarts,
                        int64_t niterations, uint64_t seed, int group_singletons,
                        int64_t *labels, double *modularity, int64_t *nclusters) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        SEV_REQUIRE(n >= 1 && labels != nullptr, "bad arguments");
        SEV_REQUIRE(nrandomstarts >= 1 && niterations >= 0,
                    "nrandomstarts must be >= 1 and niterations >= 0");
        Stream s;
        DevBuf<int64_t> dcp;
        DevBuf<int32_t> drow;
        int64_t nnz = 0;
        csc_to_device(n, n, colptr, rowval, dcp, drow, nnz, s.st);
        SEV_REQUIRE(nnz == 0 || nzval != nullptr, "nzval is NULL");
        const bool f32 = dtype == SEV_FLOAT32;
        for (int64_t p = 0; p < nnz; ++p) {
            const double x = f32 ? (double)((const float *)nzval)[p] : ((const double *)nzval)[p];
            SEV_REQUIRE(std::isfinite(x) && x >= 0.0, "graph weights must be finite and >= 0");
        }
        DevBuf<float> dw32;
        DevBuf<double> dw64;
        const void *dw = nullptr;
        if (f32) {
            upload(dw32, (const float *)nzval, (size_t)std::max<int64_t>(nnz, 1), s.st);
            dw = dw32.get();
        } else {
            upload(dw64, (const double *)nzval, (size_t)std::max<int64_t>(nnz, 1), s.st);
            dw = dw64.get();
        }
        DevBuf<int32_t> dlab;
        dlab.ensure((size_t)n);
        double q = 0.0;
        const int64_t nc = louvain_cluster(n, dcp.get(), drow.get(), dw, f32, nnz, resolution,
                                           nrandomstarts, niterations, 1e-4, seed,
                                           group_singletons != 0, dlab.get(), &q, s.st);
        std::vector<int32_t> lab((size_t)n);
        SEV_HIP(hipMemcpyAsync(lab.data(), dlab.get(), sizeof(int32_t) * (size_t)n,
                               hipMemcpyDeviceToHost, s.st));
        SEV_HIP(hipStreamSynchronize(s.st));
        for (int64_t i = 0; i < n; ++i) labels[i] = (int64_t)lab[(size_t)i] + 1;
        if (modularity) *modularity = q;
        if (nclusters) *nclusters = nc;
    });
}

int sev_snn_jaccard(int64_t n, int64_t k, const int32_t *nn_index, double prune, int dtype,
                    int64_t *nnz, int64_t *colptr, int64_t *rowval, void *nzval) {
    return guarded([&] {
        ensure_device();
        SEV_REQUIRE(dtype == SEV_FLOAT32 || dtype == SEV_FLOAT64, "bad dtype");
        SEV_REQUIRE(n >= 1 && k >= 1 && k <= kMaxK, "k must be in [1, 64]");
        SEV_REQUIRE(nnz != nullptr, "nnz is NULL");
        std::vector<int32_t> nn((size_t)(n * k));
        for (int64_t j = 0; j < n; ++j)
            for (int64_t t = 0; t < k; ++t) {
                const int32_t l = nn_index[j + t * n] - 1;
                SEV_REQUIRE(l >= 0 && l < n, "neighbour index out of range");
                nn[(size_t)(j * k + t)] = l;
            }
        Stream s;
        DevBuf<int32_t> dnn;
        upload(dnn, nn.data(), nn.size(), s.st);
        SnnWork w;
        DevBuf<int64_t> dcp;
        dcp.ensure((size_t)n + 1);
        DevBuf<int32_t> drow;
        int64_t total;
        std::vector<int32_t> rv;
        if (dtype == SEV_FLOAT32) {
            std::vector<float> wt;
            const int xmin = snn_table<float>(k, prune, wt);
            total = snn_count<float>(dnn.get(), n, k, xmin, w, dcp.get(), s.st);
            *nnz = total;
            if (!colptr) return;
            DevBuf<float> dw, dv;
            upload(dw, wt.data(), wt.size(), s.st);
            drow.ensure((size_t)total);
            dv.ensure((size_t)total);
            snn_fill<float>(dnn.get(), n, k, xmin, dw.get(), w, dcp.get(), drow.get(), dv.get(),
                            s.st);
            rv.resize((size_t)total);
            download(rv.data(), drow.get(), (size_t)total, s.st);
            download((float *)nzval, dv.get(), (size_t)total, s.st);
        } else {
            std::vector<double> wt;
            const int xmin = snn_table<double>(k, prune, wt);
            total = snn_count<double>(dnn.get(), n, k, xmin, w, dcp.get(), s.st);
            *nnz = total;
            if (!colptr) return;
            DevBuf<double> dw, dv;
            upload(dw, wt.data(), wt.size(), s.st);
            drow.ensure((size_t)total);
            dv.ensure((size_t)total);
            snn_fill<double>(dnn.get(), n, k, xmin, dw.get(), w, dcp.get(), drow.get(), dv.get(),
                             s.st);
            rv.resize((size_t)total);
            download(rv.data(), drow.get(), (size_t)total, s.st);
            download((double *)nzval, dv.get(), (size_t)total, s.st);
        }
        download(colptr, dcp.get(), (size_t)n + 1, s.st);
        SEV_HIP(hipStreamSynchronize(s.st));
        for (int64_t j = 0; j <= n; ++j) colptr[j] += 1;
        for (int64_t i = 0; i < total; ++i) rowval[i] = (int64_t)rv[(size_t)i] + 1;
    });
}

}  // extern "C"
