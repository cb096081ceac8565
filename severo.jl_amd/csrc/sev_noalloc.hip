// Allocation-free entry point: normalize_cells on a count matrix that is
// already in device memory, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The per-cell totals live in a static device array of fixed
// size instead of a buffer allocated per call.
//
// The array is a file-scope __device__ symbol; host code never passes its
// name to a kernel (on the host that name is only a stand-in).  Its address
// on each device is resolved once, with that device current, by
// hipGetSymbolAddress (from sev_device_init, or at the first call on the
// device) and kept per device.
#include "../../include/severo_hip.h"
#include "sev_kernels.hpp"

#include <mutex>

namespace sev {

namespace {

constexpr int64_t kNoAllocMaxCells = (int64_t)1 << 23;  // 8,388,608 cells: 32 MiB
constexpr int kMaxDevices = 64;

// per-cell count totals of the allocation-free normalize_cells
__device__ __align__(256) uint32_t scratch[kNoAllocMaxCells];

std::mutex g_scratch_mu;
uint32_t *g_scratch[kMaxDevices] = {};

uint32_t *scratch_on_current_device() {
    int dev = 0;
    SEV_HIP(hipGetDevice(&dev));
    SEV_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        SEV_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        SEV_REQUIRE(ptr != nullptr, "no device address for the scratch array");
        g_scratch[dev] = (uint32_t *)ptr;
    }
    return g_scratch[dev];
}

struct NormArgs {
    int64_t m, n, nnz;
    const int64_t *colptr;
    const int32_t *rowval;
    const int32_t *counts;
    int method;
    double scale_factor;
    int dtype;
    void *out;
    hipStream_t st;
};

void normalize_noalloc(void *ctx) {
    const NormArgs &a = *(const NormArgs *)ctx;
    SEV_REQUIRE(a.m >= 0 && a.n >= 0 && a.nnz >= 0, "negative size");
    SEV_REQUIRE(a.m <= kNoAllocMaxCells,
                "sev_normalize_cells_noalloc: more cells than its scratch holds (8,388,608)");
    SEV_REQUIRE(a.method == SEV_LOGNORMALIZE || a.method == SEV_RELATIVECOUNTS,
                "unknown normalization method");
    SEV_REQUIRE(a.dtype == SEV_FLOAT32 || a.dtype == SEV_FLOAT64, "bad dtype");
    SEV_REQUIRE(a.n == 0 || a.colptr, "colptr is NULL");
    SEV_REQUIRE(a.nnz == 0 || (a.rowval && a.counts && a.out), "NULL device array");
    require_device();
    uint32_t *tot = scratch_on_current_device();
    row_totals_fill(a.colptr, a.rowval, a.counts, a.nnz, a.m, a.n, tot, a.st);
    const int dolog = a.method == SEV_LOGNORMALIZE;
    if (a.dtype == SEV_FLOAT32)
        lognorm_f32(a.rowval, a.counts, tot, a.nnz, (float)a.scale_factor, dolog, (float *)a.out,
                    a.st);
    else
        lognorm_f64(a.rowval, a.counts, tot, a.nnz, a.scale_factor, dolog, (double *)a.out, a.st);
}

void setup_noalloc(void *) {
    require_device();
    (void)scratch_on_current_device();
}

}  // namespace

void noalloc_setup() { setup_noalloc(nullptr); }

}  // namespace sev

using namespace sev;

extern "C" int sev_normalize_cells_noalloc(int64_t m, int64_t n, int64_t nnz,
                                           const int64_t *colptr,
                                           const int32_t *rowval, const int32_t *counts,
                                           int method, double scale_factor, int dtype,
                                           void *out_nzval, void *stream) {
    NormArgs a{m,      n,     nnz,          colptr, rowval,   counts,
               method, scale_factor, dtype, out_nzval, (hipStream_t)stream};
    return guarded_call(normalize_noalloc, &a);
}
