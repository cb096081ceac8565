// find_variable_features(:vst) (src/variablefeatures.jl:34-50,128-161) on the
// device: per-gene count mean and sd, the loess fit of log10 sd against
// log10 mean (span 0.5 by default, degree 2), the clipped standardized
// variance per gene, and the top genes by that variance.
//
// Loess.jl, which the reference calls (`loess(xs, ys, span=...)` then
// `predict(model, xs)`), fits at kd-tree vertices and interpolates between
// them.  Here the local regression is evaluated exactly at every point:
//   q = floor(N span) (at least min(N, 3)) nearest points -- the window of q
//   consecutive points (x ascending) containing the first point equal to
//   x_i with the smallest dmax = max(x_i - x_s, x_{s+q-1} - x_i), the
//   smaller s on a tie;
//   w_j = (1 - (|x_j - x_i|/dmax)^3)^3, u_j = (x_j - x_i)/dmax;
//   y ~ a + b u + c u^2 by weighted least squares (normal equations by
//   elimination, degree lowered when a pivot vanishes); fit_i = a.
#include "sev_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace sev {

namespace {

constexpr int kBlock = 256;

// one wavefront per evaluation point
__global__ void __launch_bounds__(kBlock)
k_loess(int64_t N, const double *x, const double *y, int64_t q, double *fit) {
    const int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (i >= N) return;
    const int lane = lane_id();
    const double xi = x[i];
    // the fit depends on x_i only: tied points use the window of the first
    // of them, so their sums (and fits) are identical
    int64_t i0;
    {
        int64_t a = 0, b = i;
        while (a < b) {
            const int64_t mid = a + (b - a) / 2;
            if (x[mid] < xi) a = mid + 1;
            else b = mid;
        }
        i0 = a;
    }
    const int64_t lo = i0 - q + 1 < 0 ? 0 : i0 - q + 1;
    const int64_t hi = i0 < N - q ? i0 : N - q;
    int64_t a = lo, b = hi + 1;
    while (a < b) {
        const int64_t mid = a + (b - a) / 2;
        if (x[mid + q - 1] - xi >= xi - x[mid]) b = mid;
        else a = mid + 1;
    }
    int64_t s;
    double dmax;
    if (a > hi) {
        s = hi;
        dmax = xi - x[hi];
    } else {
        const double dr = x[a + q - 1] - xi;
        if (a > lo && xi - x[a - 1] <= dr) {
            s = a - 1;
            dmax = xi - x[a - 1];
        } else {
            s = a;
            dmax = dr;
        }
    }
    double S0 = 0, S1 = 0, S2 = 0, S3 = 0, S4 = 0, T0 = 0, T1 = 0, T2 = 0;
    for (int64_t j = s + lane; j < s + q; j += kWave) {
        double w, u;
        if (dmax > 0.0) {
            u = (x[j] - xi) / dmax;
            const double r = fabs(u);
            const double t = 1.0 - r * r * r;
            w = r < 1.0 ? t * t * t : 0.0;
        } else {
            u = 0.0;
            w = 1.0;
        }
        const double wu = w * u, wu2 = wu * u;
        S0 += w;
        S1 += wu;
        S2 += wu2;
        S3 += wu2 * u;
        S4 += wu2 * u * u;
        T0 += w * y[j];
        T1 += wu * y[j];
        T2 += wu2 * y[j];
    }
    S0 = wave_sum(S0);
    S1 = wave_sum(S1);
    S2 = wave_sum(S2);
    S3 = wave_sum(S3);
    S4 = wave_sum(S4);
    T0 = wave_sum(T0);
    T1 = wave_sum(T1);
    T2 = wave_sum(T2);
    if (lane != 0) return;
    double f;
    if (!(S0 > 0.0)) {
        f = y[i0];
    } else {
        const double l21 = S1 / S0, l31 = S2 / S0;
        const double a22 = S2 - l21 * S1, a23 = S3 - l21 * S2, a33 = S4 - l31 * S2;
        const double b2 = T1 - l21 * T0, b3 = T2 - l31 * T0;
        if (!(a22 > 1e-12 * S2)) {
            f = T0 / S0;
        } else {
            const double l32 = a23 / a22;
            const double a33b = a33 - l32 * a23, b3b = b3 - l32 * b2;
            if (!(a33b > 1e-12 * S4)) {
                const double bl = b2 / a22;
                f = (T0 - S1 * bl) / S0;
            } else {
                const double c = b3b / a33b;
                const double bb = (b2 - a23 * c) / a22;
                f = (T0 - S1 * bb - S2 * c) / S0;
            }
        }
    }
    fit[i] = f;
}

// standardized_var_clipped (src/variablefeatures.jl:21-28), one wavefront
// per gene: (sum_nz min((x - mu)/sd, vmax)^2 + (m - nnz) min(-mu/sd, vmax)^2)
// / (m - 1) for sd > 0, else 0; vmax = sqrt(m).
__global__ void __launch_bounds__(kBlock)
k_standardized_var(int64_t m, int64_t ncol, const int64_t *colptr, const int32_t *x,
                   const double *mu, const double *sd, double *out) {
    const int64_t col = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    if (col >= ncol) return;
    const int lane = lane_id();
    const double sdj = sd[col], muj = mu[col];
    if (sdj == 0.0) {
        if (lane == 0) out[col] = 0.0;
        return;
    }
    const double vmax = sqrt((double)m);
    const int64_t b = colptr[col], e = colptr[col + 1];
    double t = 0.0;
    for (int64_t p = b + lane; p < e; p += kWave) {
        double v = ((double)x[p] - muj) / sdj;
        if (v > vmax) v = vmax;
        t += v * v;
    }
    t = wave_sum(t);
    if (lane == 0) {
        double z = (0.0 - muj) / sdj;
        if (z > vmax) z = vmax;
        out[col] = (t + (double)(m - (e - b)) * (z * z)) / (double)(m - 1);
    }
}

// partialsortperm(metric, 1:nfeatures, rev=true) (src/variablefeatures.jl:159);
// ties to the lower gene index
void select_top(const std::vector<double> &mh, int64_t nfeat, int64_t *selected) {
    const int64_t G = (int64_t)mh.size();
    std::vector<int64_t> idx((size_t)G);
    std::iota(idx.begin(), idx.end(), (int64_t)0);
    std::partial_sort(idx.begin(), idx.begin() + nfeat, idx.end(), [&](int64_t a, int64_t b) {
        const double ma = mh[(size_t)a], mb = mh[(size_t)b];
        return ma > mb || (ma == mb && a < b);
    });
    for (int64_t i = 0; i < nfeat; ++i) selected[i] = idx[(size_t)i];
}

}  // namespace

void find_variable_features(int64_t m, int64_t G, const int64_t *colptr_d, const int32_t *cnt_d,
                            int64_t nfeat, double span, int64_t *selected, double *metric_out,
                            hipStream_t st) {
    SEV_REQUIRE(m >= 2 && G >= 1, "find_variable_features: need at least two cells");
    SEV_REQUIRE(nfeat >= 1 && nfeat <= G, "find_variable_features: nfeatures out of range");
    SEV_REQUIRE(span > 0.0, "find_variable_features: loess_span must be positive");
    DevBuf<double> mu, sd, xs, ys, fit, esd, met;
    mu.ensure((size_t)G);
    sd.ensure((size_t)G);
    count_mean_std(m, G, colptr_d, cnt_d, mu.get(), sd.get(), st);
    std::vector<double> muh((size_t)G), sdh((size_t)G);
    SEV_HIP(hipMemcpyAsync(muh.data(), mu.get(), sizeof(double) * G, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipMemcpyAsync(sdh.data(), sd.get(), sizeof(double) * G, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    // non-constant genes, ordered by (log10 mean, gene) for the fit
    std::vector<int64_t> nc;
    for (int64_t g = 0; g < G; ++g)
        if (sdh[(size_t)g] > 0.0) nc.push_back(g);
    const int64_t N = (int64_t)nc.size();
    std::vector<double> lx((size_t)N), ly((size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        lx[(size_t)i] = std::log10(muh[(size_t)nc[(size_t)i]]);
        ly[(size_t)i] = std::log10(sdh[(size_t)nc[(size_t)i]]);
    }
    std::vector<int64_t> ord((size_t)N);
    std::iota(ord.begin(), ord.end(), (int64_t)0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int64_t a, int64_t b) { return lx[(size_t)a] < lx[(size_t)b]; });
    std::vector<double> esdh(sdh);
    if (N > 0) {
        std::vector<double> xsh((size_t)N), ysh((size_t)N), fh((size_t)N);
        for (int64_t i = 0; i < N; ++i) {
            xsh[(size_t)i] = lx[(size_t)ord[(size_t)i]];
            ysh[(size_t)i] = ly[(size_t)ord[(size_t)i]];
        }
        int64_t q = (int64_t)std::floor((double)N * span);
        const int64_t qmin = N < 3 ? N : 3;
        if (q < qmin) q = qmin;
        if (q > N) q = N;
        xs.ensure((size_t)N);
        ys.ensure((size_t)N);
        fit.ensure((size_t)N);
        SEV_HIP(hipMemcpyAsync(xs.get(), xsh.data(), sizeof(double) * N, hipMemcpyHostToDevice,
                               st));
        SEV_HIP(hipMemcpyAsync(ys.get(), ysh.data(), sizeof(double) * N, hipMemcpyHostToDevice,
                               st));
        {
            ProfScope ps("loess", st, 0.0, 20.0 * (double)N * (double)q);
            hipLaunchKernelGGL(k_loess, dim3((unsigned)ceil_div(N * kWave, kBlock)), dim3(kBlock),
                               0, st, N, xs.get(), ys.get(), q, fit.get());
            SEV_LAUNCH_CHECK();
        }
        SEV_HIP(hipMemcpyAsync(fh.data(), fit.get(), sizeof(double) * N, hipMemcpyDeviceToHost,
                               st));
        SEV_HIP(hipStreamSynchronize(st));
        // expected sd = 10^fit; NaN -> 0 (the reference's nan2zero!, :45-46)
        for (int64_t i = 0; i < N; ++i) {
            const double e = std::pow(10.0, fh[(size_t)i]);
            esdh[(size_t)nc[(size_t)ord[(size_t)i]]] = std::isnan(e) ? 0.0 : e;
        }
    }
    esd.ensure((size_t)G);
    met.ensure((size_t)G);
    SEV_HIP(hipMemcpyAsync(esd.get(), esdh.data(), sizeof(double) * G, hipMemcpyHostToDevice, st));
    {
        ProfScope ps("std_var", st);
        hipLaunchKernelGGL(k_standardized_var, dim3((unsigned)ceil_div(G * kWave, kBlock)),
                           dim3(kBlock), 0, st, m, G, colptr_d, cnt_d, mu.get(), esd.get(),
                           met.get());
        SEV_LAUNCH_CHECK();
    }
    std::vector<double> mh((size_t)G);
    SEV_HIP(hipMemcpyAsync(mh.data(), met.get(), sizeof(double) * G, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    select_top(mh, nfeat, selected);
    if (metric_out) std::copy(mh.begin(), mh.end(), metric_out);
}

namespace {

double nan2zero(double v) { return std::isnan(v) ? 0.0 : v; }

// x with P(Z > x) = tail for a standard normal Z: Newton steps on erfc from
// above the root (the quantile of Normal() at 1 - tail)
double norm_quantile_upper(double tail) {
    double x = std::sqrt(-2.0 * std::log(tail));
    for (int it = 0; it < 100; ++it) {
        const double f = 0.5 * std::erfc(x / std::sqrt(2.0)) - tail;
        const double dfdx = -std::exp(-0.5 * x * x) / std::sqrt(2.0 * 3.141592653589793);
        const double step = f / dfdx;
        x -= step;
        if (std::fabs(step) <= 1e-16 * std::fabs(x)) break;
    }
    return x;
}

}  // namespace

// :dispersion, :meanvarplot and :saunders (src/variablefeatures.jl:60-91,
// :128-161) on row_norm(A, 1.0): the per-gene FP64 mean and variance of the
// relative counts on the device, then the O(genes) selection rules:
//   dispersion   log(var/mu), NaN -> 0 (log_VMR, src/scaling.jl:190-192);
//   meanvarplot  log1p(mu) cut into nbins bins of equal width (breaks from
//                min to max, widened by (max - min)/1000 at both ends,
//                right-closed, src/utils.jl:109-148), the dispersion's
//                z-score within its bin from the labelled Welford in gene
//                order (src/scaling.jl:90-112), NaN -> 0;
//   saunders     nolan = mean(1 ./ cell totals) (summed in cell order),
//                q = quantile(Normal(), 1 - alpha/G/2), genes with
//                var/nolan > mu + q sqrt(mu nolan / m) get
//                log10(var) - log10(mu nolan), the others 0.
void find_variable_features_method(int64_t m, int64_t G, const int64_t *colptr_d,
                                   const int32_t *row_d, const int32_t *cnt_d, int method,
                                   int64_t nfeat, int64_t nbins, double alpha, int64_t *selected,
                                   double *metric_out, hipStream_t st) {
    SEV_REQUIRE(m >= 2 && G >= 1, "find_variable_features: need at least two cells");
    SEV_REQUIRE(nfeat >= 1 && nfeat <= G, "find_variable_features: nfeatures out of range");
    SEV_REQUIRE(method >= 1 && method <= 3, "find_variable_features: unknown method");
    SEV_REQUIRE(method != 2 || nbins >= 1, "find_variable_features: num_bins must be positive");
    SEV_REQUIRE(method != 3 || alpha > 0.0, "find_variable_features: alpha_thresh must be positive");
    int64_t nnz = 0;
    SEV_HIP(hipMemcpyAsync(&nnz, colptr_d + G, sizeof nnz, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipStreamSynchronize(st));
    DevBuf<uint32_t> tot;
    DevBuf<double> mu, var;
    tot.ensure((size_t)m);
    mu.ensure((size_t)G);
    var.ensure((size_t)G);
    row_totals(colptr_d, row_d, cnt_d, nnz, m, G, tot.get(), st);
    rel_mean_var(m, G, colptr_d, row_d, cnt_d, tot.get(), mu.get(), var.get(), st);
    std::vector<double> muh((size_t)G), varh((size_t)G), mh((size_t)G, 0.0);
    std::vector<uint32_t> toth;
    SEV_HIP(hipMemcpyAsync(muh.data(), mu.get(), sizeof(double) * G, hipMemcpyDeviceToHost, st));
    SEV_HIP(hipMemcpyAsync(varh.data(), var.get(), sizeof(double) * G, hipMemcpyDeviceToHost, st));
    if (method == 3) {
        toth.resize((size_t)m);
        SEV_HIP(hipMemcpyAsync(toth.data(), tot.get(), sizeof(uint32_t) * m,
                               hipMemcpyDeviceToHost, st));
    }
    SEV_HIP(hipStreamSynchronize(st));
    if (method == 1) {
        for (int64_t g = 0; g < G; ++g)
            mh[(size_t)g] = nan2zero(std::log(varh[(size_t)g] / muh[(size_t)g]));
    } else if (method == 2) {
        std::vector<double> lmu((size_t)G), disp((size_t)G);
        double lo = 0.0, hi = 0.0;
        for (int64_t g = 0; g < G; ++g) {
            lmu[(size_t)g] = nan2zero(std::log1p(muh[(size_t)g]));
            disp[(size_t)g] = nan2zero(std::log(varh[(size_t)g] / muh[(size_t)g]));
            if (g == 0 || lmu[(size_t)g] < lo) lo = lmu[(size_t)g];
            if (g == 0 || lmu[(size_t)g] > hi) hi = lmu[(size_t)g];
        }
        const double dx = hi - lo;
        std::vector<double> br((size_t)nbins + 1);
        for (int64_t i = 0; i <= nbins; ++i)
            br[(size_t)i] = (double)(((long double)lo * (long double)(nbins - i) +
                                      (long double)hi * (long double)i) /
                                     (long double)nbins);
        br[0] = lo - dx / 1000;
        br[(size_t)nbins] = hi + dx / 1000;
        std::vector<int64_t> lab((size_t)G);
        for (int64_t g = 0; g < G; ++g) {
            const double x = lmu[(size_t)g];
            // searchsortedlast, then one down when x sits on that break
            int64_t idx = std::upper_bound(br.begin(), br.end(), x) - br.begin();
            if (idx >= 1 && br[(size_t)idx - 1] == x) idx -= 1;
            lab[(size_t)g] = idx;
        }
        std::vector<double> bm((size_t)nbins + 2, 0.0), bv((size_t)nbins + 2, 0.0);
        std::vector<int64_t> bn((size_t)nbins + 2, 0);
        for (int64_t g = 0; g < G; ++g) {
            const int64_t k = lab[(size_t)g];
            const double v = disp[(size_t)g];
            bn[(size_t)k] += 1;
            const double delta = v - bm[(size_t)k];
            bm[(size_t)k] = bm[(size_t)k] + delta / (double)bn[(size_t)k];
            bv[(size_t)k] = bv[(size_t)k] + delta * (v - bm[(size_t)k]);
        }
        for (int64_t g = 0; g < G; ++g) {
            const int64_t k = lab[(size_t)g];
            const double sd = std::sqrt(bv[(size_t)k] / (double)(bn[(size_t)k] - 1));
            mh[(size_t)g] = nan2zero((disp[(size_t)g] - bm[(size_t)k]) / sd);
        }
    } else {
        double acc = 0.0;
        for (int64_t i = 0; i < m; ++i) acc += 1.0 / (double)toth[(size_t)i];
        const double nolan = acc / (double)m;
        const double a = alpha / (double)G;
        const double p = 1.0 - a / 2.0;
        const double q = norm_quantile_upper(1.0 - p);
        for (int64_t g = 0; g < G; ++g) {
            const double mu_g = muh[(size_t)g], var_g = varh[(size_t)g];
            const double upper = mu_g + q * std::sqrt(mu_g * nolan / (double)m);
            if (var_g / nolan > upper)
                mh[(size_t)g] = std::log10(var_g) - std::log10(mu_g * nolan);
        }
    }
    select_top(mh, nfeat, selected);
    if (metric_out) std::copy(mh.begin(), mh.end(), metric_out);
}

}  // namespace sev
