// Small dense SVD for the Lanczos projected matrix (work x work, FP64, host).
//
// One-sided Jacobi (Hestenes) applied to B': rotating the columns of B' by
// plane rotations J until they are mutually orthogonal gives B' J = V S, so
// the left singular vectors of B are the accumulated rotations U = J (always
// a complete orthonormal basis, which the residual estimate
// res_i = ||F|| * U(work-1, i) relies on) and V = columns of B'J normalised.
// Columns are visited in cyclic order; sweeps stop when every pair is
// orthogonal to working precision.  Relative accuracy of small singular
// values is preserved (no B'B is formed).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "sev_kernels.hpp"

namespace sev {

void small_svd(int64_t n, const double *A, double *U, double *s, double *V) {
    const size_t nn = (size_t)n * (size_t)n;
    std::vector<double> G(nn), J(nn, 0.0);
    // G = A' (column-major): G(i, j) = A(j, i)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) G[(size_t)j * n + i] = A[(size_t)i * n + j];
    for (int64_t i = 0; i < n; ++i) J[(size_t)i * n + i] = 1.0;
    const double tol = 2.220446049250313e-16 * std::sqrt((double)n);
    for (int sweep = 0; sweep < 100; ++sweep) {
        bool rotated = false;
        for (int64_t p = 0; p + 1 < n; ++p) {
            for (int64_t q = p + 1; q < n; ++q) {
                double *gp = &G[(size_t)p * n], *gq = &G[(size_t)q * n];
                double app = 0.0, aqq = 0.0, apq = 0.0;
                for (int64_t i = 0; i < n; ++i) {
                    app += gp[i] * gp[i];
                    aqq += gq[i] * gq[i];
                    apq += gp[i] * gq[i];
                }
                if (apq == 0.0 || std::fabs(apq) <= tol * std::sqrt(app * aqq)) continue;
                rotated = true;
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = std::copysign(1.0, tau) / (std::fabs(tau) + std::hypot(1.0, tau));
                const double c = 1.0 / std::hypot(1.0, t);
                const double sn = c * t;
                for (int64_t i = 0; i < n; ++i) {
                    const double x = gp[i], y = gq[i];
                    gp[i] = c * x - sn * y;
                    gq[i] = sn * x + c * y;
                }
                double *jp = &J[(size_t)p * n], *jq = &J[(size_t)q * n];
                for (int64_t i = 0; i < n; ++i) {
                    const double x = jp[i], y = jq[i];
                    jp[i] = c * x - sn * y;
                    jq[i] = sn * x + c * y;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<double> sv((size_t)n);
    for (int64_t j = 0; j < n; ++j) {
        double t = 0.0;
        for (int64_t i = 0; i < n; ++i) t += G[(size_t)j * n + i] * G[(size_t)j * n + i];
        sv[(size_t)j] = std::sqrt(t);
    }
    std::vector<int64_t> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int64_t a, int64_t b) { return sv[(size_t)a] > sv[(size_t)b]; });
    for (int64_t c = 0; c < n; ++c) {
        const int64_t o = ord[(size_t)c];
        s[c] = sv[(size_t)o];
        std::memcpy(U + (size_t)c * n, &J[(size_t)o * n], sizeof(double) * (size_t)n);
        for (int64_t i = 0; i < n; ++i)
            V[(size_t)c * n + i] = sv[(size_t)o] > 0.0 ? G[(size_t)o * n + i] / sv[(size_t)o] : 0.0;
    }
    // complete V where a singular value vanished (Gram-Schmidt on unit vectors)
    for (int64_t c = 0; c < n; ++c) {
        if (s[c] > 0.0) continue;
        double *vc = V + (size_t)c * n;
        for (int64_t e = 0; e < n; ++e) {
            for (int64_t i = 0; i < n; ++i) vc[i] = (i == e) ? 1.0 : 0.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int64_t o = 0; o < n; ++o) {
                    if (o == c) continue;
                    const double *vo = V + (size_t)o * n;
                    if (o > c && s[o] <= 0.0) continue;
                    double d = 0.0;
                    for (int64_t i = 0; i < n; ++i) d += vo[i] * vc[i];
                    for (int64_t i = 0; i < n; ++i) vc[i] -= d * vo[i];
                }
            double nr = 0.0;
            for (int64_t i = 0; i < n; ++i) nr += vc[i] * vc[i];
            nr = std::sqrt(nr);
            if (nr > 0.5) {
                for (int64_t i = 0; i < n; ++i) vc[i] /= nr;
                break;
            }
        }
    }
}

}  // namespace sev
