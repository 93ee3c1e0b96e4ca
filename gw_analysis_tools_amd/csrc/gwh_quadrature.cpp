/*
 * gwh_quadrature.cpp -- Gauss-Legendre nodes and weights on [x1, x2] (gauleg, src/ortho_basis.cpp:14-46), the
 * rule the reference's callers build for the GAUSSLEG inner products (e.g. the log10 f grids of
 * waveform_util.cpp:3113 and apps/fisher_tool.cpp:113-126).  Host code.
 *
 * Each root z of P_n in (0, 1) starts from cos(pi (i - 1/4) / (n + 1/2)) and is polished by Newton's method on
 * the three-term recurrence (j P_j = (2j - 1) z P_(j-1) - (j - 1) P_(j-2)) with
 * P_n' = n (z P_n - P_(n-1)) / (z^2 - 1), until a step moves it by at most 1e-10 (the reference's threshold);
 * the weight is 2 / ((1 - z^2) P_n'^2) with the derivative of that last step, as the reference computes it.
 * Nodes ascend; each root also gives its mirror image.
 */
#include <cmath>

#include "../../include/gwat_hip.h"

extern "C" int gwh_gauleg(double x1, double x2, double *x, double *w, int n)
{
    if (!x || !w || n < 1) return GWH_ERR_INVALID;
    const double mid = 0.5 * (x2 + x1), half = 0.5 * (x2 - x1);
    const int roots = (n + 1) / 2;
    for (int i = 1; i <= roots; i++) {
        double z = std::cos(M_PI * (i - 0.25) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = 0.0;
            for (int j = 1; j <= n; j++) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            const double step = p0 / dp;
            z -= step;
            if (std::fabs(step) <= 1e-10) break;
        }
        x[i - 1] = mid - half * z;
        x[n - i] = mid + half * z;
        w[i - 1] = 2.0 * half / ((1.0 - z * z) * dp * dp);
        w[n - i] = w[i - 1];
    }
    return GWH_OK;
}
