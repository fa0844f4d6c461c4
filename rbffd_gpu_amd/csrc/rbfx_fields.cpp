// Real spherical harmonics on the host: the field generator of the multi-field workloads
// (seeded combinations of Y_lm of degree <= 20) and the test functions of the derivative
// checks. The reference evaluates a few fixed harmonics from closed forms
// (src/utils/spherical_harmonics.h:39-170: Y_3^2, Y_10^5, Y_20^20); this evaluator covers any
// (l, m) with the same normalisation and sign convention (orthonormal on the unit sphere,
// Condon-Shortley phase (-1)^m included):
//   m > 0: Y_lm = sqrt(2) N_lm P_l^m(cos theta) cos(m phi)
//   m < 0: Y_lm = sqrt(2) N_l|m| P_l^|m|(cos theta) sin(|m| phi)
//   m = 0: Y_l0 = N_l0 P_l(cos theta),      N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)
// with theta the polar angle of the direction of (x, y, z) and phi = atan2(y, x).
// N_lm P_l^m is computed by the normalised three-term recurrence in l (stable to high degree,
// no factorials): Pbar_mm = -sqrt((2m+1)/(2m)) sin(theta) Pbar_(m-1)(m-1), Pbar_00 = 1/sqrt(4 pi),
// Pbar_(m+1)m = sqrt(2m+3) cos(theta) Pbar_mm,
// Pbar_lm = a_lm (cos(theta) Pbar_(l-1)m - b_lm Pbar_(l-2)m),
// a_lm = sqrt((4l^2-1)/(l^2-m^2)), b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)).
#include <cmath>
#include <cstdint>

#include "rbfx.h"

namespace {
double ylm_one(int l, int m, double x, double y, double z) {
    const int am = m < 0 ? -m : m;
    const double r = std::sqrt(x * x + y * y + z * z);
    const double rho = std::sqrt(x * x + y * y);
    const double ct = r > 0 ? z / r : 1.0, st = r > 0 ? rho / r : 0.0;
    double pmm = 1.0 / std::sqrt(4.0 * M_PI);
    for (int k = 1; k <= am; k++) pmm *= -std::sqrt((2.0 * k + 1.0) / (2.0 * k)) * st;
    double p = pmm;
    if (l > am) {
        double pm2 = pmm, pm1 = std::sqrt(2.0 * am + 3.0) * ct * pmm;
        p = pm1;
        for (int ll = am + 2; ll <= l; ll++) {
            const double a = std::sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)am * am));
            const double b = std::sqrt(((ll - 1.0) * (ll - 1.0) - (double)am * am) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
            p = a * (ct * pm1 - b * pm2);
            pm2 = pm1;
            pm1 = p;
        }
    }
    if (m == 0) return p;
    const double phi = std::atan2(y, x);
    return std::sqrt(2.0) * p * (m > 0 ? std::cos(am * phi) : std::sin(am * phi));
}
}  // namespace

extern "C" int rbfx_sph_harm(int32_t l, int32_t m, int32_t n, const double* xyz, double* out) {
    if (l < 0 || l > 1000 || m < -l || m > l || n < 0 || (n > 0 && (!xyz || !out))) return RBFX_E_INVALID;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++) out[i] = ylm_one(l, m, xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]);
    return RBFX_OK;
}
