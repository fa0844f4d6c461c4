// CPU restatement of the reference's steady-Stokes solve. TEST INFRASTRUCTURE ONLY:
// tests/ and bench.py's cpu_baseline leg use it as the checker; the product never links,
// loads or calls it.
//
//   orc_stokes_assemble   src/pdes/elliptic/stokes_steady_pde.cpp:61-230 (and the same block
//                         layout in tests/util/viennacl_gmres_poisson/main_stokes.cpp:153-330):
//                         rows/cols u(0..N) v(N..2N) w(2N..3N) p(3N..4N), 4 constraint rows
//                         and columns 4N..4N+3; columns ascending per row, as the ublas
//                         compressed_matrix keeps them.
//   orc_ilu0              src/precond/ilu0.hpp:157-260 (precondition with ilu0_tag(row_start,
//                         row_end)): Saad's ILU(0), restricted to the block [row_start,
//                         row_end)^2, identity rows outside it; fill only where the input
//                         holds a nonzero value (the `input(i, j)` test).
//   orc_ilu0_apply        src/precond/ilu0.hpp:273-327 (ilu0_lu_substitute: unit-lower then
//                         upper sweep, entries in ascending column order).
//   orc_gmres             ViennaCL gmres.hpp (vendored in the reference under
//                         external/opencl_libraries/include/viennacl/linalg/gmres.hpp:166-336,
//                         Walker's Householder "simpler GMRES" with left preconditioning).
//                         Restated from that algorithm: restarts, Householder reflectors, the
//                         rho recursion and the stopping test ||M^-1 r|| / ||b|| < tol.
//
// Pinning (tests/test_oracle_stokes.py): ILU(0) equals the unpivoted LU of a dense matrix
// (no fill to drop), L*U equals A on A's pattern for a sparse one, and rows outside the
// restricted block are identity rows; the sweeps invert L*U; GMRES reaches scipy's sparse
// direct solution, returns the least-squares minimiser over the Krylov space after k
// iterations, and stops at the iteration limit; the assembled matrix applied to constant
// velocity and pressure gives the closed form (derivative weights annihilate constants).

#include "oracle.h"

#include <cmath>
#include <cstring>
#include <map>
#include <vector>

extern "C" {

int64_t orc_stokes_nnz(int32_t N, int32_t n) {
    return (int64_t)N * (3 * (2 * n + 1) + (3 * n + 1)) + 4LL * N;
}

void orc_stokes_assemble(int32_t N, int32_t n, const int32_t* idx, const double* w, double eta, int64_t* ip,
                         int32_t* ind, double* val) {
    const int64_t Nn = (int64_t)N * n;
    const double* L = w;
    const double* D[3] = {w + Nn, w + 2 * Nn, w + 3 * Nn};
    // A stencil's columns ascend once its node indices are sorted; the values follow them.
    std::vector<int> ord(n);
    int64_t e = 0;
    ip[0] = 0;
    auto sorted = [&](int32_t i) {
        for (int j = 0; j < n; ++j) ord[j] = j;
        const int32_t* s = idx + (int64_t)i * n;
        for (int a = 1; a < n; ++a)  // insertion sort by node index (n is small)
            for (int b = a; b > 0 && s[ord[b]] < s[ord[b - 1]]; --b) std::swap(ord[b], ord[b - 1]);
    };
    for (int c = 0; c < 3; ++c)
        for (int32_t i = 0; i < N; ++i) {
            sorted(i);
            const int32_t* s = idx + (int64_t)i * n;
            for (int j = 0; j < n; ++j) {
                ind[e] = s[ord[j]] + c * N;
                val[e++] = -eta * L[(int64_t)i * n + ord[j]];
            }
            for (int j = 0; j < n; ++j) {
                ind[e] = s[ord[j]] + 3 * N;
                val[e++] = D[c][(int64_t)i * n + ord[j]];
            }
            ind[e] = 4 * N + c;
            val[e++] = 1.0;
            ip[c * N + i + 1] = e;
        }
    for (int32_t i = 0; i < N; ++i) {
        sorted(i);
        const int32_t* s = idx + (int64_t)i * n;
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < n; ++j) {
                ind[e] = s[ord[j]] + c * N;
                val[e++] = D[c][(int64_t)i * n + ord[j]];
            }
        ind[e] = 4 * N + 3;
        val[e++] = 1.0;
        ip[3 * N + i + 1] = e;
    }
    for (int c = 0; c < 4; ++c) {
        for (int32_t j = 0; j < N; ++j) {
            ind[e] = c * N + j;
            val[e++] = 1.0;
        }
        ip[4 * N + c + 1] = e;
    }
}

void orc_csr_matvec(int32_t nrows, const int64_t* ip, const int32_t* ind, const double* val, const double* x,
                    double* y) {
    for (int32_t i = 0; i < nrows; ++i) {
        double s = 0.0;
        for (int64_t e = ip[i]; e < ip[i + 1]; ++e) s += val[e] * x[ind[e]];
        y[i] = s;
    }
}

int64_t orc_ilu0(int32_t nrows, const int64_t* ip, const int32_t* ind, const double* val, int32_t row_start,
                 int32_t row_end, int64_t* lip, int32_t* lind, double* lval) {
    auto in_blk = [&](int64_t k) { return k >= row_start && k < row_end; };
    auto input_at = [&](int32_t i, int32_t j) -> double {
        int64_t lo = ip[i], hi = ip[i + 1];
        while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if (ind[mid] < j) lo = mid + 1; else hi = mid;
        }
        return (lo < ip[i + 1] && ind[lo] == j) ? val[lo] : 0.0;
    };
    // Output row k kept as (column, value) pairs in ascending order; diag[k] = its diagonal.
    std::vector<double> diag(nrows, 0.0);
    lip[0] = 0;
    int64_t e = 0;
    std::map<int32_t, double> w;
    for (int32_t i = 0; i < nrows; ++i) {
        w.clear();
        for (int64_t q = ip[i]; q < ip[i + 1]; ++q) {
            if (in_blk(i)) {
                if (in_blk(ind[q])) w[ind[q]] = val[q];
            } else {
                w[i] = 1.0;
            }
        }
        for (auto it = w.begin(); it != w.end(); ++it) {
            const int32_t k = it->first;
            if (k >= i) break;
            const double t = it->second / diag[k];
            for (int64_t q = lip[k]; q < lip[k + 1]; ++q) {
                const int32_t j = lind[q];
                if (input_at(i, j) != 0.0 && j > k) w[j] -= t * lval[q];
            }
            it->second = t;
        }
        for (auto& kv : w) {
            lind[e] = kv.first;
            lval[e++] = kv.second;
            if (kv.first == i) diag[i] = kv.second;
        }
        lip[i + 1] = e;
    }
    return e;
}

void orc_ilu0_apply(int32_t nrows, const int64_t* lip, const int32_t* lind, const double* lval, double* v) {
    for (int32_t i = 0; i < nrows; ++i)
        for (int64_t q = lip[i]; q < lip[i + 1]; ++q)
            if (lind[q] < i) v[i] -= lval[q] * v[lind[q]];
    double d = 1.0;  // the upper sweep keeps the last diagonal seen, as the reference's loop does
    for (int32_t i = nrows - 1; i >= 0; --i) {
        for (int64_t q = lip[i]; q < lip[i + 1]; ++q) {
            if (lind[q] > i) v[i] -= lval[q] * v[lind[q]];
            if (lind[q] == i) d = lval[q];
        }
        v[i] /= d;
    }
}

namespace {
double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}
void reflect(std::vector<double>& x, const std::vector<double>& h, double beta) {
    const double s = beta * dot(h, x);
    for (size_t i = 0; i < x.size(); ++i) x[i] -= s * h[i];
}
}  // namespace

int orc_gmres(int32_t nrows, const int64_t* ip, const int32_t* ind, const double* val, const int64_t* lip,
              const int32_t* lind, const double* lval, const double* b, double tol, int32_t max_iter,
              int32_t krylov, double* x, int32_t* iters_out, double* error_out) {
    const size_t n = (size_t)nrows;
    std::vector<double> rhs(b, b + n), res(n), v(n), tmp(n), result(n, 0.0);
    auto A = [&](const std::vector<double>& in, std::vector<double>& out) {
        orc_csr_matvec(nrows, ip, ind, val, in.data(), out.data());
    };
    auto M = [&](std::vector<double>& vec) {
        if (lip) orc_ilu0_apply(nrows, lip, lind, lval, vec.data());
    };
    int kd = krylov;
    if ((int)n < kd) kd = (int)n;
    unsigned restarts = (unsigned)max_iter / (unsigned)krylov;
    if (restarts > 0 && restarts * (unsigned)krylov == (unsigned)max_iter) restarts -= 1;
    std::vector<std::vector<double>> R(kd, std::vector<double>(krylov)), H(kd, std::vector<double>(n));
    std::vector<double> proj(kd), betas(kd);
    int iters = 0;
    double err = 0.0;
    const double norm_rhs = std::sqrt(dot(rhs, rhs));
    auto finish = [&](int code) {
        std::memcpy(x, result.data(), n * sizeof(double));
        *iters_out = iters;
        *error_out = err;
        return code;
    };
    if (norm_rhs == 0.0) return finish(0);
    for (unsigned it = 0; it <= restarts; ++it) {
        A(result, tmp);
        for (size_t i = 0; i < n; ++i) res[i] = rhs[i] - tmp[i];
        M(res);
        const double rho0 = std::sqrt(dot(res, res));
        if (rho0 / norm_rhs < tol) {
            err = rho0 / norm_rhs;
            return finish(0);
        }
        for (auto& r : res) r /= rho0;
        double rho = 1.0;
        int k = 0;
        for (k = 0; k < kd; ++k) {
            ++iters;
            std::fill(R[k].begin(), R[k].end(), 0.0);
            std::fill(H[k].begin(), H[k].end(), 0.0);
            if (k == 0) {
                A(res, v);
                M(v);
            } else {
                std::fill(v.begin(), v.end(), 0.0);
                v[k - 1] = 1.0;
                for (int i = k - 1; i > -1; --i) reflect(v, H[i], betas[i]);
                A(v, tmp);
                M(tmp);
                v = tmp;
                for (int i = 0; i < k; ++i) reflect(v, H[i], betas[i]);
            }
            // Householder vector that zeroes v below entry k.
            double mu;
            {
                const double vk = v[k];
                for (size_t i = k + 1; i < n; ++i) H[k][i] = v[i];
                double sigma = std::sqrt(dot(H[k], H[k]));
                sigma *= sigma;
                if (sigma == 0.0) {
                    betas[k] = 0.0;
                    mu = vk;
                } else {
                    mu = std::sqrt(sigma + vk * vk);
                    const double h0 = (vk <= 0) ? (vk - mu) : (-sigma / (vk + mu));
                    betas[k] = 2.0 * h0 * h0 / (sigma + h0 * h0);
                    for (auto& h : H[k]) h /= h0;
                    H[k][k] = 1.0;
                }
            }
            for (int i = 0; i < k; ++i) R[k][i] = v[i];
            R[k][k] = mu;
            reflect(res, H[k], betas[k]);
            if (res[k] > rho) res[k] = rho;
            if (res[k] < -rho) res[k] = -rho;
            proj[k] = res[k];
            rho *= std::sin(std::acos(proj[k] / rho));
            if (std::fabs(rho * rho0 / norm_rhs) < tol) {
                err = std::fabs(rho * rho0 / norm_rhs);
                ++k;
                break;
            }
        }
        for (int i = k - 1; i > -1; --i) {
            for (int j = i + 1; j < k; ++j) proj[i] -= R[j][i] * proj[j];
            proj[i] /= R[i][i];
        }
        for (auto& r : res) r *= proj[0];
        for (int i = 0; i + 1 < k; ++i) res[i] += proj[i + 1];
        for (int i = k - 1; i >= 0; --i) reflect(res, H[i], betas[i]);
        for (size_t i = 0; i < n; ++i) result[i] += rho0 * res[i];
        err = std::fabs(rho * rho0 / norm_rhs);
        if (err < tol) return finish(0);
    }
    return finish(1);
}

}  // extern "C"
