// linear_elastic_solver.hpp -- drop-in replacement of the reference's
// linear_elastic_solver.hpp whose spectral_solver_2d / spectral_solver_3d
// run on the MI355X through the C ABI (include/mgkmc_gpu.h).
//
// Same free functions and signatures as the reference (fftfreq,
// compute_lame_2d/3d, stress_from_strain_2d/3d, green_operator_2d/3d,
// fft2d/ifft2d/fft3d/ifft3d, spectral_solver_2d/3d:
// linear_elastic_solver.hpp:22-593).  The tensor types are mgkt::Matrix2d/3d:
// Eigen's own when Eigen is included first (mgk_tensor.hpp).  The FFT helpers
// run on the host (the device solve has its own transforms inside its passes).
//
// The Gamma argument is accepted and not read: the device evaluates the Green
// operator on the fly from the mean medium (mean E, mean nu), which is what
// every caller of the reference builds Gamma from (kmc_solver.hpp:259-274,
// :1234-1254, main.cpp:555, :754).
#pragma once
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mgkmc_gpu.h"
#include "mgk_tensor.hpp"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

using GreenTensor = std::array<std::array<std::array<std::array<double, 2>, 2>, 2>, 2>;
using GreenTensor3D = std::array<std::array<std::array<std::array<double, 3>, 3>, 3>, 3>;

namespace mgk_host {

inline void check(int rc, const char* what) {
    if (rc != MGK_OK) throw std::runtime_error(std::string(what) + ": " + mgk_last_error());
}

// Device contexts of the free-function solvers, reused while the grid and
// the material fields are unchanged (the reference re-derives its Lame fields
// from E and nu on every call, :172-178).
struct SolverCache {
    mgk_ctx* ctx = nullptr;
    int dims = 0, nx = 0, ny = 0, nz = 0, plane = -1, precision = MGK_FP64;
    std::vector<double> E, nu;
    ~SolverCache() { if (ctx) mgk_destroy(ctx); }
    mgk_ctx* get(int d, int x, int y, int z, int plane_mode, const std::vector<double>& E_,
                 const std::vector<double>& nu_) {
        if (ctx && d == dims && x == nx && y == ny && z == nz && plane_mode == plane && E_ == E && nu_ == nu)
            return ctx;
        if (ctx) { mgk_destroy(ctx); ctx = nullptr; }
        check(mgk_create(d, x, y, z, 1.0, plane_mode, precision, E_.data(), nu_.data(), &ctx), "mgk_create");
        dims = d; nx = x; ny = y; nz = z; plane = plane_mode; E = E_; nu = nu_;
        return ctx;
    }
};

inline SolverCache& solver_cache() {
    thread_local SolverCache c;
    return c;
}

inline int plane_code(const std::string& plane_mode) {
    return plane_mode == "plane_stress" ? MGK_PLANE_STRESS : MGK_PLANE_STRAIN;
}

}  // namespace mgk_host

// fftfreq (linear_elastic_solver.hpp:22-33)
inline std::vector<double> fftfreq(int n, double d) {
    std::vector<double> f(n);
    const int limit = (n + 1) / 2;
    for (int i = 0; i < n; ++i) f[i] = (i < limit) ? (double)i / (n * d) : (double)(i - n) / (n * d);
    return f;
}

// compute_lame_2d (:36-45)
inline std::pair<double, double> compute_lame_2d(double E, double nu, const std::string& plane_mode) {
    const double mu = E / (2.0 * (1.0 + nu));
    const double lam = (plane_mode == "plane_stress") ? E * nu / (1.0 - nu * nu)
                                                      : E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    return {lam, mu};
}

// compute_lame_3d (:315-319)
inline std::pair<double, double> compute_lame_3d(double E, double nu) {
    return {E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))};
}

// stress_from_strain_2d/3d (:48-67, :322-341), host arithmetic
template <int D>
inline std::vector<mgkt::Matrix<D, D>> stress_from_strain_nd(const std::vector<mgkt::Matrix<D, D>>& eps,
                                                             const std::vector<double>& lam,
                                                             const std::vector<double>& mu,
                                                             const std::vector<mgkt::Matrix<D, D>>& eps_plastic) {
    const size_t N = eps.size();
    std::vector<mgkt::Matrix<D, D>> sig(N);
    const bool pl = (D == 2) ? !eps_plastic.empty() : eps_plastic.size() == N;
    for (size_t i = 0; i < N; ++i) {
        mgkt::Matrix<D, D> e = pl ? eps[i] - eps_plastic[i] : eps[i];
        sig[i] = 2.0 * mu[i] * e + lam[i] * e.trace() * mgkt::Matrix<D, D>::Identity();
    }
    return sig;
}
inline std::vector<mgkt::Matrix2d> stress_from_strain_2d(const std::vector<mgkt::Matrix2d>& eps,
                                                        const std::vector<double>& lam, const std::vector<double>& mu,
                                                        const std::vector<mgkt::Matrix2d>& eps_plastic = {}) {
    return stress_from_strain_nd<2>(eps, lam, mu, eps_plastic);
}
inline std::vector<mgkt::Matrix3d> stress_from_strain_3d(const std::vector<mgkt::Matrix3d>& eps,
                                                        const std::vector<double>& lam, const std::vector<double>& mu,
                                                        const std::vector<mgkt::Matrix3d>& eps_plastic = {}) {
    return stress_from_strain_nd<3>(eps, lam, mu, eps_plastic);
}

// green_operator_2d/3d (:70-119, :344-396), host arithmetic, for callers that
// inspect Gamma; the device solve does not read it.
template <int D, typename G>
inline std::vector<G> green_operator_nd(const int* n, const std::vector<const std::vector<double>*>& k, double lam0,
                                        double mu0) {
    long N = 1;
    for (int d = 0; d < D; ++d) N *= n[d];
    std::vector<G> Gamma(N);
    const double A = 1.0 / (4.0 * mu0);
    const double B = (lam0 + mu0) / (mu0 * (lam0 + 2.0 * mu0));
    for (long idx = 0; idx < N; ++idx) {
        double q[3] = {0, 0, 0}, k2 = 0.0;
        for (int d = 0; d < D; ++d) { q[d] = (*k[d])[idx]; k2 += q[d] * q[d]; }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) {
                        if (idx == 0) { Gamma[idx][a][b][i][j] = 0.0; continue; }
                        double t1 = 0.0;
                        if (a == i) t1 += q[b] * q[j];
                        if (b == i) t1 += q[a] * q[j];
                        if (a == j) t1 += q[b] * q[i];
                        if (b == j) t1 += q[a] * q[i];
                        t1 = A * t1 / k2;
                        Gamma[idx][a][b][i][j] = t1 - B * (q[a] * q[b] * q[i] * q[j]) / (k2 * k2);
                    }
    }
    return Gamma;
}
inline std::vector<GreenTensor> green_operator_2d(int nx, int ny, const std::vector<double>& kx,
                                                  const std::vector<double>& ky, double lam0, double mu0) {
    const int n[2] = {nx, ny};
    return green_operator_nd<2, GreenTensor>(n, {&kx, &ky}, lam0, mu0);
}
inline std::vector<GreenTensor3D> green_operator_3d(int nx, int ny, int nz, const std::vector<double>& kx,
                                                    const std::vector<double>& ky, const std::vector<double>& kz,
                                                    double lam0, double mu0) {
    const int n[3] = {nx, ny, nz};
    return green_operator_nd<3, GreenTensor3D>(n, {&kx, &ky, &kz}, lam0, mu0);
}

namespace mgk_host {

// Host DFT of any length for the FFT helpers: recursive mixed-radix
// Cooley-Tukey (decimation in time by the smallest prime factor), O(n sum p)
// per line; twiddles evaluated in long double.
inline void dft_rec(const std::complex<double>* in, long is, std::complex<double>* out, int n, int sign) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = n;
    for (int f = 2; (long)f * f <= n; ++f)
        if (n % f == 0) { p = f; break; }
    const int m = n / p;
    for (int r = 0; r < p; ++r) dft_rec(in + r * is, is * p, out + (long)r * m, m, sign);
    std::vector<std::complex<double>> tw(n), tmp(n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int e = 0; e < n; ++e) {
        const long double a = two_pi * (long double)e / (long double)n;
        tw[e] = std::complex<double>((double)cosl(a), (double)(sign * sinl(a)));
    }
    for (int k = 0; k < m; ++k)
        for (int q = 0; q < p; ++q) {
            std::complex<double> acc = 0.0;
            for (int r = 0; r < p; ++r) acc += tw[(long)r * (k + (long)m * q) % n] * out[(long)r * m + k];
            tmp[k + (long)m * q] = acc;
        }
    std::copy(tmp.begin(), tmp.end(), out);
}

// c2c over every axis of a row-major array (last axis contiguous), scaled by
// 1/sqrt(N) as the reference's ortho pocketfft calls (:130, :139, :411, :424).
inline void c2c_axes(std::vector<std::complex<double>>& a, const std::vector<int>& shape, bool forward) {
    long N = 1;
    for (int n : shape) N *= n;
    const int sign = forward ? -1 : 1;
    std::vector<std::complex<double>> line, res;
    long stride = N;
    for (size_t ax = 0; ax < shape.size(); ++ax) {
        const int n = shape[ax];
        stride /= n;   // elements between neighbours along this axis
        line.resize(n);
        res.resize(n);
        for (long outer = 0; outer < N / ((long)n * stride); ++outer)
            for (long in = 0; in < stride; ++in) {
                const long base = outer * n * stride + in;
                for (int j = 0; j < n; ++j) line[j] = a[base + (long)j * stride];
                dft_rec(line.data(), 1, res.data(), n, sign);
                for (int j = 0; j < n; ++j) a[base + (long)j * stride] = res[j];
            }
    }
    const double f = 1.0 / std::sqrt((double)N);
    for (auto& v : a) v *= f;
}

inline void fft_real_nd(const std::vector<int>& shape, const std::vector<double>& real_field,
                        std::vector<std::complex<double>>& complex_hat) {
    const size_t N = real_field.size();
    std::vector<std::complex<double>> t(N);
    for (size_t i = 0; i < N; ++i) t[i] = std::complex<double>(real_field[i], 0.0);
    c2c_axes(t, shape, true);
    complex_hat.resize(N);
    std::copy(t.begin(), t.end(), complex_hat.begin());
}

inline void ifft_real_nd(const std::vector<int>& shape, const std::vector<std::complex<double>>& complex_hat,
                         std::vector<double>& real_field) {
    std::vector<std::complex<double>> t(complex_hat.begin(), complex_hat.end());
    c2c_axes(t, shape, false);
    real_field.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i) real_field[i] = t[i].real();
}

template <int D>
inline void spectral_solve(int nx, int ny, int nz, const std::vector<double>& E, const std::vector<double>& nu,
                           const mgkt::Matrix<D, D>& eps_bar, int plane, std::vector<mgkt::Matrix<D, D>>& eps,
                           std::vector<mgkt::Matrix<D, D>>& sig_out, mgkt::Matrix<D, D>& eps_macro,
                           mgkt::Matrix<D, D>& sig_macro, int max_iter, double tol,
                           const std::vector<mgkt::Matrix<D, D>>& eps_plastic, int* iters, double* diff) {
    const size_t N = (size_t)nx * ny * nz;
    mgk_ctx* ctx = solver_cache().get(D, nx, ny, nz, plane, E, nu);
    const bool warm = eps.size() == N;   // :181, :465
    if (!warm) eps.assign(N, eps_bar);
    sig_out.resize(N);
    const bool pl = (D == 2) ? !eps_plastic.empty() : eps_plastic.size() == N;   // :56, :330
    const mgkt::Matrix<D, D> bar = eps_bar;   // eps_macro may alias eps_bar
    int it = 0;
    double df = 0.0;
    check(mgk_ls_solve(ctx, bar.data(), eps[0].data(), warm ? 1 : 0, pl ? eps_plastic[0].data() : nullptr,
                       sig_out[0].data(), eps_macro.data(), sig_macro.data(), max_iter, tol, &it, &df),
          "mgk_ls_solve");
    if (iters) *iters = it;
    if (diff) *diff = df;
}
}  // namespace mgk_host

// fft2d / ifft2d / fft3d / ifft3d (:122-143, :399-428): complex DFT of real
// data over every axis with ortho scaling; the inverse keeps the real part.
inline void fft2d(int nx, int ny, const std::vector<double>& real_field, std::vector<std::complex<double>>& complex_hat) {
    mgk_host::fft_real_nd({nx, ny}, real_field, complex_hat);
}
inline void ifft2d(int nx, int ny, const std::vector<std::complex<double>>& complex_hat, std::vector<double>& real_field) {
    mgk_host::ifft_real_nd({nx, ny}, complex_hat, real_field);
}
inline void fft3d(int nx, int ny, int nz, const std::vector<double>& real_field,
                  std::vector<std::complex<double>>& complex_hat) {
    mgk_host::fft_real_nd({nx, ny, nz}, real_field, complex_hat);
}
inline void ifft3d(int nx, int ny, int nz, const std::vector<std::complex<double>>& complex_hat,
                   std::vector<double>& real_field) {
    mgk_host::ifft_real_nd({nx, ny, nz}, complex_hat, real_field);
}

// spectral_solver_2d (:146-309) on the device
inline void spectral_solver_2d(int nx, int ny, const std::vector<double>& E, const std::vector<double>& nu,
                               const mgkt::Matrix2d& eps_bar, const std::string& plane_mode,
                               const std::vector<GreenTensor>& Gamma, std::vector<mgkt::Matrix2d>& eps,
                               std::vector<mgkt::Matrix2d>& sig_out, mgkt::Matrix2d& eps_macro,
                               mgkt::Matrix2d& sig_macro, int max_iter = 200, double tol = 1e-6,
                               bool verbose = false, const std::vector<mgkt::Matrix2d>& eps_plastic = {}) {
    (void)Gamma;
    int it = 0;
    double diff = 0.0;
    mgk_host::spectral_solve<2>(nx, ny, 1, E, nu, eps_bar, mgk_host::plane_code(plane_mode), eps, sig_out, eps_macro,
                                sig_macro, max_iter, tol, eps_plastic, &it, &diff);
    if (verbose) std::printf("  LS solve: %d iterations, diff = %g\n", it, diff);
}

// spectral_solver_3d (:431-593) on the device
inline void spectral_solver_3d(int nx, int ny, int nz, const std::vector<double>& E, const std::vector<double>& nu,
                               const mgkt::Matrix3d& eps_bar, const std::vector<GreenTensor3D>& Gamma,
                               std::vector<mgkt::Matrix3d>& eps, std::vector<mgkt::Matrix3d>& sig_out,
                               mgkt::Matrix3d& eps_macro, mgkt::Matrix3d& sig_macro, int max_iter = 200,
                               double tol = 1e-6, bool verbose = false,
                               const std::vector<mgkt::Matrix3d>& eps_plastic = {}) {
    (void)Gamma;
    int it = 0;
    double diff = 0.0;
    mgk_host::spectral_solve<3>(nx, ny, nz, E, nu, eps_bar, MGK_PLANE_STRAIN, eps, sig_out, eps_macro, sig_macro,
                                max_iter, tol, eps_plastic, &it, &diff);
    if (verbose) std::printf("  LS solve: %d iterations, diff = %g\n", it, diff);
}
