// mgkmc_oracle.cpp -- CPU restatement of mgkmc_cpp's KMC + Lippmann-Schwinger
// hot path.  TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP
// path is compared against (tests/, the smoke check and the CPU-baseline leg of
// the benchmark).  It is never linked into, loaded by, or called from the
// product library, which fails loudly when its own HIP build is missing.
//
// Every function cites the reference lines it restates.  The arithmetic is
// IEEE FP64 in the reference's evaluation order where that order is visible in
// the source (Eigen's internal reduction order for squaredNorm/trace is not,
// so sums of 2x2/3x3 entries use plain left-to-right order).
//
// FFT: by default a self-written mixed-radix complex FFT (radix 4/2/3/5 plus a
// direct DFT for other prime factors) with the reference's "ortho" scaling
// 1/sqrt(N) on both directions (pocketfft c2c with fct = 1/sqrt(N),
// pocketfft_hdronly.h:3383-3394, :3122-3161).  Built with
// -DMGK_ORACLE_POCKETFFT and -I<reference>/third_party (the `ref` target of
// oracle/Makefile) the same calls go to the reference's vendored pocketfft
// instead; that build lives in oracle/_ref/ and pins this FFT in tests.
//
// RNG: the same libstdc++ std::mt19937 / std::normal_distribution /
// std::uniform_real_distribution objects, created and drawn in the same order
// as the reference, so the random stream is identical for a given seed.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../include/mgkmc_gpu.h"
#include "../include/mgkmc_host.h"

#ifdef MGK_ORACLE_POCKETFFT
#include "pocketfft_hdronly.h"
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace orc {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// FFT
// ---------------------------------------------------------------------------

#ifndef MGK_ORACLE_POCKETFFT
// Recursive decimation-in-time over the smallest factor; the combine step for
// a factor p reads out[k + m*q] for q < p and writes the same p slots, so it is
// in place.  tw[j] = exp(sign * 2 pi i j / n0) for the top-level length n0.
static void dft_rec(const cd* in, cd* out, int n, long stride, const cd* tw, int n0) {
    if (n == 1) { out[0] = in[0]; return; }
    int p;
    if (n % 4 == 0) p = 4;
    else if (n % 2 == 0) p = 2;
    else if (n % 3 == 0) p = 3;
    else if (n % 5 == 0) p = 5;
    else {
        p = n;
        for (int f = 7; f * f <= n; f += 2) if (n % f == 0) { p = f; break; }
    }
    int m = n / p;
    for (int q = 0; q < p; ++q) dft_rec(in + q * stride, out + q * m, m, stride * p, tw, n0);
    const int step = n0 / n;          // W_n^j = tw[j * step]
    const int pstep = n0 / p;         // W_p^j = tw[j * pstep]
    cd t[64];
    std::vector<cd> tbig;
    cd* tp = t;
    if (p > 64) { tbig.resize(p); tp = tbig.data(); }
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            long e = (long)q * k * step % n0;
            tp[q] = (q == 0) ? out[k] : out[k + q * m] * tw[e];
        }
        if (p == 2) {
            out[k] = tp[0] + tp[1];
            out[k + m] = tp[0] - tp[1];
        } else if (p == 4) {
            cd a = tp[0] + tp[2], b = tp[0] - tp[2];
            cd c = tp[1] + tp[3], d = tp[1] - tp[3];
            cd dj = d * tw[pstep];   // d * W_4^1 (= -i*d forward, +i*d backward)
            out[k] = a + c;
            out[k + m] = b + dj;
            out[k + 2 * m] = a - c;
            out[k + 3 * m] = b - dj;
        } else {
            for (int s = 0; s < p; ++s) {
                cd acc = 0.0;
                for (int q = 0; q < p; ++q) acc += tp[q] * tw[((long)q * s % p) * pstep];
                out[k + s * m] = acc;
            }
        }
    }
}

struct TwCache {
    int n = 0;
    int sign = 0;
    std::vector<cd> tw;
};

static const cd* twiddles(int n, int sign) {
    thread_local std::map<std::pair<int, int>, std::vector<cd>> cache;
    auto key = std::make_pair(n, sign);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second.data();
    std::vector<cd> tw(n);
    for (int j = 0; j < n; ++j) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)n;
        tw[j] = cd((double)cosl(a), (double)(sign * sinl(a)));
    }
    auto& ref = cache[key];
    ref = std::move(tw);
    return ref.data();
}

// In-place 1D c2c along one line (contiguous), sign -1 forward, +1 backward.
static void fft_line(cd* line, int n, int sign, std::vector<cd>& buf) {
    buf.assign(line, line + n);
    dft_rec(buf.data(), line, n, 1, twiddles(n, sign), n);
}
#endif

// c2c over all axes of a row-major array with shape `shape` (last axis
// fastest), scaled by fct: restates pocketfft::c2c as called at
// linear_elastic_solver.hpp:127-130, :134-139, :404-411, :416-424.
void c2c(const std::vector<int>& shape, bool forward, const cd* in, cd* out, double fct) {
#ifdef MGK_ORACLE_POCKETFFT
    pocketfft::shape_t sh;
    for (int s : shape) sh.push_back((size_t)s);
    pocketfft::stride_t st(shape.size());
    ptrdiff_t acc = sizeof(cd);
    for (int d = (int)shape.size() - 1; d >= 0; --d) { st[d] = acc; acc *= shape[d]; }
    pocketfft::shape_t axes;
    for (size_t d = 0; d < shape.size(); ++d) axes.push_back(d);
    pocketfft::c2c(sh, st, st, axes, forward ? pocketfft::FORWARD : pocketfft::BACKWARD, in, out, fct);
#else
    long total = 1;
    for (int s : shape) total *= s;
    if (out != in) std::copy(in, in + total, out);
    const int sign = forward ? -1 : +1;
    std::vector<cd> line, buf;
    const int nd = (int)shape.size();
    for (int ax = 0; ax < nd; ++ax) {
        int n = shape[ax];
        long inner = 1;
        for (int d = ax + 1; d < nd; ++d) inner *= shape[d];
        long outer = total / (inner * n);
        line.resize(n);
        for (long o = 0; o < outer; ++o) {
            for (long i = 0; i < inner; ++i) {
                cd* base = out + o * n * inner + i;
                for (int j = 0; j < n; ++j) line[j] = base[j * inner];
                fft_line(line.data(), n, sign, buf);
                for (int j = 0; j < n; ++j) base[j * inner] = line[j];
            }
        }
    }
    for (long i = 0; i < total; ++i) out[i] *= fct;
#endif
}

// ---------------------------------------------------------------------------
// Elasticity (linear_elastic_solver.hpp)
// ---------------------------------------------------------------------------

// fftfreq (linear_elastic_solver.hpp:22-33)
static std::vector<double> fftfreq(int n, double d) {
    std::vector<double> f(n);
    int limit = (n + 1) / 2;
    for (int i = 0; i < n; ++i) f[i] = (i < limit) ? (double)i / (n * d) : (double)(i - n) / (n * d);
    return f;
}

// compute_lame_2d (:36-45) and compute_lame_3d (:315-319)
static void lame(int D, double E, double nu, bool plane_stress, double& lam, double& mu) {
    mu = E / (2.0 * (1.0 + nu));
    if (D == 2 && plane_stress) lam = E * nu / (1.0 - nu * nu);
    else lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
}

template <int D> using Mat = std::array<double, D * D>;   // column-major: a[i + D*j] = (i, j)

template <int D> static inline double tr(const Mat<D>& a) {
    double t = 0.0;
    for (int i = 0; i < D; ++i) t += a[i + D * i];
    return t;
}

// sigma = 2 mu eps_eff + lam tr(eps_eff) I (:48-67, :322-341)
template <int D> static inline Mat<D> hooke(const Mat<D>& e, double lam, double mu) {
    Mat<D> s;
    double tm = 2.0 * mu, lt = lam * tr<D>(e);
    for (int j = 0; j < D; ++j)
        for (int i = 0; i < D; ++i) s[i + D * j] = tm * e[i + D * j] + lt * (i == j ? 1.0 : 0.0);
    return s;
}

template <int D> static inline double sqnorm(const Mat<D>& a) {
    double s = 0.0;
    for (int k = 0; k < D * D; ++k) s += a[k] * a[k];
    return s;
}

// green_operator_2d/3d (:70-119, :344-396) with the wave vectors of
// KmcSimulation*::init_gamma (kmc_solver.hpp:259-274, :1234-1254):
// xi = 2 pi fftfreq(n, L/n).  Storage: G[idx][((k*D + h)*D + i)*D + j].
template <int D>
static std::vector<double> green(const int* n, double pixel, double lam0, double mu0) {
    long N = 1;
    for (int d = 0; d < D; ++d) N *= n[d];
    std::vector<double> G((size_t)N * D * D * D * D, 0.0);
    std::vector<double> f[3];
    for (int d = 0; d < D; ++d) f[d] = fftfreq(n[d], (n[d] * pixel) / n[d]);
    double A = 1.0 / (4.0 * mu0);
    double B = (lam0 + mu0) / (mu0 * (lam0 + 2.0 * mu0));
    for (long idx = 0; idx < N; ++idx) {
        int c[3] = {0, 0, 0};
        long r = idx;
        for (int d = D - 1; d >= 0; --d) { c[d] = (int)(r % n[d]); r /= n[d]; }
        if (idx == 0) continue;
        double q[3];
        for (int d = 0; d < D; ++d) q[d] = 2.0 * M_PI * f[d][c[d]];
        double k2 = 0.0;
        for (int d = 0; d < D; ++d) k2 += q[d] * q[d];
        double* g = &G[(size_t)idx * D * D * D * D];
        for (int k = 0; k < D; ++k)
            for (int h = 0; h < D; ++h)
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) {
                        double t1 = 0.0;
                        if (k == i) t1 += q[h] * q[j];
                        if (h == i) t1 += q[k] * q[j];
                        if (k == j) t1 += q[h] * q[i];
                        if (h == j) t1 += q[k] * q[i];
                        t1 = A * t1 / k2;
                        double t2 = B * (q[k] * q[h] * q[i] * q[j]) / (k2 * k2);
                        g[((k * D + h) * D + i) * D + j] = t1 - t2;
                    }
    }
    return G;
}

// spectral_solver_2d / spectral_solver_3d (:146-309, :431-593).
// n = {nx, ny[, nz]}; eps in/out (warm start when warm); eps_plastic may be
// null (empty vector); G is the precomputed Green operator.
template <int D>
static int ls_solve(const int* n, const double* E, const double* nu, bool plane_stress,
                    const double* eps_bar_in, const std::vector<double>& G,
                    double* eps_io, bool warm, const double* eps_plastic,
                    double* sig_out, double* eps_macro, double* sig_macro,
                    int max_iter, double tol, double* diff_out) {
    long N = 1;
    for (int d = 0; d < D; ++d) N *= n[d];
    const int C = D * D;
    Mat<D> eps_bar;
    for (int k = 0; k < C; ++k) eps_bar[k] = eps_bar_in[k];   // copy: eps_macro may alias
    double E_avg = 0.0, nu_avg = 0.0;
    for (long i = 0; i < N; ++i) { E_avg += E[i]; nu_avg += nu[i]; }
    E_avg /= N;
    nu_avg /= N;
    double lam0, mu0;
    lame(D, E_avg, nu_avg, plane_stress, lam0, mu0);
    std::vector<double> lam(N), mu(N);
    for (long i = 0; i < N; ++i) lame(D, E[i], nu[i], plane_stress, lam[i], mu[i]);

    Mat<D>* eps = reinterpret_cast<Mat<D>*>(eps_io);
    const Mat<D>* ep = reinterpret_cast<const Mat<D>*>(eps_plastic);
    if (warm) {   // :181-191
        Mat<D> mean{};
        for (long i = 0; i < N; ++i)
            for (int k = 0; k < C; ++k) mean[k] += eps[i][k];
        for (int k = 0; k < C; ++k) mean[k] /= N;
        Mat<D> md;
        for (int k = 0; k < C; ++k) md[k] = eps_bar[k] - mean[k];
        for (long i = 0; i < N; ++i)
            for (int k = 0; k < C; ++k) eps[i][k] += md[k];
    } else {
        for (long i = 0; i < N; ++i) eps[i] = eps_bar;
    }

    std::vector<int> shape(n, n + D);
    const double fct = 1.0 / std::sqrt((double)N);
    std::vector<cd> tmp(N), that[9], ehat[9];
    std::vector<double> et[9];
    for (int c = 0; c < C; ++c) { that[c].resize(N); ehat[c].resize(N); et[c].resize(N); }
    std::vector<Mat<D>> eps_new(N), tau(N);
    int it = 0;
    double diff = 0.0;
    for (it = 0; it < max_iter; ++it) {
        for (long i = 0; i < N; ++i) {   // :205-213, :496-503
            Mat<D> ee = eps[i];
            if (ep) for (int k = 0; k < C; ++k) ee[k] = eps[i][k] - ep[i][k];
            Mat<D> s = hooke<D>(ee, lam[i], mu[i]);
            Mat<D> s0 = hooke<D>(eps[i], lam0, mu0);
            for (int k = 0; k < C; ++k) tau[i][k] = s[k] - s0[k];
        }
        // forward FFT of tau(r, s) in the order xx, xy, yx, yy (:215-227, :506-511)
        for (int r = 0; r < D; ++r)
            for (int s = 0; s < D; ++s) {
                int comp = r * D + s;
                for (long i = 0; i < N; ++i) tmp[i] = cd(tau[i][r + D * s], 0.0);
                c2c(shape, true, tmp.data(), that[comp].data(), fct);
            }
        // eps_hat_kh = -sum_rs G_khrs tau_hat_rs (:229-255, :513-536)
        for (long i = 0; i < N; ++i) {
            const double* g = &G[(size_t)i * C * C];
            for (int k = 0; k < D; ++k)
                for (int h = 0; h < D; ++h) {
                    cd v = 0.0;
                    for (int r = 0; r < D; ++r)
                        for (int s = 0; s < D; ++s) v += g[((k * D + h) * D + r) * D + s] * that[r * D + s][i];
                    ehat[k * D + h][i] = -v;
                }
        }
        // inverse FFT, keep the real part (:258-261, :539-543)
        for (int comp = 0; comp < C; ++comp) {
            c2c(shape, false, ehat[comp].data(), tmp.data(), fct);
            for (long i = 0; i < N; ++i) et[comp][i] = tmp[i].real();
        }
        // mean subtraction, update, relative increment (:263-296, :545-581)
        double mc[9];
        for (int comp = 0; comp < C; ++comp) {
            double sum = 0.0;
            for (long i = 0; i < N; ++i) sum += et[comp][i];
            mc[comp] = sum / N;
        }
        double diff_sq = 0.0, norm_sq = 0.0;
        for (long i = 0; i < N; ++i) {
            Mat<D> d;
            for (int r = 0; r < D; ++r)
                for (int s = 0; s < D; ++s) {
                    int comp = r * D + s;
                    eps_new[i][r + D * s] = eps_bar[r + D * s] + (et[comp][i] - mc[comp]);
                }
            for (int k = 0; k < C; ++k) d[k] = eps_new[i][k] - eps[i][k];
            diff_sq += sqnorm<D>(d);
            norm_sq += sqnorm<D>(eps[i]);
        }
        diff = std::sqrt(diff_sq) / (std::sqrt(norm_sq) + 1e-20);
        for (long i = 0; i < N; ++i) eps[i] = eps_new[i];
        if (diff < tol) { ++it; break; }
    }
    // final stress and macros (:299-308, :583-592)
    Mat<D> em{}, sm{};
    for (long i = 0; i < N; ++i) {
        Mat<D> ee = eps[i];
        if (ep) for (int k = 0; k < C; ++k) ee[k] = eps[i][k] - ep[i][k];
        Mat<D> s = hooke<D>(ee, lam[i], mu[i]);
        if (sig_out) for (int k = 0; k < C; ++k) sig_out[i * C + k] = s[k];
        for (int k = 0; k < C; ++k) { em[k] += eps[i][k]; sm[k] += s[k]; }
    }
    for (int k = 0; k < C; ++k) { eps_macro[k] = em[k] / N; sig_macro[k] = sm[k] / N; }
    if (diff_out) *diff_out = diff;
    return it;
}

// ---------------------------------------------------------------------------
// Material field generator (material_generator.hpp:121-258)
// ---------------------------------------------------------------------------

// box_muller_normal (:121-128)
static double box_muller_normal(std::mt19937& rng) {
    double u1 = (rng() - rng.min()) / double(rng.max() - rng.min() + 1.0);
    double u2 = (rng() - rng.min()) / double(rng.max() - rng.min() + 1.0);
    while (u1 <= 1e-15) u1 = (rng() - rng.min()) / double(rng.max() - rng.min() + 1.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// generate_correlated_field_cpp (:130-258)
// noise != nullptr replaces step 1's N draws (for pinning the filter with
// chosen inputs, e.g. a unit impulse)
static void correlated_field(int nx, int ny, int nz, bool is3d, double mean_t, double std_t, double corr,
                             bool has_min, double cmin, bool has_max, double cmax, int seed, double* out,
                             const double* noise = nullptr) {
    if (!is3d) nz = 1;
    long N = (long)nx * ny * nz;
    std::vector<double> field(N);
    std::mt19937 rng(seed);
    for (long i = 0; i < N; ++i) field[i] = noise ? noise[i] : box_muller_normal(rng);
    std::vector<int> shape = is3d ? std::vector<int>{nx, ny, nz} : std::vector<int>{nx, ny};
    std::vector<cd> tmp(N), hat(N);
    for (long i = 0; i < N; ++i) tmp[i] = cd(field[i], 0.0);
    double fct = 1.0 / std::sqrt((double)N);
    c2c(shape, true, tmp.data(), hat.data(), fct);
    auto freq = [](int n) {
        std::vector<double> f(n);
        int lim = (n + 1) / 2;
        for (int i = 0; i < n; ++i) f[i] = (i < lim ? i : i - n) / double(n);
        return f;
    };
    std::vector<double> fx = freq(nx), fy = freq(ny), fz = freq(nz);
    double coeff = -2.0 * M_PI * M_PI * corr * corr;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                long idx = ((long)x * ny + y) * nz + z;
                double f2 = fx[x] * fx[x] + fy[y] * fy[y] + (is3d ? fz[z] * fz[z] : 0.0);
                hat[idx] *= std::exp(coeff * f2);
            }
    c2c(shape, false, hat.data(), tmp.data(), fct);
    for (long i = 0; i < N; ++i) field[i] = tmp[i].real();
    double mean = 0.0;
    for (double v : field) mean += v;
    mean /= N;
    double var = 0.0;
    for (double v : field) var += (v - mean) * (v - mean);
    double sd = std::sqrt(var / N);
    if (sd < 1e-15) sd = 1.0;
    for (long i = 0; i < N; ++i) field[i] = (field[i] - mean) / sd;
    for (long i = 0; i < N; ++i) field[i] = mean_t + std_t * field[i];
    if (has_min) for (long i = 0; i < N; ++i) if (field[i] < cmin) field[i] = cmin;
    if (has_max) for (long i = 0; i < N; ++i) if (field[i] > cmax) field[i] = cmax;
    std::copy(field.begin(), field.end(), out);
}

// ---------------------------------------------------------------------------
// KMC (kmc_solver.hpp)
// ---------------------------------------------------------------------------

// BarrierGenerator::generate (:35-95)
static double barrier_generate(const mgk_sim_config& c, std::mt19937& rng) {
    double val = 0.0;
    const double mean = c.b_mean, sd = c.b_std;
    auto mod_rayleigh = [&]() {
        double a = std::max(0.0, mean - 6.0 * sd);
        double b = mean + 6.0 * sd;
        double x_max = (mean + std::sqrt(mean * mean + 4.0 * sd * sd)) / 2.0;
        double c_max = x_max * std::exp(-std::pow(x_max - mean, 2.0) / (2.0 * sd * sd));
        std::uniform_real_distribution<double> dist_x(a, b);
        std::uniform_real_distribution<double> dist_u(0.0, c_max);
        while (true) {
            double x = dist_x(rng);
            double u = dist_u(rng);
            double pdf = x * std::exp(-std::pow(x - mean, 2.0) / (2.0 * sd * sd));
            if (u < pdf) return x;
        }
    };
    if (c.barrier_type == MGK_BARRIER_GAUSSIAN) {
        std::normal_distribution<double> dist(mean, sd);
        val = dist(rng);
    } else if (c.barrier_type == MGK_BARRIER_RAYLEIGH) {
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        double u = dist(rng);
        val = c.b_loc + c.b_scale * std::sqrt(-2.0 * std::log(1.0 - u));
    } else if (c.barrier_type == MGK_BARRIER_MODIFIED_RAYLEIGH) {
        val = mod_rayleigh();
    } else if (c.barrier_type == MGK_BARRIER_MODIFIED_RAYLEIGH_EXP) {
        std::uniform_real_distribution<double> dist_u(0.0, 1.0);
        double u_mix = dist_u(rng);
        if (u_mix < c.b_ratio) {
            std::uniform_real_distribution<double> dist_exp(0.0, 1.0);
            val = -c.b_epsilon * std::log(1.0 - dist_exp(rng));
        } else {
            val = mod_rayleigh();
        }
    }
    if (c.b_min_cutoff >= 0.0 && val < c.b_min_cutoff) val = c.b_min_cutoff;
    if (c.b_max_cutoff >= 0.0 && val > c.b_max_cutoff) val = c.b_max_cutoff;
    return val;
}

// Fixed shear catalog (extension; no counterpart in the reference).
// 2D: mode m is the pure shear at angle theta_m = m*pi/M,
//     (g/2) [[cos 2t, sin 2t], [sin 2t, -cos 2t]].
// 3D: 12 deviatoric directions -- the 6 signed simple shears xy, xz, yz and
//     the 6 signed axial shears xx-yy, yy-zz, zz-xx -- each normalised to
//     Frobenius norm 1/sqrt(2) and scaled by g (the normalisation of
//     stz_mode_glass_3d, kmc_solver.hpp:1062-1073).
template <int D> static Mat<D> fixed_mode(int m, int M, double g) {
    Mat<D> a{};
    if (D == 2) {
        double t = 2.0 * M_PI * (double)m / (double)(2 * M);
        double c = std::cos(2.0 * t), s = std::sin(2.0 * t);
        a[0] = 0.5 * g * c;
        a[3] = -0.5 * g * c;
        a[1] = a[2] = 0.5 * g * s;
    } else {
        int k = m % 12;
        double sgn = (k % 2 == 0) ? 1.0 : -1.0;
        int kind = k / 2;
        auto at = [&](int i, int j) -> double& { return a[i + 3 * j]; };
        if (kind < 3) {   // simple shears: norm of e_ij + e_ji is sqrt(2)
            int i = (kind == 0) ? 0 : (kind == 1 ? 0 : 1);
            int j = (kind == 0) ? 1 : (kind == 1 ? 2 : 2);
            at(i, j) = at(j, i) = sgn * g * 0.5;
        } else {           // axial shears: norm of e_ii - e_jj is sqrt(2)
            int i = kind - 3, j = (kind - 3 + 1) % 3;
            at(i, i) = sgn * g * 0.5;
            at(j, j) = -sgn * g * 0.5;
        }
    }
    return a;
}

template <int D>
struct Sim {
    mgk_sim_config c;
    int n[3];
    long N;
    int M;
    double volume, tau;
    std::vector<double> E, nu;
    std::vector<Mat<D>> eps, sig, eps_p, prev_dir, catalog;
    std::vector<std::array<double, 2>> soft;
    std::vector<double> last_t, Q, Q0;
    Mat<D> eps_macro{}, sig_macro{};
    double time = 0.0;
    std::vector<double> G;
    std::mt19937 rng;
    std::vector<mgk_event_record> events;
    long long ls_solves = 0, ls_iters = 0;
    int last_iters = 0;
    // fast patching state (kmc_solver.hpp:154-160, :276-332)
    std::vector<std::vector<Mat<D>>> patch_kernels;   // [basis][window cell]
    std::vector<Mat<D>> patch_missing_mean;           // [basis]
    Mat<D> sigma_macro_unit{};
    bool sigma_macro_unit_initialized = false;
    int flips_since_sync = 0;

    long vidx(int x, int y, int z) const { return D == 2 ? (long)x * n[1] + y : ((long)x * n[1] + y) * n[2] + z; }

    // stz_mode_glass (:1042-1052) -- the 2D draw of :232-237 is the same
    // sequence of two normals from one distribution object.
    Mat<D> mode_planar(std::normal_distribution<double>& g) {
        Mat<D> a{};
        double gxx = 0.5 * c.gamma0 * g(rng);
        double gxy = 0.5 * c.gamma0 * g(rng);
        a[0] = gxx;
        a[D + 1] = -gxx;
        a[D] = gxy;   // (0,1)
        a[1] = gxy;   // (1,0)
        return a;
    }
    // stz_mode_glass_3d (:1054-1074)
    Mat<D> mode_glass_3d() {
        std::normal_distribution<double> gaussian(0.0, 1.0);
        double m[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m[i][j] = gaussian(rng);
        double s[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) s[i][j] = 0.5 * (m[i][j] + m[j][i]);
        double t = s[0][0] + s[1][1] + s[2][2];
        double g[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) g[i][j] = s[i][j] - (1.0 / 3.0) * t * (i == j ? 1.0 : 0.0);
        double nsq = 0.0;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) nsq += g[i][j] * g[i][j];
        double norm = std::sqrt(nsq);
        Mat<D> a{};
        if (norm < 1e-12) return a;
        double f = 1.0 / (std::sqrt(2.0) * norm);
        std::uniform_real_distribution<double> ud(1e-15, 1.0);
        double u = ud(rng);
        double mag = c.gamma0 * std::sqrt(-2.0 * std::log(u));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) a[i + 3 * j] = (g[i][j] * f) * mag;
        return a;
    }
    Mat<D> draw_mode(std::normal_distribution<double>* shared, int m) {
        if (c.catalog_mode == MGK_CATALOG_FIXED) return fixed_mode<D>(m, M, c.catalog_magnitude);
        if (D == 3 && c.use_3d_barriers) return mode_glass_3d();
        if (D == 2 && shared) return mode_planar(*shared);
        std::normal_distribution<double> g(0.0, 1.0);   // fresh per mode (:1043)
        return mode_planar(g);
    }

    // constructor (:165-257, :1142-1232)
    Sim(const mgk_sim_config& cfg, const double* E_, const double* nu_) : c(cfg), rng(cfg.seed) {
        n[0] = c.nx; n[1] = c.ny; n[2] = (D == 3) ? c.nz : 1;
        N = (long)n[0] * n[1] * n[2];
        M = c.M;
        volume = c.pixel * c.pixel * c.pixel;
        E.assign(E_, E_ + N);
        nu.assign(nu_, nu_ + N);
        eps.assign(N, Mat<D>{});
        sig.assign(N, Mat<D>{});
        eps_p.assign(N, Mat<D>{});
        soft.assign(N, {0.0, 0.0});
        last_t.assign(N, -std::numeric_limits<double>::infinity());
        if (c.temperature > 0.0) {
            double kB = 8.617e-5;
            tau = 1.0 / (c.nu0 * std::exp(-c.q_act_temp / (kB * c.temperature)));
        } else {
            tau = std::numeric_limits<double>::infinity();
        }
        Q.assign(N * M, 0.0);
        Q0.resize(N * M);
        catalog.resize(N * M);
        prev_dir.assign(N, Mat<D>{});
        std::normal_distribution<double> standard_normal(0.0, 1.0);   // :223 (2D only)
        for (long v = 0; v < N; ++v)
            for (int m = 0; m < M; ++m) {
                long mi = v * M + m;
                Q0[mi] = barrier_generate(c, rng);
                catalog[mi] = draw_mode(D == 2 ? &standard_normal : nullptr, m);
            }
        double Es = 0.0, ns = 0.0;
        for (long i = 0; i < N; ++i) { Es += E[i]; ns += nu[i]; }
        init_gamma(Es / N, ns / N);
        if (c.fast_patching) precompute_patch_kernels();
    }

    // --- fast patching (kmc_solver.hpp:276-332, 3D :1256-1331) -------------
    // Deviatoric unit bases: 2D xx = -yy, xy; 3D xx = -yy, yy = -zz, xy, xz, yz.
    static int n_bases() { return D == 2 ? 2 : 5; }
    static Mat<D> basis(int b) {
        Mat<D> a{};
        auto at = [&](int i, int j) -> double& { return a[i + D * j]; };
        if (D == 2) {
            if (b == 0) { at(0, 0) = 1.0; at(1, 1) = -1.0; }
            else { at(0, 1) = 1.0; at(1, 0) = 1.0; }
        } else {
            if (b == 0) { at(0, 0) = 1.0; at(1, 1) = -1.0; }
            else if (b == 1) { at(1, 1) = 1.0; at(2, 2) = -1.0; }
            else if (b == 2) { at(0, 1) = 1.0; at(1, 0) = 1.0; }
            else if (b == 3) { at(0, 2) = 1.0; at(2, 0) = 1.0; }
            else { at(1, 2) = 1.0; at(2, 1) = 1.0; }
        }
        return a;
    }
    // Coefficients of a flipped strain on the bases (:885-886, :1896-1900).
    static void patch_coefs(const Mat<D>& C, double* a) {
        if (D == 2) { a[0] = C[0]; a[1] = C[D]; }
        else { a[0] = C[0]; a[1] = -C[8]; a[2] = C[3]; a[3] = C[6]; a[4] = C[7]; }
    }
    int side() const { return 2 * c.patch_radius + 1; }
    int side_z() const { return (D == 3 && n[2] > 1) ? side() : 1; }
    int zr() const { return (D == 3 && n[2] > 1) ? c.patch_radius : 0; }

    void precompute_patch_kernels() {
        const int R = c.patch_radius, sd = side(), sz = side_z(), Z = zr();
        const int cx = n[0] / 2, cy = n[1] / 2, cz = (D == 3) ? n[2] / 2 : 0;
        patch_kernels.assign(n_bases(), std::vector<Mat<D>>((size_t)sd * sd * sz, Mat<D>{}));
        patch_missing_mean.assign(n_bases(), Mat<D>{});
        for (int b = 0; b < n_bases(); ++b) {
            std::vector<Mat<D>> ep(N, Mat<D>{}), e(N, Mat<D>{}), sg(N, Mat<D>{});
            ep[vidx(cx, cy, cz)] = basis(b);
            Mat<D> bar{}, em{}, sm{};
            double diff;
            ls_solve<D>(n, E.data(), nu.data(), c.plane_mode == MGK_PLANE_STRESS, bar.data(), G, e[0].data(), false,
                        ep[0].data(), sg[0].data(), em.data(), sm.data(), 200, 1e-6, &diff);
            Mat<D> sum_crop{};
            for (int dx = -R; dx <= R; ++dx)
                for (int dy = -R; dy <= R; ++dy)
                    for (int dz = -Z; dz <= Z; ++dz) {
                        const int sx = (cx + dx + n[0]) % n[0], sy = (cy + dy + n[1]) % n[1];
                        const int szz = (D == 3) ? (cz + dz + n[2]) % n[2] : 0;
                        const Mat<D>& val = sg[vidx(sx, sy, szz)];
                        patch_kernels[b][((size_t)(dx + R) * sd + (dy + R)) * sz + (dz + Z)] = val;
                        for (int k = 0; k < D * D; ++k) sum_crop[k] += val[k];
                    }
            Mat<D> sum_sig{};
            for (long i = 0; i < N; ++i)
                for (int k = 0; k < D * D; ++k) sum_sig[k] += sg[i][k];
            for (int k = 0; k < D * D; ++k) patch_missing_mean[b][k] = sum_sig[k] / N - sum_crop[k] / N;
        }
    }

    // sigma_macro_unit (:867-878): macro stress of a unit strain in the driven
    // component (cold solve, tol 1e-6, 200 iterations)
    void init_sigma_macro_unit(const Mat<D>& strain_unit) {
        std::vector<Mat<D>> e(N, Mat<D>{}), sg(N, Mat<D>{});
        Mat<D> em{}, sm{};
        double diff;
        ls_solve<D>(n, E.data(), nu.data(), c.plane_mode == MGK_PLANE_STRESS, strain_unit.data(), G, e[0].data(),
                    false, nullptr, sg[0].data(), em.data(), sm.data(), 200, 1e-6, &diff);
        sigma_macro_unit = sm;
        sigma_macro_unit_initialized = true;
    }
    // sig += sigma_macro_unit * (strain_rate * dt) everywhere (:889-893, :944-948)
    void predictor(double dt) {
        const double x = c.strain_rate * dt;
        for (long i = 0; i < N; ++i)
            for (int k = 0; k < D * D; ++k) sig[i][k] += sigma_macro_unit[k] * x;
    }
    // kernel superposition around voxel v and the missing-mean shift (:895-913)
    void apply_patch(long v, const Mat<D>& C) {
        double a[5];
        patch_coefs(C, a);
        const int R = c.patch_radius, sd = side(), sz = side_z(), Z = zr();
        int x, y, z = 0;
        if (D == 2) { y = (int)(v % n[1]); x = (int)(v / n[1]); }
        else { z = (int)(v % n[2]); long t = v / n[2]; y = (int)(t % n[1]); x = (int)(t / n[1]); }
        for (int dx = -R; dx <= R; ++dx)
            for (int dy = -R; dy <= R; ++dy)
                for (int dz = -Z; dz <= Z; ++dz) {
                    const int px = (x + dx + n[0]) % n[0], py = (y + dy + n[1]) % n[1];
                    const int pz = (D == 3) ? (z + dz + n[2]) % n[2] : 0;
                    const size_t ci = ((size_t)(dx + R) * sd + (dy + R)) * sz + (dz + Z);
                    Mat<D>& s_ = sig[vidx(px, py, pz)];
                    for (int k = 0; k < D * D; ++k) {
                        double pv = a[0] * patch_kernels[0][ci][k];
                        for (int b = 1; b < n_bases(); ++b) pv += a[b] * patch_kernels[b][ci][k];
                        s_[k] += pv;
                    }
                }
        Mat<D> shift{};
        for (int k = 0; k < D * D; ++k) {
            double m = a[0] * patch_missing_mean[0][k];
            for (int b = 1; b < n_bases(); ++b) m += a[b] * patch_missing_mean[b][k];
            shift[k] = m;
        }
        for (long i = 0; i < N; ++i)
            for (int k = 0; k < D * D; ++k) sig[i][k] += shift[k];
    }

    // init_gamma (:259-274, :1234-1254)
    void init_gamma(double E_avg, double nu_avg) {
        double lam0, mu0;
        lame(D, E_avg, nu_avg, c.plane_mode == MGK_PLANE_STRESS, lam0, mu0);
        G = green<D>(n, c.pixel, lam0, mu0);
    }

    // elastic_run (:334-339, :1333-1338) with the configurable tolerance
    void elastic_run() {
        double diff;
        last_iters = ls_solve<D>(n, E.data(), nu.data(), c.plane_mode == MGK_PLANE_STRESS, eps_macro.data(), G,
                                 eps[0].data(), true, eps_p[0].data(), sig[0].data(), eps_macro.data(),
                                 sig_macro.data(), c.ls_max_iter, c.ls_tol, &diff);
        ls_solves++;
        ls_iters += last_iters;
    }

    // update_barriers (:341-410, :1340-1411)
    void update_barriers() {
        bool directional = c.softening_directional != 0;
        const double GPa_nm3_to_eV = 6.241509;
        for (long v = 0; v < N; ++v) {
            double g_t = soft[v][1];
            double t_last = last_t[v];
            double g_t_curr = 0.0;
            if (g_t > 0.0) {
                if (std::isinf(tau)) g_t_curr = g_t;
                else if (t_last == -std::numeric_limits<double>::infinity()) g_t_curr = 0.0;
                else {
                    double dt = time - t_last;
                    if (dt < 0.0) dt = 0.0;
                    g_t_curr = g_t * std::exp(-dt / tau);
                }
            }
            double g_base = soft[v][0] + g_t_curr;
            if (c.softening_cap > 0.0 && g_base > c.softening_cap) g_base = c.softening_cap;
            for (int m = 0; m < M; ++m) {
                long mi = v * M + m;
                double modifier = 1.0;
                if (directional) {
                    double dot = 0.0, nm = 0.0, np = 0.0;
                    for (int i = 0; i < D; ++i)
                        for (int j = 0; j < D; ++j) {
                            double a = catalog[mi][i + D * j], b = prev_dir[v][i + D * j];
                            dot += a * b;
                            nm += a * a;
                            np += b * b;
                        }
                    double npv = std::sqrt(np), nmv = std::sqrt(nm);
                    if (npv < 1e-12 || nmv < 1e-12) modifier = 1.0;
                    else {
                        double ct = dot / (nmv * npv);
                        modifier = std::pow(1.0 + ct, 2.0) / 4.0;
                    }
                }
                double g_eff = modifier * g_base;
                double w_sum = 0.0;
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j) w_sum += sig[v][i + D * j] * catalog[mi][i + D * j];
                double w = 0.5 * volume * (w_sum / 1e9) * GPa_nm3_to_eV;
                Q[mi] = Q0[mi] * std::exp(-g_eff) - w;
            }
        }
    }

    // find_unstable_indices (:412-421)
    std::vector<int> find_unstable() const {
        std::vector<int> u;
        for (long i = 0; i < (long)Q.size(); ++i)
            if (Q[i] < c.stability_threshold) u.push_back((int)i);
        return u;
    }

    // apply_flip_soa_2d/3d (:423-462, :1424-1471)
    void apply_flip(long v, int m) {
        long mi = v * M + m;
        for (int k = 0; k < D * D; ++k) eps_p[v][k] += catalog[mi][k];
        const Mat<D>& a = catalog[mi];
        double sum_sq;
        if (D == 2) {
            double e11 = a[0], e22 = a[3], e12 = a[2];
            sum_sq = (e12 * e12) + (e22 * e22 + e11 * e11 + std::pow(e11 - e22, 2.0)) / 6.0;
        } else {
            double e11 = a[0], e22 = a[4], e33 = a[8], e12 = a[3], e13 = a[6], e23 = a[7];
            sum_sq = (e12 * e12 + e23 * e23 + e13 * e13) +
                     (std::pow(e22 - e33, 2.0) + std::pow(e33 - e11, 2.0) + std::pow(e11 - e22, 2.0)) / 6.0;
        }
        double gp = soft[v][0] + c.jp * sum_sq;
        if (c.softening_cap > 0.0 && gp > c.softening_cap) gp = c.softening_cap;
        soft[v][0] = gp;
        soft[v][1] = c.jt * sum_sq;
        last_t[v] = time;
        if (c.neighbor_softening_fraction > 0.0) {
            int x, y, z = 0;
            if (D == 2) { y = (int)(v % n[1]); x = (int)(v / n[1]); }
            else { z = (int)(v % n[2]); long t = v / n[2]; y = (int)(t % n[1]); x = (int)(t / n[1]); }
            int zr = (D == 3) ? 1 : 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -zr; dz <= zr; ++dz) {
                        if (dx == 0 && dy == 0 && dz == 0) continue;
                        int xn = (x + dx + n[0]) % n[0], yn = (y + dy + n[1]) % n[1];
                        int zn = (D == 3) ? (z + dz + n[2]) % n[2] : 0;
                        long nb = vidx(xn, yn, zn);
                        double g = soft[nb][0] + c.neighbor_softening_fraction * c.jp * sum_sq;
                        if (c.softening_cap > 0.0 && g > c.softening_cap) g = c.softening_cap;
                        soft[nb][0] = g;
                        soft[nb][1] += c.neighbor_softening_fraction * c.jt * sum_sq;
                    }
        }
    }

    // redraw after a flip (:839-864, :1854-1875; cascade :499-523, :1512-1533)
    void redraw(long v, long flat) {
        std::normal_distribution<double> standard_normal(0.0, 1.0);   // fresh per event (:840, :499)
        std::normal_distribution<double>* sn = (D == 2) ? &standard_normal : nullptr;
        if (c.redraw_directions || c.redraw_barriers) {
            for (int mm = 0; mm < M; ++mm) {
                long mi = v * M + mm;
                if (c.redraw_directions) catalog[mi] = draw_mode(sn, mm);
                if (c.redraw_barriers) Q0[mi] = barrier_generate(c, rng);
            }
        } else {
            catalog[flat] = draw_mode(sn, (int)(flat % M));
            Q0[flat] = barrier_generate(c, rng);
        }
    }

    // run_cascade (:464-564, :1473-1590)
    std::pair<int, int> run_cascade() {
        int total = 0, local = 0;
        long limit = N;
        while (true) {
            update_barriers();
            std::vector<int> u = find_unstable();
            if (u.empty()) break;
            std::sort(u.begin(), u.end(), [this](int a, int b) { return Q[a] < Q[b]; });
            std::vector<bool> flipped(N, false);
            int batch = 0;
            for (int k = 0; k < (int)u.size(); ++k) {
                long mi = u[k];
                int m = (int)(mi % M);
                long v = mi / M;
                if (flipped[v]) continue;
                flipped[v] = true;
                batch++;
                Mat<D> C = catalog[mi];
                apply_flip(v, m);
                prev_dir[v] = C;
                redraw(v, mi);
                if (c.fast_patching) apply_patch(v, C);
            }
            if (!c.fast_patching) elastic_run();
            total += batch;
            local++;
            if (local > limit) break;
        }
        if (c.fast_patching && total > 0) elastic_run();
        return {local, total};
    }

    struct RateResult { double total; std::vector<double> rates; std::vector<int> indices; };

    // rate loop (:783-806)
    RateResult rates() const {
        RateResult r;
        double eff_volume = c.scale_rate_by_volume ? volume : 1.0;
        double kB = 8.617e-5;
        double beta = (c.temperature > 0.0) ? (1.0 / (kB * c.temperature)) : std::numeric_limits<double>::infinity();
        r.total = 0.0;
        for (long i = 0; i < (long)Q.size(); ++i) {
            double q = Q[i];
            if (c.instability_kmc || q > 0.0) {
                double rv = (q <= 0.0) ? eff_volume * c.nu0 : eff_volume * c.nu0 * std::exp(-q * beta);
                r.rates.push_back(rv);
                r.total += rv;
                r.indices.push_back((int)i);
            }
        }
        return r;
    }

    // selection (:816-826)
    static int select(const RateResult& r, double target) {
        double cur = 0.0;
        int sel = (int)r.rates.size() - 1;
        for (size_t i = 0; i < r.rates.size(); ++i) {
            cur += r.rates[i];
            if (cur >= target) { sel = (int)i; break; }
        }
        return sel;
    }

    // accepted-event body (:811-937); returns is_instab
    int accept_event(const RateResult& rr, double t_wait, const Mat<D>& strain_unit,
                     std::uniform_real_distribution<double>& uniform_dist) {
        time += t_wait;
        for (int k = 0; k < D * D; ++k) eps_macro[k] += strain_unit[k] * (c.strain_rate * t_wait);
        double target = uniform_dist(rng) * rr.total;
        int sel = select(rr, target);
        int flat = rr.indices[sel];
        int m = flat % M;
        long v = flat / M;
        int is_instab = (Q[flat] <= c.stability_threshold) ? 1 : 0;
        Mat<D> C = catalog[flat];
        apply_flip(v, m);
        prev_dir[v] = C;
        redraw(v, flat);
        last_iters = 0;
        if (c.fast_patching) {
            // :866-913, 3D :1876-1937
            if (!sigma_macro_unit_initialized) init_sigma_macro_unit(strain_unit);
            if (t_wait > 0.0) predictor(t_wait);
            apply_patch(v, C);
            flips_since_sync++;
            if (flips_since_sync >= c.sync_interval) {
                elastic_run();
                flips_since_sync = 0;
            }
        } else {
            elastic_run();
        }
        mgk_event_record e{};
        e.flat_idx = flat;
        e.is_instab = is_instab;
        e.ls_iters = last_iters;
        e.time = time;
        e.total_rate = rr.total;
        for (int k = 0; k < D * D; ++k) { e.eps_macro[k] = eps_macro[k]; e.sig_macro[k] = sig_macro[k]; }
        events.push_back(e);
        return is_instab;
    }

    // run_simulation (:721-1037, :1735-2062): logging and VTU off; max_events
    // is the event-cap extension.
    void run(const mgk_run_config& r, mgk_run_stats& st) {
        std::memset(&st, 0, sizeof(st));
        int elastic_done = 0;
        long long total_kmc = 0, cascades = 0, cflips = 0;
        int seq = 0;
        int max_seq = static_cast<int>(r.max_kmc_steps_pct * N);
        elastic_run();
        Mat<D> su{};
        su[r.drv_i + D * r.drv_j] = 1.0;
        double dt_step = std::abs(r.step_size) / c.strain_rate;
        double remaining = 0.0;
        std::uniform_real_distribution<double> uniform_dist(0.0, 1.0);
        auto finish = [&]() {
            st.kmc_events = total_kmc;
            st.cascades = cascades;
            st.cascade_flips = cflips;
            st.elastic_steps = elastic_done;
            st.ls_solves = ls_solves;
            st.ls_iters = ls_iters;
            st.time = time;
        };
        while (elastic_done < r.n_global_steps) {
            remaining += dt_step;
            while (remaining > 0.0) {
                update_barriers();
                if (!c.instability_kmc) {
                    if (!find_unstable().empty()) {
                        auto cs = run_cascade();
                        if (cs.second > 0) { cascades++; cflips += cs.second; }
                        continue;
                    }
                }
                RateResult rr = rates();
                if (rr.total > 0.0) {
                    double r_val = uniform_dist(rng);
                    double t_wait = -std::log(r_val) / rr.total;
                    if (t_wait < remaining) {
                        remaining -= t_wait;   // the reference subtracts after advancing eps_macro (:812-814)
                        int inst = accept_event(rr, t_wait, su, uniform_dist);
                        if (inst) seq++; else seq = 0;
                        if (seq > max_seq) { st.terminated = 1; finish(); return; }
                        total_kmc++;
                        if (r.max_events > 0 && total_kmc >= r.max_events) { st.event_cap_reached = 1; finish(); return; }
                        continue;
                    }
                }
                if (c.fast_patching && sigma_macro_unit_initialized) predictor(remaining);   // :941-949
                for (int k = 0; k < D * D; ++k) eps_macro[k] += su[k] * (c.strain_rate * remaining);
                time += remaining;
                remaining = 0.0;
            }
            // mixed boundary conditions (:955-987)
            double Es = 0.0, ns = 0.0;
            for (double v : E) Es += v;
            for (double v : nu) ns += v;
            double E_avg = Es / E.size(), nu_avg = ns / nu.size();
            for (int it = 0; it < r.mixed_max_iter; ++it) {
                elastic_run();
                Mat<D> err{};
                double err_max = 0.0;
                // std::map iteration order: sorted by (i, j)
                std::vector<int> order(r.n_targets);
                for (int t = 0; t < r.n_targets; ++t) order[t] = t;
                std::sort(order.begin(), order.end(), [&](int a, int b) {
                    return std::make_pair(r.target_i[a], r.target_j[a]) < std::make_pair(r.target_i[b], r.target_j[b]);
                });
                for (int t : order) {
                    int mi = r.target_i[t], mj = r.target_j[t];
                    double e = r.target_val[t] - sig_macro[mi + D * mj];
                    err[mi + D * mj] = e;
                    err_max = std::max(err_max, std::abs(e));
                }
                if (err_max < r.mixed_tol) break;
                double trs = tr<D>(err);
                for (int t : order) {
                    int mi = r.target_i[t], mj = r.target_j[t];
                    double d = (err[mi + D * mj] - nu_avg * trs * (mi == mj ? 1.0 : 0.0)) / E_avg;
                    eps_macro[mi + D * mj] += d;
                }
            }
            elastic_done++;
        }
        finish();
    }

    // forced event steps (bench workload); see mgk_sim_event_steps
    void event_steps(int count, int di, int dj, mgk_run_stats& st) {
        Mat<D> su{};
        su[di + D * dj] = 1.0;
        std::uniform_real_distribution<double> uniform_dist(0.0, 1.0);
        for (int s = 0; s < count; ++s) {
            update_barriers();
            if (!c.instability_kmc && !find_unstable().empty()) {
                auto cs = run_cascade();
                if (cs.second > 0) { st.cascades++; st.cascade_flips += cs.second; }
                continue;
            }
            RateResult rr = rates();
            if (!(rr.total > 0.0)) break;
            double r_val = uniform_dist(rng);
            double t_wait = -std::log(r_val) / rr.total;
            accept_event(rr, t_wait, su, uniform_dist);
            st.kmc_events++;
        }
        st.ls_solves = ls_solves;
        st.ls_iters = ls_iters;
        st.time = time;
    }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface (used by tests and by the benchmark's CPU-baseline leg)
// ---------------------------------------------------------------------------

using namespace orc;

template <int D> static double* orc_field_ptr(Sim<D>* S, int which) {
    switch (which) {
        case MGK_FIELD_EPS: return S->eps[0].data();
        case MGK_FIELD_SIG: return S->sig[0].data();
        case MGK_FIELD_EPS_PLASTIC: return S->eps_p[0].data();
        case MGK_FIELD_PREV_DIR: return S->prev_dir[0].data();
        case MGK_FIELD_Q: return S->Q.data();
        case MGK_FIELD_Q0: return S->Q0.data();
        case MGK_FIELD_CATALOG: return S->catalog[0].data();
        case MGK_FIELD_SOFT: return S->soft[0].data();
        case MGK_FIELD_LAST_EVENT_TIME: return S->last_t.data();
        case MGK_FIELD_E: return S->E.data();
        case MGK_FIELD_NU: return S->nu.data();
    }
    return nullptr;
}

struct orc_sim {
    int dims;
    Sim<2>* s2 = nullptr;
    Sim<3>* s3 = nullptr;
};

extern "C" {

void orc_fft_c2c(int ndim, const int* shape, int forward, const double* in, double* out, double fct) {
    std::vector<int> sh(shape, shape + ndim);
    c2c(sh, forward != 0, reinterpret_cast<const cd*>(in), reinterpret_cast<cd*>(out), fct);
}

void orc_fftfreq(int n, double d, double* out) {
    auto f = fftfreq(n, d);
    std::copy(f.begin(), f.end(), out);
}

void orc_lame(int dims, double E, double nu, int plane_stress, double* lam, double* mu) {
    lame(dims, E, nu, plane_stress != 0, *lam, *mu);
}

// Green operator of init_gamma (kmc_solver.hpp:259-274, :1234-1254)
void orc_green(int dims, const int* n, double pixel, double lam0, double mu0, double* out) {
    if (dims == 2) { auto g = green<2>(n, pixel, lam0, mu0); std::copy(g.begin(), g.end(), out); }
    else { auto g = green<3>(n, pixel, lam0, mu0); std::copy(g.begin(), g.end(), out); }
}

int orc_ls_solve(int dims, const int* n, double pixel, int plane_stress, const double* E, const double* nu,
                 const double* eps_bar, double* eps, int warm, const double* eps_plastic, double* sig_out,
                 double* eps_macro, double* sig_macro, int max_iter, double tol, double* diff_out) {
    long N = 1;
    for (int d = 0; d < dims; ++d) N *= n[d];
    double Es = 0.0, ns = 0.0;
    for (long i = 0; i < N; ++i) { Es += E[i]; ns += nu[i]; }
    double lam0, mu0;
    lame(dims, Es / N, ns / N, plane_stress != 0, lam0, mu0);
    if (dims == 2) {
        auto G = green<2>(n, pixel, lam0, mu0);
        return ls_solve<2>(n, E, nu, plane_stress != 0, eps_bar, G, eps, warm != 0, eps_plastic, sig_out,
                           eps_macro, sig_macro, max_iter, tol, diff_out);
    }
    auto G = green<3>(n, pixel, lam0, mu0);
    return ls_solve<3>(n, E, nu, false, eps_bar, G, eps, warm != 0, eps_plastic, sig_out, eps_macro, sig_macro,
                       max_iter, tol, diff_out);
}

void orc_correlated_field(int nx, int ny, int nz, int is3d, double mean, double sd, double corr, int has_min,
                          double cmin, int has_max, double cmax, int seed, double* out) {
    correlated_field(nx, ny, nz, is3d != 0, mean, sd, corr, has_min != 0, cmin, has_max != 0, cmax, seed, out);
}

void orc_correlated_field_noise(int nx, int ny, int nz, int is3d, double mean, double sd, double corr,
                                const double* noise, double* out) {
    correlated_field(nx, ny, nz, is3d != 0, mean, sd, corr, false, 0.0, false, 0.0, 0, out, noise);
}

// libstdc++ stream probes: k draws of the given kind from mt19937(seed)
// kind 0: raw 32-bit words, 1: uniform_real_distribution<double>(0,1),
// 2: one normal_distribution<double>(0,1) object.
void orc_rng_stream(int seed, int kind, int k, double* out) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<double> g(0.0, 1.0);
    for (int i = 0; i < k; ++i) out[i] = (kind == 0) ? (double)rng() : (kind == 1 ? u(rng) : g(rng));
}

void orc_barrier_draws(const mgk_sim_config* c, int k, double* out) {
    std::mt19937 rng(c->seed);
    for (int i = 0; i < k; ++i) out[i] = barrier_generate(*c, rng);
}

int orc_sim_create(const mgk_sim_config* c, const double* E, const double* nu, orc_sim** out) {
    orc_sim* s = new orc_sim();
    s->dims = c->dims;
    if (c->dims == 2) s->s2 = new Sim<2>(*c, E, nu);
    else s->s3 = new Sim<3>(*c, E, nu);
    *out = s;
    return 0;
}

int orc_sim_destroy(orc_sim* s) {
    delete s->s2;
    delete s->s3;
    delete s;
    return 0;
}

#define ORC_DISPATCH(s, expr) ((s)->dims == 2 ? ([&](auto* S) { return expr; })((s)->s2) : ([&](auto* S) { return expr; })((s)->s3))

int orc_sim_run(orc_sim* s, const mgk_run_config* r, mgk_run_stats* st) {
    return ORC_DISPATCH(s, (S->run(*r, *st), 0));
}

int orc_sim_event_steps(orc_sim* s, int count, int di, int dj, mgk_run_stats* st) {
    std::memset(st, 0, sizeof(*st));
    return ORC_DISPATCH(s, (S->event_steps(count, di, dj, *st), 0));
}

int orc_sim_elastic_run(orc_sim* s) { return ORC_DISPATCH(s, (S->elastic_run(), S->last_iters)); }
int orc_sim_update_barriers(orc_sim* s) { return ORC_DISPATCH(s, (S->update_barriers(), 0)); }

// run_cascade on the current state: {local_step, total_flips}
int orc_sim_run_cascade(orc_sim* s, int* local_step, int* total_flips) {
    return ORC_DISPATCH(s, ([&]() {
        auto r = S->run_cascade();
        *local_step = r.first;
        *total_flips = r.second;
        return 0;
    })());
}

// rates + total + selection for a given u2 (used to pin the selection rule)
int orc_sim_rates(orc_sim* s, double* total, int* n_included, double* rates_out, int* idx_out) {
    return ORC_DISPATCH(s, ([&]() {
        auto rr = S->rates();
        *total = rr.total;
        *n_included = (int)rr.rates.size();
        if (rates_out) std::copy(rr.rates.begin(), rr.rates.end(), rates_out);
        if (idx_out) std::copy(rr.indices.begin(), rr.indices.end(), idx_out);
        return 0;
    })());
}

int orc_select_sequential(const double* rates, int n, double target) {
    double cur = 0.0;
    int sel = n - 1;
    for (int i = 0; i < n; ++i) {
        cur += rates[i];
        if (cur >= target) { sel = i; break; }
    }
    return sel;
}

// kernel superposition of a flipped strain at a voxel (fast patching); the
// strain is D x D column-major; fails (-1) without fast patching
int orc_sim_patch(orc_sim* s, long voxel, const double* strain) {
    return ORC_DISPATCH(s, ([&]() {
        if (!S->c.fast_patching) return -1;
        auto C = S->sigma_macro_unit;   // a Mat<D>
        for (int k = 0; k < (int)C.size(); ++k) C[k] = strain[k];
        S->apply_patch(voxel, C);
        return 0;
    })());
}

int orc_sim_apply_flip(orc_sim* s, int flat, double time) {
    return ORC_DISPATCH(s, ([&]() {
        S->time = time;
        long v = flat / S->M;
        auto C = S->catalog[flat];
        S->apply_flip(v, flat % S->M);
        S->prev_dir[v] = C;
        return 0;
    })());
}

static long orc_field_size(orc_sim* s, int which) {
    long N = s->dims == 2 ? s->s2->N : s->s3->N;
    int M = s->dims == 2 ? s->s2->M : s->s3->M;
    int C = s->dims * s->dims;
    switch (which) {
        case MGK_FIELD_EPS: case MGK_FIELD_SIG: case MGK_FIELD_EPS_PLASTIC: case MGK_FIELD_PREV_DIR: return N * C;
        case MGK_FIELD_Q: case MGK_FIELD_Q0: return N * M;
        case MGK_FIELD_CATALOG: return N * M * C;
        case MGK_FIELD_SOFT: return N * 2;
        case MGK_FIELD_LAST_EVENT_TIME: case MGK_FIELD_E: case MGK_FIELD_NU: return N;
    }
    return -1;
}

int orc_sim_get_field(orc_sim* s, int which, double* out) {
    long n = orc_field_size(s, which);
    if (n < 0) return 1;
    double* p = ORC_DISPATCH(s, orc_field_ptr(S, which));
    std::copy(p, p + n, out);
    return 0;
}

int orc_sim_set_field(orc_sim* s, int which, const double* in) {
    long n = orc_field_size(s, which);
    if (n < 0) return 1;
    double* p = ORC_DISPATCH(s, orc_field_ptr(S, which));
    std::copy(in, in + n, p);
    return 0;
}

int orc_sim_get_macro(orc_sim* s, double* eps_macro, double* sig_macro, double* time) {
    return ORC_DISPATCH(s, ([&]() {
        int C = s->dims * s->dims;
        for (int k = 0; k < C; ++k) { eps_macro[k] = S->eps_macro[k]; sig_macro[k] = S->sig_macro[k]; }
        *time = S->time;
        return 0;
    })());
}

int orc_sim_set_macro(orc_sim* s, const double* eps_macro, double time) {
    return ORC_DISPATCH(s, ([&]() {
        for (int k = 0; k < s->dims * s->dims; ++k) S->eps_macro[k] = eps_macro[k];
        S->time = time;
        return 0;
    })());
}

double orc_sim_tau(orc_sim* s) { return ORC_DISPATCH(s, S->tau); }

long long orc_sim_num_events(orc_sim* s) { return ORC_DISPATCH(s, (long long)S->events.size()); }

int orc_sim_get_events(orc_sim* s, mgk_event_record* out, long long cap) {
    return ORC_DISPATCH(s, ([&]() {
        long long n = std::min<long long>(cap, (long long)S->events.size());
        for (long long i = 0; i < n; ++i) out[i] = S->events[i];
        return (int)n;
    })());
}

// Fixed catalog entry (extension), for tests and the device driver's check.
void orc_fixed_mode(int dims, int m, int M, double g, double* out) {
    if (dims == 2) { auto a = fixed_mode<2>(m, M, g); std::copy(a.begin(), a.end(), out); }
    else { auto a = fixed_mode<3>(m, M, g); std::copy(a.begin(), a.end(), out); }
}

}  // extern "C"
