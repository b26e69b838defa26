// A caller of the drop-in headers written the way the reference's main.cpp
// calls its own: the pure-elastic branch (green_operator_3d, then
// spectral_solver_3d with a warm-started field vector, main.cpp:540-580) and
// the KMC branch (BarrierGenerator, the KmcSimulation3D constructor with the
// positional arguments main.cpp:379-386 passes, run_simulation with a stress
// target map, :388-395).  Built twice: on the drop-in's default tensor type,
// and with -DCALLER_LAYOUT_TYPE on a second layout-identical type
// (layout_type.hpp), standing for Eigen's.
//
//   caller host                      host-only functions (no device)
//   caller device <in> <out>         in: nx ny nz then E (N doubles) as raw doubles;
//                                    out: sigM (9), sig_field (9N), then the KMC
//                                    run's sig_macro (9), time, sum of eps_plastic (9)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#ifdef CALLER_LAYOUT_TYPE
#include "layout_type.hpp"
#define MGK_TENSOR_TEMPLATE ::layout::Mat
using Matrix3d = layout::Mat<3, 3>;
#else
#include "mgk_tensor.hpp"
using Matrix3d = mgkt::Matrix3d;
#endif
#include "kmc_solver.hpp"
#include "linear_elastic_solver.hpp"

// kmc_solver.hpp declares it and the reference's main.cpp defines it (:18-24)
std::string get_current_timestamp() { return "2026-01-01 00:00:00"; }

static int host_only() {
    // fftfreq, Lame constants, the Green operator, the FFT helpers and the
    // stress helper, with no device
    const int nx = 6, ny = 5, nz = 4, N = nx * ny * nz;
    std::vector<double> fx = fftfreq(6, 1.0);
    if (fx[3] != -0.5 || fx[1] != 1.0 / 6.0) return 10;
    auto [lam0, mu0] = compute_lame_3d(85e9, 0.35);
    std::vector<double> kx(N), ky(N), kz(N);
    std::vector<double> kx1 = fftfreq(nx, 1.0), ky1 = fftfreq(ny, 1.0), kz1 = fftfreq(nz, 1.0);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int idx = x * (ny * nz) + y * nz + z;
                kx[idx] = 2.0 * M_PI * kx1[x];
                ky[idx] = 2.0 * M_PI * ky1[y];
                kz[idx] = 2.0 * M_PI * kz1[z];
            }
    std::vector<GreenTensor3D> Gamma = green_operator_3d(nx, ny, nz, kx, ky, kz, lam0, mu0);
    // Gamma(xi) : C0 : Gamma(xi) = Gamma(xi) (a projector) at one wave vector
    const auto& G = Gamma[1 * (ny * nz) + 2 * nz + 1];
    double err = 0.0, mag = 0.0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double s = 0.0;
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) {
                            double cg = 0.0;   // (C0 : Gamma)_pqij
                            for (int r = 0; r < 3; ++r) cg += lam0 * (p == q) * G[r][r][i][j];
                            cg += mu0 * (G[p][q][i][j] + G[q][p][i][j]);
                            s += G[a][b][p][q] * cg;
                        }
                    err = std::max(err, std::abs(s - G[a][b][i][j]));
                    mag = std::max(mag, std::abs(G[a][b][i][j]));
                }
    if (!(err < 1e-12 * mag)) return 11;
    // fft3d / ifft3d: a round trip, and one known coefficient
    std::vector<double> f(N), g(N);
    for (int i = 0; i < N; ++i) f[i] = std::cos(0.3 * i) + 0.01 * i;
    std::vector<std::complex<double>> hat(N);
    fft3d(nx, ny, nz, f, hat);
    double dc = 0.0;
    for (double v : f) dc += v;
    if (std::abs(hat[0].real() - dc / std::sqrt((double)N)) > 1e-12 * std::abs(dc)) return 12;
    ifft3d(nx, ny, nz, hat, g);
    for (int i = 0; i < N; ++i)
        if (std::abs(g[i] - f[i]) > 1e-13) return 13;
    // stress_from_strain_3d on the caller's tensor type
    std::vector<Matrix3d> eps(2, Matrix3d::Zero()), epl(2, Matrix3d::Zero());
    eps[1](0, 0) = 1e-3;
    epl[1](0, 1) = 2e-4;
    epl[1](1, 0) = 2e-4;
    std::vector<double> lam(2, lam0), mu(2, mu0);
    std::vector<Matrix3d> sig = stress_from_strain_3d(eps, lam, mu, epl);
    if (std::abs(sig[1](0, 0) - (2.0 * mu0 + lam0) * 1e-3) > 1e-6 || std::abs(sig[1](0, 1) + 2.0 * mu0 * 2e-4) > 1e-6)
        return 14;
    std::printf("host-only checks passed\n");
    return 0;
}

static int device(const char* in_path, const char* out_path) {
    std::ifstream in(in_path, std::ios::binary);
    double hdr[3];
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    const int nx = (int)hdr[0], ny = (int)hdr[1], nz = (int)hdr[2], N = nx * ny * nz;
    std::vector<double> E_field(N), nu_field(N, 0.35);
    in.read(reinterpret_cast<char*>(E_field.data()), N * sizeof(double));
    if (!in) return 20;
    std::ofstream out(out_path, std::ios::binary);
    auto put = [&](const double* p, size_t n) { out.write(reinterpret_cast<const char*>(p), n * sizeof(double)); };

    // --- elastic branch (main.cpp:540-580) ---
    double E_avg = 0.0, nu_avg = 0.0;
    for (int i = 0; i < N; ++i) { E_avg += E_field[i]; nu_avg += nu_field[i]; }
    E_avg /= N;
    nu_avg /= N;
    auto [lam0, mu0] = compute_lame_3d(E_avg, nu_avg);
    const double pixel = 1.0, Lx = nx * pixel, Ly = ny * pixel, Lz = nz * pixel;
    std::vector<double> kx(N), ky(N), kz(N);
    std::vector<double> kx_1d = fftfreq(nx, Lx / nx), ky_1d = fftfreq(ny, Ly / ny), kz_1d = fftfreq(nz, Lz / nz);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const int idx = x * (ny * nz) + y * nz + z;
                kx[idx] = 2.0 * M_PI * kx_1d[x];
                ky[idx] = 2.0 * M_PI * ky_1d[y];
                kz[idx] = 2.0 * M_PI * kz_1d[z];
            }
    std::vector<GreenTensor3D> Gamma = green_operator_3d(nx, ny, nz, kx, ky, kz, lam0, mu0);
    Matrix3d current_eps_bar = Matrix3d::Zero();
    current_eps_bar(0, 0) = 1e-3;
    std::vector<Matrix3d> eps;
    std::vector<Matrix3d> sig_field;
    Matrix3d epsM, sigM;
    spectral_solver_3d(nx, ny, nz, E_field, nu_field, current_eps_bar, Gamma, eps, sig_field, epsM, sigM, 200, 1e-8,
                       false);
    put(sigM.data(), 9);
    put(sig_field[0].data(), 9 * (size_t)N);

    // --- KMC branch (main.cpp:379-395) ---
    BarrierGenerator bg;
    bg.type = "gaussian";
    bg.mean = 1.0;
    bg.std_val = 0.1;
    std::map<std::pair<int, int>, double> stress_targets;
    stress_targets[{1, 1}] = 0.0;
    stress_targets[{2, 2}] = 0.0;
    const int M_val = 6, n_steps = 4;
    KmcSimulation3D sim(nx, ny, nz, M_val, 0.02, E_field, nu_field, pixel, bg, "isotropic", 2.0, 10.0, 30.0, 0.0,
                        0.37, "caller_out", 330.0, 1.0, 0.0, 1e13, false, 5, 100, "kmc", "none", true, true, true,
                        false, 42);
    sim.run_simulation(n_steps, 1e-4, {0, 0}, stress_targets, 1e6, 50, "none", "checkpoint", "none", true, false,
                       false, "summary_log.txt", false, false, 0.3);
    put(sim.sig_macro.data(), 9);
    put(&sim.time, 1);
    Matrix3d ep_sum = Matrix3d::Zero();
    for (const Matrix3d& e : sim.eps_plastic) ep_sum += e;
    put(ep_sum.data(), 9);
    std::printf("device run written\n");
    return out ? 0 : 21;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "host") return host_only();
        if (argc >= 4 && std::string(argv[1]) == "device") return device(argv[2], argv[3]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    std::fprintf(stderr, "usage: caller host | caller device <in> <out>\n");
    return 1;
}
