// kmc_solver.hpp -- drop-in replacement of the reference's kmc_solver.hpp:
// KmcSimulation2D / KmcSimulation3D with the same constructors, public
// members and methods, whose per-voxel work runs on the MI355X through the C
// ABI (include/mgkmc_gpu.h).
//
// What stays on the host, as in the reference: the std::mt19937 stream and
// every draw from it (so a seed gives the same stream of u1, u2, catalog and
// barrier values), the time / strain bookkeeping of run_simulation, the
// cascade ordering (std::sort of the unstable list by Q), and the mixed
// boundary-condition update.  What runs on the device: the LS solve
// (elastic_run), update_barriers + rates + total rate, event selection,
// apply_flip + redraw upload, and the unstable-mode compaction.
//
// Field members (eps_field, sig_field, Q, ...) are host mirrors of the device
// state: sync_from_device() refreshes them (run_simulation does so before it
// returns), sync_to_device() pushes edits made on the host.
//
// Extensions (defaults reproduce the reference): ls_tol / ls_max_iter
// (hard-coded 1e-6 / 200 in elastic_run, kmc_solver.hpp:337), max_events
// (stop after N accepted KMC events), precision (MGK_FP64 or
// MGK_FP32_REFINE), and a fixed shear catalog (catalog_fixed).
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <cstring>
#include <limits>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mgkmc_gpu.h"
#include "linear_elastic_solver.hpp"
#include "mgk_tensor.hpp"
#include "vtu_writer.hpp"

std::string get_current_timestamp();

// BarrierGenerator (kmc_solver.hpp:24-96)
struct BarrierGenerator {
    std::string type;
    double mean = 2.0;
    double std_val = 0.6;
    double min_cutoff = 0.1;
    double max_cutoff = -1.0;
    double loc = 1.0;
    double scale = 0.5;
    double epsilon = 0.1;
    double ratio = 0.8;

    double modified_rayleigh(std::mt19937& rng) const {
        const double a = std::max(0.0, mean - 6.0 * std_val);
        const double b = mean + 6.0 * std_val;
        const double x_max = (mean + std::sqrt(mean * mean + 4.0 * std_val * std_val)) / 2.0;
        const double c_max = x_max * std::exp(-std::pow(x_max - mean, 2.0) / (2.0 * std_val * std_val));
        std::uniform_real_distribution<double> dist_x(a, b);
        std::uniform_real_distribution<double> dist_u(0.0, c_max);
        for (;;) {
            const double x = dist_x(rng);
            const double u = dist_u(rng);
            if (u < x * std::exp(-std::pow(x - mean, 2.0) / (2.0 * std_val * std_val))) return x;
        }
    }

    double generate(std::mt19937& rng) const {
        double val = 0.0;
        if (type == "gaussian") {
            std::normal_distribution<double> dist(mean, std_val);
            val = dist(rng);
        } else if (type == "rayleigh") {
            std::uniform_real_distribution<double> dist(0.0, 1.0);
            const double u = dist(rng);
            val = loc + scale * std::sqrt(-2.0 * std::log(1.0 - u));
        } else if (type == "modified_rayleigh") {
            val = modified_rayleigh(rng);
        } else if (type == "modified_rayleigh_with_exponential") {
            std::uniform_real_distribution<double> dist_u(0.0, 1.0);
            if (dist_u(rng) < ratio) {
                std::uniform_real_distribution<double> dist_exp(0.0, 1.0);
                val = -epsilon * std::log(1.0 - dist_exp(rng));
            } else {
                val = modified_rayleigh(rng);
            }
        }
        if (min_cutoff >= 0.0 && val < min_cutoff) val = min_cutoff;
        if (max_cutoff >= 0.0 && val > max_cutoff) val = max_cutoff;
        return val;
    }
};

namespace mgk_host {

template <int D> using Mat = mgkt::Matrix<D, D>;

// Fixed shear catalog (extension).  2D: mode m is the pure shear at
// theta_m = m*pi/M, (g/2)[[cos 2t, sin 2t], [sin 2t, -cos 2t]].  3D: the six
// signed simple shears (xy, xz, yz) and six signed axial shears (xx-yy,
// yy-zz, zz-xx), each of Frobenius norm g/sqrt(2) as stz_mode_glass_3d
// normalises its modes (kmc_solver.hpp:1062-1073).
template <int D> inline Mat<D> fixed_shear(int m, int M, double g) {
    Mat<D> a = Mat<D>::Zero();
    if constexpr (D == 2) {
        const double t = M_PI * (double)m / (double)M;
        const double c = std::cos(2.0 * t), s = std::sin(2.0 * t);
        a(0, 0) = 0.5 * g * c;
        a(1, 1) = -0.5 * g * c;
        a(0, 1) = a(1, 0) = 0.5 * g * s;
    } else {
        const int k = m % 12;
        const double sg = (k % 2 == 0) ? 1.0 : -1.0;
        const int kind = k / 2;
        if (kind < 3) {
            const int i = (kind == 2) ? 1 : 0, j = (kind == 0) ? 1 : 2;
            a(i, j) = a(j, i) = sg * 0.5 * g;
        } else {
            const int i = kind - 3, j = (kind - 2) % 3;
            a(i, i) = sg * 0.5 * g;
            a(j, j) = -sg * 0.5 * g;
        }
    }
    return a;
}

struct RunCounters {
    long long kmc_events = 0, cascades = 0, cascade_flips = 0, elastic_steps = 0;
    long long ls_solves = 0, ls_iters = 0, ls_iters_fp32 = 0;
    long long exact_selections = 0;   // not part of checkpoints (a diagnostic)
    bool terminated = false, event_cap_reached = false;
};

struct EventRecord {
    int flat_idx;
    int is_instab;
    int ls_iters;
    double time;
    double total_rate;
    double eps_macro[9];
    double sig_macro[9];
};

template <int D>
class KmcSimulationT {
public:
    int nx, ny, nz;
    int M;
    double gamma0;
    double pixel;
    double volume;
    std::string output_dir;

    std::string softening_scheme;
    double softening_cap;
    double jp, jt;
    double neighbor_softening_fraction;
    double temperature;
    double strain_rate;
    double stability_threshold;
    double nu0;
    double q_act_temp;
    std::string plane_mode;

    bool fast_patching_enabled;
    int patch_radius;
    int sync_interval;
    int flips_since_sync;

    std::string instability_mode;
    std::string cascade_timing;
    bool scale_rate_by_volume;
    bool redraw_directions;
    bool redraw_barriers;
    double tau;
    bool use_3d_barriers;

    std::vector<double> E_field;
    std::vector<double> nu_field;
    std::vector<Mat<D>> eps_field;
    std::vector<Mat<D>> sig_field;
    std::vector<Mat<D>> eps_plastic;
    std::vector<mgkt::Vector2d> soft_prop;
    std::vector<double> last_event_time;
    Mat<D> eps_macro;
    Mat<D> sig_macro;
    double time;

    std::vector<double> Q;
    std::vector<double> Q0;
    std::vector<Mat<D>> catalog;
    std::vector<Mat<D>> prev_strain_dir;

    BarrierGenerator barrier_gen;
    std::mt19937 rng;

    std::ofstream f_summary;
    std::ofstream f_global;

    // extensions
    double ls_tol = 1e-6;
    int ls_max_iter = 200;
    long long max_events = -1;
    bool catalog_fixed = false;
    double catalog_magnitude = 0.1;
    bool device_draws = false;   // constructor draws on the device (mgk_kmc_draw): the reference's distributions, not its stream
    RunCounters counters;
    std::vector<EventRecord> events;
    bool record_events = false;

    mgk_ctx* ctx = nullptr;

    ~KmcSimulationT() {
        close_logs();
        if (ctx) mgk_destroy(ctx);
    }
    KmcSimulationT(const KmcSimulationT&) = delete;
    KmcSimulationT& operator=(const KmcSimulationT&) = delete;

    long N() const { return (long)nx * ny * nz; }

protected:
    // Common constructor body (kmc_solver.hpp:165-257, :1142-1232).
    KmcSimulationT(int nx_, int ny_, int nz_, int M_, double gamma0_, const std::vector<double>& E_,
                   const std::vector<double>& nu_, double pixel_, const BarrierGenerator& bg,
                   std::string softening_scheme_, double cap, double jp_, double jt_, double nfrac, double q_act,
                   std::string output_dir_, double T, double rate, double thr, double nu0_, std::string plane_mode_,
                   bool fp_enabled, int fp_radius, int fp_sync, std::string instab, std::string ctiming,
                   bool scale_vol, bool redraw_d, bool redraw_b, bool use3d, int seed, int precision,
                   bool fixed_catalog, double fixed_mag, bool device_draws_ = false)
        : nx(nx_), ny(ny_), nz(nz_), M(M_), gamma0(gamma0_), pixel(pixel_), volume(pixel_ * pixel_ * pixel_),
          output_dir(output_dir_), softening_scheme(softening_scheme_), softening_cap(cap), jp(jp_), jt(jt_),
          neighbor_softening_fraction(nfrac), temperature(T), strain_rate(rate), stability_threshold(thr),
          nu0(nu0_), q_act_temp(q_act), plane_mode(plane_mode_), fast_patching_enabled(fp_enabled),
          patch_radius(fp_radius), sync_interval(fp_sync), flips_since_sync(0), instability_mode(instab),
          cascade_timing(ctiming), scale_rate_by_volume(scale_vol), redraw_directions(redraw_d),
          redraw_barriers(redraw_b), use_3d_barriers(use3d), barrier_gen(bg), rng(seed),
          catalog_fixed(fixed_catalog), catalog_magnitude(fixed_mag), device_draws(device_draws_) {
        E_field = E_;
        nu_field = nu_;
        const long n = N();
        eps_field.assign(n, Mat<D>::Zero());
        sig_field.assign(n, Mat<D>::Zero());
        eps_plastic.assign(n, Mat<D>::Zero());
        soft_prop.assign(n, mgkt::Vector2d::Zero());
        last_event_time.assign(n, -std::numeric_limits<double>::infinity());
        eps_macro = Mat<D>::Zero();
        sig_macro = Mat<D>::Zero();
        time = 0.0;
        if (temperature > 0.0) {
            const double kB = 8.617e-5;
            tau = 1.0 / (nu0 * std::exp(-q_act_temp / (kB * temperature)));
        } else {
            tau = std::numeric_limits<double>::infinity();
        }
        Q.assign(n * M, 0.0);
        Q0.resize(n * M);
        catalog.resize(n * M);
        prev_strain_dir.assign(n, Mat<D>::Zero());
        std::normal_distribution<double> standard_normal(0.0, 1.0);   // :223 (2D draws share it)
        if (!device_draws) {
            for (long v = 0; v < n; ++v)
                for (int m = 0; m < M; ++m) {
                    const long mi = v * M + m;
                    Q0[mi] = barrier_gen.generate(rng);
                    catalog[mi] = draw_mode(D == 2 ? &standard_normal : nullptr, m);
                }
        } else if (catalog_fixed) {
            for (long v = 0; v < n; ++v)
                for (int m = 0; m < M; ++m) catalog[v * M + m] = fixed_shear<D>(m, M, catalog_magnitude);
        }
        check(mgk_create(D, nx, ny, nz, pixel, plane_code(plane_mode), precision, E_field.data(), nu_field.data(),
                         &ctx),
              "mgk_create");
        mgk_kmc_params kp{};
        kp.M = M;
        kp.softening_directional = (softening_scheme == "directional") ? 1 : 0;
        kp.softening_cap = softening_cap;
        kp.jp = jp;
        kp.jt = jt;
        kp.neighbor_softening_fraction = neighbor_softening_fraction;
        kp.tau = tau;
        kp.volume = volume;
        kp.temperature = temperature;
        kp.nu0 = nu0;
        kp.stability_threshold = stability_threshold;
        kp.instability_kmc = (instability_mode == "kmc") ? 1 : 0;
        kp.scale_rate_by_volume = scale_rate_by_volume ? 1 : 0;
        check(mgk_kmc_init(ctx, &kp), "mgk_kmc_init");
        if (!device_draws) {
            check(mgk_set_field(ctx, MGK_FIELD_Q0, Q0.data()), "mgk_set_field(Q0)");
            check(mgk_set_field(ctx, MGK_FIELD_CATALOG, catalog[0].data()), "mgk_set_field(catalog)");
        } else {
            // the reference's distributions on the device; host mirrors follow
            if (catalog_fixed) check(mgk_set_field(ctx, MGK_FIELD_CATALOG, catalog[0].data()), "mgk_set_field(catalog)");
            mgk_draw_params dp{};
            dp.barrier_type = barrier_code(barrier_gen.type);
            dp.b_mean = barrier_gen.mean;
            dp.b_std = barrier_gen.std_val;
            dp.b_min_cutoff = barrier_gen.min_cutoff;
            dp.b_max_cutoff = barrier_gen.max_cutoff;
            dp.b_loc = barrier_gen.loc;
            dp.b_scale = barrier_gen.scale;
            dp.b_epsilon = barrier_gen.epsilon;
            dp.b_ratio = barrier_gen.ratio;
            dp.gamma0 = gamma0;
            dp.use_3d_barriers = use_3d_barriers ? 1 : 0;
            check(mgk_kmc_draw(ctx, &dp, (unsigned long long)(unsigned)seed, catalog_fixed ? 0 : 1), "mgk_kmc_draw");
            check(mgk_get_field(ctx, MGK_FIELD_Q0, Q0.data()), "get Q0");
            if (!catalog_fixed) check(mgk_get_field(ctx, MGK_FIELD_CATALOG, catalog[0].data()), "get catalog");
        }
        if (fast_patching_enabled) precompute_patch_kernels();
    }
    static int barrier_code(const std::string& t) {
        if (t == "gaussian") return 0;
        if (t == "rayleigh") return 1;
        if (t == "modified_rayleigh") return 2;
        if (t == "modified_rayleigh_with_exponential") return 3;
        return 4;
    }

public:
    // --- fast patching (kmc_solver.hpp:276-332, :1256-1331) ----------------
    // The kernels live on the device (mgk_patch_init); sigma_macro_unit is
    // solved there before the first patched event (mgk_patch_unit) and counts
    // as initialised, as in the reference, once an event has used it.
    bool sigma_macro_unit_initialized = false;
    bool unit_on_device_ = false;
    void precompute_patch_kernels() {
        const char* who = (D == 2) ? "KmcSimulation2D" : "KmcSimulation3D";
        std::cout << "\n [" << who << "] Pre-computing Stress Patches (Radius=" << patch_radius << ")..." << std::endl;
        check(mgk_patch_init(ctx, patch_radius), "mgk_patch_init");
        std::cout << " [" << who << "] Fast Patching Kernels Ready." << std::endl;
    }
    void ensure_unit_stress(std::pair<int, int> component) {
        if (unit_on_device_) return;
        Mat<D> su = Mat<D>::Zero();
        su(component.first, component.second) = 1.0;
        check(mgk_patch_unit(ctx, su.data()), "mgk_patch_unit");
        unit_on_device_ = true;
    }

public:
    // --- mode draws (kmc_solver.hpp:232-237, :1042-1074) -------------------
    Mat<D> mode_planar(std::normal_distribution<double>& g) {
        Mat<D> a = Mat<D>::Zero();
        const double gxx = 0.5 * gamma0 * g(rng);
        const double gxy = 0.5 * gamma0 * g(rng);
        a(0, 0) = gxx;
        a(1, 1) = -gxx;
        a(0, 1) = gxy;
        a(1, 0) = gxy;
        return a;
    }
    Mat<D> mode_glass_3d() {
        std::normal_distribution<double> gaussian(0.0, 1.0);
        Mat<D> m;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) m(i, j) = gaussian(rng);
        Mat<D> s = 0.5 * (m + m.transpose());
        Mat<D> g = s - (1.0 / 3.0) * s.trace() * Mat<D>::Identity();
        const double nrm = g.norm();
        if (nrm < 1e-12) return Mat<D>::Zero();
        g = g * (1.0 / (std::sqrt(2.0) * nrm));
        std::uniform_real_distribution<double> ud(1e-15, 1.0);
        const double u = ud(rng);
        return g * (gamma0 * std::sqrt(-2.0 * std::log(u)));
    }
    Mat<D> draw_mode(std::normal_distribution<double>* shared, int m) {
        if (catalog_fixed) return fixed_shear<D>(m, M, catalog_magnitude);
        if constexpr (D == 3) {
            if (use_3d_barriers) return mode_glass_3d();
            std::normal_distribution<double> g(0.0, 1.0);   // fresh per mode (:1043)
            return mode_planar(g);
        } else {
            return mode_planar(*shared);
        }
    }

    // --- device <-> host mirrors -----------------------------------------
    void sync_from_device() {
        check(mgk_get_field(ctx, MGK_FIELD_EPS, eps_field[0].data()), "get eps");
        check(mgk_get_field(ctx, MGK_FIELD_SIG, sig_field[0].data()), "get sig");
        check(mgk_get_field(ctx, MGK_FIELD_EPS_PLASTIC, eps_plastic[0].data()), "get eps_plastic");
        check(mgk_get_field(ctx, MGK_FIELD_SOFT, soft_prop[0].data()), "get soft_prop");
        check(mgk_get_field(ctx, MGK_FIELD_LAST_EVENT_TIME, last_event_time.data()), "get last_event_time");
        check(mgk_get_field(ctx, MGK_FIELD_Q, Q.data()), "get Q");
        check(mgk_get_field(ctx, MGK_FIELD_Q0, Q0.data()), "get Q0");
        check(mgk_get_field(ctx, MGK_FIELD_CATALOG, catalog[0].data()), "get catalog");
        check(mgk_get_field(ctx, MGK_FIELD_PREV_DIR, prev_strain_dir[0].data()), "get prev_strain_dir");
    }
    void sync_to_device() {
        check(mgk_set_field(ctx, MGK_FIELD_EPS, eps_field[0].data()), "set eps");
        check(mgk_set_field(ctx, MGK_FIELD_SIG, sig_field[0].data()), "set sig");
        check(mgk_set_field(ctx, MGK_FIELD_EPS_PLASTIC, eps_plastic[0].data()), "set eps_plastic");
        check(mgk_set_field(ctx, MGK_FIELD_SOFT, soft_prop[0].data()), "set soft_prop");
        check(mgk_set_field(ctx, MGK_FIELD_LAST_EVENT_TIME, last_event_time.data()), "set last_event_time");
        check(mgk_set_field(ctx, MGK_FIELD_Q0, Q0.data()), "set Q0");
        check(mgk_set_field(ctx, MGK_FIELD_CATALOG, catalog[0].data()), "set catalog");
        check(mgk_set_field(ctx, MGK_FIELD_PREV_DIR, prev_strain_dir[0].data()), "set prev_strain_dir");
    }

    // --- the reference's methods -------------------------------------------
    // elastic_run (kmc_solver.hpp:334-339): eps_macro in, field mean out
    int elastic_run() {
        int it = 0, inner = 0;
        double diff = 0.0;
        check(mgk_elastic_run(ctx, eps_macro.data(), sig_macro.data(), ls_max_iter, ls_tol, &it, &diff),
              "mgk_elastic_run");
        mgk_last_solve_stats(ctx, nullptr, &inner);
        counters.ls_solves++;
        counters.ls_iters += it;
        counters.ls_iters_fp32 += inner;
        last_iters_ = it;
        return it;
    }

    // update_barriers (kmc_solver.hpp:341-410) fused with the rate loop
    // (:783-806): Q on the device, the FP64 total rate here.
    void update_barriers() {
        check(mgk_update_rates(ctx, time, &total_rate_, &n_unstable_, &n_included_), "mgk_update_rates");
    }
    double total_rate() const { return total_rate_; }
    int n_unstable() const { return n_unstable_; }

    // find_unstable_indices (kmc_solver.hpp:412-421); also fills unstable_q_
    std::vector<int> find_unstable_indices() {
        int n = 0;
        check(mgk_find_unstable(ctx, nullptr, nullptr, 0, &n), "mgk_find_unstable");
        std::vector<int> idx(n);
        unstable_q_.assign(n, 0.0);
        if (n > 0) check(mgk_find_unstable(ctx, idx.data(), unstable_q_.data(), n, &n), "mgk_find_unstable");
        return idx;
    }

    void apply_flip(long flat) {
        check(mgk_apply_flip(ctx, (int)flat, time, nullptr, nullptr, 0, 0), "mgk_apply_flip");
    }

    // redraw after a flip (kmc_solver.hpp:839-864, cascade :499-523) + flip
    // The redraw of flipping mode `flat` of `voxel` (:839-864, cascade :499-523),
    // drawn from the host stream in the reference's order into cat[0..n) and
    // q0[0..n): all M modes of the voxel when a redraw switch is on, else the
    // flipped mode alone (first = its mode).
    void draw_redraw(long voxel, long flat, Mat<D>* cat, double* q0, int& first, int& n, bool& has_cat,
                     bool& has_q0) {
        std::normal_distribution<double> standard_normal(0.0, 1.0);   // fresh per event (:840, :499)
        std::normal_distribution<double>* sn = (D == 2) ? &standard_normal : nullptr;
        if (redraw_directions || redraw_barriers) {
            for (int mm = 0; mm < M; ++mm) {
                const long mi = voxel * M + mm;
                cat[mm] = catalog[mi];
                q0[mm] = Q0[mi];
                if (redraw_directions) cat[mm] = draw_mode(sn, mm);
                if (redraw_barriers) q0[mm] = barrier_gen.generate(rng);
            }
            first = 0;
            n = M;
            has_cat = redraw_directions;
            has_q0 = redraw_barriers;
        } else {
            cat[0] = draw_mode(sn, (int)(flat % M));
            q0[0] = barrier_gen.generate(rng);
            first = (int)(flat % M);
            n = 1;
            has_cat = has_q0 = true;
        }
    }

    // The host copies of a redrawn voxel's entries.
    void keep_redraw(long voxel, const Mat<D>* cat, const double* q0, int first, int n, bool has_cat, bool has_q0) {
        for (int j = 0; j < n; ++j) {
            const long mi = voxel * M + first + j;
            if (has_cat) catalog[mi] = cat[j];
            if (has_q0) Q0[mi] = q0[j];
        }
    }

    void flip_and_redraw(long voxel, long flat) {
        const Mat<D> flipped = catalog[flat];   // the strain before its redraw (:488, :1501)
        redraw_cat_.resize((size_t)M);
        redraw_q0_.resize((size_t)M);
        int first = 0, n = 0;
        bool has_cat = false, has_q0 = false;
        draw_redraw(voxel, flat, redraw_cat_.data(), redraw_q0_.data(), first, n, has_cat, has_q0);
        check(mgk_apply_flip(ctx, (int)flat, time, has_cat ? redraw_cat_[0].data() : nullptr,
                             has_q0 ? redraw_q0_.data() : nullptr, first, n),
              "mgk_apply_flip");
        keep_redraw(voxel, redraw_cat_.data(), redraw_q0_.data(), first, n, has_cat, has_q0);
        if (fast_patching_enabled) check(mgk_patch_apply(ctx, (int)voxel, flipped.data(), 0.0), "mgk_patch_apply");
    }

    // A cascade round's flips (distinct voxels, in order): every redraw drawn
    // first, in the flips' order, then one device call for all of them.
    void flip_round(const std::vector<int>& flips) {
        const bool all = redraw_directions || redraw_barriers;
        const size_t per = all ? (size_t)M : 1;
        round_cat_.resize(flips.size() * per);
        round_q0_.resize(flips.size() * per);
        int first = 0, n = 0;
        bool has_cat = false, has_q0 = false;
        for (size_t i = 0; i < flips.size(); ++i)
            draw_redraw(flips[i] / M, flips[i], &round_cat_[i * per], &round_q0_[i * per], first, n, has_cat, has_q0);
        check(mgk_apply_flips(ctx, (int)flips.size(), flips.data(), time, has_cat ? round_cat_[0].data() : nullptr,
                              has_q0 ? round_q0_.data() : nullptr, all ? 1 : 0),
              "mgk_apply_flips");
        for (size_t i = 0; i < flips.size(); ++i)
            keep_redraw(flips[i] / M, &round_cat_[i * per], &round_q0_[i * per], all ? 0 : flips[i] % M, (int)per,
                        has_cat, has_q0);
    }

    // run_cascade (kmc_solver.hpp:464-564, :1473-1590)
    std::pair<int, int> run_cascade(int step, std::pair<int, int> component) {
        (void)step;
        (void)component;
        int total_flips = 0, local_step = 0;
        const long limit = N();
        while (true) {
            update_barriers();
            if (n_unstable_ == 0) break;
            std::vector<int> unstable = find_unstable_indices();
            if (unstable.empty()) break;
            std::map<int, double> qv;
            for (size_t k = 0; k < unstable.size(); ++k) qv[unstable[k]] = unstable_q_[k];
            std::sort(unstable.begin(), unstable.end(), [&](int a, int b) { return qv[a] < qv[b]; });
            std::vector<bool> flipped(N(), false);
            std::vector<int> flips;
            for (int mi : unstable) {
                const long v = mi / M;
                if (flipped[v]) continue;
                flipped[v] = true;
                flips.push_back(mi);
            }
            const int batch = (int)flips.size();
            if (fast_patching_enabled) {
                // each flip patches the stress before the next one
                for (int mi : flips) flip_and_redraw(mi / M, mi);
            } else if (batch > 0) {
                flip_round(flips);
            }
            if (!fast_patching_enabled) elastic_run();
            total_flips += batch;
            local_step++;
            if (local_step > limit) break;
        }
        if (fast_patching_enabled && total_flips > 0) elastic_run();
        return {local_step, total_flips};
    }

    // init_logs (kmc_solver.hpp:566-602): the summary and global log headers.
    void init_logs(std::string summary_filename, bool enable_summary_log, bool enable_global_log) {
        if (!enable_summary_log && !enable_global_log) return;
        std::filesystem::create_directories(output_dir);
        if (enable_summary_log) {
            f_summary.open(output_dir + "/" + summary_filename);
            f_summary << kSummaryHeader << std::string(std::strlen(kSummaryHeader) - 1, '-') << "\n";
        }
        if (enable_global_log) {
            f_global.open(output_dir + "/global_log.txt");
            f_global << std::left << std::setw(10) << "GlobalStep" << " " << std::setw(12) << "ElasticStep" << " "
                     << std::setw(10) << "KMCStep" << " " << std::setw(15) << "Eps_xx" << " " << std::setw(15)
                     << "Eps_yy" << " " << std::setw(15) << "Eps_zz" << " " << std::setw(15) << "Eps_xy" << " "
                     << std::setw(15) << "Eps_xz" << " " << std::setw(15) << "Eps_yz" << " " << std::setw(15)
                     << "Sig_xx(GPa)" << " " << std::setw(15) << "Sig_yy(GPa)" << " " << std::setw(15)
                     << "Sig_zz(GPa)" << " " << std::setw(15) << "Sig_xy(GPa)" << " " << std::setw(15)
                     << "Sig_xz(GPa)" << " " << std::setw(15) << "Sig_yz(GPa)" << " " << std::setw(14)
                     << "CascadeSteps" << " " << std::setw(17) << "TotalCascadeFlips" << " "
                     << "SimTime(s)\n";
        }
    }
    void close_logs() {
        wait_vtu();
        if (f_summary.is_open()) f_summary.close();
        if (f_global.is_open()) f_global.close();
    }

    static constexpr const char* kSummaryHeader =
        "Timestamp              Elapsed(s)   Step     Type            Eps_xx       Sig_xx(GPa)     KMC      "
        "Cascade  Flips    SimTime(s)\n";

    // --- VTU output ----------------------------------------------------------
    // The fields are copied from the device into the host mirrors, and a
    // writer thread formats the file from a snapshot of them, so the run goes
    // on while the ASCII is written (at most two files in flight).
    std::vector<std::thread> vtu_jobs_;
    std::string vtu_error_;

    void export_vtu(const std::string& name) {
        check(mgk_get_field(ctx, MGK_FIELD_EPS, eps_field[0].data()), "get eps");
        check(mgk_get_field(ctx, MGK_FIELD_SIG, sig_field[0].data()), "get sig");
        check(mgk_get_field(ctx, MGK_FIELD_EPS_PLASTIC, eps_plastic[0].data()), "get eps_plastic");
        check(mgk_get_field(ctx, MGK_FIELD_SOFT, soft_prop[0].data()), "get soft_prop");
        while (vtu_jobs_.size() >= 2) {
            vtu_jobs_.front().join();
            vtu_jobs_.erase(vtu_jobs_.begin());
        }
        vtu_jobs_.emplace_back([this, name, e = eps_field, s = sig_field, ep = eps_plastic, sp = soft_prop,
                                E = E_field, nu = nu_field]() {
            try {
                if constexpr (D == 2) export_to_vtu_2d(name, nx, ny, e, s, E, nu, pixel, {}, ep, sp);
                else export_to_vtu(name, nx, ny, nz, e, s, E, nu, pixel, {}, ep, sp);
            } catch (const std::exception& ex) {
                vtu_error_ = ex.what();
            }
        });
    }
    void wait_vtu() {
        for (auto& t : vtu_jobs_) t.join();
        vtu_jobs_.clear();
        if (!vtu_error_.empty()) {
            const std::string e = vtu_error_;
            vtu_error_.clear();
            throw std::runtime_error(e);
        }
    }
    std::string vtu_step_name(int step, const char* suffix = "") const {
        std::ostringstream ss;
        ss << output_dir << "/vtk_step_" << std::setfill('0') << std::setw(6) << step << suffix << ".vtu";
        return ss.str();
    }

    // do_logging (kmc_solver.hpp:606-719, 3D :1630-1733): the summary and
    // global lines and the VTU files vtk_interval_str asks for ("none",
    // "current", "last" or an interval of logged steps, counted in elastic
    // steps when vtk_elastic_only).
    void do_logging(int global_step, int elastic_steps_done, int total_kmc_steps, std::string step_type,
                    int cascade_steps, int cascade_flips, std::pair<int, int> component, bool enable_console_log,
                    std::chrono::high_resolution_clock::time_point start_time_total,
                    std::string vtk_interval_str = "none", bool vtk_elastic_only = true) {
        if (step_type == "INIT") {
            if (vtk_interval_str == "current") export_vtu(output_dir + "/step.vtu");
            else if (vtk_interval_str != "none" && vtk_interval_str != "last") export_vtu(vtu_step_name(0));
            return;
        }
        log_lines(global_step, elastic_steps_done, total_kmc_steps, step_type, cascade_steps, cascade_flips,
                  component, enable_console_log, start_time_total);
        if (vtk_interval_str == "none") return;
        if (vtk_interval_str == "current") {
            export_vtu(output_dir + "/step.vtu");
            return;
        }
        int interval = 0;
        try {
            interval = std::stoi(vtk_interval_str);
        } catch (...) {
            return;
        }
        if (interval <= 0) return;
        const bool due = vtk_elastic_only ? (step_type == "ELASTIC" && elastic_steps_done % interval == 0)
                                          : (global_step % interval == 0);
        if (due) export_vtu(vtu_step_name(global_step));
    }

    void log_lines(int global_step, int elastic_steps_done, int total_kmc_steps, const std::string& step_type,
                   int cascade_steps, int cascade_flips, std::pair<int, int> component, bool enable_console_log,
                   std::chrono::high_resolution_clock::time_point start_time_total) {
        if (!f_summary.is_open() && !f_global.is_open() && !enable_console_log) return;
        const double elapsed =
            std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - start_time_total).count();
        char buf[256];
        std::snprintf(buf, sizeof(buf), "%-22s %-12.2f %-8d %-15s %-12.6f %-15.3f %-8d %-8d %-8d %-15.6e\n",
                      get_current_timestamp().c_str(), elapsed, global_step, step_type.c_str(),
                      eps_macro(component.first, component.second),
                      sig_macro(component.first, component.second) / 1e9, total_kmc_steps, cascade_steps,
                      cascade_flips, time);
        if (f_summary.is_open()) f_summary << buf;
        if (enable_console_log) std::cout << buf;
        if (f_global.is_open()) {
            double e[6] = {eps_macro(0, 0), eps_macro(1, 1), 0.0, eps_macro(0, 1), 0.0, 0.0};
            double s[6] = {sig_macro(0, 0), sig_macro(1, 1), 0.0, sig_macro(0, 1), 0.0, 0.0};
            if constexpr (D == 3) {
                e[2] = eps_macro(2, 2); e[4] = eps_macro(0, 2); e[5] = eps_macro(1, 2);
                s[2] = sig_macro(2, 2); s[4] = sig_macro(0, 2); s[5] = sig_macro(1, 2);
            } else if (plane_mode == "plane_strain") {
                double nsum = 0.0;
                for (double v : nu_field) nsum += v;
                s[2] = (nsum / nu_field.size()) * (s[0] + s[1]);
            }
            char g[1024];
            std::snprintf(g, sizeof(g),
                          "%-10d %-12d %-10d %-15.6f %-15.6f %-15.6f %-15.6f %-15.6f %-15.6f %-15.3f %-15.3f "
                          "%-15.3f %-15.3f %-15.3f %-15.3f %-14d %-17d %-15.6e\n",
                          global_step, elastic_steps_done, total_kmc_steps, e[0], e[1], e[2], e[3], e[4], e[5],
                          s[0] / 1e9, s[1] / 1e9, s[2] / 1e9, s[3] / 1e9, s[4] / 1e9, s[5] / 1e9, cascade_steps,
                          cascade_flips, time);
            f_global << g;
        }
    }

    // Random draws of one attempt of run_simulation's inner loop, taken
    // before its outcome is known, and the generator states to rewind to.
    struct EventDraw {
        std::mt19937 rng_before, rng_after_u1;
        double r_val = 0.0;
        int first = 0, n = 0;
        bool has_cat = false, has_q0 = false;
    };

    // First half of an attempt: u1, u2 and the redraw values in the
    // reference's order (kmc_solver.hpp:787-864), and the device arguments.
    // The redraw values stay in this object until complete_event.
    void prepare_event(double remaining, std::pair<int, int> component,
                       std::uniform_real_distribution<double>& uniform_dist, EventDraw& d, mgk_event_in& in) {
        d.rng_before = rng;
        d.r_val = uniform_dist(rng);
        d.rng_after_u1 = rng;
        const double u2 = uniform_dist(rng);
        // the redraw values come next in the stream (:839-864) and do not
        // depend on which voxel the device selects
        std::normal_distribution<double> standard_normal(0.0, 1.0);   // fresh per event (:840)
        std::normal_distribution<double>* sn = (D == 2) ? &standard_normal : nullptr;
        redraw_cat_.resize((size_t)M);
        redraw_q0_.resize((size_t)M);
        if (redraw_directions || redraw_barriers) {
            for (int mm = 0; mm < M; ++mm) {
                if (redraw_directions) redraw_cat_[mm] = draw_mode(sn, mm);
                if (redraw_barriers) redraw_q0_[mm] = barrier_gen.generate(rng);
            }
            d.first = 0;
            d.n = M;
            d.has_cat = redraw_directions && !catalog_fixed;   // a fixed catalog redraws to itself
            d.has_q0 = redraw_barriers;
        } else {
            redraw_cat_[0] = catalog_fixed ? Mat<D>::Zero() : draw_mode(sn, 0);
            redraw_q0_[0] = barrier_gen.generate(rng);
            d.first = -1;   // the selected mode only
            d.n = 1;
            d.has_cat = !catalog_fixed;
            d.has_q0 = true;
        }
        std::memset(&in, 0, sizeof(in));
        in.time = time;
        in.neg_log_u1 = -std::log(d.r_val);
        in.remaining = remaining;
        in.strain_rate = strain_rate;
        in.drv_i = component.first;
        in.drv_j = component.second;
        in.u2 = u2;
        in.new_catalog = d.has_cat ? redraw_cat_[0].data() : nullptr;
        in.new_Q0 = d.has_q0 ? redraw_q0_.data() : nullptr;
        in.first_mode = d.first;
        in.n_modes = d.n;
        in.max_iter = ls_max_iter;
        in.tol = ls_tol;
        in.cascade_check = (instability_mode == "cascade") ? 1 : 0;
        if (fast_patching_enabled) {
            // patch; the flip that reaches sync_interval also solves (:908-912)
            ensure_unit_stress(component);
            in.patch_mode = (flips_since_sync + 1 >= sync_interval) ? 2 : 1;
        }
    }

    // Second half: the generator is rewound to where the reference would
    // stand when the attempt has no event (no rate or cascade pending: before
    // u1; t_wait too long: after u1); on an event the accepted-event body
    // (:811-937) follows: time, statistics, host mirrors, records.  eps_macro
    // and sig_macro already hold what the device call returned.
    int complete_event(const EventDraw& d, const mgk_event_out& out, int* is_instab, double* t_wait_out) {
        total_rate_ = out.total_rate;
        n_unstable_ = out.n_unstable;
        n_included_ = out.n_included;
        if (out.outcome == MGK_EVENT_UNSTABLE || out.outcome == MGK_EVENT_NO_RATE) {
            rng = d.rng_before;
            return out.outcome;
        }
        if (out.outcome == MGK_EVENT_REJECTED) {
            rng = d.rng_after_u1;
            return out.outcome;
        }
        if (out.outcome != MGK_EVENT_DONE) {
            rng = d.rng_before;
            return out.outcome;
        }
        // an event always names a mode of this grid; anything else is a device
        // error, reported before any host array is indexed with it
        if (out.flat_idx < 0 || (long)out.flat_idx >= (long)Q0.size())
            throw std::runtime_error("mgk_kmc_event: event reported without a valid mode (flat index " +
                                     std::to_string(out.flat_idx) + ")");
        const double t_wait = -std::log(d.r_val) / total_rate_;   // the reference's expression (= out.t_wait)
        time += t_wait;
        const bool solved = !fast_patching_enabled || flips_since_sync + 1 >= sync_interval;
        if (fast_patching_enabled) {
            sigma_macro_unit_initialized = true;
            flips_since_sync = solved ? 0 : flips_since_sync + 1;
        }
        if (solved) {
            int inner = 0;
            mgk_last_solve_stats(ctx, nullptr, &inner);
            counters.ls_solves++;
            counters.ls_iters += out.iters;
            counters.ls_iters_fp32 += inner;
        }
        last_iters_ = out.iters;
        if (out.exact_selection) counters.exact_selections++;
        const long flat = out.flat_idx;
        const long v = flat / M;
        const int inst = (out.q_sel <= stability_threshold) ? 1 : 0;
        // host copies of the redrawn entries
        for (int j = 0; j < d.n; ++j) {
            const long mi = (d.first < 0) ? flat : v * M + d.first + j;
            if (d.has_cat) catalog[mi] = redraw_cat_[j];
            if (d.has_q0) Q0[mi] = redraw_q0_[j];
        }
        if (record_events) {
            EventRecord e{};
            e.flat_idx = (int)flat;
            e.is_instab = inst;
            e.ls_iters = last_iters_;
            e.time = time;
            e.total_rate = total_rate_;
            for (int k = 0; k < D * D; ++k) { e.eps_macro[k] = eps_macro.data()[k]; e.sig_macro[k] = sig_macro.data()[k]; }
            events.push_back(e);
        }
        if (is_instab) *is_instab = inst;
        if (t_wait_out) *t_wait_out = t_wait;
        return MGK_EVENT_DONE;
    }

    // One attempt of run_simulation's inner loop (kmc_solver.hpp:763-952,
    // 3D :1776-1980) on the device with one round trip (mgk_kmc_event).
    int try_event(double remaining, std::pair<int, int> component,
                  std::uniform_real_distribution<double>& uniform_dist, int* is_instab, double* t_wait_out) {
        EventDraw d;
        mgk_event_in in;
        prepare_event(remaining, component, uniform_dist, d, in);
        mgk_event_out out{};
        check(mgk_kmc_event(ctx, &in, eps_macro.data(), sig_macro.data(), &out), "mgk_kmc_event");
        return complete_event(d, out, is_instab, t_wait_out);
    }

    // --- checkpoint and resume (no counterpart in the reference: main.cpp:252-258
    // parses output.checkpoint_interval and ignores it) -----------------------
    // A checkpoint holds the device state (mgk_state_save), the host state the
    // trajectory depends on (time, macros, the mt19937 stream, counters, the
    // event history) and the loop position of run_simulation.  run_simulation
    // continues a loaded checkpoint where it was written, bit for bit.
    struct LoopState {
        bool valid = false;
        int elastic_steps_done = 0, total_kmc_steps = 0, step = 1, seq = 0;
        bool mid_step = false;    // written inside a strain step (event cap, termination)
        double remaining = 0.0;   // time left in that strain step
    };
    LoopState resume_;

    void save_checkpoint(const std::string& path, const LoopState& at = LoopState{}) {
        const long long nb = mgk_state_size(ctx);
        if (nb < 0) throw std::runtime_error("mgk_state_size failed");
        std::vector<char> dev((size_t)nb);
        check(mgk_state_save(ctx, dev.data(), nb), "mgk_state_save");
        std::ostringstream rs;
        rs << rng;
        const std::string rng_text = rs.str();
        const std::string tmp = path + ".tmp";
        const std::filesystem::path dir = std::filesystem::path(path).parent_path();
        if (!dir.empty()) std::filesystem::create_directories(dir);
        {
            std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
            if (!f) throw std::runtime_error("cannot write checkpoint " + tmp);
            auto put = [&](const void* p, size_t n) { f.write(static_cast<const char*>(p), (std::streamsize)n); };
            const char magic[8] = {'M', 'G', 'K', 'C', 'H', 'K', 'P', 'T'};
            const int hdr[9] = {1, D, nx, ny, nz, M, at.valid ? 1 : 0, flips_since_sync,
                                sigma_macro_unit_initialized ? 1 : 0};
            put(magic, 8);
            put(hdr, sizeof(hdr));
            const int loop[5] = {at.elastic_steps_done, at.total_kmc_steps, at.step, at.seq, at.mid_step ? 1 : 0};
            put(loop, sizeof(loop));
            put(&at.remaining, sizeof(double));
            put(&time, sizeof(time));
            put(eps_macro.data(), sizeof(double) * D * D);
            put(sig_macro.data(), sizeof(double) * D * D);
            const long long cnt[7] = {counters.kmc_events, counters.cascades, counters.cascade_flips,
                                      counters.elastic_steps, counters.ls_solves, counters.ls_iters,
                                      counters.ls_iters_fp32};
            const int flags[2] = {counters.terminated ? 1 : 0, counters.event_cap_reached ? 1 : 0};
            put(cnt, sizeof(cnt));
            put(flags, sizeof(flags));
            const long long nr = (long long)rng_text.size();
            put(&nr, sizeof(nr));
            put(rng_text.data(), rng_text.size());
            const long long ne = (long long)events.size();
            put(&ne, sizeof(ne));
            if (ne) put(events.data(), sizeof(EventRecord) * events.size());
            put(&nb, sizeof(nb));
            put(dev.data(), dev.size());
            if (!f) throw std::runtime_error("write error on checkpoint " + tmp);
        }
        std::filesystem::rename(tmp, path);
    }

    void load_checkpoint(const std::string& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot read checkpoint " + path);
        auto get = [&](void* p, size_t n) {
            f.read(static_cast<char*>(p), (std::streamsize)n);
            if (!f) throw std::runtime_error("truncated checkpoint " + path);
        };
        char magic[8];
        int hdr[9], loop[5];
        double rem = 0.0;
        get(magic, 8);
        if (std::memcmp(magic, "MGKCHKPT", 8) != 0) throw std::runtime_error(path + " is not a checkpoint");
        get(hdr, sizeof(hdr));
        if (hdr[0] != 1 || hdr[1] != D || hdr[2] != nx || hdr[3] != ny || hdr[4] != nz || hdr[5] != M)
            throw std::runtime_error("checkpoint " + path + " is for another version, grid or mode count");
        get(loop, sizeof(loop));
        get(&rem, sizeof(rem));
        double t;
        Mat<D> em, sm;
        get(&t, sizeof(t));
        get(em.data(), sizeof(double) * D * D);
        get(sm.data(), sizeof(double) * D * D);
        long long cnt[7];
        int flags[2];
        get(cnt, sizeof(cnt));
        get(flags, sizeof(flags));
        long long nr = 0;
        get(&nr, sizeof(nr));
        if (nr < 0 || nr > (1 << 20)) throw std::runtime_error("corrupt checkpoint " + path);
        std::string rng_text((size_t)nr, '\0');
        get(&rng_text[0], (size_t)nr);
        long long ne = 0;
        get(&ne, sizeof(ne));
        if (ne < 0) throw std::runtime_error("corrupt checkpoint " + path);
        std::vector<EventRecord> ev((size_t)ne);
        if (ne) get(ev.data(), sizeof(EventRecord) * ev.size());
        long long nb = 0;
        get(&nb, sizeof(nb));
        if (nb <= 0) throw std::runtime_error("corrupt checkpoint " + path);
        std::vector<char> dev((size_t)nb);
        get(dev.data(), dev.size());
        check(mgk_state_load(ctx, dev.data(), nb), "mgk_state_load");
        std::istringstream rs(rng_text);
        rs >> rng;
        if (!rs) throw std::runtime_error("corrupt generator state in " + path);
        time = t;
        eps_macro = em;
        sig_macro = sm;
        flips_since_sync = hdr[7];
        sigma_macro_unit_initialized = hdr[8] != 0;
        if (sigma_macro_unit_initialized && fast_patching_enabled) unit_on_device_ = false;   // solved again on use
        counters.kmc_events = cnt[0]; counters.cascades = cnt[1]; counters.cascade_flips = cnt[2];
        counters.elastic_steps = cnt[3]; counters.ls_solves = cnt[4]; counters.ls_iters = cnt[5];
        counters.ls_iters_fp32 = cnt[6];
        counters.terminated = flags[0] != 0;
        counters.event_cap_reached = flags[1] != 0;
        events = std::move(ev);
        resume_ = LoopState{};
        if (hdr[6]) {
            resume_.valid = true;
            resume_.elastic_steps_done = loop[0];
            resume_.total_kmc_steps = loop[1];
            resume_.step = loop[2];
            resume_.seq = loop[3];
            resume_.mid_step = loop[4] != 0;
            resume_.remaining = rem;
        }
        sync_from_device();
    }

    // The file a checkpoint of run_simulation goes to: checkpoint_path under
    // output_dir (unless absolute), ".mgkc" appended, "{step}" replaced by the
    // elastic step count.
    std::string checkpoint_file(const std::string& checkpoint_path, int elastic_steps_done) const {
        std::string name = checkpoint_path;
        const size_t k = name.find("{step}");
        if (k != std::string::npos) name.replace(k, 6, std::to_string(elastic_steps_done));
        std::filesystem::path p(name + ".mgkc");
        if (p.is_relative()) p = std::filesystem::path(output_dir) / p;
        return p.string();
    }

    // checkpoint_interval: "none" (default), a positive integer k (a checkpoint
    // after every k-th elastic step) or "end" (one when the run returns).
    static int checkpoint_every(const std::string& interval) {
        std::string v = interval;
        std::transform(v.begin(), v.end(), v.begin(), ::tolower);
        if (v == "none" || v.empty()) return 0;
        if (v == "end") return -1;
        try {
            size_t used = 0;
            const int k = std::stoi(v, &used);
            if (used == v.size() && k > 0) return k;
        } catch (...) {
        }
        throw std::runtime_error("checkpoint_interval must be 'none', 'end' or a positive integer: " + interval);
    }

    // run_simulation (kmc_solver.hpp:721-1037, :1735-2062), fast patching off.
    void run_simulation(int n_global_steps, double step_size, std::pair<int, int> component,
                        const std::map<std::pair<int, int>, double>& stress_targets, double mixed_tol = 1e-4,
                        int mixed_max_iter = 50, std::string checkpoint_interval = "none",
                        std::string checkpoint_path = "checkpoint", std::string vtk_interval = "none",
                        bool vtk_elastic_only = true, bool track_cascades = false, bool enable_console_log = true,
                        std::string summary_filename = "summary_log.txt", bool enable_summary_log = true,
                        bool enable_global_log = true, double max_kmc_steps_pct = 0.3) {
        (void)track_cascades;
        const int ck_every = checkpoint_every(checkpoint_interval);
        init_logs(summary_filename, enable_summary_log, enable_global_log);
        const auto start = std::chrono::high_resolution_clock::now();
        int elastic_steps_done = 0, total_kmc_steps = 0, step = 1, seq = 0;
        double remaining = 0.0;
        bool mid_step = false;
        const bool resumed = resume_.valid;
        if (resumed) {
            // a loaded checkpoint: its counters, loop position and solved state
            elastic_steps_done = resume_.elastic_steps_done;
            total_kmc_steps = resume_.total_kmc_steps;
            step = resume_.step;
            seq = resume_.seq;
            mid_step = resume_.mid_step;
            remaining = resume_.remaining;
            resume_ = LoopState{};
        } else {
            counters = RunCounters{};
        }
        const int max_seq = static_cast<int>(max_kmc_steps_pct * N());
        if (!vtk_interval.empty() && vtk_interval != "none") std::filesystem::create_directories(output_dir);
        if (!resumed) {
            elastic_run();
            if (enable_console_log) std::cout << kSummaryHeader << std::string(std::strlen(kSummaryHeader) - 1, '-') << "\n";
            do_logging(0, elastic_steps_done, total_kmc_steps, "INIT", 0, 0, component, enable_console_log, start,
                       vtk_interval, vtk_elastic_only);
        }
        Mat<D> su = Mat<D>::Zero();
        su(component.first, component.second) = 1.0;
        const double dt_step = std::abs(step_size) / strain_rate;
        std::uniform_real_distribution<double> uniform_dist(0.0, 1.0);
        auto loop_state = [&](bool inside) {
            LoopState l;
            l.valid = true;
            l.elastic_steps_done = elastic_steps_done;
            l.total_kmc_steps = total_kmc_steps;
            l.step = step;
            l.seq = seq;
            l.mid_step = inside;
            l.remaining = inside ? remaining : 0.0;
            return l;
        };
        auto finish = [&](bool inside) {
            counters.kmc_events = total_kmc_steps;
            counters.elastic_steps = elastic_steps_done;
            sync_from_device();
            close_logs();
            if (ck_every < 0) save_checkpoint(checkpoint_file(checkpoint_path, elastic_steps_done), loop_state(inside));
        };
        while (elastic_steps_done < n_global_steps) {
            if (!mid_step) remaining += dt_step;
            mid_step = false;
            while (remaining > 0.0) {
                int inst = 0;
                double t_wait = 0.0;
                const int outcome = try_event(remaining, component, uniform_dist, &inst, &t_wait);
                if (outcome == MGK_EVENT_UNSTABLE) {
                    auto cs = run_cascade(step, component);
                    if (cs.second > 0) {
                        counters.cascades++;
                        counters.cascade_flips += cs.second;
                        do_logging(step, elastic_steps_done, total_kmc_steps, "CASCADE", cs.first, cs.second,
                                   component, enable_console_log, start, vtk_interval, vtk_elastic_only);
                        step++;
                    }
                    continue;
                }
                if (outcome == MGK_EVENT_DONE) {
                    remaining -= t_wait;
                    seq = inst ? seq + 1 : 0;
                    if (seq > max_seq) {
                        std::cout << "\n[TERMINATE] KMC instability sequence limit reached! " << seq
                                  << " steps." << std::endl;
                        counters.terminated = true;
                        finish(true);
                        return;
                    }
                    do_logging(step, elastic_steps_done, total_kmc_steps, inst ? "KMC_INSTAB" : "KMC", 0, 1,
                               component, enable_console_log, start, vtk_interval, vtk_elastic_only);
                    total_kmc_steps++;
                    step++;
                    if (max_events > 0 && total_kmc_steps >= max_events) {
                        counters.event_cap_reached = true;
                        finish(true);
                        return;
                    }
                    continue;
                }
                // no event (no rate, or t_wait >= remaining): to the end of the strain step
                if (fast_patching_enabled && sigma_macro_unit_initialized) {   // :941-949
                    ensure_unit_stress(component);
                    check(mgk_patch_apply(ctx, -1, nullptr, strain_rate * remaining), "mgk_patch_apply");
                }
                eps_macro += su * (strain_rate * remaining);
                time += remaining;
                remaining = 0.0;
            }
            // mixed boundary conditions (:955-987)
            double E_sum = 0.0, nu_sum = 0.0;
            for (double v : E_field) E_sum += v;
            for (double v : nu_field) nu_sum += v;
            const double E_avg = E_sum / E_field.size(), nu_avg = nu_sum / nu_field.size();
            for (int it = 0; it < mixed_max_iter; ++it) {
                elastic_run();
                Mat<D> err = Mat<D>::Zero();
                double err_max = 0.0;
                for (const auto& t : stress_targets) {
                    const double e = t.second - sig_macro(t.first.first, t.first.second);
                    err(t.first.first, t.first.second) = e;
                    err_max = std::max(err_max, std::abs(e));
                }
                if (err_max < mixed_tol) break;
                const double tr_sig = err.trace();
                const Mat<D> d_eps = (err - nu_avg * tr_sig * Mat<D>::Identity()) / E_avg;
                for (const auto& t : stress_targets)
                    eps_macro(t.first.first, t.first.second) += d_eps(t.first.first, t.first.second);
            }
            elastic_steps_done++;
            do_logging(step, elastic_steps_done, total_kmc_steps, "ELASTIC", 0, 0, component, enable_console_log,
                       start, vtk_interval, vtk_elastic_only);
            step++;
            if (ck_every > 0 && elastic_steps_done % ck_every == 0) {
                counters.kmc_events = total_kmc_steps;
                counters.elastic_steps = elastic_steps_done;
                save_checkpoint(checkpoint_file(checkpoint_path, elastic_steps_done), loop_state(false));
            }
        }
        // the last VTU unless the interval already wrote it (:998-1015)
        bool final_vtu = (vtk_interval == "last");
        if (!final_vtu && vtk_interval != "none") {
            try {
                const int k = std::stoi(vtk_interval);
                final_vtu = k > 0 && (step - 1) % k != 0;
            } catch (...) {
            }
        }
        if (final_vtu) export_vtu(vtu_step_name(step - 1, "_final"));
        // run time (:1017-1035)
        const double dur = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - start).count();
        const int h = static_cast<int>(dur / 3600.0);
        const int m = static_cast<int>((dur - h * 3600.0) / 60.0);
        const double sec = dur - h * 3600.0 - m * 60.0;
        std::ostringstream fin;
        fin << "\nSimulation Finish Time: " << get_current_timestamp() << "\n"
            << "Total Duration: " << std::fixed << std::setprecision(2) << dur << " seconds (" << h << "h "
            << std::setw(2) << std::setfill('0') << m << "m " << std::setw(2) << std::setfill('0')
            << static_cast<int>(sec) << "s)\n";
        if (f_summary.is_open()) f_summary << fin.str();
        if (enable_console_log) std::cout << fin.str();
        finish(false);
    }

    // Forced event steps: the accepted-event body without the remaining-time
    // test (the per-event hot path used for throughput measurement).
    void event_steps(int count, std::pair<int, int> component) {
        std::uniform_real_distribution<double> uniform_dist(0.0, 1.0);
        const double forever = std::numeric_limits<double>::infinity();
        for (int s = 0; s < count; ++s) {
            const int outcome = try_event(forever, component, uniform_dist, nullptr, nullptr);
            if (outcome == MGK_EVENT_UNSTABLE) {
                auto cs = run_cascade(0, component);
                if (cs.second > 0) { counters.cascades++; counters.cascade_flips += cs.second; }
                continue;
            }
            if (outcome != MGK_EVENT_DONE) break;
            counters.kmc_events++;
        }
    }

protected:
    double total_rate_ = 0.0;
    int n_unstable_ = 0, n_included_ = 0;
    int last_iters_ = 0;
    std::vector<double> unstable_q_;
    std::vector<Mat<D>> redraw_cat_;
    std::vector<Mat<D>> round_cat_;   // a cascade round's redraws
    std::vector<double> round_q0_;
    std::vector<double> redraw_q0_;
};

// Forced event steps (as KmcSimulationT::event_steps) of several
// trajectories advanced together through a device batch built from their
// contexts (mgk_batch_kmc_event): one launch sequence per step covers every
// trajectory still running.  A trajectory whose attempt finds unstable modes
// runs its cascade alone and goes on; one without rate stops, and so does
// one that has made caps[b] events in this call (caps may be null).
template <int D>
void batch_event_steps(mgk_batch* batch, const std::vector<KmcSimulationT<D>*>& sims, int count,
                       std::pair<int, int> component, const long long* caps = nullptr) {
    const int B = (int)sims.size();
    if (B != mgk_batch_size(batch)) throw std::runtime_error("batch size differs from the trajectory count");
    std::vector<typename KmcSimulationT<D>::EventDraw> draws(B);
    std::vector<mgk_event_in> in(B);
    std::vector<mgk_event_out> out(B);
    std::vector<int> active(B, 1);
    std::vector<double> em((size_t)B * D * D), sm((size_t)B * D * D);
    std::uniform_real_distribution<double> uniform_dist(0.0, 1.0);
    const double forever = std::numeric_limits<double>::infinity();
    std::vector<long long> made(B, 0);
    for (int b = 0; b < B; ++b)
        if (caps && caps[b] <= 0) active[b] = 0;
    for (int step = 0; step < count; ++step) {
        bool any = false;
        for (int b = 0; b < B; ++b) {
            if (!active[b]) continue;
            sims[b]->prepare_event(forever, component, uniform_dist, draws[b], in[b]);
            for (int k = 0; k < D * D; ++k) em[(size_t)b * D * D + k] = sims[b]->eps_macro.data()[k];
            any = true;
        }
        if (!any) break;
        check(mgk_batch_kmc_event(batch, in.data(), active.data(), em.data(), sm.data(), out.data()),
              "mgk_batch_kmc_event");
        for (int b = 0; b < B; ++b) {
            if (!active[b]) continue;
            KmcSimulationT<D>& sim = *sims[b];
            if (out[b].outcome == MGK_EVENT_DONE)
                for (int k = 0; k < D * D; ++k) {
                    sim.eps_macro.data()[k] = em[(size_t)b * D * D + k];
                    sim.sig_macro.data()[k] = sm[(size_t)b * D * D + k];
                }
            const int outcome = sim.complete_event(draws[b], out[b], nullptr, nullptr);
            if (outcome == MGK_EVENT_UNSTABLE) {
                auto cs = sim.run_cascade(0, component);
                if (cs.second > 0) { sim.counters.cascades++; sim.counters.cascade_flips += cs.second; }
            } else if (outcome == MGK_EVENT_DONE) {
                sim.counters.kmc_events++;
                if (caps && ++made[b] >= caps[b]) active[b] = 0;
            } else {
                active[b] = 0;
            }
        }
    }
}

}  // namespace mgk_host

// KmcSimulation2D (kmc_solver.hpp:98-1038): same constructor signature.
class KmcSimulation2D : public mgk_host::KmcSimulationT<2> {
public:
    KmcSimulation2D(int nx_, int ny_, int M_, double gamma0_, const std::vector<double>& E_field_,
                    const std::vector<double>& nu_field_, double pixel_ = 1.0,
                    const BarrierGenerator& barrier_gen_ = BarrierGenerator(),
                    std::string softening_scheme_ = "isotropic", double softening_cap_ = 2.0, double jp_ = 10.0,
                    double jt_ = 30.0, double neighbor_softening_fraction_ = 0.0, double q_act_temp_ = 0.37,
                    std::string output_dir_ = "output", double temperature_ = 0.0, double strain_rate_ = 1.0,
                    double stability_threshold_ = 0.0, double nu0_ = 1e13,
                    std::string plane_mode_ = "plane_strain", bool fast_patching_enabled_ = false,
                    int patch_radius_ = 5, int sync_interval_ = 100, std::string instability_mode_ = "cascade",
                    std::string cascade_timing_ = "none", bool scale_rate_by_volume_ = true,
                    bool redraw_directions_ = true, bool redraw_barriers_ = true, int seed_ = 42,
                    int precision_ = MGK_FP64, bool catalog_fixed_ = false, double catalog_magnitude_ = 0.1,
                    bool device_draws_ = false)
        : KmcSimulationT<2>(nx_, ny_, 1, M_, gamma0_, E_field_, nu_field_, pixel_, barrier_gen_, softening_scheme_,
                            softening_cap_, jp_, jt_, neighbor_softening_fraction_, q_act_temp_, output_dir_,
                            temperature_, strain_rate_, stability_threshold_, nu0_, plane_mode_,
                            fast_patching_enabled_, patch_radius_, sync_interval_, instability_mode_,
                            cascade_timing_, scale_rate_by_volume_, redraw_directions_, redraw_barriers_, false,
                            seed_, precision_, catalog_fixed_, catalog_magnitude_, device_draws_) {}
    void apply_flip_soa_2d(int x, int y, int m) { apply_flip(((long)x * ny + y) * M + m); }
};

// KmcSimulation3D (kmc_solver.hpp:1040-2063): same constructor signature.
class KmcSimulation3D : public mgk_host::KmcSimulationT<3> {
public:
    KmcSimulation3D(int nx_, int ny_, int nz_, int M_, double gamma0_, const std::vector<double>& E_field_,
                    const std::vector<double>& nu_field_, double pixel_ = 1.0,
                    const BarrierGenerator& barrier_gen_ = BarrierGenerator(),
                    std::string softening_scheme_ = "isotropic", double softening_cap_ = 2.0, double jp_ = 10.0,
                    double jt_ = 30.0, double neighbor_softening_fraction_ = 0.0, double q_act_temp_ = 0.37,
                    std::string output_dir_ = "output", double temperature_ = 0.0, double strain_rate_ = 1.0,
                    double stability_threshold_ = 0.0, double nu0_ = 1e13, bool fast_patching_enabled_ = false,
                    int patch_radius_ = 5, int sync_interval_ = 100, std::string instability_mode_ = "cascade",
                    std::string cascade_timing_ = "none", bool scale_rate_by_volume_ = true,
                    bool redraw_directions_ = true, bool redraw_barriers_ = true, bool use_3d_barriers_ = false,
                    int seed_ = 42, int precision_ = MGK_FP64, bool catalog_fixed_ = false,
                    double catalog_magnitude_ = 0.1, bool device_draws_ = false)
        : KmcSimulationT<3>(nx_, ny_, nz_, M_, gamma0_, E_field_, nu_field_, pixel_, barrier_gen_,
                            softening_scheme_, softening_cap_, jp_, jt_, neighbor_softening_fraction_, q_act_temp_,
                            output_dir_, temperature_, strain_rate_, stability_threshold_, nu0_, "plane_strain",
                            fast_patching_enabled_, patch_radius_, sync_interval_, instability_mode_,
                            cascade_timing_, scale_rate_by_volume_, redraw_directions_, redraw_barriers_,
                            use_3d_barriers_, seed_, precision_, catalog_fixed_, catalog_magnitude_, device_draws_) {}
    void apply_flip_soa_3d(int x, int y, int z, int m) { apply_flip((((long)x * ny + y) * nz + z) * M + m); }
};
