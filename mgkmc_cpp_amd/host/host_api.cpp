// host_api.cpp -- C façade (include/mgkmc_host.h) over the drop-in
// KmcSimulation2D/3D classes, for language bindings and benchmarks.
#include <chrono>
#include <cstring>
#include <ctime>
#include <iomanip>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

// only the C ABI is exported (the library is built with hidden visibility):
// the header-only C++ classes it instantiates stay private to it, so a caller
// that instantiates them with its own tensor types gets no clash
#pragma GCC visibility push(default)
#include "../../include/mgkmc_gpu.h"
#include "../../include/mgkmc_host.h"
#pragma GCC visibility pop
#include "kmc_solver.hpp"

// get_current_timestamp (declared by kmc_solver.hpp:22, defined in main.cpp:18-24)
std::string get_current_timestamp() {
    auto now = std::chrono::system_clock::now();
    auto t = std::chrono::system_clock::to_time_t(now);
    std::stringstream ss;
    ss << std::put_time(std::localtime(&t), "%Y-%m-%d %H:%M:%S");
    return ss.str();
}

namespace {
thread_local std::string g_err;

template <typename F> int guard(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return MGK_E_ARG;
    }
}

BarrierGenerator barrier_from(const mgk_sim_config& c) {
    BarrierGenerator b;
    static const char* names[] = {"gaussian", "rayleigh", "modified_rayleigh", "modified_rayleigh_with_exponential",
                                  "none"};
    b.type = names[(c.barrier_type >= 0 && c.barrier_type <= 4) ? c.barrier_type : 4];
    b.mean = c.b_mean;
    b.std_val = c.b_std;
    b.min_cutoff = c.b_min_cutoff;
    b.max_cutoff = c.b_max_cutoff;
    b.loc = c.b_loc;
    b.scale = c.b_scale;
    b.epsilon = c.b_epsilon;
    b.ratio = c.b_ratio;
    return b;
}
}  // namespace

struct mgk_sim {
    int dims = 2;
    std::string checkpoint_interval = "none", checkpoint_path = "checkpoint";
    std::string vtk_interval = "none", summary_filename = "summary_log.txt";
    bool vtk_elastic_only = true, enable_summary_log = false, enable_global_log = false;
    std::unique_ptr<KmcSimulation2D> s2;
    std::unique_ptr<KmcSimulation3D> s3;
    template <typename F> auto with(F&& f) { return dims == 2 ? f(*s2) : f(*s3); }
};

extern "C" {

const char* mgk_host_last_error(void) { return g_err.c_str(); }

void mgk_sim_config_defaults(mgk_sim_config* c) {
    std::memset(c, 0, sizeof(*c));
    // constructor defaults (kmc_solver.hpp:165-181) and main.cpp's YAML defaults (:279-317)
    c->dims = 2;
    c->nx = c->ny = c->nz = 32;
    c->pixel = 1.0;
    c->plane_mode = MGK_PLANE_STRAIN;
    c->M = 20;
    c->gamma0 = 0.14;
    c->barrier_type = MGK_BARRIER_GAUSSIAN;
    c->b_mean = 2.0;
    c->b_std = 0.6;
    c->b_min_cutoff = 0.1;
    c->b_max_cutoff = -1.0;
    c->b_loc = 1.0;
    c->b_scale = 0.5;
    c->b_epsilon = 0.1;
    c->b_ratio = 0.8;
    c->softening_directional = 0;
    c->softening_cap = 2.0;
    c->jp = 10.0;
    c->jt = 30.0;
    c->neighbor_softening_fraction = 0.0;
    c->q_act_temp = 0.37;
    c->temperature = 0.0;
    c->strain_rate = 1.0;
    c->stability_threshold = 0.0;
    c->nu0 = 1e13;
    c->instability_kmc = 0;
    c->scale_rate_by_volume = 1;
    c->redraw_directions = 1;
    c->redraw_barriers = 1;
    c->use_3d_barriers = 0;
    c->seed = 42;
    c->catalog_mode = MGK_CATALOG_RANDOM;
    c->catalog_magnitude = 0.1;
    c->ls_max_iter = 200;
    c->ls_tol = 1e-6;
    c->precision = MGK_FP64;
    c->fast_patching = 0;
    c->patch_radius = 5;
    c->sync_interval = 100;
    c->device_draws = 0;
}

void mgk_run_config_defaults(mgk_run_config* r) {
    std::memset(r, 0, sizeof(*r));
    r->n_global_steps = 500;
    r->step_size = 1e-4;
    r->drv_i = r->drv_j = 0;
    r->mixed_tol = 1e6;   // 1 MPa (main.cpp:280, :390)
    r->mixed_max_iter = 50;
    r->max_kmc_steps_pct = 0.3;
    r->max_events = -1;
}

int mgk_sim_create(const mgk_sim_config* c, const double* E, const double* nu, mgk_sim** out) {
    return guard([&]() -> int {
        if (!c || !E || !nu || !out) throw std::invalid_argument("null argument");
        const long N = (long)c->nx * c->ny * (c->dims == 3 ? c->nz : 1);
        std::vector<double> Ev(E, E + N), nv(nu, nu + N);
        auto s = std::make_unique<mgk_sim>();
        s->dims = c->dims;
        const BarrierGenerator bg = barrier_from(*c);
        const std::string scheme = c->softening_directional ? "directional" : "isotropic";
        const std::string imode = c->instability_kmc ? "kmc" : "cascade";
        const bool fixed = c->catalog_mode == MGK_CATALOG_FIXED;
        if (c->dims == 2) {
            s->s2 = std::make_unique<KmcSimulation2D>(
                c->nx, c->ny, c->M, c->gamma0, Ev, nv, c->pixel, bg, scheme, c->softening_cap, c->jp, c->jt,
                c->neighbor_softening_fraction, c->q_act_temp, "output", c->temperature, c->strain_rate,
                c->stability_threshold, c->nu0, c->plane_mode == MGK_PLANE_STRESS ? "plane_stress" : "plane_strain",
                c->fast_patching != 0, c->patch_radius, c->sync_interval, imode, "none", c->scale_rate_by_volume != 0, c->redraw_directions != 0,
                c->redraw_barriers != 0, c->seed, c->precision, fixed, c->catalog_magnitude, c->device_draws != 0);
        } else if (c->dims == 3) {
            s->s3 = std::make_unique<KmcSimulation3D>(
                c->nx, c->ny, c->nz, c->M, c->gamma0, Ev, nv, c->pixel, bg, scheme, c->softening_cap, c->jp, c->jt,
                c->neighbor_softening_fraction, c->q_act_temp, "output", c->temperature, c->strain_rate,
                c->stability_threshold, c->nu0, c->fast_patching != 0, c->patch_radius, c->sync_interval, imode, "none", c->scale_rate_by_volume != 0,
                c->redraw_directions != 0, c->redraw_barriers != 0, c->use_3d_barriers != 0, c->seed, c->precision,
                fixed, c->catalog_magnitude, c->device_draws != 0);
        } else {
            throw std::invalid_argument("dims must be 2 or 3");
        }
        s->with([&](auto& sim) {
            sim.ls_tol = c->ls_tol;
            sim.ls_max_iter = c->ls_max_iter;
            sim.record_events = true;
            return 0;
        });
        *out = s.release();
        return MGK_OK;
    });
}

int mgk_sim_destroy(mgk_sim* s) {
    delete s;
    return MGK_OK;
}

static void fill_stats(const mgk_host::RunCounters& k, double time, mgk_run_stats* st) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->kmc_events = k.kmc_events;
    st->cascades = k.cascades;
    st->cascade_flips = k.cascade_flips;
    st->elastic_steps = k.elastic_steps;
    st->ls_solves = k.ls_solves;
    st->ls_iters = k.ls_iters;
    st->ls_iters_fp32 = k.ls_iters_fp32;
    st->exact_selections = k.exact_selections;
    st->terminated = k.terminated ? 1 : 0;
    st->event_cap_reached = k.event_cap_reached ? 1 : 0;
    st->time = time;
}

int mgk_sim_run(mgk_sim* s, const mgk_run_config* r, mgk_run_stats* stats) {
    return guard([&]() -> int {
        std::map<std::pair<int, int>, double> targets;
        for (int t = 0; t < r->n_targets && t < 9; ++t) targets[{r->target_i[t], r->target_j[t]}] = r->target_val[t];
        return s->with([&](auto& sim) {
            sim.max_events = r->max_events;
            sim.run_simulation(r->n_global_steps, r->step_size, {r->drv_i, r->drv_j}, targets, r->mixed_tol,
                               r->mixed_max_iter, s->checkpoint_interval, s->checkpoint_path, s->vtk_interval,
                               s->vtk_elastic_only, false, false, s->summary_filename, s->enable_summary_log,
                               s->enable_global_log, r->max_kmc_steps_pct);
            fill_stats(sim.counters, sim.time, stats);
            return MGK_OK;
        });
    });
}

int mgk_sim_set_checkpoint(mgk_sim* s, const char* interval, const char* path) {
    return guard([&]() -> int {
        if (!s || !interval || !path) throw std::runtime_error("null argument");
        return s->with([&](auto& sim) {
            sim.checkpoint_every(interval);   // validates
            s->checkpoint_interval = interval;
            s->checkpoint_path = path;
            return MGK_OK;
        });
    });
}

int mgk_sim_set_output(mgk_sim* s, const char* output_dir, const char* vtk_interval, int vtk_elastic_only,
                       int enable_summary_log, int enable_global_log) {
    return guard([&]() -> int {
        if (!s || !output_dir || !vtk_interval) throw std::runtime_error("null argument");
        return s->with([&](auto& sim) {
            sim.output_dir = output_dir;
            s->vtk_interval = vtk_interval;
            s->vtk_elastic_only = vtk_elastic_only != 0;
            s->enable_summary_log = enable_summary_log != 0;
            s->enable_global_log = enable_global_log != 0;
            return MGK_OK;
        });
    });
}

int mgk_sim_export_vtu(mgk_sim* s, const char* path) {
    return guard([&]() -> int {
        if (!s || !path) throw std::runtime_error("null argument");
        return s->with([&](auto& sim) {
            sim.export_vtu(path);
            sim.wait_vtu();
            return MGK_OK;
        });
    });
}

int mgk_vtu_write(const char* path, int dims, int nx, int ny, int nz, double pixel, const double* eps,
                  const double* sig, const double* E, const double* nu, const double* eps_plastic,
                  const double* soft) {
    return guard([&]() -> int {
        if (!path || !eps || !sig || !E || !nu) throw std::runtime_error("null argument");
        if (dims != 2 && dims != 3) throw std::runtime_error("dims must be 2 or 3");
        if (dims == 2) nz = 1;
        const long N = (long)nx * ny * nz;
        std::vector<double> Ev(E, E + N), nuv(nu, nu + N);
        std::vector<mgkt::Vector2d> sp;
        if (soft) {
            sp.resize(N);
            std::memcpy(sp.data(), soft, sizeof(double) * 2 * N);
        }
        auto tensors = [N](const double* a, auto tag) {
            using T = decltype(tag);
            std::vector<T> v;
            if (a) {
                v.resize(N);
                std::memcpy(v.data(), a, sizeof(T) * N);
            }
            return v;
        };
        if (dims == 2)
            export_to_vtu_2d(path, nx, ny, tensors(eps, mgkt::Matrix2d{}), tensors(sig, mgkt::Matrix2d{}), Ev, nuv,
                             pixel, {}, tensors(eps_plastic, mgkt::Matrix2d{}), sp);
        else
            export_to_vtu(path, nx, ny, nz, tensors(eps, mgkt::Matrix3d{}), tensors(sig, mgkt::Matrix3d{}), Ev, nuv,
                          pixel, {}, tensors(eps_plastic, mgkt::Matrix3d{}), sp);
        return MGK_OK;
    });
}

int mgk_sim_save(mgk_sim* s, const char* path) {
    return guard([&]() -> int {
        if (!s || !path) throw std::runtime_error("null argument");
        return s->with([&](auto& sim) {
            sim.save_checkpoint(path);
            return MGK_OK;
        });
    });
}

int mgk_sim_load(mgk_sim* s, const char* path) {
    return guard([&]() -> int {
        if (!s || !path) throw std::runtime_error("null argument");
        return s->with([&](auto& sim) {
            sim.load_checkpoint(path);
            return MGK_OK;
        });
    });
}

int mgk_sim_event_steps(mgk_sim* s, int n, int drv_i, int drv_j, mgk_run_stats* stats) {
    return guard([&]() -> int {
        return s->with([&](auto& sim) {
            sim.counters = mgk_host::RunCounters{};
            sim.event_steps(n, {drv_i, drv_j});
            fill_stats(sim.counters, sim.time, stats);
            return MGK_OK;
        });
    });
}

int mgk_sim_elastic_run(mgk_sim* s) {
    return guard([&]() -> int { return s->with([&](auto& sim) { return sim.elastic_run(); }); });
}

int mgk_sim_get_field(mgk_sim* s, int which, double* host) {
    return guard([&]() -> int {
        return s->with([&](auto& sim) {
            mgk_host::check(mgk_sync(sim.ctx), "sync");
            mgk_host::check(mgk_get_field(sim.ctx, which, host), "mgk_get_field");
            return MGK_OK;
        });
    });
}

int mgk_sim_set_field(mgk_sim* s, int which, const double* host) {
    return guard([&]() -> int {
        return s->with([&](auto& sim) {
            mgk_host::check(mgk_set_field(sim.ctx, which, host), "mgk_set_field");
            const long N = sim.N();
            if (which == MGK_FIELD_Q0) std::memcpy(sim.Q0.data(), host, sizeof(double) * N * sim.M);
            if (which == MGK_FIELD_CATALOG) std::memcpy(sim.catalog[0].data(), host, sizeof(sim.catalog[0]) * N * sim.M);
            return MGK_OK;
        });
    });
}

int mgk_sim_get_macro(mgk_sim* s, double* eps_macro, double* sig_macro, double* time) {
    return guard([&]() -> int {
        return s->with([&](auto& sim) {
            const int F = (int)(sizeof(sim.eps_macro) / sizeof(double));
            for (int k = 0; k < F; ++k) { eps_macro[k] = sim.eps_macro.data()[k]; sig_macro[k] = sim.sig_macro.data()[k]; }
            *time = sim.time;
            return MGK_OK;
        });
    });
}

int mgk_sim_set_macro(mgk_sim* s, const double* eps_macro, double time) {
    return guard([&]() -> int {
        return s->with([&](auto& sim) {
            const int F = (int)(sizeof(sim.eps_macro) / sizeof(double));
            for (int k = 0; k < F; ++k) sim.eps_macro.data()[k] = eps_macro[k];
            sim.time = time;
            return MGK_OK;
        });
    });
}

long long mgk_sim_num_events(mgk_sim* s) {
    return s->with([&](auto& sim) { return (long long)sim.events.size(); });
}

int mgk_sim_get_events(mgk_sim* s, mgk_event_record* out, long long cap) {
    return s->with([&](auto& sim) {
        long long n = std::min<long long>(cap, (long long)sim.events.size());
        for (long long i = 0; i < n; ++i) {
            const auto& e = sim.events[i];
            std::memset(&out[i], 0, sizeof(out[i]));
            out[i].flat_idx = e.flat_idx;
            out[i].is_instab = e.is_instab;
            out[i].ls_iters = e.ls_iters;
            out[i].time = e.time;
            out[i].total_rate = e.total_rate;
            std::memcpy(out[i].eps_macro, e.eps_macro, sizeof(e.eps_macro));
            std::memcpy(out[i].sig_macro, e.sig_macro, sizeof(e.sig_macro));
        }
        return (int)n;
    });
}

int mgk_sim_device_ctx(mgk_sim* s, mgk_ctx** out) {
    *out = s->with([&](auto& sim) { return sim.ctx; });
    return MGK_OK;
}

}  // extern "C"

struct mgk_sim_batch {
    int dims = 2;
    std::vector<mgk_sim*> sims;
    mgk_batch* dev = nullptr;
    ~mgk_sim_batch() {
        if (dev) mgk_batch_destroy(dev);
    }
};

extern "C" {

int mgk_sim_batch_create(mgk_sim* const* sims, int B, mgk_sim_batch** out) {
    return guard([&]() -> int {
        if (!sims || !out || B < 1) throw std::runtime_error("bad argument");
        *out = nullptr;
        auto b = std::make_unique<mgk_sim_batch>();
        b->dims = sims[0]->dims;
        std::vector<mgk_ctx*> ctxs;
        for (int i = 0; i < B; ++i) {
            if (!sims[i] || sims[i]->dims != b->dims) throw std::runtime_error("the simulations of a batch share dims");
            b->sims.push_back(sims[i]);
            ctxs.push_back(sims[i]->with([&](auto& sim) { return sim.ctx; }));
        }
        if (mgk_batch_create(ctxs.data(), B, &b->dev) != MGK_OK) throw std::runtime_error(mgk_last_error());
        *out = b.release();
        return MGK_OK;
    });
}

int mgk_sim_batch_destroy(mgk_sim_batch* b) {
    delete b;
    return MGK_OK;
}

int mgk_sim_batch_set_member(mgk_sim_batch* b, int slot, mgk_sim* sim) {
    return guard([&]() -> int {
        if (!b || !sim) throw std::runtime_error("null argument");
        if (slot < 0 || slot >= (int)b->sims.size()) throw std::runtime_error("slot out of range");
        if (sim->dims != b->dims) throw std::runtime_error("the simulations of a batch share dims");
        mgk_ctx* ctx = sim->with([&](auto& s) { return s.ctx; });
        if (mgk_batch_set_member(b->dev, slot, ctx) != MGK_OK) throw std::runtime_error(mgk_last_error());
        b->sims[slot] = sim;
        return MGK_OK;
    });
}

int mgk_sim_batch_event_steps(mgk_sim_batch* b, int n, int drv_i, int drv_j, const long long* caps,
                              mgk_run_stats* stats) {
    return guard([&]() -> int {
        if (!b) throw std::runtime_error("null batch");
        auto run = [&](auto tag) {
            constexpr int D = decltype(tag)::value;
            std::vector<mgk_host::KmcSimulationT<D>*> v;
            for (mgk_sim* s : b->sims) {
                mgk_host::KmcSimulationT<D>* p;
                if constexpr (D == 2) p = s->s2.get(); else p = s->s3.get();
                p->counters = mgk_host::RunCounters{};
                v.push_back(p);
            }
            mgk_host::batch_event_steps<D>(b->dev, v, n, {drv_i, drv_j}, caps);
            for (size_t i = 0; i < v.size(); ++i) fill_stats(v[i]->counters, v[i]->time, stats ? stats + i : nullptr);
        };
        if (b->dims == 2) run(std::integral_constant<int, 2>{});
        else run(std::integral_constant<int, 3>{});
        return MGK_OK;
    });
}

}  // extern "C"
