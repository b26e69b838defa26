/*
 * mgkmc_host.h -- C façade of the native host driver (the drop-in
 * KmcSimulation2D/3D classes of include/mgkmc/kmc_solver.hpp running over
 * the device C ABI of mgkmc_gpu.h).  Used by language bindings (ctypes),
 * benchmarks and tests; C++ callers use the classes directly.
 *
 * The configuration mirrors the constructor arguments of
 * KmcSimulation2D/3D (kmc_solver.hpp:165-181, :1142-1158) and of
 * run_simulation (kmc_solver.hpp:721-729, :1735-1743), plus the keys this
 * project adds (catalog mode, LS tolerance, event cap, precision); their
 * defaults reproduce the reference.
 */
#ifndef MGKMC_HOST_H
#define MGKMC_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

enum mgk_barrier_type {
    MGK_BARRIER_GAUSSIAN = 0,                 /* "gaussian" */
    MGK_BARRIER_RAYLEIGH = 1,                 /* "rayleigh" */
    MGK_BARRIER_MODIFIED_RAYLEIGH = 2,        /* "modified_rayleigh" */
    MGK_BARRIER_MODIFIED_RAYLEIGH_EXP = 3,    /* "modified_rayleigh_with_exponential" */
    MGK_BARRIER_NONE = 4                      /* any other string: 0 then clipped (kmc_solver.hpp:35-95) */
};

enum mgk_catalog_mode {
    MGK_CATALOG_RANDOM = 0,   /* reference: per-(voxel, mode) random STZ strains (kmc_solver.hpp:232-237, :1042-1074) */
    MGK_CATALOG_FIXED = 1     /* extension: M fixed shears of engineering magnitude catalog_magnitude, shared by all voxels */
};

typedef struct mgk_sim_config {
    int    dims;                 /* 2 or 3 */
    int    nx, ny, nz;           /* nz ignored in 2D */
    double pixel;                /* nm */
    int    plane_mode;           /* MGK_PLANE_STRAIN / MGK_PLANE_STRESS (2D) */
    int    M;
    double gamma0;
    /* BarrierGenerator (kmc_solver.hpp:24-33) */
    int    barrier_type;
    double b_mean, b_std, b_min_cutoff, b_max_cutoff, b_loc, b_scale, b_epsilon, b_ratio;
    /* physics */
    int    softening_directional;
    double softening_cap, jp, jt, neighbor_softening_fraction, q_act_temp;
    double temperature, strain_rate, stability_threshold, nu0;
    /* dynamics */
    int    instability_kmc;      /* 0: "cascade", 1: "kmc" */
    int    scale_rate_by_volume;
    int    redraw_directions;
    int    redraw_barriers;
    int    use_3d_barriers;      /* 3D: stz_mode_glass_3d instead of stz_mode_glass */
    int    seed;
    /* extensions */
    int    catalog_mode;         /* mgk_catalog_mode */
    double catalog_magnitude;    /* engineering shear magnitude of the fixed catalog */
    int    ls_max_iter;          /* reference: 200 (kmc_solver.hpp:337) */
    double ls_tol;               /* reference: 1e-6 (kmc_solver.hpp:337) */
    int    precision;            /* mgk_precision (device driver only) */
    /* fast patching (dynamics.fast_patching, main.cpp:310-312; kmc_solver.hpp:276-332):
     * stress update by superposed unit-eigenstrain kernels, full solve every
     * sync_interval flips */
    int    fast_patching;        /* reference default: 0 */
    int    patch_radius;         /* reference default: 5 */
    int    sync_interval;        /* reference default: 100 */
    /* extension: constructor draws (Q0, random catalog) on the device with
     * Philox4x32-10 (mgk_kmc_draw) instead of the host std::mt19937 -- the
     * reference's distributions, not its stream; the generator stays seeded
     * for the redraws that follow */
    int    device_draws;
} mgk_sim_config;

typedef struct mgk_run_config {
    int    n_global_steps;
    double step_size;
    int    drv_i, drv_j;
    int    n_targets;            /* mixed stress targets, at most 9 */
    int    target_i[9];
    int    target_j[9];
    double target_val[9];        /* Pa */
    double mixed_tol;            /* Pa */
    int    mixed_max_iter;
    double max_kmc_steps_pct;
    long long max_events;        /* extension: stop after this many KMC events (<= 0: no cap) */
} mgk_run_config;

typedef struct mgk_run_stats {
    long long kmc_events;        /* total_kmc_steps */
    long long cascades;          /* cascade_event_count */
    long long cascade_flips;
    long long elastic_steps;     /* elastic_steps_done */
    long long ls_solves;
    long long ls_iters;          /* FP64 iterations summed over solves */
    long long ls_iters_fp32;     /* FP32 inner iterations summed over solves */
    int    terminated;           /* instability termination (kmc_solver.hpp:925-929) */
    int    event_cap_reached;
    double time;
    long long exact_selections;  /* events whose selection the left-to-right pass decided
                                    (targets in the rounding window of the parallel prefix) */
} mgk_run_stats;

/* One record per accepted KMC event (kmc_solver.hpp:811-937). */
typedef struct mgk_event_record {
    int    flat_idx;
    int    is_instab;
    int    ls_iters;
    int    pad;
    double time;
    double total_rate;
    double eps_macro[9];         /* column-major D x D, after the event's solve */
    double sig_macro[9];
} mgk_event_record;

typedef struct mgk_sim mgk_sim;
struct mgk_ctx;

const char* mgk_host_last_error(void);

void mgk_sim_config_defaults(mgk_sim_config* c);
void mgk_run_config_defaults(mgk_run_config* r);

int  mgk_sim_create(const mgk_sim_config* c, const double* E, const double* nu, mgk_sim** out);
int  mgk_sim_destroy(mgk_sim* s);
/* run_simulation with logging and VTU output off */
int  mgk_sim_run(mgk_sim* s, const mgk_run_config* r, mgk_run_stats* stats);
/* n forced KMC event steps on the current state: update_barriers, cascade
 * if unstable modes exist (cascade mode), else rates, u1/t_wait (time and
 * eps_macro advance), u2 selection, flip, redraw and the warm-started LS
 * solve -- the body of the accepted-event branch of run_simulation without
 * the remaining-time test. */
int  mgk_sim_event_steps(mgk_sim* s, int n, int drv_i, int drv_j, mgk_run_stats* stats);
/* Checkpoints (no counterpart in the reference, which parses
 * output.checkpoint_interval and ignores it, main.cpp:252-258).
 * mgk_sim_set_checkpoint: the checkpoint_interval and checkpoint_path that
 * mgk_sim_run passes to run_simulation ("none", "end" or a positive integer
 * k: a checkpoint after every k-th elastic step; the file is checkpoint_path
 * under the output directory, ".mgkc" appended, "{step}" replaced by the
 * elastic step count).  mgk_sim_save / mgk_sim_load write and read one file
 * now; a run_simulation after a load of a checkpoint it wrote continues where
 * it stopped, and the trajectory equals the uninterrupted one bit for bit. */
int  mgk_sim_set_checkpoint(mgk_sim* s, const char* interval, const char* path);
/* Outputs of mgk_sim_run (off by default): run_simulation's output directory,
 * vtk_interval and vtk_elastic_only (kmc_solver.hpp:721-729; "none",
 * "current", "last" or an interval), and the summary and global logs
 * (do_logging, :566-719).  VTU files are written by a host thread from a copy
 * of the fields. */
int  mgk_sim_set_output(mgk_sim* s, const char* output_dir, const char* vtk_interval, int vtk_elastic_only,
                        int enable_summary_log, int enable_global_log);
/* The current fields as one VTU file (export_to_vtu[_2d], vtu_writer.hpp:169-393). */
int  mgk_sim_export_vtu(mgk_sim* s, const char* path);
/* export_to_vtu / export_to_vtu_2d on host arrays (no device): eps, sig and
 * eps_plastic (nullable) N x D x D column-major per cell, E and nu N values,
 * soft (nullable) N x 2; voxel order x*ny*nz + y*nz + z. */
int  mgk_vtu_write(const char* path, int dims, int nx, int ny, int nz, double pixel, const double* eps,
                   const double* sig, const double* E, const double* nu, const double* eps_plastic,
                   const double* soft);
int  mgk_sim_save(mgk_sim* s, const char* path);
int  mgk_sim_load(mgk_sim* s, const char* path);
int  mgk_sim_elastic_run(mgk_sim* s);
int  mgk_sim_get_field(mgk_sim* s, int which, double* host);
int  mgk_sim_set_field(mgk_sim* s, int which, const double* host);
int  mgk_sim_get_macro(mgk_sim* s, double* eps_macro, double* sig_macro, double* time);
int  mgk_sim_set_macro(mgk_sim* s, const double* eps_macro, double time);
long long mgk_sim_num_events(mgk_sim* s);
int  mgk_sim_get_events(mgk_sim* s, mgk_event_record* out, long long cap);
/* the device context the simulation runs on (owned by the simulation) */
int  mgk_sim_device_ctx(mgk_sim* s, struct mgk_ctx** out);

/* Simulations advanced together (no counterpart in the reference): forced
 * event steps (as mgk_sim_event_steps) of B simulations of one shape,
 * precision and mode count, one device launch sequence per step for all of
 * them (mgk_batch in mgkmc_gpu.h).  A simulation whose attempt finds
 * unstable modes runs its cascade alone; one without rate stops, and so
 * does one that has made caps[b] events in the call (caps may be NULL).
 * Each trajectory equals its own mgk_sim_event_steps run (FP64 bit for
 * bit).  The simulations must outlive the batch; stats holds B entries. */
typedef struct mgk_sim_batch mgk_sim_batch;
int  mgk_sim_batch_create(mgk_sim* const* sims, int B, mgk_sim_batch** out);
int  mgk_sim_batch_destroy(mgk_sim_batch* b);
int  mgk_sim_batch_event_steps(mgk_sim_batch* b, int n, int drv_i, int drv_j, const long long* caps,
                               mgk_run_stats* stats);
/* Puts simulation `sim` (same dims, shape, precision and mode count, not in
 * another row) into row `slot` in place of the one there, which the batch no
 * longer uses (no counterpart in the reference: the serving loop's
 * continuous batching; mgk_batch_set_member). */
int  mgk_sim_batch_set_member(mgk_sim_batch* b, int slot, mgk_sim* sim);

#ifdef __cplusplus
}
#endif
#endif /* MGKMC_HOST_H */
