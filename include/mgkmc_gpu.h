/*
 * mgkmc_gpu.h -- C ABI of the MI355X (gfx950) kinetic Monte Carlo + FFT
 * Lippmann-Schwinger hot path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (MGK_OK == 0); mgk_last_error() gives the message of the last failure on
 * the calling thread.  A context (mgk_ctx) owns one HIP stream and all
 * device memory of one grid/trajectory; calls on one context are not
 * thread-safe, distinct contexts may be used from distinct threads.
 *
 * Tensor layout at this boundary is the reference's own: an array of N
 * column-major D x D matrices (Eigen::Matrix2d / Matrix3d storage), i.e.
 * 4 doubles per voxel in 2D ([xx, yx, xy, yy]) and 9 in 3D
 * ([xx, yx, zx, xy, yy, zy, xz, yz, zz]).  Voxel order is the reference's:
 * idx = x*ny + y (2D) and idx = x*ny*nz + y*nz + z (3D)
 * (linear_elastic_solver.hpp:81, :357).  So
 * reinterpret_cast<const double*>(std::vector<Eigen::Matrix2d>::data())
 * can be passed directly.
 */
#ifndef MGKMC_GPU_H
#define MGKMC_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define MGK_ABI_VERSION 1

enum mgk_status {
    MGK_OK = 0,
    MGK_E_ARG = 1,          /* invalid argument (null pointer, bad size, ...) */
    MGK_E_HIP = 2,          /* a HIP runtime call failed */
    MGK_E_UNSUPPORTED = 3,  /* valid for the reference, not built here (e.g. FFT length with a prime factor > 5) */
    MGK_E_STATE = 4         /* call order violated (e.g. KMC call before mgk_kmc_init) */
};

enum mgk_plane_mode { MGK_PLANE_STRAIN = 0, MGK_PLANE_STRESS = 1 };

/* Arithmetic of the Lippmann-Schwinger fixed point.
 * MGK_FP64: the reference's basic scheme, every iteration in FP64.
 * MGK_FP32_REFINE: FP32 inner iterations on the correction, FP64 residual
 * and convergence test (iterative refinement). */
enum mgk_precision { MGK_FP64 = 0, MGK_FP32_REFINE = 1 };

typedef struct mgk_ctx mgk_ctx;

/* Device-resident fields, for mgk_set_field / mgk_get_field.
 * Sizes (doubles): EPS, SIG, EPS_PLASTIC, PREV_DIR: N*D*D (AoS col-major);
 * Q, Q0: N*M; CATALOG: N*M*D*D; SOFT: N*2 ([g_p, g_t] per voxel,
 * kmc_solver.hpp:139); LAST_EVENT_TIME: N; E, NU: N (read-only). */
enum mgk_field {
    MGK_FIELD_EPS = 0,
    MGK_FIELD_SIG = 1,
    MGK_FIELD_EPS_PLASTIC = 2,
    MGK_FIELD_Q = 3,
    MGK_FIELD_Q0 = 4,
    MGK_FIELD_CATALOG = 5,
    MGK_FIELD_SOFT = 6,
    MGK_FIELD_LAST_EVENT_TIME = 7,
    MGK_FIELD_PREV_DIR = 8,
    MGK_FIELD_E = 9,
    MGK_FIELD_NU = 10
};

/* Per-trajectory KMC physics on the device (the members of
 * KmcSimulation2D/3D that update_barriers, the rate loop and
 * apply_flip_soa_* read: kmc_solver.hpp:100-131, :1077-1108). */
typedef struct mgk_kmc_params {
    int    M;                            /* modes per voxel */
    int    softening_directional;        /* softening_scheme == "directional" */
    double softening_cap;
    double jp;
    double jt;
    double neighbor_softening_fraction;
    double tau;                          /* softening decay time (s); +inf when T == 0 (kmc_solver.hpp:207-214) */
    double volume;                       /* pixel^3, also in 2D (kmc_solver.hpp:182) */
    double temperature;                  /* K; <= 0 gives beta = +inf (kmc_solver.hpp:786) */
    double nu0;                          /* attempt frequency (1/s) */
    double stability_threshold;          /* eV */
    int    instability_kmc;              /* instability_mode == "kmc": every mode is rated */
    int    scale_rate_by_volume;
} mgk_kmc_params;

const char* mgk_last_error(void);
int mgk_abi_version(void);

/* Context lifecycle.  dims = 2 or 3; nz is ignored (taken as 1) for dims = 2.
 * E and nu are N doubles in Pa and 1.  Replaces the per-call E/nu arguments
 * of spectral_solver_2d/3d (linear_elastic_solver.hpp:146-160, :431-444) and
 * the E_field/nu_field members of KmcSimulation2D/3D (kmc_solver.hpp:134-135). */
int mgk_create(int dims, int nx, int ny, int nz, double pixel, int plane_mode,
               int precision, const double* E, const double* nu, mgk_ctx** out);
int mgk_destroy(mgk_ctx* ctx);
int mgk_set_precision(mgk_ctx* ctx, int precision);

/* spectral_solver_2d / spectral_solver_3d
 * (linear_elastic_solver.hpp:146-309, :431-593), with host arrays.
 *   eps_bar      D*D  imposed macro strain
 *   eps          N*D*D in/out; warm_start != 0 means "eps.size() == N" in the
 *                reference (mean shifted to eps_bar), else it is filled with eps_bar
 *   eps_plastic  N*D*D or NULL (empty vector in the reference)
 *   sig_out      N*D*D or NULL
 *   eps_macro, sig_macro  D*D outputs (may alias eps_bar)
 *   iters_out    iterations run (== max_iter when not converged; the
 *                reference continues silently, so this is not an error)
 *   diff_out     last relative increment
 * The Green operator argument of the reference is not passed: it is
 * evaluated on the fly from the same mean medium (mean E, mean nu) that the
 * reference's callers build it from (kmc_solver.hpp:259-274, main.cpp:555). */
int mgk_ls_solve(mgk_ctx* ctx, const double* eps_bar, double* eps, int warm_start,
                 const double* eps_plastic, double* sig_out,
                 double* eps_macro, double* sig_macro,
                 int max_iter, double tol, int* iters_out, double* diff_out);

/* Device-resident fields. */
int mgk_set_field(mgk_ctx* ctx, int which, const double* host);
int mgk_get_field(mgk_ctx* ctx, int which, double* host);

/* KMC state on the device: allocates Q, Q0, catalog, softening state,
 * last_event_time, prev_strain_dir, eps_plastic (all zero; last_event_time
 * = -inf), as the constructor does (kmc_solver.hpp:196-220). */
int mgk_kmc_init(mgk_ctx* ctx, const mgk_kmc_params* p);

/* KmcSimulation*::elastic_run (kmc_solver.hpp:334-339, :1333-1338) on the
 * device-resident eps/sig/eps_plastic: eps_macro is the imposed strain on
 * entry and the field mean on return (the reference aliases the two). */
int mgk_elastic_run(mgk_ctx* ctx, double* eps_macro, double* sig_macro,
                    int max_iter, double tol, int* iters_out, double* diff_out);

/* update_barriers + find_unstable_indices count + rate loop
 * (kmc_solver.hpp:341-410, :412-421, :783-806; 3D :1340-1411, :1796-1818).
 * Writes Q on the device.  total_rate is the FP64 sum of the rates of the
 * included modes; n_unstable counts Q < stability_threshold; n_included
 * counts the rated modes. */
int mgk_update_rates(mgk_ctx* ctx, double time, double* total_rate,
                     int* n_unstable, int* n_included);

/* Event selection (kmc_solver.hpp:816-834, :1828-1849): with
 * target = u2 * total_rate, the first included mode, in flat order, whose
 * sequential FP64 prefix sum of rates is >= target (the last included mode if
 * none is).  Uses the rates of the last mgk_update_rates.  The device decides
 * from block sums and an in-block scan; when the target is within the
 * rounding bound of a prefix boundary it reruns the reference's exact
 * left-to-right sum (total and prefixes), so the index is the reference's.
 * Outputs the flat mode index and Q of that mode (is_instab = q_sel <= thr). */
int mgk_select(mgk_ctx* ctx, double u2, int* flat_idx, double* q_sel);

/* The reference's selection computed exactly as written (kmc_solver.hpp:808-826,
 * 3D :1820-1838): the left-to-right FP64 total, target = u2 * total and the
 * first included mode whose left-to-right prefix reaches it, with no parallel
 * shortcut.  Same outputs as mgk_select; one wave, O(N*M) sequential adds. */
int mgk_select_exact(mgk_ctx* ctx, double u2, int* flat_idx, double* q_sel);

/* apply_flip_soa_2d/3d + prev_strain_dir update + redraw upload
 * (kmc_solver.hpp:423-462, :835-864; 3D :1424-1471, :1850-1875).
 * new_catalog (n_modes*D*D, AoS) and new_Q0 (n_modes) may be NULL; when
 * given they overwrite modes [first_mode, first_mode + n_modes) of the
 * flipped voxel after the flip has used the old catalog entry. */
int mgk_apply_flip(mgk_ctx* ctx, int flat_idx, double time,
                   const double* new_catalog, const double* new_Q0,
                   int first_mode, int n_modes);

/* The flips of one cascade round in their order (run_cascade's loop over the
 * sorted unstable modes, kmc_solver.hpp:488-523, 3D :1500-1533): flip i is
 * flat_idx[i], with new_catalog (n * n_modes * D*D, AoS) and new_Q0
 * (n * n_modes) holding its redraw, n_modes = M (all_modes: modes 0..M-1 of
 * the voxel) or 1 (the flipped mode).  Each flip sees the softening of the
 * flips before it, as mgk_apply_flip called n times; one upload and one
 * launch per 64 flips. */
int mgk_apply_flips(mgk_ctx* ctx, int n, const int* flat_idx, double time, const double* new_catalog,
                    const double* new_Q0, int all_modes);

/* find_unstable_indices (kmc_solver.hpp:412-421): flat indices with
 * Q < stability_threshold in ascending order, and their Q values.
 * Writes at most cap entries; n_out is the full count. */
int mgk_find_unstable(mgk_ctx* ctx, int* idx_out, double* q_out, int cap, int* n_out);

/* Selection and flip chained on the device with no host round trip: the
 * accepted-event body of run_simulation from the u2 draw to the redraw
 * (kmc_solver.hpp:816-864, :1828-1875), for hosts that draw the redraw
 * values before knowing the selected voxel (the draws do not depend on it).
 * The redraw covers modes [first_mode, first_mode + n_modes) of the
 * selected voxel, or with first_mode = -1 and n_modes = 1 only the selected
 * mode.  With a shared (fixed) catalog a redraw must reproduce it.  The
 * chosen index is read back with mgk_last_selection, which syncs only if the
 * next mgk_elastic_run has not already done so. */
int mgk_select_flip(mgk_ctx* ctx, double u2, double time, const double* new_catalog,
                    const double* new_Q0, int first_mode, int n_modes);
int mgk_last_selection(mgk_ctx* ctx, int* flat_idx, double* q_sel);

/* One attempt of run_simulation's inner loop (kmc_solver.hpp:763-952,
 * 3D :1776-1980) with one host round trip: update_barriers and the FP64 total
 * rate R, the residence time t_wait = neg_log_u1 / R (the reference's
 * -log(u1) / total_rate), and, when the attempt is an event, time and
 * eps_macro(drv_i, drv_j) advanced by t_wait and strain_rate * t_wait,
 * selection with u2, flip + redraw upload (as mgk_select_flip) and the LS
 * solve (elastic_run).  The caller draws u1, u2 and the redraw values before
 * the outcome is known, in the reference's order, and rewinds its generator
 * when the outcome says the reference would not have drawn them:
 *   MGK_EVENT_DONE      event applied; eps_macro/sig_macro hold the
 *                       post-solve macros (eps_macro aliased to the mean)
 *   MGK_EVENT_UNSTABLE  cascade_check and unstable modes exist (the
 *                       reference runs run_cascade first; nothing changed)
 *   MGK_EVENT_NO_RATE   R is not > 0 (nothing drawn in the reference)
 *   MGK_EVENT_REJECTED  t_wait >= remaining (only u1 drawn in the reference)
 * remaining = +inf accepts every finite t_wait.
 * Selection: on an FP64 context exactly as mgk_select (the reference's index).
 * On an FP32+refine context the parallel prefix decides alone: targets within
 * the rounding window of a boundary (about 1e-9 of R at 3e6 modes, where
 * mgk_select reruns the left-to-right pass) are left to it, since that
 * context's barriers already differ from the FP64 ones by the solve's
 * tolerance.  The choice is deterministic either way. */
enum mgk_event_outcome {
    MGK_EVENT_DONE = 0,
    MGK_EVENT_UNSTABLE = 1,
    MGK_EVENT_NO_RATE = 2,
    MGK_EVENT_REJECTED = 3,
    MGK_EVENT_SKIPPED = 4        /* batch only: the trajectory was not active */
};
typedef struct mgk_event_in {
    double time;                 /* simulation time before the attempt */
    double neg_log_u1;           /* -log(u1) */
    double remaining;            /* time left in the strain step */
    double strain_rate;
    int drv_i, drv_j;            /* driven component of eps_macro */
    double u2;
    const double* new_catalog;   /* redraw, as mgk_select_flip (may be NULL) */
    const double* new_Q0;
    int first_mode, n_modes;
    int max_iter;                /* LS solve (>= 1) */
    double tol;
    int cascade_check;           /* instability_mode == "cascade" */
    int patch_mode;              /* 0: LS solve after the flip; 1: fast patching instead of the solve;
                                    2: fast patching, then the solve (a sync); see mgk_patch_init */
} mgk_event_in;
typedef struct mgk_event_out {
    int outcome;                 /* mgk_event_outcome */
    int n_unstable, n_included;
    int flat_idx;                /* selected flat mode, -1 without an event */
    double total_rate, t_wait, q_sel;
    int iters;                   /* LS iterations of the solve */
    int exact_selection;         /* 1: the selection was within the rounding window and the
                                    left-to-right pass (mgk_select_exact's) decided it (FP64 only) */
    double diff;
} mgk_event_out;
int mgk_kmc_event(mgk_ctx* ctx, const mgk_event_in* in, double* eps_macro, double* sig_macro,
                  mgk_event_out* out);

/* Batches (no counterpart in the reference, which runs one trajectory per
 * process): B contexts of one shape, precision and mode count whose event
 * attempts run together, each launch covering all B trajectories.  The
 * contexts keep their state and their own entry points (cascades, field
 * reads) between batch calls; a context belongs to at most one batch at a
 * time and must outlive it.  Per trajectory y, in[y], eps_macro + y*D*D,
 * sig_macro + y*D*D and out[y] mean what they mean for mgk_kmc_event, and
 * the results equal B separate mgk_kmc_event calls (FP64 bit for bit; in
 * FP32+refine the batch uses the largest inner-step count of its members).
 * in[y].max_iter, tol and the redraw mode range must agree across y;
 * active (may be NULL) switches trajectories off (MGK_EVENT_SKIPPED). */
typedef struct mgk_batch mgk_batch;
int mgk_batch_create(mgk_ctx* const* ctxs, int B, mgk_batch** out);
int mgk_batch_destroy(mgk_batch* b);
int mgk_batch_size(const mgk_batch* b);
/* on != 0 (the default): the launches of an event call that do not depend
 * on its outcome are captured once per launch structure into a HIP graph and
 * replayed; 0: launched one by one (same results). */
int mgk_batch_set_graphs(mgk_batch* b, int on);
/* Replaces the context of one row (no counterpart in the reference: the
 * serving loop's continuous batching).  The new context has the batch's
 * shape, precision and mode count and is not in another row; the old one is
 * no longer used by the batch and may be destroyed.  Captured graphs stay
 * valid: a row's parameter blocks are rebuilt from its context on every
 * event call. */
int mgk_batch_set_member(mgk_batch* b, int slot, mgk_ctx* ctx);
int mgk_batch_kmc_event(mgk_batch* b, const mgk_event_in* in, const int* active, double* eps_macro,
                        double* sig_macro, mgk_event_out* out);

/* Stream synchronisation and timing of the context's stream (HIP events). */
int mgk_sync(mgk_ctx* ctx);
int mgk_timer_start(mgk_ctx* ctx);
int mgk_timer_stop(mgk_ctx* ctx, double* ms);

/* Measurement helper (no counterpart in the reference): average duration in
 * ms of one pass of the LS pipeline (0: tau + r2c along the last axis,
 * 1: forward c2c along the middle axis (3D), 2: c2c + Green operator +
 * inverse c2c along the first axis, 3: inverse c2c along the middle axis,
 * 4: c2r along the last axis + update, 5: pass 0 with the FP32 CG vector
 * update, FP32+refine contexts only; the plane passes of 3D grids with square
 * nm = nl planes, FP32 only: 6 forward plane with the CG vector update
 * (FP32+refine contexts), 7 pass 2 on the plane layout, 8 inverse plane
 * without its reduction; the FP64 plane passes of such grids with planes of
 * up to 64 points, FP64 only: 9 forward plane, 10 pass 2 on the tiled FP64
 * spectrum, 11 inverse plane without its reduction), FP64 (fp32 == 0) or FP32, launched
 * reps times back to back on the context's stream between two HIP events.
 * Leaves the spectral buffer, the FP32 CG vectors and (pass 4 in FP64) the
 * strain field undefined: a measurement context, not a simulation one. */
int mgk_time_ls_pass(mgk_ctx* ctx, int pass, int fp32, int reps, double* avg_ms);

/* Persistent CG cycles (no counterpart in the reference, whose
 * spectral_solver_3d runs the basic scheme, linear_elastic_solver.hpp:494-581):
 * on an FP32+refine context whose grid is cubic with 64-point planes, each
 * refinement cycle's CG steps run as one launch whose blocks are all resident
 * and keep the CG vectors on chip (iterates equal to the separately launched
 * plane passes, bit for bit; the environment variable MGK_LS_CYCLE=0 turns it
 * off).  One such cycle runs at a time per process: a context's solve waits
 * while another context's cycle is in flight.
 * mgk_cycle_supported: 1 when this context's solves use it (the grid runs
 * it and the autotuner, or mgk_set_tuning, kept it on).
 * mgk_cycle_timer(on): measurement helper; resets the counters and, when on,
 * times every cycle launch of the context's solves with HIP events on the
 * context's stream (a solve whose refinement batch holds one cycle);
 * mgk_cycle_timer_read: the summed duration in ms, the timed launches and the
 * CG steps they ran. */
int mgk_cycle_supported(mgk_ctx* ctx, int* on);

/* Autotuner (no counterpart in the reference): the first context of a grid
 * and precision in a process times the block-shape variants of its LS passes
 * (pass A's line pairs and the M passes' columns, for the FP64 iterations and
 * the five-pass FP32 CG steps, on 3D lines of 32 points or more) and, where
 * the persistent CG cycle applies, a cold solve with it against one with the
 * launched plane passes; later contexts of that grid reuse the choice.  No
 * choice changes a result.  MGK_AUTOTUNE=0 keeps the default shapes.
 * choices (6 ints): pass A variant for FP64 iterations, for FP32 CG steps,
 * M-pass variant for FP64, for FP32 (0 default, 1 half-size blocks), the
 * persistent cycle (1 on, 0 off, -1 not applicable), 1 if the grid was timed
 * in this process.  ms (10 doubles, nullable): the timings behind them (the
 * two variants of each of the four passes, then the cold solve with and
 * without the cycle; 0 where not timed). */
int mgk_tuning(mgk_ctx* ctx, int* choices, double* ms);
/* Sets this context's variants (choices as mgk_tuning's first five; the
 * cycle entry 0 or 1), for measurements and tests. */
int mgk_set_tuning(mgk_ctx* ctx, const int* choices);
int mgk_cycle_timer(mgk_ctx* ctx, int on);
int mgk_cycle_timer_read(mgk_ctx* ctx, double* total_ms, long long* launches, long long* steps);

/* Statistics of the last LS solve: FP64 iterations, FP32 inner iterations. */
int mgk_last_solve_stats(mgk_ctx* ctx, int* fp64_iters, int* fp32_iters);

/* Measurement helper (no counterpart in the reference, whose verbose mode
 * prints the same quantity every 10 iterations, linear_elastic_solver.hpp:294):
 * the relative increment of each iteration of the last solve in the order run
 * (FP64 iterations and FP32 inner steps interleaved; at most 32), n_out the
 * count; at most cap values are written. */
int mgk_last_solve_history(mgk_ctx* ctx, double* diffs, int cap, int* n_out);

/* Device memory (no counterpart in the reference): the bytes a context
 * holds, and the bytes a context of this shape and mode count will hold
 * after mgk_create + mgk_kmc_init (M = 0: no KMC state) and a catalog upload
 * (per_voxel_catalog = 0: every voxel holds the same M strains).  Used to
 * admit work by HBM footprint. */
long long mgk_device_bytes(const mgk_ctx* ctx);
long long mgk_estimate_bytes(int dims, int nx, int ny, int nz, int M, int per_voxel_catalog);

/* Fast patching (dynamics.fast_patching; kmc_solver.hpp:276-332, :1256-1331):
 * an approximate stress update that replaces most LS solves.
 * mgk_patch_init: the stress responses of unit deviatoric eigenstrains at the
 *   grid centre (2 bases in 2D: xx = -yy, xy; 5 in 3D: xx = -yy, yy = -zz, xy,
 *   xz, yz), each an FP64 cold solve with tol 1e-6 and 200 iterations, cropped
 *   to a (2R+1)^d window (z window 1 when nz = 1) plus the mean the window
 *   misses.  0 <= radius <= every grid length.
 * mgk_patch_unit: sigma_macro_unit, the macro stress of a unit strain in the
 *   driven component (D x D column-major; cold solve, tol 1e-6, 200).
 * mgk_patch_apply: sig += sigma_macro_unit * pred (pred != 0), then the window
 *   superposition for a strain flipped at `voxel` (coefficients xx, xy in 2D;
 *   xx, -zz, xy, xz, yz in 3D) and the missing-mean shift; voxel < 0 applies
 *   the predictor alone.  Only the symmetric part of sigma_macro_unit is
 *   stored (the antisymmetric part does not enter any barrier).
 * mgk_kmc_event with patch_mode 1 or 2 applies the same update after the
 * device-decided flip with pred = strain_rate * t_wait (t_wait > 0). */
int mgk_patch_init(mgk_ctx* ctx, int radius);
int mgk_patch_unit(mgk_ctx* ctx, const double* strain_unit);
int mgk_patch_apply(mgk_ctx* ctx, int voxel, const double* strain, double pred);

/* Input generation on the device (no counterpart in the reference, which
 * draws on the host: constructor draws kmc_solver.hpp:216-240, :1194-1215;
 * generate_correlated_field_cpp material_generator.hpp:130-258).  Random
 * numbers are Philox4x32-10 (counter-based, keyed by seed), so the draws follow
 * the reference's distributions but not its std::mt19937 stream: for
 * statistically equivalent trajectories (batches, serving), not for parity.
 * mgk_kmc_draw: Q0 of every (voxel, mode) from the BarrierGenerator and, with
 *   draw_catalog, the per-(voxel, mode) STZ strains (planar, or
 *   stz_mode_glass_3d with use_3d_barriers) into the context's KMC state.
 * mgk_correlated_field: the reference's field (noise, Gaussian filter
 *   exp(-2 pi^2 corr^2 |f|^2), standardise, scale, clip) into host `out`;
 *   noise = N host normals (the result is then the reference's field to
 *   rounding) or NULL for device normals. */
typedef struct mgk_draw_params {
    int    barrier_type;         /* mgk_barrier_type values of include/mgkmc_host.h */
    double b_mean, b_std, b_min_cutoff, b_max_cutoff, b_loc, b_scale, b_epsilon, b_ratio;
    double gamma0;
    int    use_3d_barriers;
} mgk_draw_params;
int mgk_kmc_draw(mgk_ctx* ctx, const mgk_draw_params* p, unsigned long long seed, int draw_catalog);
int mgk_correlated_field(int dims, int nx, int ny, int nz, double mean, double std, double corr, int has_min,
                         double clip_min, int has_max, double clip_max, unsigned long long seed, const double* noise,
                         double* out);

/* mgk_correlated_field without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops that admit requests without hipMalloc).  noise_dev
 * (N device doubles, or NULL for device normals keyed by seed) and out_dev
 * (N device doubles) are the caller's; the work is enqueued on `stream` (a
 * hipStream_t, NULL for the null stream) and the call returns without waiting.
 * Every other buffer (the filter taps, the ping-pong field, the moments) lives
 * in one static device array of about 16 MiB per GPU, held until the process
 * ends.  Limits: N <= 2^21 voxels (128^3) and every axis <= 2048; larger grids
 * return MGK_E_ARG.  Calls on one GPU share that array: calls that overlap on
 * one GPU race.  The taps are computed on the device, so the result equals
 * mgk_correlated_field's to about 1e-13 relative. */
int mgk_correlated_field_noalloc(int dims, int nx, int ny, int nz, double mean, double std, double corr,
                                 int has_min, double clip_min, int has_max, double clip_max, unsigned long long seed,
                                 const double* noise_dev, double* out_dev, void* stream);

/* Trajectory state snapshot (no counterpart in the reference, whose
 * output.checkpoint_interval is parsed and ignored, main.cpp:252-258): the
 * strain, stress and plastic-strain fields, the KMC fields (Q, Q0, catalog,
 * softening, last event times, previous directions) and the solver's
 * warm-start and CG-tuning state, as one host buffer of mgk_state_size bytes.
 * Loading needs a context of the same grid, material (checked by hash) and
 * mode count; a trajectory continued from a loaded state repeats the saved
 * one bit for bit.  For checkpoints, preemption and eviction. */
long long mgk_state_size(const mgk_ctx* ctx);
int mgk_state_save(mgk_ctx* ctx, void* buf, long long bytes);
int mgk_state_load(mgk_ctx* ctx, const void* buf, long long bytes);

#ifdef __cplusplus
}
#endif
#endif /* MGKMC_GPU_H */
