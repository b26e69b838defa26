"""ctypes mirrors of the C ABI structs and constants (include/mgkmc_gpu.h,
include/mgkmc_host.h)."""
import ctypes as C

MGK_OK = 0
MGK_E_ARG = 1
MGK_E_HIP = 2
MGK_E_UNSUPPORTED = 3
MGK_E_STATE = 4

MGK_PLANE_STRAIN = 0
MGK_PLANE_STRESS = 1

MGK_FP64 = 0
MGK_FP32_REFINE = 1

FIELD_EPS = 0
FIELD_SIG = 1
FIELD_EPS_PLASTIC = 2
FIELD_Q = 3
FIELD_Q0 = 4
FIELD_CATALOG = 5
FIELD_SOFT = 6
FIELD_LAST_EVENT_TIME = 7
FIELD_PREV_DIR = 8
FIELD_E = 9
FIELD_NU = 10

BARRIER_GAUSSIAN = 0
BARRIER_RAYLEIGH = 1
BARRIER_MODIFIED_RAYLEIGH = 2
BARRIER_MODIFIED_RAYLEIGH_EXP = 3
BARRIER_NONE = 4

CATALOG_RANDOM = 0
CATALOG_FIXED = 1

EVENT_DONE = 0
EVENT_UNSTABLE = 1
EVENT_NO_RATE = 2
EVENT_REJECTED = 3
EVENT_SKIPPED = 4


class DrawParams(C.Structure):
    """mgk_draw_params (include/mgkmc_gpu.h)."""
    _fields_ = [
        ("barrier_type", C.c_int),
        ("b_mean", C.c_double), ("b_std", C.c_double), ("b_min_cutoff", C.c_double), ("b_max_cutoff", C.c_double),
        ("b_loc", C.c_double), ("b_scale", C.c_double), ("b_epsilon", C.c_double), ("b_ratio", C.c_double),
        ("gamma0", C.c_double),
        ("use_3d_barriers", C.c_int),
    ]


class EventIn(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("neg_log_u1", C.c_double),
        ("remaining", C.c_double),
        ("strain_rate", C.c_double),
        ("drv_i", C.c_int), ("drv_j", C.c_int),
        ("u2", C.c_double),
        ("new_catalog", C.POINTER(C.c_double)),
        ("new_Q0", C.POINTER(C.c_double)),
        ("first_mode", C.c_int), ("n_modes", C.c_int),
        ("max_iter", C.c_int),
        ("tol", C.c_double),
        ("cascade_check", C.c_int),
        ("patch_mode", C.c_int),
    ]


class EventOut(C.Structure):
    _fields_ = [
        ("outcome", C.c_int),
        ("n_unstable", C.c_int), ("n_included", C.c_int),
        ("flat_idx", C.c_int),
        ("total_rate", C.c_double), ("t_wait", C.c_double), ("q_sel", C.c_double),
        ("iters", C.c_int), ("exact_selection", C.c_int),
        ("diff", C.c_double),
    ]


class KmcParams(C.Structure):
    _fields_ = [
        ("M", C.c_int),
        ("softening_directional", C.c_int),
        ("softening_cap", C.c_double),
        ("jp", C.c_double),
        ("jt", C.c_double),
        ("neighbor_softening_fraction", C.c_double),
        ("tau", C.c_double),
        ("volume", C.c_double),
        ("temperature", C.c_double),
        ("nu0", C.c_double),
        ("stability_threshold", C.c_double),
        ("instability_kmc", C.c_int),
        ("scale_rate_by_volume", C.c_int),
    ]


class SimConfig(C.Structure):
    _fields_ = [
        ("dims", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("pixel", C.c_double),
        ("plane_mode", C.c_int),
        ("M", C.c_int),
        ("gamma0", C.c_double),
        ("barrier_type", C.c_int),
        ("b_mean", C.c_double), ("b_std", C.c_double), ("b_min_cutoff", C.c_double),
        ("b_max_cutoff", C.c_double), ("b_loc", C.c_double), ("b_scale", C.c_double),
        ("b_epsilon", C.c_double), ("b_ratio", C.c_double),
        ("softening_directional", C.c_int),
        ("softening_cap", C.c_double), ("jp", C.c_double), ("jt", C.c_double),
        ("neighbor_softening_fraction", C.c_double), ("q_act_temp", C.c_double),
        ("temperature", C.c_double), ("strain_rate", C.c_double),
        ("stability_threshold", C.c_double), ("nu0", C.c_double),
        ("instability_kmc", C.c_int),
        ("scale_rate_by_volume", C.c_int),
        ("redraw_directions", C.c_int),
        ("redraw_barriers", C.c_int),
        ("use_3d_barriers", C.c_int),
        ("seed", C.c_int),
        ("catalog_mode", C.c_int),
        ("catalog_magnitude", C.c_double),
        ("ls_max_iter", C.c_int),
        ("ls_tol", C.c_double),
        ("precision", C.c_int),
        ("fast_patching", C.c_int),
        ("patch_radius", C.c_int),
        ("sync_interval", C.c_int),
        ("device_draws", C.c_int),
    ]


class RunConfig(C.Structure):
    _fields_ = [
        ("n_global_steps", C.c_int),
        ("step_size", C.c_double),
        ("drv_i", C.c_int), ("drv_j", C.c_int),
        ("n_targets", C.c_int),
        ("target_i", C.c_int * 9),
        ("target_j", C.c_int * 9),
        ("target_val", C.c_double * 9),
        ("mixed_tol", C.c_double),
        ("mixed_max_iter", C.c_int),
        ("max_kmc_steps_pct", C.c_double),
        ("max_events", C.c_longlong),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("kmc_events", C.c_longlong),
        ("cascades", C.c_longlong),
        ("cascade_flips", C.c_longlong),
        ("elastic_steps", C.c_longlong),
        ("ls_solves", C.c_longlong),
        ("ls_iters", C.c_longlong),
        ("ls_iters_fp32", C.c_longlong),
        ("terminated", C.c_int),
        ("event_cap_reached", C.c_int),
        ("time", C.c_double),
        ("exact_selections", C.c_longlong),
    ]


class EventRecord(C.Structure):
    _fields_ = [
        ("flat_idx", C.c_int),
        ("is_instab", C.c_int),
        ("ls_iters", C.c_int),
        ("pad", C.c_int),
        ("time", C.c_double),
        ("total_rate", C.c_double),
        ("eps_macro", C.c_double * 9),
        ("sig_macro", C.c_double * 9),
    ]


def default_sim_config(**kw):
    """SimConfig with the reference's defaults (kmc_solver.hpp:165-181,
    main.cpp:279-317), then overridden by keyword arguments."""
    c = SimConfig()
    vals = dict(
        dims=2, nx=32, ny=32, nz=32, pixel=1.0, plane_mode=MGK_PLANE_STRAIN, M=20, gamma0=0.14,
        barrier_type=BARRIER_GAUSSIAN, b_mean=2.0, b_std=0.6, b_min_cutoff=0.1, b_max_cutoff=-1.0,
        b_loc=1.0, b_scale=0.5, b_epsilon=0.1, b_ratio=0.8, softening_directional=0,
        softening_cap=2.0, jp=10.0, jt=30.0, neighbor_softening_fraction=0.0, q_act_temp=0.37,
        temperature=0.0, strain_rate=1.0, stability_threshold=0.0, nu0=1e13, instability_kmc=0,
        scale_rate_by_volume=1, redraw_directions=1, redraw_barriers=1, use_3d_barriers=0, seed=42,
        catalog_mode=CATALOG_RANDOM, catalog_magnitude=0.1, ls_max_iter=200, ls_tol=1e-6,
        precision=MGK_FP64, fast_patching=0, patch_radius=5, sync_interval=100, device_draws=0,
    )
    vals.update(kw)
    for k, v in vals.items():
        setattr(c, k, v)
    return c


def default_run_config(**kw):
    r = RunConfig()
    vals = dict(n_global_steps=500, step_size=1e-4, drv_i=0, drv_j=0, n_targets=0, mixed_tol=1e6,
                mixed_max_iter=50, max_kmc_steps_pct=0.3, max_events=-1)
    targets = kw.pop("targets", None)
    vals.update(kw)
    for k, v in vals.items():
        setattr(r, k, v)
    if targets:
        r.n_targets = len(targets)
        for t, ((i, j), val) in enumerate(sorted(targets.items())):
            r.target_i[t] = i
            r.target_j[t] = j
            r.target_val[t] = val
    return r
