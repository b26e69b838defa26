"""Python mirror of the reference interface for the hot path, over the C ABI.

* ``spectral_solver_2d`` / ``spectral_solver_3d`` follow
  linear_elastic_solver.hpp:146-160 / :431-444 (tensor fields are numpy
  arrays of shape (N, D, D); the Gamma argument is not taken: the device
  evaluates the Green operator on the fly).
* ``Context`` is one device context (mgk_ctx): device-resident fields and the
  per-event KMC primitives.
* ``Simulation`` is the native drop-in KmcSimulation2D/3D (include/mgkmc_host.h).
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import DP, IP, P, check, check_host, gpu, host

_PLANE = {"plane_strain": abi.MGK_PLANE_STRAIN, "plane_stress": abi.MGK_PLANE_STRESS}


def _dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


def _aos(x, N, D):
    """(N, D, D) row-major numpy -> contiguous column-major AoS doubles."""
    x = np.asarray(x, dtype=np.float64).reshape(N, D, D)
    return np.ascontiguousarray(np.transpose(x, (0, 2, 1))).reshape(-1)


def _from_aos(buf, N, D):
    return np.ascontiguousarray(np.transpose(buf.reshape(N, D, D), (0, 2, 1)))


def _mat_aos(m, D):
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(D, D).T).reshape(-1)


def _mat_from(buf, D):
    return np.asarray(buf, dtype=np.float64).reshape(D, D).T.copy()


def _state_save(h):
    n = int(gpu().mgk_state_size(h))
    buf = C.create_string_buffer(n)
    check(gpu().mgk_state_save(h, buf, n))
    return buf.raw


def _state_load(h, data):
    buf = C.create_string_buffer(bytes(data), len(data))
    check(gpu().mgk_state_load(h, buf, len(data)))


class Context:
    """One trajectory's device state (mgk_ctx)."""

    def __init__(self, shape, E, nu, pixel=1.0, plane_mode="plane_strain", precision=abi.MGK_FP64):
        self.shape = tuple(int(s) for s in shape)
        self.dims = len(self.shape)
        self.D = self.dims
        self.N = int(np.prod(self.shape))
        E = np.ascontiguousarray(np.asarray(E, dtype=np.float64).reshape(-1))
        nu = np.ascontiguousarray(np.asarray(nu, dtype=np.float64).reshape(-1))
        if E.size != self.N or nu.size != self.N:
            raise ValueError("E and nu must have one value per voxel")
        pm = _PLANE[plane_mode] if isinstance(plane_mode, str) else int(plane_mode)
        nx, ny = self.shape[0], self.shape[1]
        nz = self.shape[2] if self.dims == 3 else 1
        h = P()
        check(gpu().mgk_create(self.dims, nx, ny, nz, float(pixel), pm, int(precision), _dp(E), _dp(nu), C.byref(h)))
        self._h = h
        self.M = None

    def close(self):
        if getattr(self, "_h", None):
            gpu().mgk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_precision(self, precision):
        check(gpu().mgk_set_precision(self._h, int(precision)))

    # ---- LS solve with host arrays (spectral_solver_2d/3d) ------------------
    def ls_solve(self, eps_bar, eps=None, eps_plastic=None, max_iter=200, tol=1e-6, want_sig=True):
        D, N = self.D, self.N
        bar = _mat_aos(eps_bar, D)
        warm = eps is not None
        eps_buf = _aos(eps, N, D) if warm else np.zeros(N * D * D)
        ep = _aos(eps_plastic, N, D) if eps_plastic is not None else None
        sig = np.zeros(N * D * D) if want_sig else None
        em = np.zeros(D * D)
        sm = np.zeros(D * D)
        it = C.c_int(0)
        diff = C.c_double(0.0)
        check(gpu().mgk_ls_solve(self._h, _dp(bar), _dp(eps_buf), 1 if warm else 0, _dp(ep), _dp(sig), _dp(em),
                                 _dp(sm), int(max_iter), float(tol), C.byref(it), C.byref(diff)))
        return dict(eps=_from_aos(eps_buf, N, D), sig=_from_aos(sig, N, D) if want_sig else None,
                    eps_macro=_mat_from(em, D), sig_macro=_mat_from(sm, D), iters=it.value, diff=diff.value)

    # ---- device-resident fields --------------------------------------------
    def _field_shape(self, which):
        D, N, M = self.D, self.N, self.M
        if which in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_PREV_DIR):
            return (N, D, D), True
        if which in (abi.FIELD_Q, abi.FIELD_Q0):
            return (N * M,), False
        if which == abi.FIELD_CATALOG:
            return (N * M, D, D), True
        if which == abi.FIELD_SOFT:
            return (N, 2), False
        return (N,), False

    def get_field(self, which):
        shape, tensor = self._field_shape(which)
        buf = np.zeros(int(np.prod(shape)))
        check(gpu().mgk_get_field(self._h, which, _dp(buf)))
        if tensor:
            return _from_aos(buf, shape[0], self.D)
        return buf.reshape(shape)

    def set_field(self, which, value):
        shape, tensor = self._field_shape(which)
        if tensor:
            buf = _aos(value, shape[0], self.D)
        else:
            buf = np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(-1))
        check(gpu().mgk_set_field(self._h, which, _dp(buf)))

    # ---- KMC primitives ----------------------------------------------------
    def kmc_init(self, **kw):
        p = abi.KmcParams()
        vals = dict(M=6, softening_directional=0, softening_cap=2.0, jp=10.0, jt=30.0,
                    neighbor_softening_fraction=0.0, tau=float("inf"), volume=1.0, temperature=0.0, nu0=1e13,
                    stability_threshold=0.0, instability_kmc=0, scale_rate_by_volume=1)
        vals.update(kw)
        for k, v in vals.items():
            setattr(p, k, v)
        check(gpu().mgk_kmc_init(self._h, C.byref(p)))
        self.M = int(p.M)

    def elastic_run(self, eps_macro, max_iter=200, tol=1e-6):
        D = self.D
        em = _mat_aos(eps_macro, D)
        sm = np.zeros(D * D)
        it = C.c_int(0)
        diff = C.c_double(0.0)
        check(gpu().mgk_elastic_run(self._h, _dp(em), _dp(sm), int(max_iter), float(tol), C.byref(it), C.byref(diff)))
        return dict(eps_macro=_mat_from(em, D), sig_macro=_mat_from(sm, D), iters=it.value, diff=diff.value)

    def update_rates(self, time):
        tot = C.c_double(0.0)
        nu = C.c_int(0)
        ni = C.c_int(0)
        check(gpu().mgk_update_rates(self._h, float(time), C.byref(tot), C.byref(nu), C.byref(ni)))
        return tot.value, nu.value, ni.value

    def select(self, u2, exact=False):
        """Selected flat mode index and its Q; exact=True runs the reference's
        left-to-right pass itself (mgk_select_exact)."""
        idx = C.c_int(-1)
        q = C.c_double(0.0)
        f = gpu().mgk_select_exact if exact else gpu().mgk_select
        check(f(self._h, float(u2), C.byref(idx), C.byref(q)))
        return idx.value, q.value

    def apply_flip(self, flat, time, new_catalog=None, new_Q0=None, first_mode=0):
        D = self.D
        cat = None
        n = 0
        if new_catalog is not None:
            new_catalog = np.asarray(new_catalog, dtype=np.float64)
            n = new_catalog.shape[0]
            cat = _aos(new_catalog, n, D)
        q0 = None
        if new_Q0 is not None:
            q0 = np.ascontiguousarray(np.asarray(new_Q0, dtype=np.float64).reshape(-1))
            n = q0.size
        check(gpu().mgk_apply_flip(self._h, int(flat), float(time), _dp(cat), _dp(q0), int(first_mode), int(n)))

    def select_flip(self, u2, time, new_catalog=None, new_Q0=None, first_mode=0):
        """Selection and flip chained on the device (no host round trip)."""
        D = self.D
        cat = None
        n = 0
        if new_catalog is not None:
            new_catalog = np.asarray(new_catalog, dtype=np.float64)
            n = new_catalog.shape[0]
            cat = _aos(new_catalog, n, D)
        q0 = None
        if new_Q0 is not None:
            q0 = np.ascontiguousarray(np.asarray(new_Q0, dtype=np.float64).reshape(-1))
            n = q0.size
        check(gpu().mgk_select_flip(self._h, float(u2), float(time), _dp(cat), _dp(q0), int(first_mode), int(n)))

    def kmc_event(self, eps_macro, time, neg_log_u1, u2, remaining=float("inf"), strain_rate=1.0, drv=(0, 0),
                  new_catalog=None, new_Q0=None, first_mode=0, max_iter=200, tol=1e-6, cascade_check=False):
        """One inner-loop attempt decided and applied on the device
        (mgk_kmc_event).  Returns (EventOut, eps_macro, sig_macro)."""
        D = self.D
        ev = abi.EventIn()
        ev.time, ev.neg_log_u1, ev.remaining, ev.strain_rate = float(time), float(neg_log_u1), float(remaining), \
            float(strain_rate)
        ev.drv_i, ev.drv_j = int(drv[0]), int(drv[1])
        ev.u2 = float(u2)
        n = 0
        cat = q0 = None
        if new_catalog is not None:
            new_catalog = np.asarray(new_catalog, dtype=np.float64)
            n = new_catalog.shape[0]
            cat = _aos(new_catalog, n, D)
        if new_Q0 is not None:
            q0 = np.ascontiguousarray(np.asarray(new_Q0, dtype=np.float64).reshape(-1))
            n = q0.size
        ev.new_catalog, ev.new_Q0 = _dp(cat), _dp(q0)
        ev.first_mode, ev.n_modes = int(first_mode), int(n)
        ev.max_iter, ev.tol, ev.cascade_check = int(max_iter), float(tol), 1 if cascade_check else 0
        em = _mat_aos(eps_macro, D)
        sm = np.zeros(D * D)
        out = abi.EventOut()
        check(gpu().mgk_kmc_event(self._h, C.byref(ev), _dp(em), _dp(sm), C.byref(out)))
        return out, _mat_from(em, D), _mat_from(sm, D)

    def last_selection(self):
        idx = C.c_int(-1)
        q = C.c_double(0.0)
        check(gpu().mgk_last_selection(self._h, C.byref(idx), C.byref(q)))
        return idx.value, q.value

    def find_unstable(self):
        n = C.c_int(0)
        check(gpu().mgk_find_unstable(self._h, None, None, 0, C.byref(n)))
        idx = np.zeros(max(n.value, 1), dtype=np.int32)
        q = np.zeros(max(n.value, 1))
        if n.value:
            check(gpu().mgk_find_unstable(self._h, idx.ctypes.data_as(IP), _dp(q), n.value, C.byref(n)))
        return idx[: n.value], q[: n.value]

    def sync(self):
        check(gpu().mgk_sync(self._h))

    def timer_start(self):
        check(gpu().mgk_timer_start(self._h))

    def timer_stop(self):
        ms = C.c_double(0.0)
        check(gpu().mgk_timer_stop(self._h, C.byref(ms)))
        return ms.value

    def time_ls_pass(self, pass_id, fp32=False, reps=20):
        ms = C.c_double(0.0)
        check(gpu().mgk_time_ls_pass(self._h, int(pass_id), 1 if fp32 else 0, int(reps), C.byref(ms)))
        return ms.value

    def device_bytes(self):
        return int(gpu().mgk_device_bytes(self._h))

    def kmc_draw(self, params, seed, draw_catalog=True):
        """Q0 and (draw_catalog) the STZ catalog drawn on the device
        (mgk_kmc_draw); params: abi.DrawParams."""
        check(gpu().mgk_kmc_draw(self._h, C.byref(params), int(seed) & (2 ** 64 - 1), int(bool(draw_catalog))))

    def state_save(self):
        """The trajectory state (mgk_state_save) as bytes."""
        return _state_save(self._h)

    def state_load(self, data):
        _state_load(self._h, data)

    def last_solve_stats(self):
        a = C.c_int(0)
        b = C.c_int(0)
        check(gpu().mgk_last_solve_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def last_solve_history(self):
        """Relative increment of every iteration of the last solve, in order
        (FP64 iterations and FP32 inner steps interleaved)."""
        return solve_history(self._h)


def spectral_solver_2d(nx, ny, E, nu, eps_bar, plane_mode, eps=None, max_iter=200, tol=1e-6, eps_plastic=None):
    """spectral_solver_2d (linear_elastic_solver.hpp:146-309) on the device.
    eps: None for a cold start, else the (N, 2, 2) warm-start field."""
    ctx = Context((nx, ny), E, nu, plane_mode=plane_mode)
    try:
        return ctx.ls_solve(eps_bar, eps=eps, eps_plastic=eps_plastic, max_iter=max_iter, tol=tol)
    finally:
        ctx.close()


def spectral_solver_3d(nx, ny, nz, E, nu, eps_bar, eps=None, max_iter=200, tol=1e-6, eps_plastic=None):
    """spectral_solver_3d (linear_elastic_solver.hpp:431-593) on the device."""
    ctx = Context((nx, ny, nz), E, nu)
    try:
        return ctx.ls_solve(eps_bar, eps=eps, eps_plastic=eps_plastic, max_iter=max_iter, tol=tol)
    finally:
        ctx.close()


def cycle_supported(ctx_handle):
    """True when the context's FP32+refine solves run each CG cycle as one
    persistent launch (mgk_cycle_supported)."""
    on = C.c_int(0)
    check(gpu().mgk_cycle_supported(ctx_handle, C.byref(on)))
    return bool(on.value)


def tuning(ctx_handle):
    """The autotuner's choices for the context's grid (mgk_tuning): a dict of
    the pass variants, the CG pipeline, whether the grid was timed, and the
    timings in ms."""
    ch = (C.c_int * 6)()
    ms = np.zeros(10)
    check(gpu().mgk_tuning(ctx_handle, ch, _dp(ms)))
    return dict(fwd_fp64=ch[0], fwd_fp32=ch[1], mid_fp64=ch[2], mid_fp32=ch[3], cycle=ch[4], tuned=bool(ch[5]),
                ms=dict(fwd_fp64=ms[0:2].tolist(), mid_fp64=ms[2:4].tolist(), fwd_fp32=ms[4:6].tolist(),
                        mid_fp32=ms[6:8].tolist(), solve_cycle=float(ms[8]), solve_plane=float(ms[9])))


def set_tuning(ctx_handle, fwd_fp64=0, fwd_fp32=0, mid_fp64=0, mid_fp32=0, cycle=1):
    """Forces the context's block-shape variants and CG pipeline (mgk_set_tuning)."""
    ch = (C.c_int * 5)(fwd_fp64, fwd_fp32, mid_fp64, mid_fp32, cycle)
    check(gpu().mgk_set_tuning(ctx_handle, ch))


def cycle_timer(ctx_handle, on=True):
    """Reset the persistent-cycle counters and time (on) every cycle launch
    of the context's solves with HIP events on its stream."""
    check(gpu().mgk_cycle_timer(ctx_handle, 1 if on else 0))


def cycle_timer_read(ctx_handle):
    """(summed ms, timed launches, CG steps they ran) since cycle_timer."""
    ms = C.c_double(0.0)
    n = C.c_longlong(0)
    st = C.c_longlong(0)
    check(gpu().mgk_cycle_timer_read(ctx_handle, C.byref(ms), C.byref(n), C.byref(st)))
    return ms.value, int(n.value), int(st.value)


def solve_history(ctx_handle):
    buf = np.zeros(32)
    n = C.c_int(0)
    check(gpu().mgk_last_solve_history(ctx_handle, _dp(buf), 32, C.byref(n)))
    return buf[: n.value].copy()


class Simulation:
    """The native drop-in KmcSimulation2D/3D (libmgkmc_host.so)."""

    def __init__(self, config, E, nu):
        self.config = config
        self.dims = config.dims
        self.D = config.dims
        shape = (config.nx, config.ny) + ((config.nz,) if config.dims == 3 else ())
        self.N = int(np.prod(shape))
        self.M = config.M
        E = np.ascontiguousarray(np.asarray(E, dtype=np.float64).reshape(-1))
        nu = np.ascontiguousarray(np.asarray(nu, dtype=np.float64).reshape(-1))
        h = P()
        check_host(host().mgk_sim_create(C.byref(config), _dp(E), _dp(nu), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            host().mgk_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, run_config):
        st = abi.RunStats()
        check_host(host().mgk_sim_run(self._h, C.byref(run_config), C.byref(st)))
        return st

    def set_checkpoint(self, interval, path):
        """checkpoint_interval ("none", "end" or k) and checkpoint_path of run()."""
        check_host(host().mgk_sim_set_checkpoint(self._h, str(interval).encode(), str(path).encode()))

    def set_output(self, output_dir, vtk_interval="none", vtk_elastic_only=True, summary_log=False,
                   global_log=False):
        """Outputs of run(): output directory, VTU interval ("none",
        "current", "last" or k) and the summary / global logs."""
        check_host(host().mgk_sim_set_output(self._h, str(output_dir).encode(), str(vtk_interval).encode(),
                                             int(bool(vtk_elastic_only)), int(bool(summary_log)),
                                             int(bool(global_log))))

    def export_vtu(self, path):
        check_host(host().mgk_sim_export_vtu(self._h, str(path).encode()))

    def save(self, path):
        check_host(host().mgk_sim_save(self._h, str(path).encode()))

    def load(self, path):
        """Restore a checkpoint; a following run() continues where it was written."""
        check_host(host().mgk_sim_load(self._h, str(path).encode()))

    def event_steps(self, n, drv=(0, 0)):
        st = abi.RunStats()
        check_host(host().mgk_sim_event_steps(self._h, int(n), drv[0], drv[1], C.byref(st)))
        return st

    def elastic_run(self):
        rc = host().mgk_sim_elastic_run(self._h)
        if rc < 0:
            check_host(rc)
        return rc

    def _shape(self, which):
        return Context._field_shape(self, which)

    def get_field(self, which):
        shape, tensor = self._shape(which)
        buf = np.zeros(int(np.prod(shape)))
        check_host(host().mgk_sim_get_field(self._h, which, _dp(buf)))
        return _from_aos(buf, shape[0], self.D) if tensor else buf.reshape(shape)

    def set_field(self, which, value):
        shape, tensor = self._shape(which)
        buf = _aos(value, shape[0], self.D) if tensor else np.ascontiguousarray(
            np.asarray(value, dtype=np.float64).reshape(-1))
        check_host(host().mgk_sim_set_field(self._h, which, _dp(buf)))

    def macro(self):
        D = self.D
        e = np.zeros(D * D)
        s = np.zeros(D * D)
        t = C.c_double(0.0)
        check_host(host().mgk_sim_get_macro(self._h, _dp(e), _dp(s), C.byref(t)))
        return _mat_from(e, D), _mat_from(s, D), t.value

    def set_macro(self, eps_macro, time):
        check_host(host().mgk_sim_set_macro(self._h, _dp(_mat_aos(eps_macro, self.D)), float(time)))

    def events(self):
        n = host().mgk_sim_num_events(self._h)
        arr = (abi.EventRecord * max(n, 1))()
        host().mgk_sim_get_events(self._h, arr, n)
        return list(arr)[:n]

    def device_context_handle(self):
        h = P()
        host().mgk_sim_device_ctx(self._h, C.byref(h))
        return h

    def device_bytes(self):
        """Device bytes this simulation's context holds."""
        return int(gpu().mgk_device_bytes(self.device_context_handle()))

    def patch(self, voxel, strain, pred=0.0):
        """Fast-patching stress update (mgk_patch_apply) for a strain (D, D)
        flipped at a voxel; voxel < 0: the predictor sig += sigma_unit * pred."""
        h = self.device_context_handle()
        D = self.config.dims
        st = _mat_aos(strain, D) if strain is not None else None
        check(gpu().mgk_patch_apply(h, int(voxel), _dp(st), float(pred)))

    def state_save(self):
        """The device state of the trajectory (mgk_state_save) as bytes; the
        host state (generator, macros, time) is in save()/load() files."""
        return _state_save(self.device_context_handle())

    def state_load(self, data):
        _state_load(self.device_context_handle(), data)

    def time_ls_pass(self, pass_id, fp32=False, reps=20):
        ms = C.c_double(0.0)
        check(gpu().mgk_time_ls_pass(self.device_context_handle(), int(pass_id), 1 if fp32 else 0, int(reps),
                                     C.byref(ms)))
        return ms.value


class SimulationBatch:
    """Simulations of one shape, precision and mode count whose forced event
    steps run together on the device (mgk_sim_batch_event_steps): one launch
    sequence per step for all of them.  The simulations stay usable on their
    own and must outlive the batch."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (P * len(self.sims))(*[s._h for s in self.sims])
        h = P()
        check_host(host().mgk_sim_batch_create(arr, len(self.sims), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            host().mgk_sim_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_member(self, slot, sim):
        """Puts `sim` (same shape, precision and mode count) into row `slot`
        in place of the simulation there, which the batch no longer uses
        (mgk_sim_batch_set_member): the batch and its captured graphs stay."""
        check_host(host().mgk_sim_batch_set_member(self._h, int(slot), sim._h))
        self.sims[slot] = sim

    def event_steps(self, n, drv=(0, 0), caps=None):
        """Up to n attempts for every simulation, simulation b stopping after
        caps[b] events (caps=None: no cap); one RunStats per simulation."""
        st = (abi.RunStats * len(self.sims))()
        cp = None
        if caps is not None:
            cp = (C.c_longlong * len(self.sims))(*[int(x) for x in caps])
        check_host(host().mgk_sim_batch_event_steps(self._h, int(n), drv[0], drv[1], cp, st))
        return list(st)


def correlated_field_device(shape, mean, std, corr, seed, clip_min=None, clip_max=None, noise=None):
    """generate_correlated_field_cpp (material_generator.hpp:130-258) on the
    device: noise (N normals, e.g. fields.box_muller_normals for the
    reference's own) or Philox normals keyed by seed, the Gaussian filter as
    three circular convolutions, standardisation, scale and clips."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    N = int(np.prod(shape))
    out = np.empty(N)
    nz_ = np.ascontiguousarray(np.asarray(noise, dtype=np.float64).reshape(-1)) if noise is not None else None
    check(gpu().mgk_correlated_field(D, shape[0], shape[1], shape[2] if D == 3 else 1, float(mean), float(std),
                                     float(corr), int(clip_min is not None), float(clip_min or 0.0),
                                     int(clip_max is not None), float(clip_max or 0.0), int(seed) & (2 ** 64 - 1),
                                     _dp(nz_), _dp(out)))
    return out


def write_vtu(path, shape, eps, sig, E, nu, pixel=1.0, eps_plastic=None, soft=None):
    """export_to_vtu / export_to_vtu_2d (vtu_writer.hpp:169-393) on host
    arrays: eps, sig, eps_plastic (N, D, D); E, nu (N,); soft (N, 2)."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    N = int(np.prod(shape))
    nz = shape[2] if D == 3 else 1
    e = _aos(eps, N, D)
    s = _aos(sig, N, D)
    ep = _aos(eps_plastic, N, D) if eps_plastic is not None else None
    E = np.ascontiguousarray(np.asarray(E, dtype=np.float64).reshape(-1))
    nu = np.ascontiguousarray(np.asarray(nu, dtype=np.float64).reshape(-1))
    sp = np.ascontiguousarray(np.asarray(soft, dtype=np.float64).reshape(N, 2)) if soft is not None else None
    check_host(host().mgk_vtu_write(str(path).encode(), D, shape[0], shape[1], nz, float(pixel), _dp(e), _dp(s),
                                    _dp(E), _dp(nu), _dp(ep), _dp(sp)))


def estimate_bytes(shape, M, per_voxel_catalog=False):
    """Device bytes a context of this grid and mode count holds once its KMC
    state and catalog are set (mgk_estimate_bytes)."""
    dims = len(shape)
    nz = shape[2] if dims == 3 else 1
    return int(gpu().mgk_estimate_bytes(dims, shape[0], shape[1], nz, int(M), 1 if per_voxel_catalog else 0))

