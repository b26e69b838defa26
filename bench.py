"""KMC events/s at 64^3 (3D heterogeneous LS-KMC, config c2), FP32+refine vs
FP64, on one MI355X, with the HBM roofline of the dominant kernel and the
CPU restatement timed on the host cores beside it.

A step is one KMC event of the per-event hot path on a pre-loaded state:
update_barriers + rates + FP64 total, u1 residence time, u2 selection, flip
+ redraw, and the warm-started Lippmann-Schwinger solve at tolerance 1e-8
(the accepted-event branch of run_simulation, kmc_solver.hpp:1776-1965).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  Synthetic inputs: E is a Gaussian random field (mean
85 GPa, CoV 0.15, correlation length 4 voxels, seed 42) generated on the host,
nu = 0.35; 12 fixed deviatoric shears of magnitude 0.1; barriers N(1.0, 0.1)
eV; T = 300 K; the state is pre-loaded at 1% uniaxial strain (eps_xx = 0.01).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "KMC events/s at 64^3, FP32+refine vs FP64; p50/p99 latency of mixed requests"
SHAPE = (64, 64, 64)
PRELOAD_STRAIN = 0.01


PASS_NAMES = {0: "k_fwd_last", 1: "k_mid<fwd>", 2: "k_first_gamma", 3: "k_mid<inv>", 4: "k_inv_last",
              5: "k_fwd_last+cg_update", 6: "k_fwd_plane+cg_update", 7: "k_first_gamma<plane>", 8: "k_inv_plane",
              9: "k_fwd_plane64", 10: "k_first_gamma<plane>", 11: "k_inv_plane64"}


def plane_passes():
    """The FP32 CG steps run as plane passes on the 64^3 grid (square
    64 x 64 planes) unless MGK_LS_PLANE=0."""
    return os.environ.get("MGK_LS_PLANE", "1") != "0"


def plane64_passes():
    """The FP64 iterations run as plane passes on the 64^3 grid unless
    MGK_LS_PLANE64=0."""
    return os.environ.get("MGK_LS_PLANE64", "1") != "0"


def _passes(fp32):
    """Pipeline passes of one precision as the event loop runs them."""
    if not fp32:
        return [9, 10, 11] if plane64_passes() else [0, 1, 2, 3, 4]
    return [6, 7, 8] if plane_passes() else [0, 1, 2, 3, 4, 5]


def _algorithmic_bytes_per_voxel(pass_id, fp32, nl):
    """Bytes per voxel one launch of a pipeline pass must move (read + write):
    symmetric 6-component tensors, half spectrum nl/2+1 of nl bins."""
    h = (nl // 2 + 1) / nl
    cs = 8 if fp32 else 16           # complex bytes
    spec = 6 * h * cs                # one half-spectrum tensor field per voxel
    if pass_id in (0, 9):            # y, dlam, dmu | eps, eps_p, lam, mu  -> spectrum
        return ((6 * 4 + 2 * 4) if fp32 else (6 * 8 + 6 * 8 + 2 * 8)) + spec
    if pass_id in (5, 6):            # r w p s x read, p s x r written, dlam dmu -> spectrum
        return 5 * 6 * 4 + 4 * 6 * 4 + 2 * 4 + spec
    if pass_id in (4, 8, 11):        # spectrum -> w (reads r) | eps_new (reads eps, writes the FP32 increment)
        return spec + (2 * 6 * 4 if fp32 else 2 * 6 * 8 + 6 * 4)
    return 2 * spec                  # B, C, D: read and write the spectrum


# The algorithmic bytes of the survey's three-pass model of one LS iteration
# (bytes per voxel, 3D, C = 6 strain components, s = 4 (FP32) or 8 (FP64)):
#   pass A  tau and the forward transforms of the first two axes: reads the
#           iterate (C s), eps_p (C, FP64), lam and mu (2 s), writes C s;
#   pass B  the transform along the third axis, Gamma0 and its inverse:
#           reads and writes C s;
#   pass C  the inverse transforms and the update: reads C s and the old
#           iterate (C s), writes C s.
# b_iter = (7C + 2) s + 8C: 224 B (FP32), 400 B (FP64).  The pipeline's passes
# map onto the model as: FP32 plane passes 6 -> A, 7 -> B, 8 -> C, FP64
# plane passes 9 -> A, 10 -> B, 11 -> C; the five-pass pipeline 0 + 1 -> A,
# 2 -> B, 3 + 4 -> C (the model's bytes of a split pass go to its first
# kernel and 0 to the second, so a per-kernel figure is only quoted for the
# plane passes, 2 and the whole-event figure below).
def _model_bytes_per_voxel(pass_id, fp32):
    Cc, sz = 6, (4 if fp32 else 8)
    a = (2 * Cc + 2) * sz + 8 * Cc
    b = 2 * Cc * sz
    c = 3 * Cc * sz
    return {0: a, 1: 0, 2: b, 3: 0, 4: c, 5: a, 6: a, 7: b, 8: c, 9: a, 10: b, 11: c}[pass_id]


def _model_event_bytes(info32, N, M=12):
    """The survey's bytes of one event: K b_iter(FP32) N + R b_iter(FP64) N
    + b_solve_end N + b_rate, with K the CG steps and R the FP64 iterations
    per event, b_solve_end = 8C (the final FP64 stress), b_rate = N (8C + 8M
    + 24) for a shared catalog (stress, Q0, softening state)."""
    Cc = 6
    K = info32["fp32_iters_per_event"]
    R = info32["ls_iters_per_event"]
    b = K * ((7 * Cc + 2) * 4 + 8 * Cc) + R * ((7 * Cc + 2) * 8 + 8 * Cc) + 8 * Cc + (8 * Cc + 8 * M + 24)
    return b * N


def _launches_per_event(pass_id, fp32, info):
    """Launches of a pass per event in the FP32+refine run."""
    cycles = info["ls_iters_per_event"]          # FP64 iterations = refinement cycles
    steps = info["fp32_iters_per_event"]         # CG steps, the first of each cycle without update
    if not fp32:
        return cycles
    if pass_id == 0:
        return cycles
    if pass_id in (5, 6):
        return max(steps - cycles, 0.0)
    return steps


def _preload(sim, D):
    e = np.zeros((D, D))
    e[0, 0] = PRELOAD_STRAIN
    sim.set_macro(e, 0.0)
    sim.elastic_run()


def run_gpu(precision, steps, warmup, E, nu):
    from mgkmc_cpp_amd import Simulation, abi, configs
    cfg = configs.sim_config(SHAPE, precision=precision, seed=42)
    sim = Simulation(cfg, E, nu)
    _preload(sim, 3)
    if warmup:
        sim.event_steps(warmup)
    t0 = time.perf_counter()
    st = sim.event_steps(steps)
    dt = time.perf_counter() - t0
    ev = st.kmc_events
    info = dict(events=int(ev), seconds=dt, events_per_s=ev / dt if dt > 0 else 0.0,
                ls_iters_per_event=st.ls_iters / max(st.ls_solves, 1),
                fp32_iters_per_event=st.ls_iters_fp32 / max(st.ls_solves, 1),
                cascades=int(st.cascades))
    return sim, info


# Bytes per voxel the persistent CG cycle moves: per step the spectrum out
# (F), read and written by pass C (G) and read back (I): 4 x 6 components x
# (nl/2+1)/nl bins x 8 B; per launch r, the tau factor's dlam and dmu (read
# by the plane blocks of their components) in, r and x out.
def _cycle_kernel_bytes(steps_per_launch, nl):
    h = (nl // 2 + 1) / nl
    return steps_per_launch * 4 * 6 * h * 8 + (6 * 4 + 6 * 4 + 4) + 2 * 6 * 4


def cycle_timing(E, nu, events=40, warmup=5):
    """The persistent CG cycle in situ: `events` FP32+refine events of the
    same workload with every cycle launch timed by HIP events on the context's
    stream (mgk_cycle_timer); None when the grid does not run it."""
    from mgkmc_cpp_amd import Simulation, abi, configs
    from mgkmc_cpp_amd.solver import cycle_supported, cycle_timer, cycle_timer_read
    cfg = configs.sim_config(SHAPE, precision=abi.MGK_FP32_REFINE, seed=42)
    sim = Simulation(cfg, E, nu)
    h = sim.device_context_handle()
    if not cycle_supported(h):
        sim.close()
        return None
    _preload(sim, 3)
    sim.event_steps(warmup)
    cycle_timer(h, True)
    st = sim.event_steps(events)
    ms, n, steps = cycle_timer_read(h)
    cycle_timer(h, False)
    sim.close()
    if n == 0:
        return None
    return dict(ms=ms / n, launches=n, steps=steps, events=int(st.kmc_events))


def kernel_roofline(sim, info32, cyc=None):
    """Average launch time of every pipeline pass (HIP events on the
    context's stream), weighted by its launches per event in the FP32+refine
    run; the pass with the largest time per event is reported.  `achieved`
    uses the survey model's bytes of that pass (_model_bytes_per_voxel); the
    bytes the kernel itself has to move (CG vectors included) are reported
    beside it as kernel_bytes_per_launch / kernel_gbs."""
    nl = SHAPE[2]
    N = int(np.prod(SHAPE))
    rows = []
    for fp32 in (False, True):
        for pass_id in _passes(fp32):
            ms = sim.time_ls_pass(pass_id, fp32=fp32, reps=20)
            b = _algorithmic_bytes_per_voxel(pass_id, fp32, nl) * N
            mb = _model_bytes_per_voxel(pass_id, fp32) * N
            n = _launches_per_event(pass_id, fp32, info32)
            rows.append(dict(kernel=PASS_NAMES[pass_id] + ("<float>" if fp32 else "<double>"), pass_id=pass_id,
                             fp32=fp32, ms=round(ms, 5), model_bytes=int(mb), bytes=int(b),
                             launches_per_event=round(n, 2), ms_per_event=round(ms * n, 5),
                             model_gbs=round(mb / (ms * 1e-3) / 1e9, 1), gbs=round(b / (ms * 1e-3) / 1e9, 1)))
    if cyc is not None:
        # the persistent cycle replaces the plane passes' launches: survey
        # bytes of its CG steps (b_iter per step) per launch
        spl = cyc["steps"] / cyc["launches"]
        mb = spl * ((7 * 6 + 2) * 4 + 8 * 6) * N
        b = _cycle_kernel_bytes(spl, nl) * N
        n = cyc["launches"] / max(cyc["events"], 1)
        ms = cyc["ms"]
        rows = [r for r in rows if r["pass_id"] not in (6, 7, 8)]
        rows.append(dict(kernel="k_cg_cycle<64>", pass_id="cycle", fp32=True, ms=round(ms, 5), model_bytes=int(mb),
                         bytes=int(b), launches_per_event=round(n, 3), ms_per_event=round(ms * n, 5),
                         steps_per_launch=round(spl, 3), model_gbs=round(mb / (ms * 1e-3) / 1e9, 1),
                         gbs=round(b / (ms * 1e-3) / 1e9, 1)))
    top = max(rows, key=lambda r: r["ms_per_event"])
    achieved = top["model_bytes"] / (top["ms"] * 1e-3) / 1e9
    own = top["bytes"] / (top["ms"] * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=_measured_traffic(top["kernel"]),
                traffic_source="profiles/traffic_64cube.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same "
                               "pass, calibrated on a device copy (not measured in this run)",
                bytes_model=("survey three-pass model: b_iter 224 B per voxel per CG step, "
                             f"{top['steps_per_launch']} steps per launch" if top["pass_id"] == "cycle" else
                             "survey three-pass model, bytes per voxel " + str(_model_bytes_per_voxel(top["pass_id"], top["fp32"]))),
                kernel_bytes_per_launch=int(top["bytes"]), kernel_gbs=round(own, 1),
                kernel_frac=round(own / HBM_PEAK_GBS, 4))
    return top, rows, roof


def _measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a device copy), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_64cube.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def fp64_rates(shapes, events=200, min_s=0.0):
    """Plain FP64 single-trajectory events/s per grid shape: the sequential
    baseline of the serving and batch workloads.  Each shape runs the served
    requests' own workload (instability mode "kmc", inputs drawn on the
    device) for `events` events after 2 of warm-up (longer if min_s asks);
    the cascades and exact-selection passes of the sample are reported."""
    from mgkmc_cpp_amd import abi
    from mgkmc_cpp_amd.serving import Request, request_simulation
    rates, detail = {}, {}
    for shape in shapes:
        sim = request_simulation(Request(rid=-1, shape=shape, events=events, seed=99, precision=abi.MGK_FP64,
                                         instability_kmc=1, device_inputs=True))
        sim.elastic_run()
        sim.event_steps(2)
        n = casc = exact = 0
        t0 = time.perf_counter()
        while n < events or time.perf_counter() - t0 < min_s:
            st = sim.event_steps(min(events, 50))
            n += int(st.kmc_events)
            casc += int(st.cascades)
            exact += int(st.exact_selections)
            if st.kmc_events == 0:
                break
        dt = time.perf_counter() - t0
        rates[shape] = n / dt if dt > 0 else 0.0
        detail["x".join(map(str, shape))] = dict(events=n, seconds=round(dt, 3), events_per_s=round(rates[shape], 1),
                                                 cascades=casc, exact_selections=exact)
        sim.close()
    return rates, detail


def serve(requests, budget_gb, label):
    """Runs `requests` through the serving loop; latency summary plus the
    sequential plain FP64 time of the same events."""
    from mgkmc_cpp_amd import serving
    print(f"[bench] {label}: {len(requests)} requests, {sum(r.events for r in requests)} events", file=sys.stderr,
          flush=True)
    srv = serving.Server(int(budget_gb * 2 ** 30),
                         progress=lambda t: print(f"[bench] {label}: {t}", file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    srv.run(requests)
    wall = time.perf_counter() - t0
    summ = serving.summary(requests)
    rates, detail = fp64_rates(sorted({r.shape for r in requests}, key=lambda x: (len(x), x)))
    seq = sum(r.events_done / rates[r.shape] for r in requests if rates[r.shape] > 0)
    summ.update(wall_s=round(wall, 3), sequential_fp64_s=round(seq, 3),
                speedup_vs_sequential_fp64=round(seq / summ["span_s"], 2) if summ["span_s"] else None,
                fp64_single=detail)
    for k in ("p50_latency_s", "p99_latency_s", "events_per_s", "span_s"):
        if summ.get(k) is not None:
            summ[k] = round(summ[k], 4)
    print(f"[bench] {label}: {summ}", file=sys.stderr, flush=True)
    return summ


def probe():
    """Launches for the counter runs: one 256 MiB device copy (calibration)
    then 5 launches of every pass at 64^3 in both precisions."""
    import torch
    from mgkmc_cpp_amd import Context, abi
    from mgkmc_cpp_amd import configs
    x = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print("[probe] device copy of 268435456 bytes", flush=True)
    E, nu = configs.material(SHAPE, 4.0, 42)
    ctx = Context(SHAPE, E, nu, precision=abi.MGK_FP32_REFINE)
    for fp32 in (False, True):
        for pass_id in _passes(fp32):
            ctx.time_ls_pass(pass_id, fp32=fp32, reps=4)
            print(f"[probe] pass {pass_id} {'fp32' if fp32 else 'fp64'}: 5 launches", flush=True)
    ctx.close()
    # the persistent CG cycle runs inside events: 10 FP32+refine events
    from mgkmc_cpp_amd import Simulation
    sim = Simulation(configs.sim_config(SHAPE, precision=abi.MGK_FP32_REFINE, seed=42), E, nu)
    _preload(sim, 3)
    sim.event_steps(10)
    sim.close()
    print("[probe] 10 FP32+refine events", flush=True)


def cpu_baseline(E, nu, budget_s=15.0):
    """The oracle (CPU restatement; FFTs through the reference's own pocketfft
    when oracle/_ref is built) on one host core, same workload and state, for
    a bounded number of events."""
    from mgkmc_cpp_amd import abi, configs
    from oracle import pyoracle as po
    ref = os.path.exists(po.REF_LIB)
    cfg = configs.sim_config(SHAPE, precision=abi.MGK_FP64, seed=42)
    osim = po.OracleSim(cfg, E, nu, ref=ref)
    e = np.zeros((3, 3))
    e[0, 0] = PRELOAD_STRAIN
    osim.set_macro(e, 0.0)
    osim.elastic_run()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s and n < 50:
        st = osim.event_steps(1)
        n += int(st.kmc_events)
        if st.kmc_events == 0:
            break
    dt = time.perf_counter() - t0
    try:
        cpu = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in cpu.splitlines() if l.startswith("Model name")), "")
    except Exception:
        model = ""
    return dict(value=round(n / dt, 4), unit="events/s", cores=1, kind="port",
                sample=f"{n} FP64 KMC events of the same 64^3 workload after the same 1% pre-load, "
                       f"{dt:.1f} s on one core ({model}); FFTs via "
                       f"{'the reference pocketfft (oracle/_ref)' if ref else 'the oracle FFT'}")


def field_errors(E, nu, events=40):
    """FP32+refine against the FP64 path on the same workload and state: the
    same number of events from the same pre-loaded state, then the largest
    difference of each field relative to the largest FP64 value (eps, sigma,
    eps_p, Q), and of the total rate R and the event times over the events."""
    from mgkmc_cpp_amd import abi
    sims = {}
    for prec in (abi.MGK_FP64, abi.MGK_FP32_REFINE):
        sims[prec], _ = run_gpu(prec, events, 0, E, nu)
    a, b = sims[abi.MGK_FP32_REFINE], sims[abi.MGK_FP64]
    ea, eb = a.events(), b.events()
    same = 0
    for x, y in zip(ea, eb):
        if x.flat_idx != y.flat_idx:
            break
        same += 1

    def rel(fa, fb):
        return float(np.abs(fa - fb).max() / max(np.abs(fb).max(), 1e-300))

    out = dict(events=events, same_events=same, norm="max |fp32_refine - fp64| / max |fp64|")
    if same == len(ea) == len(eb):
        out["rel_err"] = {name: rel(a.get_field(f), b.get_field(f)) for name, f in
                          (("eps", abi.FIELD_EPS), ("sig", abi.FIELD_SIG), ("eps_plastic", abi.FIELD_EPS_PLASTIC),
                           ("Q", abi.FIELD_Q))}
        out["rel_err"]["R"] = max(abs(x.total_rate - y.total_rate) / y.total_rate for x, y in zip(ea, eb))
        out["rel_err"]["time"] = max(abs(x.time - y.time) / abs(y.time) for x, y in zip(ea, eb))
        out["rel_err"] = {k: float("%.3g" % v) for k, v in out["rel_err"].items()}
    for sim in sims.values():
        sim.close()
    return out


def whole_run(E, nu, events=1000, steps=1000):
    """run_simulation on the 64^3 workload (FP32+refine, strain steps of 1e-4
    along xx as configs.uniaxial_run, logging off) until `events` KMC events
    or `steps` elastic steps: events/s with the elastic steps included."""
    from mgkmc_cpp_amd import Simulation, abi, configs
    cfg = configs.sim_config(SHAPE, precision=abi.MGK_FP32_REFINE, seed=42)
    sim = Simulation(cfg, E, nu)
    run = abi.default_run_config(n_global_steps=steps, **configs.uniaxial_run(events))
    t0 = time.perf_counter()
    st = sim.run(run)
    dt = time.perf_counter() - t0
    sim.close()
    return dict(events=int(st.kmc_events), elastic_steps=int(st.elastic_steps), cascades=int(st.cascades),
                seconds=round(dt, 3), events_per_s=round(st.kmc_events / dt, 1) if dt > 0 else None,
                ls_solves=int(st.ls_solves), workload=f"run_simulation, 64^3 FP32+refine, strain steps of 1e-4 "
                f"along xx from zero, up to {events} events or {steps} elastic steps, logging off")


def cpu_worker(seed, budget_s):
    """One CPU-baseline trajectory (the oracle on one core) for the aggregate
    leg: prints {events, seconds}.  Touches no GPU."""
    from mgkmc_cpp_amd import abi, configs
    from oracle import pyoracle as po
    E, nu = configs.material(SHAPE, 4.0, 42)
    cfg = configs.sim_config(SHAPE, precision=abi.MGK_FP64, seed=seed)
    osim = po.OracleSim(cfg, E, nu, ref=os.path.exists(po.REF_LIB))
    e = np.zeros((3, 3))
    e[0, 0] = PRELOAD_STRAIN
    osim.set_macro(e, 0.0)
    osim.elastic_run()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < budget_s and n < 50:
        st = osim.event_steps(1)
        n += int(st.kmc_events)
        if st.kmc_events == 0:
            break
    print(json.dumps(dict(events=n, seconds=time.perf_counter() - t0)), flush=True)


def cpu_aggregate(budget_s):
    """P independent oracle trajectories (seeds 42..42+P-1), one per core, run
    at once: aggregate CPU events/s.  P = the host's cores, at most 16 (a GPU
    box's CPU share)."""
    P = max(1, min(16, os.cpu_count() or 1))
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(42 + i),
                               "--cpu-budget", str(budget_s)], stdout=subprocess.PIPE, text=True,
                              env=dict(os.environ, OMP_NUM_THREADS="1")) for i in range(P)]
    t0 = time.perf_counter()
    res = []
    for pr in procs:
        outp, _ = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError("CPU worker failed")
        res.append(json.loads(outp.strip().splitlines()[-1]))
    wall = time.perf_counter() - t0
    n = sum(r["events"] for r in res)
    return dict(value=round(n / wall, 4), unit="events/s", cores=P, kind="port",
                sample=f"{P} independent FP64 trajectories of the same 64^3 workload (seeds 42..{41 + P}), "
                       f"one per core, {n} events in {wall:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--probe", action="store_true", help="only the launches of the counter runs")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c3-shape"], default="c2",
                    help="c2: the headline line (with a bounded c4 sample); c3 / c4: the full batch / serving runs; "
                         "c3-shape: the c3 trajectories of one grid (--shape) as one batch, one thread")
    ap.add_argument("--shape", default="48x48x48", help="grid of the c3-shape workload")
    ap.add_argument("--seconds", type=float, default=3.0, help="timed seconds of the c3-shape workload")
    ap.add_argument("--serving-requests", type=int, default=48, help="requests of the bounded c4 sample")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="leave out the full c3 batch run of the default line")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="HBM admission budget of the serving loop")
    ap.add_argument("--cpu-worker", type=int, default=None, help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.cpu_worker is not None:
        cpu_worker(a.cpu_worker, a.cpu_budget)
        return
    if a.probe:
        probe()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}); "
              "trajectories shard across GPUs without communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    from mgkmc_cpp_amd import abi, configs
    if a.workload == "c3-shape":
        return main_c3_shape(a)
    if a.workload in ("c3", "c4"):
        return main_many(a)
    E, nu = configs.material(SHAPE, 4.0, 42)
    print(f"[bench] FP64 run: {a.steps} events after {a.warmup} warm-up", file=sys.stderr, flush=True)
    sim64, i64 = run_gpu(abi.MGK_FP64, a.steps, a.warmup, E, nu)
    print(f"[bench] FP64 {i64}", file=sys.stderr, flush=True)
    sim64.close()
    print(f"[bench] FP32+refine run", file=sys.stderr, flush=True)
    sim32, i32 = run_gpu(abi.MGK_FP32_REFINE, a.steps, a.warmup, E, nu)
    print(f"[bench] FP32+refine {i32}", file=sys.stderr, flush=True)
    cyc = cycle_timing(E, nu)
    print(f"[bench] persistent CG cycle {cyc}", file=sys.stderr, flush=True)
    top, rows, roof = kernel_roofline(sim32, i32, cyc)
    print(f"[bench] pass timings {json.dumps(rows)}", file=sys.stderr, flush=True)
    roof["kernel"] = top["kernel"]
    roof["bytes_per_launch"] = int(top["model_bytes"])
    roof["avg_launch_ms"] = round(top["ms"], 5)
    N = int(np.prod(SHAPE))
    ev_bytes = _model_event_bytes(i32, N)
    ev_gbs = ev_bytes / (i32["seconds"] / max(i32["events"], 1)) / 1e9
    roof["per_event"] = dict(model_bytes=int(ev_bytes), achieved=round(ev_gbs, 1), frac=round(ev_gbs / HBM_PEAK_GBS, 4),
                             note="survey bytes of a whole event (CG steps, FP64 iterations, final stress, rate "
                                  "pass) over the measured time per event")
    sim32.close()
    print("[bench] FP32+refine against FP64: per-field errors", file=sys.stderr, flush=True)
    errs = field_errors(E, nu)
    print(f"[bench] field errors {errs}", file=sys.stderr, flush=True)
    print("[bench] whole run with elastic steps", file=sys.stderr, flush=True)
    whole = whole_run(E, nu)
    print(f"[bench] whole run {whole}", file=sys.stderr, flush=True)
    cpu = cpu_all = None
    if rank == 0 and not a.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(E, nu, a.cpu_budget)
        print(f"[bench] CPU baseline {cpu}", file=sys.stderr, flush=True)
        cpu_all = cpu_aggregate(a.cpu_budget)
        print(f"[bench] CPU aggregate {cpu_all}", file=sys.stderr, flush=True)
    c3 = None
    if not a.no_serving and not a.no_c3:
        reqs = configs.c3_requests(256, events=2000)
        c3 = serve(reqs, a.budget_gb, "c3")
        c3["workload"] = ("c3: 256 trajectories on grids drawn from 32^2, 64^2, 128^2, 32^3, 48^3 (seed 7), "
                          "2000 events each, FP32+refine, batched by grid")
    serving_sample = None
    if not a.no_serving:
        # bounded c4 sample: Poisson arrivals at 50/s, grids up to 64^3, events
        # per request one tenth of c4's 500-5000
        reqs = configs.c4_requests(a.serving_requests, rate=50.0, seed=11, events=(50, 500),
                                   grids=configs.C4_GRIDS[:-1])
        serving_sample = serve(reqs, a.budget_gb, "c4 sample")
        serving_sample["workload"] = (f"c4 sample: {a.serving_requests} requests, Poisson arrivals 50/s, grids "
                                      "32^2..64^3, 50-500 events each, FP32+refine")
    value = i32["events_per_s"]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * i32["seconds"] / max(i32["events"], 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-refine",
        "data": "synthetic",
        "config": {"workload": "c2: 3D 64^3 heterogeneous E (CoV 0.15, corr 4), 12 fixed shears |0.1|, "
                               "T=300K, LS tol 1e-8, one KMC event per step, FP32 iterations + FP64 refinement",
                   "grid": list(SHAPE), "modes": 12, "ls_tol": 1e-8, "preload_strain_xx": PRELOAD_STRAIN},
        "fp64_events_per_s": round(i64["events_per_s"], 3),
        "fp32_refine_speedup_vs_fp64": round(value / i64["events_per_s"], 3) if i64["events_per_s"] else None,
        "ls_iters_per_event": {"fp64": round(i64["ls_iters_per_event"], 2),
                               "fp32_refine_fp64": round(i32["ls_iters_per_event"], 2),
                               "fp32_refine_fp32": round(i32["fp32_iters_per_event"], 2)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_aggregate": cpu_all,
        "fp32_refine_errors_vs_fp64": errs,
        "whole_run": whole,
        "serving": serving_sample,
        "c3": c3,
    }
    if rank == 0:
        print(json.dumps(out), flush=True)


def main_c3_shape(a):
    """The c3 trajectories of one grid advanced as one batch from one thread
    (no other shape beside it): set-up, one untimed event step, then event
    steps of 8 events for about `a.seconds`.  Their single-trajectory
    FP32+refine rate (the first trajectory alone, 100 events) is reported
    beside the batch's."""
    from mgkmc_cpp_amd import SimulationBatch, configs
    from mgkmc_cpp_amd.serving import request_simulation
    shape = tuple(int(x) for x in a.shape.split("x"))
    reqs = [r for r in configs.c3_requests(256, events=2000) if r.shape == shape]
    if not reqs:
        sys.exit(f"no c3 trajectory has the grid {a.shape}")
    print(f"[bench] c3-shape {a.shape}: {len(reqs)} trajectories", file=sys.stderr, flush=True)
    sims = [request_simulation(r) for r in reqs]
    for s in sims:
        s.elastic_run()
    one = sims[0]
    one.event_steps(5)
    t0 = time.perf_counter()
    st = one.event_steps(100)
    single = st.kmc_events / (time.perf_counter() - t0)
    batch = SimulationBatch(sims)
    batch.event_steps(8)
    n = rounds = fp64 = fp32 = solves = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < a.seconds:
        stats = batch.event_steps(8)
        n += sum(int(x.kmc_events) for x in stats)
        fp64 += sum(int(x.ls_iters) for x in stats)
        fp32 += sum(int(x.ls_iters_fp32) for x in stats)
        solves += sum(int(x.ls_solves) for x in stats)
        rounds += 1
    dt = time.perf_counter() - t0
    batch.close()
    for s in sims:
        s.close()
    out = {"metric": "aggregate KMC events/s of one grid's c3 trajectories in one batch", "value": round(n / dt, 2),
           "unit": "events/s", "n_gpus": 1, "steps": n, "warmup": 8, "ms_per_step": round(1e3 * dt / max(n, 1), 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64-refine",
           "data": "synthetic",
           "config": {"workload": f"c3 trajectories on {a.shape} (seed 7 draw), one batch of {len(sims)}, "
                                  "8 events per step, FP32+refine, inputs drawn on the device"},
           "trajectories": len(sims), "single_trajectory_events_per_s": round(single, 2),
           "ms_per_batch_step": round(1e3 * dt / max(rounds, 1), 3),
           "ls_iters_per_event": {"fp64": round(fp64 / max(solves, 1), 2), "fp32": round(fp32 / max(solves, 1), 2)}}
    print(json.dumps(out), flush=True)


def main_many(a):
    """The batch (c3) and serving (c4) workloads in full: one JSON line."""
    from mgkmc_cpp_amd import configs
    if a.workload == "c3":
        reqs = configs.c3_requests(256, events=2000)
        summ = serve(reqs, a.budget_gb, "c3")
        metric, workload = ("aggregate KMC events/s, 256 independent trajectories",
                            "c3: 256 trajectories on grids drawn from 32^2, 64^2, 128^2, 32^3, 48^3 (seed 7), "
                            "2000 events each, FP32+refine, batched by grid")
    else:
        reqs = configs.c4_requests(256, rate=50.0, seed=11)
        summ = serve(reqs, a.budget_gb, "c4")
        metric, workload = ("p50/p99 latency of mixed requests and aggregate KMC events/s",
                            "c4: 256 requests, Poisson arrivals 50/s, grids 32^2..128^3, 500-5000 events each, "
                            "FP32+refine, admitted by HBM footprint")
    out = {"metric": metric, "value": summ["events_per_s"], "unit": "events/s", "n_gpus": 1, "steps": summ["events"],
           "warmup": 0, "ms_per_step": round(1e3 * summ["span_s"] / max(summ["events"], 1), 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64-refine",
           "data": "synthetic", "config": {"workload": workload}, "serving": summ}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
