"""Persistent CG cycles (one launch per refinement cycle whose blocks are all
resident, 64^3 planes) against the separately launched plane passes they
replace: the same iterates bit for bit, in a cold LS solve with several
refinement cycles and in a KMC event trajectory; the in-situ cycle timer
counts what the solves ran; contexts that run cycles in several threads, or
beside a batch on another stream, give their sequential results."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Context, Simulation, abi, configs
from mgkmc_cpp_amd.solver import cycle_supported, cycle_timer, cycle_timer_read, set_tuning

pytestmark = pytest.mark.gpu

SHAPE = (64, 64, 64)


def _field(shape, seed, cov=0.15, mean=85e9):
    rng = np.random.default_rng(seed)
    N = int(np.prod(shape))
    return mean * (1.0 + cov * rng.standard_normal(N)), np.full(N, 0.35)


@pytest.mark.parametrize("tol", [1e-8, 1e-10])
def test_cycle_solve_equals_plane_passes(tol, monkeypatch):
    E, nu = _field(SHAPE, 31)
    bar = np.array([[1e-3, 2e-4, 0.0], [2e-4, -4e-4, 1e-4], [0.0, 1e-4, 3e-4]])
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("MGK_LS_CYCLE", mode)
        ctx = Context(SHAPE, E, nu, precision=abi.MGK_FP32_REFINE)
        set_tuning(ctx.handle, cycle=1)   # whatever the autotuner chose for the grid
        assert cycle_supported(ctx.handle) == (mode == "1")
        got = ctx.ls_solve(bar, tol=tol)
        out[mode] = (got, ctx.last_solve_stats(), ctx.last_solve_history())
        print(mode, "iters", got["iters"], out[mode][1])
        ctx.close()
    (a, sa, ha), (b, sb, hb) = out["1"], out["0"]
    assert sa == sb and sa[0] >= 2          # a cold solve runs more than one refinement cycle
    assert np.array_equal(ha, hb)
    for k in ("eps", "sig", "eps_macro", "sig_macro"):
        assert np.array_equal(a[k], b[k]), k


def test_cycle_events_equal_plane_passes(monkeypatch):
    E, nu = configs.material(SHAPE, 4, 42)
    cfg = configs.sim_config(SHAPE, precision=abi.MGK_FP32_REFINE, seed=42)
    res = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("MGK_LS_CYCLE", mode)
        sim = Simulation(cfg, E, nu)
        set_tuning(sim.device_context_handle(), cycle=1)
        e = np.zeros((3, 3))
        e[0, 0] = 0.01
        sim.set_macro(e, 0.0)
        sim.elastic_run()
        st = sim.event_steps(12)
        ev = [(x.flat_idx, x.time, x.ls_iters) for x in sim.events()]
        fields = [sim.get_field(w) for w in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_Q)]
        res[mode] = (st.kmc_events, st.ls_iters, st.ls_iters_fp32, ev, fields, sim.macro())
        sim.close()
    a, b = res["1"], res["0"]
    assert a[0] == b[0] == 12 and a[1:4] == b[1:4]
    for fa, fb in zip(a[4], b[4]):
        assert np.array_equal(fa, fb)
    for ma, mb in zip(a[5], b[5]):
        assert np.array_equal(np.asarray(ma), np.asarray(mb))


def test_cycle_timer_counts_the_cycles():
    E, nu = _field(SHAPE, 37)
    ctx = Context(SHAPE, E, nu, precision=abi.MGK_FP32_REFINE)
    set_tuning(ctx.handle, cycle=1)
    assert cycle_supported(ctx.handle)
    bar = np.diag([1e-3, -3e-4, -3e-4])
    ctx.ls_solve(bar, tol=1e-8)
    cycle_timer(ctx.handle, True)
    steps = 0
    for k in range(4):
        bar[0, 0] += 1e-6
        ctx.ls_solve(bar, tol=1e-8, eps=ctx.get_field(abi.FIELD_EPS))
        steps += ctx.last_solve_stats()[1]
    ms, n, st = cycle_timer_read(ctx.handle)
    print("timed", n, "cycles,", st, "steps,", ms, "ms")
    assert 1 <= n and st <= steps and st >= n
    assert 0.0 < ms / n < 50.0
    cycle_timer(ctx.handle, False)
    assert cycle_timer_read(ctx.handle) == (0.0, 0, 0)
    ctx.close()


def _run_events(sim, n):
    e = np.zeros((3, 3))
    e[0, 0] = 0.01
    sim.set_macro(e, 0.0)
    sim.elastic_run()
    sim.event_steps(n)
    return ([(x.flat_idx, x.time, x.ls_iters) for x in sim.events()],
            [sim.get_field(w) for w in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC)])


def _same(a, b):
    assert a[0] == b[0]
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("beside", ["cycle", "batch"])
def test_cycles_beside_other_work_equal_sequential_runs(beside):
    # One persistent cycle is in flight per process: a second context's cycle
    # waits for it instead of competing for the same CU slots.  Work of other
    # streams (a batch of 48^3 trajectories) only delays a cycle's blocks.
    import threading
    from mgkmc_cpp_amd import SimulationBatch
    E, nu = configs.material(SHAPE, 4, 42)
    cfg = configs.sim_config(SHAPE, precision=abi.MGK_FP32_REFINE, seed=42)
    ref = Simulation(cfg, E, nu)
    set_tuning(ref.device_context_handle(), cycle=1)
    assert cycle_supported(ref.device_context_handle()) == 1
    want = _run_events(ref, 8)
    ref.close()
    if beside == "cycle":
        E2, nu2 = configs.material(SHAPE, 4, 7)
        cfg2 = configs.sim_config(SHAPE, precision=abi.MGK_FP32_REFINE, seed=7)
        ref2 = Simulation(cfg2, E2, nu2)
        set_tuning(ref2.device_context_handle(), cycle=1)
        want2 = _run_events(ref2, 8)
        ref2.close()
    else:
        shape2 = (48, 48, 48)
        cfgs = [configs.sim_config(shape2, precision=abi.MGK_FP32_REFINE, seed=s) for s in (3, 4)]
        mats = [configs.material(shape2, 4, s) for s in (3, 4)]
        alone = [Simulation(c, *m) for c, m in zip(cfgs, mats)]
        for a in alone:
            a.elastic_run()
            a.event_steps(6)
        want2 = [([x.flat_idx for x in a.events()], a.get_field(abi.FIELD_SIG)) for a in alone]
        for a in alone:
            a.close()
    sim = Simulation(cfg, E, nu)
    set_tuning(sim.device_context_handle(), cycle=1)
    got, errs = {}, []

    def first():
        try:
            got[1] = _run_events(sim, 8)
        except BaseException as e:   # reported below
            errs.append(e)

    def second():
        try:
            if beside == "cycle":
                sim2 = Simulation(cfg2, E2, nu2)
                set_tuning(sim2.device_context_handle(), cycle=1)
                got[2] = _run_events(sim2, 8)
                sim2.close()
            else:
                sims = [Simulation(c, *m) for c, m in zip(cfgs, mats)]
                for x in sims:
                    x.elastic_run()
                batch = SimulationBatch(sims)
                batch.event_steps(6)
                got[2] = [([x.flat_idx for x in y.events()], y.get_field(abi.FIELD_SIG)) for y in sims]
                batch.close()
                for x in sims:
                    x.close()
        except BaseException as e:   # reported below
            errs.append(e)

    ts = [threading.Thread(target=first), threading.Thread(target=second)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    sim.close()
    assert not errs, errs
    _same(got[1], want)
    if beside == "cycle":
        _same(got[2], want2)
    else:
        for (gi, gs), (wi, ws) in zip(got[2], want2):
            assert gi == wi and np.array_equal(gs, ws)
