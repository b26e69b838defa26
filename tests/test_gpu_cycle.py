"""Persistent CG cycles (one cooperative launch per refinement cycle, 64^3
planes) against the separately launched plane passes they replace: the same
iterates bit for bit, in a cold LS solve with several refinement cycles and in
a KMC event trajectory; the in-situ cycle timer counts what the solves ran."""
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
