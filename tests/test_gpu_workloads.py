"""The batch (c3) and serving (c4) workloads at their real grid sizes.

c3: FP32+refine trajectories batched on 48^3 and 128^2 grids (two of c3's five
    grid shapes, the largest 3D and 2D ones) equal the same trajectories run
    one by one, bit for bit, and stay within the FP32+refine bound of the FP64
    oracle: stress within 1e-6 of max |sigma| of the converged solution of the
    same event sequence, or within twice the FP64 path's own error at tol 1e-8
    where that is larger.
c4: a served mix of FP32+refine requests that includes a 64^3 and a 128^3
    request (the largest c4 grids) gives each request the trajectory it gives
    alone, bit for bit, within a footprint budget."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Simulation, SimulationBatch, abi, configs
from mgkmc_cpp_amd.serving import Request, Server, request_footprint, request_simulation, summary
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("shape,B,events", [((48, 48, 48), 2, 6), ((128, 128), 3, 10)])
def test_c3_batches_equal_separate_runs_and_oracle(shape, B, events):
    corr = 4.0 if len(shape) == 3 else 3.0
    seeds = [1000 + i for i in range(B)]
    cfgs = [configs.sim_config(shape, precision=abi.MGK_FP32_REFINE, seed=s, instability_kmc=1) for s in seeds]
    mats = [configs.material(shape, corr, s) for s in seeds]
    single = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    batched = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    pre = np.zeros((len(shape), len(shape)))
    pre[0, 0] = 0.01
    for s in single + batched:
        s.set_macro(pre, 0.0)
        s.elastic_run()
    st_single = [s.event_steps(events) for s in single]
    batch = SimulationBatch(batched)
    st_batch = batch.event_steps(events)
    for a, b, sa, sb in zip(single, batched, st_single, st_batch):
        assert sa.kmc_events == sb.kmc_events == events
        assert (sa.ls_iters, sa.ls_iters_fp32) == (sb.ls_iters, sb.ls_iters_fp32)
        assert [e.flat_idx for e in a.events()] == [e.flat_idx for e in b.events()]
        assert [e.time for e in a.events()] == [e.time for e in b.events()]
        for f in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_Q, abi.FIELD_Q0):
            assert np.array_equal(a.get_field(f), b.get_field(f)), f
    batch.close()
    # the first trajectory against the FP64 oracle (tol 1e-8) and the
    # converged oracle solution of the same events (tol 1e-12)
    runs = {}
    for tol in (1e-8, 1e-12):
        c = configs.sim_config(shape, precision=abi.MGK_FP64, seed=seeds[0], instability_kmc=1, ls_tol=tol)
        o = po.OracleSim(c, *mats[0])
        o.set_macro(pre, 0.0)
        o.elastic_run()
        o.event_steps(events)
        runs[tol] = o
    flats = [e.flat_idx for e in batched[0].events()]
    assert [e.flat_idx for e in runs[1e-8].events()] == flats == [e.flat_idx for e in runs[1e-12].events()]
    conv = runs[1e-12].get_field(abi.FIELD_SIG)
    err32 = _rel(batched[0].get_field(abi.FIELD_SIG), conv)
    err64 = _rel(runs[1e-8].get_field(abi.FIELD_SIG), conv)
    print(shape, "FP32+refine", err32, "FP64", err64)
    assert err32 < max(1e-6, 2.0 * err64), (err32, err64)
    for s in single + batched:
        s.close()


def _alone(req):
    sim = request_simulation(req)
    sim.elastic_run()
    made = 0
    while made < req.events:
        st = sim.event_steps(1)
        made += st.kmc_events
        if st.kmc_events == 0 and st.cascades == 0:
            break
    out = (made, sim.get_field(abi.FIELD_SIG), sim.get_field(abi.FIELD_EPS_PLASTIC),
           [e.flat_idx for e in sim.events()], [e.time for e in sim.events()])
    sim.close()
    return out


def test_c4_mix_with_largest_grids_equals_standalone_runs():
    shapes = [(64, 64, 64), (32, 32), (128, 128, 128), (48, 48, 48), (128, 128), (32, 32), (64, 64, 64)]
    evs = [24, 40, 12, 20, 30, 16, 18]
    reqs = [Request(rid=i, shape=s, events=evs[i], seed=70 + i, arrival=0.01 * i, precision=abi.MGK_FP32_REFINE,
                    instability_kmc=1, device_inputs=True) for i, s in enumerate(shapes)]
    got = {}

    def keep(req, sim):
        got[req.rid] = (req.events_done, sim.get_field(abi.FIELD_SIG), sim.get_field(abi.FIELD_EPS_PLASTIC),
                        [e.flat_idx for e in sim.events()], [e.time for e in sim.events()])

    budget = 4 * request_footprint(reqs[2])
    srv = Server(budget, step_events=4, on_complete=keep)
    srv.run(reqs)
    assert srv.peak_used <= budget
    s = summary(reqs)
    assert s["completed"] == len(reqs)
    for r in reqs:
        made, sig, epsp, idx, times = _alone(r)
        g_made, g_sig, g_epsp, g_idx, g_times = got[r.rid]
        assert g_made == made == r.events_done == r.events, (r.rid, r.shape)
        assert g_idx == idx and g_times == times, (r.rid, r.shape)
        assert np.array_equal(g_sig, sig) and np.array_equal(g_epsp, epsp), (r.rid, r.shape)
