"""Batched and served trajectories on the device.

* the device bytes a simulation holds equal the admission estimate;
* requests served concurrently (batched by grid, admitted by footprint,
  shapes in their own worker threads) produce exactly the trajectory each
  request gives when run alone (FP64: bit for bit), and the footprint budget
  is never exceeded."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Simulation, abi, configs
from mgkmc_cpp_amd.serving import Request, Server, request_footprint, request_simulation, summary
from mgkmc_cpp_amd.solver import estimate_bytes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,M", [((16, 16), 6), ((8, 8, 8), 12)])
def test_device_bytes_match_estimate(shape, M):
    E, nu = configs.material(shape, 2.0, 3)
    sim = Simulation(configs.sim_config(shape, seed=3), E, nu)                 # fixed catalog: one shared copy
    assert sim.device_bytes() == estimate_bytes(shape, M, per_voxel_catalog=False)
    sim.close()
    sim = Simulation(configs.sim_config(shape, seed=3, catalog_mode=abi.CATALOG_RANDOM), E, nu)
    assert sim.device_bytes() == estimate_bytes(shape, M, per_voxel_catalog=True)
    sim.close()


def _alone(req):
    """The request's trajectory run by itself, attempt by attempt."""
    sim = request_simulation(req)
    sim.elastic_run()
    made = 0
    while made < req.events:
        st = sim.event_steps(1)
        made += st.kmc_events
        if st.kmc_events == 0 and st.cascades == 0:
            break
    out = (made, sim.get_field(abi.FIELD_SIG), sim.get_field(abi.FIELD_EPS_PLASTIC),
           [e.flat_idx for e in sim.events()])
    sim.close()
    return out


@pytest.mark.parametrize("device_inputs", [False, True])
def test_served_requests_equal_standalone_runs(device_inputs):
    shapes = [(16, 16), (8, 8, 8), (16, 16), (32, 32), (8, 8, 8), (16, 16)]
    reqs = [Request(rid=i, shape=s, events=12 + 3 * i, seed=50 + i, arrival=0.02 * i, precision=abi.MGK_FP64,
                    device_inputs=device_inputs) for i, s in enumerate(shapes)]
    got = {}

    def keep(req, sim):
        got[req.rid] = (req.events_done, sim.get_field(abi.FIELD_SIG), sim.get_field(abi.FIELD_EPS_PLASTIC),
                        [e.flat_idx for e in sim.events()])

    # room for the largest request and one 16x16 at a time
    budget = max(request_footprint(r) for r in reqs) + request_footprint(reqs[0])
    srv = Server(budget, step_events=4, on_complete=keep)
    srv.run(reqs)
    assert srv.peak_used <= budget
    s = summary(reqs)
    assert s["completed"] == len(reqs) and s["p99_latency_s"] >= s["p50_latency_s"] > 0
    for r in reqs:
        assert r.finished >= r.started >= r.admitted >= r.arrival - 1e-3
        made, sig, epsp, idx = _alone(r)
        g_made, g_sig, g_epsp, g_idx = got[r.rid]
        assert g_made == made == r.events_done
        assert g_idx == idx
        assert np.array_equal(g_sig, sig) and np.array_equal(g_epsp, epsp)


def test_fp32_refine_serving_reuses_parked_rows_and_equals_standalone_runs():
    # A long request keeps its shape's batch running while short ones arrive
    # one after another: each short request's row is parked when it finishes
    # and taken by the next (the batch and its graphs stay), and the set-up
    # solves run beside the batch.  Every request's trajectory equals its
    # standalone FP32+refine run bit for bit.
    specs = [((16, 16), 300)] + [((16, 16), 12 + 2 * i) for i in range(6)] + [((8, 8, 8), 10), ((8, 8, 8), 14)]
    reqs = [Request(rid=i, shape=s, events=n, seed=80 + i, arrival=0.0 if i == 0 else 0.05 + 0.06 * i,
                    precision=abi.MGK_FP32_REFINE, instability_kmc=1) for i, (s, n) in enumerate(specs)]
    got = {}

    def keep(req, sim):
        got[req.rid] = (req.events_done, sim.get_field(abi.FIELD_SIG), sim.get_field(abi.FIELD_EPS_PLASTIC),
                        [e.flat_idx for e in sim.events()])

    budget = sum(request_footprint(r) for r in reqs)
    srv = Server(budget, step_events=4, on_complete=keep)
    srv.run(reqs)
    print("serving stats", srv.stats(), summary(reqs))
    assert srv.stats()["reuses"] >= 1
    assert srv.peak_used <= budget and srv.used == 0
    for r in reqs:
        made, sig, epsp, idx = _alone(r)
        g_made, g_sig, g_epsp, g_idx = got[r.rid]
        assert g_made == made == r.events_done
        assert g_idx == idx
        assert np.array_equal(g_sig, sig) and np.array_equal(g_epsp, epsp)
