"""The reference-scale CPU workloads through the whole device driver against
the oracle: run_simulation (kmc_solver.hpp:721-1037) with strain steps of
1e-4 along xx and sigma_yy = 0 (mixed tolerance 1 MPa, main.cpp:280).

c0: 32 x 32, homogeneous E = 85 GPa, fixed catalog, 2000 events (the cap is
    reached near step 197, after ~1660 cascades; the oracle takes ~6 s).
c1: 32 x 32, E a correlated field (CoV 0.15, correlation length 3), the
    first 775 events (320 strain steps, no cascade yet; the oracle takes
    ~10 s).  From about the 780th event on, every attempt runs a cascade to
    its N-round limit (nx*ny rounds of an LS solve each, kmc_solver.hpp:467,
    :556): the oracle then needs hours for the config's 5000 events.

Bar: the same counters and event sequence (voxel and mode of every flip,
LS iteration counts), event times within 1e-12 relative, final stress field
and macro stress within 1e-9 relative."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Simulation, abi, configs
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name,steps,cap", [("c0", 500, None), ("c1", 100000, 775)])
def test_config_run_matches_oracle(name, steps, cap):
    cfg, (E, nu), events = getattr(configs, name)()
    run = abi.default_run_config(n_global_steps=steps, step_size=1e-4, drv_i=0, drv_j=0,
                                 max_events=cap or events, mixed_tol=1e6, targets={(1, 1): 0.0})
    osim = po.OracleSim(cfg, E, nu)
    so = osim.run(run)
    dsim = Simulation(cfg, E, nu)
    sd = dsim.run(run)
    print(name, "oracle", so.kmc_events, so.cascades, so.cascade_flips, so.elastic_steps,
          "device", sd.kmc_events, sd.cascades, sd.cascade_flips, sd.elastic_steps)
    assert (sd.kmc_events, sd.cascades, sd.cascade_flips, sd.elastic_steps) == \
           (so.kmc_events, so.cascades, so.cascade_flips, so.elastic_steps)
    eo, ed = osim.events(), dsim.events()
    assert len(eo) == len(ed) and len(eo) > 100
    assert [e.flat_idx for e in ed] == [e.flat_idx for e in eo]
    assert [e.ls_iters for e in ed] == [e.ls_iters for e in eo]
    for a, b in zip(ed, eo):
        assert abs(a.time - b.time) <= 1e-12 * abs(b.time)
    assert _rel(dsim.get_field(abi.FIELD_SIG), osim.get_field(abi.FIELD_SIG)) < 1e-9
    _, s_o, t_o = osim.macro()
    _, s_d, t_d = dsim.macro()
    assert _rel(s_d, s_o) < 1e-9 and abs(t_d - t_o) <= 1e-12 * abs(t_o)
    dsim.close()
