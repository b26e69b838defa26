"""The device KMC driver on the reference's DAMASK comparison case: one
strain step (elastic solve plus the mixed-BC loop, kmc_solver.hpp:1735-2012)
on the 32 x 32 x 1 grid, against DAMASK's stress field and against the
oracle (same strain step: 1e-9 relative in FP64, 1e-6 in FP32+refine)."""
import numpy as np
import pytest

import damask_case as dc
from mgkmc_cpp_amd import Simulation, abi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,bound", [(abi.MGK_FP64, 1e-9), (abi.MGK_FP32_REFINE, 1e-6)])
def test_device_matches_damask_and_oracle(precision, bound):
    case = dc.load()
    osim = po.OracleSim(dc.sim_config(), case["E"], case["nu"])
    osim.run(dc.run_config(case))
    dsim = Simulation(dc.sim_config(precision), case["E"], case["nu"])
    st = dsim.run(dc.run_config(case))
    assert (st.kmc_events, st.cascades, st.elastic_steps) == (0, 0, 1)
    sig = dsim.get_field(abi.FIELD_SIG)
    d = dc.assert_matches_damask(sig, case)
    ref = osim.get_field(abi.FIELD_SIG)
    rel = np.abs(sig - ref).max() / np.abs(ref).max()
    print(precision, d, "vs oracle", rel)
    assert rel < bound
    e_d, s_d, _ = dsim.macro()
    e_o, s_o, _ = osim.macro()
    assert np.abs(e_d - e_o).max() < bound * np.abs(e_o).max()
    dsim.close()
