"""Writes tests/golden/c1_5000_events.npz: the oracle's run of workload c1
(32 x 32 heterogeneous E, 6 fixed shears, T = 300 K, LS tol 1e-8, uniaxial
tension in 1e-4 strain steps with sigma_yy = 0) to its 5000-event cap.

    python tests/golden/make_c1_fixture.py

The oracle (oracle/liboracle.so, the CPU restatement of kmc_solver.hpp and
linear_elastic_solver.hpp) takes minutes for this run, because later strain
steps enter long cascade chains; the GPU test compares the device driver with
this file instead of rerunning it.  Stored: the run counters, every event's
flat mode index, LS iteration count, time and instability flag, the macro
stress after each event, and the final stress field and macros."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mgkmc_cpp_amd import abi, configs  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

STEPS = 100000


def run_config(events):
    return abi.default_run_config(n_global_steps=STEPS, step_size=1e-4, drv_i=0, drv_j=0, max_events=events,
                                  mixed_tol=1e6, targets={(1, 1): 0.0})


def main():
    cfg, (E, nu), events = configs.c1()
    t0 = time.time()
    osim = po.OracleSim(cfg, E, nu)
    st = osim.run(run_config(events))
    ev = osim.events()
    eps_m, sig_m, t_end = osim.macro()
    out = os.path.join(HERE, "c1_5000_events.npz")
    np.savez_compressed(
        out,
        counters=np.array([st.kmc_events, st.cascades, st.cascade_flips, st.elastic_steps, st.event_cap_reached,
                           st.terminated], dtype=np.int64),
        flat_idx=np.array([e.flat_idx for e in ev], dtype=np.int32),
        ls_iters=np.array([e.ls_iters for e in ev], dtype=np.int32),
        is_instab=np.array([e.is_instab for e in ev], dtype=np.int8),
        time=np.array([e.time for e in ev]),
        sig_macro=np.array([list(e.sig_macro) for e in ev]),
        sig=osim.get_field(abi.FIELD_SIG),
        eps_macro_end=eps_m, sig_macro_end=sig_m, time_end=np.array(t_end),
    )
    print(f"{out}: {st.kmc_events} events, {st.cascades} cascades ({st.cascade_flips} flips), "
          f"{st.elastic_steps} strain steps, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
