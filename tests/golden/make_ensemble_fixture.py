"""Writes tests/golden/ensemble_16x16.npz: the oracle's ensemble of
stress-strain curves for the statistical trajectory test.

    python tests/golden/make_ensemble_fixture.py

Workload: 16 x 16 plane strain, E a correlated field (mean 85 GPa, CoV 0.15,
correlation length 3) drawn per seed, nu = 0.35, 6 fixed shears of magnitude
0.1, barriers N(1.0, 0.1) eV, T = 300 K, LS tolerance 1e-8: the c0/c1
physics on a smaller grid, in instability mode "kmc" (unstable modes are rated
events, as in the batch and serving workloads; in cascade mode some seeds
enter cascade chains of up to N rounds that take the oracle minutes).  Each trajectory is strained along xx
with sigma_yy = 0 (mixed tolerance 1 MPa) in 20 chunks of 10 strain steps of
1e-4 (to 2% strain: elastic loading, yield and flow), with a budget of 300
events per trajectory: a trajectory that spends it (or meets the instability
termination) stops, and its later points are missing (NaN); the same rule
applies to the ensembles compared with this one. (run_simulation, kmc_solver.hpp:721-1037, called once per chunk on the
running state); after each chunk eps_xx, sigma_xx (the do_logging quantities,
:606-719) and the event count are recorded.

Stored: the reference ensemble (FP64 oracle, seeds 1000..1063) and a mutant
ensemble with barriers N(0.97, 0.1) eV (seeds 3000..3063) that the test must
tell apart from the device ensemble, to show the comparison has power."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mgkmc_cpp_amd import abi, configs  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

SHAPE = (16, 16)
CHUNKS, PER, STEP = 20, 10, 1e-4
SEEDS = 64


BUDGET = 300   # events per trajectory: a few seeds enter avalanches of ~3e4 events


def run_config(max_events):
    return abi.default_run_config(n_global_steps=PER, step_size=STEP, drv_i=0, drv_j=0, mixed_tol=1e6,
                                  targets={(1, 1): 0.0}, max_events=max_events)


def curve(sim):
    """(eps_xx, sigma_xx, events so far) after each chunk; a trajectory that
    spends its event budget inside a chunk stops there (NaN from that chunk on).
    sim: an OracleSim or a device Simulation (the same run/macro surface)."""
    out = np.full((CHUNKS, 3), np.nan)
    ev = 0
    for k in range(CHUNKS):
        st = sim.run(run_config(BUDGET - ev))
        ev += st.kmc_events
        if st.event_cap_reached or st.terminated:
            break
        e, s, _ = sim.macro()
        out[k] = (e[0, 0], s[0, 0], ev)
    return out


def one(args):
    seed, kw = args
    cfg = configs.sim_config(SHAPE, seed=seed, instability_kmc=1, **kw)
    E, nu = configs.material(SHAPE, 3.0, seed)
    return curve(po.OracleSim(cfg, E, nu))


def ensemble(seed0, **kw):
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(min(6, os.cpu_count() or 1)) as pool:
        return np.array(pool.map(one, [(seed0 + i, kw) for i in range(SEEDS)]))


def main():
    t0 = time.time()
    ref = ensemble(1000)
    mut = ensemble(3000, b_mean=0.97)
    out = os.path.join(HERE, "ensemble_16x16.npz")
    np.savez_compressed(out, ref=ref, mutant=mut, chunks=CHUNKS, per=PER, step=STEP)
    print(f"{out}: {SEEDS} + {SEEDS} trajectories, mean events {ref[:, -1, 2].mean():.1f} / "
          f"{mut[:, -1, 2].mean():.1f}, {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
