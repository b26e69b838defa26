"""Statistical parity of whole trajectories (the device path against the
reference's algorithm in distribution, where it does not follow the
reference's random stream).

The oracle's ensemble (tests/golden/ensemble_16x16.npz, written by
tests/golden/make_ensemble_fixture.py) holds 64 FP64 trajectories of a 16x16
heterogeneous grid strained along xx with sigma_yy = 0 to 2% strain, the
stress-strain pairs that do_logging records (kmc_solver.hpp:606-719) after
each of 20 chunks of 10 strain steps, and the event counts.  The device runs
64 FP32+refine trajectories of the same workload whose E fields, barriers and
nothing else are drawn on the device (Philox4x32-10 instead of std::mt19937:
the same distributions, another stream), under the same 300-event budget.

Bar, at every chunk with at least 20 finished trajectories on both sides:
the ensemble means of sigma_xx and of the event count agree within 4 combined
standard errors, and so do the fractions of trajectories that spent their
budget.  The fixture's mutant ensemble (barrier mean 0.97 instead of 1.0 eV)
must fall outside the band at some chunk: the comparison has the power to see
a 3% shift of the barrier distribution."""
import os
import sys

import numpy as np
import pytest

from mgkmc_cpp_amd import Simulation, abi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_ensemble_fixture as ens  # noqa: E402

pytestmark = pytest.mark.gpu
Z = 4.0


def _zmax(a, b):
    """Largest |difference of means| / combined standard error over the chunks
    (sigma_xx and event counts), and that of the budget-spent fractions."""
    zs = []
    for k in range(a.shape[1]):
        for col in (1, 2):
            x, y = a[:, k, col], b[:, k, col]
            x, y = x[~np.isnan(x)], y[~np.isnan(y)]
            if len(x) < 20 or len(y) < 20:
                continue
            se = np.sqrt(x.var(ddof=1) / len(x) + y.var(ddof=1) / len(y))
            d = abs(x.mean() - y.mean())
            zs.append(d / se if se > 0 else (0.0 if d == 0 else np.inf))
    pa, pb = np.isnan(a[:, -1, 1]).mean(), np.isnan(b[:, -1, 1]).mean()
    p = 0.5 * (pa + pb)
    se = np.sqrt(max(p * (1 - p), 1e-12) * (1.0 / len(a) + 1.0 / len(b)))
    return max(zs), abs(pa - pb) / se


def test_device_ensemble_matches_oracle_stress_strain_curves():
    fx = np.load(os.path.join(ROOT, "tests", "golden", "ensemble_16x16.npz"))
    ref, mut = fx["ref"], fx["mutant"]
    dev = []
    for i in range(ref.shape[0]):
        seed = 5000 + i
        cfg = configs.sim_config(ens.SHAPE, precision=abi.MGK_FP32_REFINE, seed=seed, instability_kmc=1,
                                 device_draws=1)
        E, nu = configs.material(ens.SHAPE, 3.0, seed, device=True)
        sim = Simulation(cfg, E, nu)
        dev.append(ens.curve(sim))
        sim.close()
        if i % 16 == 15:
            print(f"[ensemble] {i + 1} device trajectories", flush=True)
    dev = np.array(dev)
    z, zp = _zmax(dev, ref)
    zm, zpm = _zmax(dev, mut)
    print(f"device vs oracle: max z {z:.2f} (spent budgets {zp:.2f}); device vs mutant: max z {zm:.2f}")
    assert z < Z and zp < Z, (z, zp)
    assert max(zm, zpm) > Z, (zm, zpm)
