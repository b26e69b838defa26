"""The oracle against the reference's DAMASK comparison (known answer for the
elastic solve plus the mixed boundary conditions, run through the KMC
driver's strain step: kmc_solver.hpp:1735-2012) and against the closed
relation of thin 3D grids to 2D plane strain.  CPU only."""
import numpy as np
import pytest

import damask_case as dc
from mgkmc_cpp_amd import abi
from oracle import pyoracle as po


def test_fixture_is_consistent():
    case = dc.load()
    assert case["E"].shape == (1024,) and abs(case["E"].mean() / 70e9 - 1) < 1e-12
    # DAMASK's converged mixed BC: P_yy = P_zz = 0 on average, F_xx = 1.001
    assert abs(case["F"][:, 0, 0].mean() - 1.001) < 1e-7
    assert abs(case["P"][:, 1, 1].mean()) < 1e-5 * abs(case["P"][:, 0, 0].mean())
    # S = F^-1 P is symmetric and sigma is its push-forward
    S = np.linalg.solve(case["F"], case["P"])
    assert np.abs(S - S.transpose(0, 2, 1)).max() < 1e-6 * np.abs(S).max()
    assert np.abs(dc.push_forward(case["S"], case["F"]) - case["sigma"]).max() < 1e-5 * np.abs(case["sigma"]).max()
    assert abs(case["sigma"][:, 0, 0].mean() / dc.MEAN_SIGMA_XX - 1) < 1e-6


def test_oracle_matches_damask():
    case = dc.load()
    sim = po.OracleSim(dc.sim_config(), case["E"], case["nu"])
    st = sim.run(dc.run_config(case))
    assert (st.kmc_events, st.cascades, st.elastic_steps) == (0, 0, 1)
    eps_macro, sig_macro, _ = sim.macro()
    assert abs(eps_macro[0, 0] - case["green_xx"]) < 1e-15
    assert abs(sig_macro[1, 1]) < 1.0 and abs(sig_macro[2, 2]) < 1.0
    d = dc.assert_matches_damask(sim.get_field(abi.FIELD_SIG), case)
    print(d)


def test_mixed_bc_tolerance_matters():
    # the yaml's 1 MPa stress tolerance stops the loop early: the mean then
    # misses DAMASK by more than the 1e-4 bar, so the bar is a real check
    case = dc.load()
    sim = po.OracleSim(dc.sim_config(), case["E"], case["nu"])
    sim.run(dc.run_config(case, mixed_tol=1e6))
    d = dc.compare(sim.get_field(abi.FIELD_SIG), case)
    assert d["mean_S"] > dc.MEAN_RTOL


@pytest.mark.parametrize("thin", [(32, 16, 1), (32, 1, 16), (1, 32, 16)])
def test_oracle_thin_3d_equals_plane_strain(thin):
    keep = [a for a in range(3) if thin[a] > 1]
    shape2 = tuple(thin[a] for a in keep)
    rng = np.random.default_rng(5)
    N = int(np.prod(shape2))
    E = 85e9 * (1 + 0.15 * rng.standard_normal(N))
    nu = np.full(N, 0.35)
    b2 = np.array([[1e-3, 2e-4], [2e-4, -3e-4]])
    b3 = np.zeros((3, 3))
    b3[np.ix_(keep, keep)] = b2
    r2 = po.ls_solve(shape2, E, nu, b2, tol=1e-10)
    r3 = po.ls_solve(thin, E, nu, b3, tol=1e-10)
    assert r2["iters"] == r3["iters"]
    e3 = r3["eps"][:, keep][:, :, keep]
    assert np.abs(e3 - r2["eps"]).max() <= 1e-14 * np.abs(r2["eps"]).max()
    out = [a for a in range(3) if a not in keep][0]
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    want = lam * (r2["eps"][:, 0, 0] + r2["eps"][:, 1, 1])
    assert np.abs(r3["sig"][:, out, out] - want).max() <= 1e-12 * np.abs(r2["sig"]).max()
