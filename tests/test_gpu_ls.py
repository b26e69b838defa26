"""Lippmann-Schwinger solve on the device against the CPU oracle
(spectral_solver_2d/3d, linear_elastic_solver.hpp:146-309, :431-593).

Bar: same iteration count and stress within 1e-10 relative (FP64 path)."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Context, abi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _field(shape, seed, cov=0.15, mean=85e9):
    rng = np.random.default_rng(seed)
    N = int(np.prod(shape))
    return mean * (1.0 + cov * rng.standard_normal(N)), np.full(N, 0.35)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _bar(D, seed=1, sym=True):
    rng = np.random.default_rng(seed)
    b = 1e-3 * rng.standard_normal((D, D))
    return 0.5 * (b + b.T) if sym else b


CASES = [(8, 8), (16, 16), (32, 32), (48, 32), (64, 64), (128, 128), (8, 8, 8), (16, 16, 16), (32, 32, 32),
         (16, 48, 32), (64, 64, 64)]


@pytest.mark.parametrize("shape", CASES)
def test_cold_solve_matches_oracle(shape):
    D = len(shape)
    E, nu = _field(shape, 3)
    bar = _bar(D)
    tol = 1e-8
    ref = po.ls_solve(shape, E, nu, bar, tol=tol)
    ctx = Context(shape, E, nu)
    got = ctx.ls_solve(bar, tol=tol)
    print(shape, "iters", ref["iters"], got["iters"], "diff", ref["diff"], got["diff"])
    assert got["iters"] == ref["iters"]
    assert _rel(got["sig"], ref["sig"]) < RTOL
    assert _rel(got["eps"], ref["eps"]) < RTOL
    assert _rel(got["sig_macro"], ref["sig_macro"]) < RTOL


@pytest.mark.parametrize("shape,plane", [((32, 32), "plane_stress"), ((16, 32), "plane_strain"),
                                         ((16, 16, 16), "plane_strain")])
def test_warm_start_with_plastic_strain(shape, plane):
    D = len(shape)
    N = int(np.prod(shape))
    E, nu = _field(shape, 5)
    rng = np.random.default_rng(7)
    ep = np.zeros((N, D, D))
    for v in rng.choice(N, size=max(3, N // 50), replace=False):
        a = 0.05 * rng.standard_normal((D, D))
        a = 0.5 * (a + a.T)
        a -= np.trace(a) / D * np.eye(D)
        ep[v] = a
    bar0 = _bar(D, 11)
    ps = plane == "plane_stress"
    first = po.ls_solve(shape, E, nu, bar0, tol=1e-8, plane_stress=ps)
    bar1 = bar0 + np.diag([1e-4] + [0.0] * (D - 1))
    ref = po.ls_solve(shape, E, nu, bar1, eps=first["eps"], eps_plastic=ep, tol=1e-8, plane_stress=ps)
    ctx = Context(shape, E, nu, plane_mode=plane)
    got = ctx.ls_solve(bar1, eps=first["eps"], eps_plastic=ep, tol=1e-8)
    assert got["iters"] == ref["iters"]
    assert _rel(got["sig"], ref["sig"]) < RTOL
    assert _rel(got["eps_macro"], ref["eps_macro"]) < RTOL


def test_antisymmetric_imposed_strain():
    # driving component "xy" makes eps_macro non-symmetric (main.cpp:336, kmc_solver.hpp:752-753)
    shape = (16, 16)
    E, nu = _field(shape, 9)
    bar = np.array([[2e-4, 1e-3], [0.0, -1e-4]])
    ref = po.ls_solve(shape, E, nu, bar, tol=1e-8)
    got = Context(shape, E, nu).ls_solve(bar, tol=1e-8)
    assert got["iters"] == ref["iters"]
    assert _rel(got["sig"], ref["sig"]) < RTOL
    assert _rel(got["eps"], ref["eps"]) < RTOL


def test_iteration_cap_and_zero_iterations():
    shape = (16, 16)
    E, nu = _field(shape, 13, cov=0.3)
    bar = _bar(2)
    for max_iter in (0, 1, 3):
        ref = po.ls_solve(shape, E, nu, bar, tol=1e-14, max_iter=max_iter)
        got = Context(shape, E, nu).ls_solve(bar, tol=1e-14, max_iter=max_iter)
        assert got["iters"] == ref["iters"] == max_iter
        assert _rel(got["sig"], ref["sig"]) < RTOL


def test_fp32_refine_error_bound():
    shape = (32, 32, 32)
    E, nu = _field(shape, 21)
    bar = _bar(3)
    ref = po.ls_solve(shape, E, nu, bar, tol=1e-8)
    ctx = Context(shape, E, nu, precision=abi.MGK_FP32_REFINE)
    got = ctx.ls_solve(bar, tol=1e-8)
    err = _rel(got["sig"], ref["sig"])
    print("fp32+refine iters", got["iters"], ctx.last_solve_stats(), "rel err", err)
    assert err < 1e-6
