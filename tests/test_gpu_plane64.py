"""FP64 Moulinec-Suquet iterations as plane passes: one launch per
(component, plane) for the local stress and both in-plane forward
transforms, pass C on the tiled FP64 spectrum (transform along the first
axis, Green operator, inverse), one launch per (component, plane) for both
inverse in-plane transforms and the update -- three launches per iteration
instead of the five-pass pipeline's five.  Square planes of 16 to 64 points.

Bars: against the oracle (spectral_solver_3d, linear_elastic_solver.hpp:
431-593) the FP64 bar -- the same iteration count, stress and strain within
1e-10 relative; against the five-pass pipeline (MGK_LS_PLANE64=0) the same
iteration count and iterates within 1e-12 (the transforms round in a
different order); FP32+refine solves, whose CG cycles start from the plane
passes' Cartesian increment, within the FP32+refine error bound."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Context, abi
from mgkmc_cpp_amd._lib import MgkError
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

RTOL = 1e-10
SHAPES = [(16, 16, 16), (32, 32, 32), (48, 48, 48), (64, 64, 64), (32, 48, 48), (16, 64, 64), (64, 32, 32)]


def _field(shape, seed, cov=0.15, mean=85e9):
    rng = np.random.default_rng(seed)
    N = int(np.prod(shape))
    return mean * (1.0 + cov * rng.standard_normal(N)), np.full(N, 0.35)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _bar(seed=1):
    rng = np.random.default_rng(seed)
    b = 1e-3 * rng.standard_normal((3, 3))
    return 0.5 * (b + b.T)


def _plane64_on(shape, E, nu):
    """Whether FP64 solves on this grid run the plane passes: pass C of the
    FP64 plane passes (it touches only the spectrum) is timed on a context
    of its own, which refuses it where the path is off."""
    ctx = Context(shape, E, nu)
    try:
        ctx.time_ls_pass(10, fp32=False, reps=1)
        return True
    except MgkError:
        return False
    finally:
        ctx.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_plane64_cold_solve_matches_oracle(shape):
    E, nu = _field(shape, 3)
    assert _plane64_on(shape, E, nu)
    bar = _bar()
    ref = po.ls_solve(shape, E, nu, bar, tol=1e-8)
    ctx = Context(shape, E, nu)
    got = ctx.ls_solve(bar, tol=1e-8)
    print(shape, "iters", ref["iters"], got["iters"], "diff", ref["diff"], got["diff"], flush=True)
    assert got["iters"] == ref["iters"]
    assert _rel(got["sig"], ref["sig"]) < RTOL
    assert _rel(got["eps"], ref["eps"]) < RTOL
    assert _rel(got["sig_macro"], ref["sig_macro"]) < RTOL
    ctx.close()


@pytest.mark.parametrize("shape", [(16, 16, 16), (48, 48, 48), (64, 64, 64)])
def test_plane64_equals_five_pass_pipeline(shape, monkeypatch):
    E, nu = _field(shape, 29, cov=0.3)
    bar = _bar(5)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("MGK_LS_PLANE64", mode)
        assert _plane64_on(shape, E, nu) == (mode == "1")
        ctx = Context(shape, E, nu)
        got = ctx.ls_solve(bar, tol=1e-10)
        out[mode] = (got, ctx.last_solve_history())
        print(shape, "plane64", mode, "iters", got["iters"], flush=True)
        ctx.close()
    (a, ha), (b, hb) = out["1"], out["0"]
    assert a["iters"] == b["iters"]
    assert np.allclose(ha, hb, rtol=1e-6, atol=0.0)
    for k in ("eps", "sig", "sig_macro"):
        assert _rel(a[k], b[k]) < 1e-12, k


@pytest.mark.parametrize("shape", [(16, 16, 16), (48, 48, 48)])
def test_plane64_warm_start_with_plastic_strain(shape):
    # warm start from a converged field, plastic strain on 2% of the voxels
    # and a changed macro strain: the pass A stress subtracts eps_p
    N = int(np.prod(shape))
    E, nu = _field(shape, 5)
    assert _plane64_on(shape, E, nu)
    rng = np.random.default_rng(7)
    ep = np.zeros((N, 3, 3))
    for v in rng.choice(N, size=max(3, N // 50), replace=False):
        a = 0.05 * rng.standard_normal((3, 3))
        a = 0.5 * (a + a.T)
        a -= np.trace(a) / 3 * np.eye(3)
        ep[v] = a
    bar0 = _bar(11)
    first = po.ls_solve(shape, E, nu, bar0, tol=1e-8)
    bar1 = bar0 + np.diag([1e-4, 0.0, 0.0])
    ref = po.ls_solve(shape, E, nu, bar1, eps=first["eps"], eps_plastic=ep, tol=1e-8)
    ctx = Context(shape, E, nu)
    got = ctx.ls_solve(bar1, eps=first["eps"], eps_plastic=ep, tol=1e-8)
    print(shape, "iters", ref["iters"], got["iters"], flush=True)
    assert got["iters"] == ref["iters"]
    assert _rel(got["sig"], ref["sig"]) < RTOL
    assert _rel(got["eps_macro"], ref["eps_macro"]) < RTOL
    ctx.close()


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 64, 64)])
def test_fp32_refine_on_plane64_iterations(shape, monkeypatch):
    # FP32+refine solves: the FP64 iterations on the plane passes hand the
    # CG cycle its start as a Cartesian increment; both paths stay within the
    # error bound of the converged solution (the oracle at tol 1e-11)
    E, nu = _field(shape, 23)
    bar = _bar(4)
    ref = po.ls_solve(shape, E, nu, bar, tol=1e-11)
    c64 = Context(shape, E, nu)
    e64 = _rel(c64.ls_solve(bar, tol=1e-8)["sig"], ref["sig"])
    c64.close()
    bound = max(1e-6, 2.0 * e64)
    for mode in ("1", "0"):
        monkeypatch.setenv("MGK_LS_PLANE64", mode)
        ctx = Context(shape, E, nu, precision=abi.MGK_FP32_REFINE)
        got = ctx.ls_solve(bar, tol=1e-8)
        err = _rel(got["sig"], ref["sig"])
        print(shape, "plane64", mode, "iters", got["iters"], ctx.last_solve_stats(), "rel err", err, flush=True)
        assert err < bound
        assert _rel(got["sig_macro"], ref["sig_macro"]) < bound
        ctx.close()
