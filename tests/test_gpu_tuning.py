"""The autotuner: the first context of a grid times the block-shape variants
of its passes (and, on 64^3 FP32+refine, the persistent CG cycle against the
launched plane passes), later contexts reuse the choice, and no variant
changes a result: solves with every variant forced equal each other bit for
bit (a line's or a column's arithmetic does not depend on the block shape)."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Context, abi
from mgkmc_cpp_amd.solver import set_tuning, tuning

pytestmark = pytest.mark.gpu


def _field(shape, seed, cov=0.15, mean=85e9):
    rng = np.random.default_rng(seed)
    N = int(np.prod(shape))
    return mean * (1.0 + cov * rng.standard_normal(N)), np.full(N, 0.35)


# FP64 on 48x32x48: planes that are not square keep the five-pass pipeline
# (a square-plane grid runs its FP64 iterations as plane passes, which have no
# block-shape variants)
@pytest.mark.parametrize("shape,precision", [((48, 32, 48), abi.MGK_FP64), ((32, 48, 32), abi.MGK_FP32_REFINE),
                                             ((64, 64, 64), abi.MGK_FP32_REFINE)])
def test_variants_do_not_change_results(shape, precision):
    E, nu = _field(shape, 41)
    bar = np.array([[1e-3, 1e-4, 0.0], [1e-4, -3e-4, 2e-4], [0.0, 2e-4, -2e-4]])
    ctx = Context(shape, E, nu, precision=precision)
    info = tuning(ctx.handle)
    print(shape, precision, info)
    assert info["tuned"]
    if shape[1] != shape[2]:   # FP64 iterations on the five-pass pipeline: pass A's variants were timed
        assert info["ms"]["fwd_fp64"][0] > 0.0
    if shape == (64, 64, 64):
        assert info["cycle"] in (0, 1) and info["ms"]["solve_cycle"] > 0.0 and info["ms"]["solve_plane"] > 0.0
    ctx.close()
    # later contexts of the grid reuse the choice; a fresh context per
    # variant (the CG cycle lengths a context predicts depend on its history)
    out = []
    for v in (0, 1):
        c = Context(shape, E, nu, precision=precision)
        i2 = tuning(c.handle)
        assert {k: i2[k] for k in ("fwd_fp64", "fwd_fp32", "mid_fp64", "mid_fp32", "cycle")} == \
               {k: info[k] for k in ("fwd_fp64", "fwd_fp32", "mid_fp64", "mid_fp32", "cycle")}
        set_tuning(c.handle, v, v, v, v, 1 - v)
        got = c.ls_solve(bar, tol=1e-9)
        out.append((got, c.last_solve_stats()))
        c.close()
    (a, sa), (b, sb) = out
    assert sa == sb
    for k in ("eps", "sig", "eps_macro", "sig_macro"):
        assert np.array_equal(a[k], b[k]), k
