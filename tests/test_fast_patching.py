"""Fast patching in the oracle (kmc_solver.hpp:276-332, :1256-1331; use
:525-547, :866-913): the restatement pinned by what the mathematics fixes.

In a homogeneous medium the Lippmann-Schwinger response to one eigenstrain is
translation invariant and the solve is exact at the first iteration, so a
window that covers the whole periodic grid (2R + 1 = n) reproduces the full
solve after a flip anywhere; a smaller window misses the far field but, with
the missing-mean shift, keeps the mean stress exact.  A wrong basis
coefficient (e.g. +zz for -zz in 3D), window index or wrap fails these.
CPU only."""
import numpy as np
import pytest

from mgkmc_cpp_amd import abi
from oracle import pyoracle as po


def _patched_vs_solved(shape, R, voxel, seed=1):
    D = len(shape)
    N = int(np.prod(shape))
    kw = dict(dims=D, nx=shape[0], ny=shape[1], M=4, fast_patching=1, patch_radius=R, ls_tol=1e-12, ls_max_iter=500)
    if D == 3:
        kw["nz"] = shape[2]
    o = po.OracleSim(abi.default_sim_config(**kw), np.full(N, 70e9), np.full(N, 0.3))
    o.elastic_run()
    rng = np.random.default_rng(seed)
    C = 1e-2 * rng.standard_normal((D, D))
    C = 0.5 * (C + C.T)
    C -= np.trace(C) / D * np.eye(D)
    ep = o.get_field(abi.FIELD_EPS_PLASTIC)
    ep[voxel] = C
    o.set_field(abi.FIELD_EPS_PLASTIC, ep)
    o.patch(voxel, C)
    patched = o.get_field(abi.FIELD_SIG)
    o.elastic_run()
    return patched, o.get_field(abi.FIELD_SIG)


@pytest.mark.parametrize("shape,R,voxel", [((9, 9), 4, 7), ((9, 9), 4, 40), ((5, 5, 5), 2, 7), ((5, 5, 5), 2, 93),
                                           ((5, 5, 1), 2, 11)])
def test_full_window_equals_solve_homogeneous(shape, R, voxel):
    p, s = _patched_vs_solved(shape, R, voxel)
    assert np.abs(p - s).max() <= 1e-13 * np.abs(s).max()


@pytest.mark.parametrize("shape,R", [((9, 9), 1), ((6, 5, 4), 1), ((8, 8), 2)])
def test_small_window_keeps_the_mean(shape, R):
    p, s = _patched_vs_solved(shape, R, 5)
    assert np.abs(p - s).max() > 1e-3 * np.abs(s).max()           # the far field is missing
    assert np.abs(p.mean(0) - s.mean(0)).max() <= 1e-14 * np.abs(s).max()


def test_patched_run_solves_every_sync_interval():
    rng = np.random.default_rng(23)
    E = 85e9 * (1 + 0.15 * rng.standard_normal(256))
    nu = np.full(256, 0.35)
    c = abi.default_sim_config(dims=2, nx=16, ny=16, M=4, b_mean=1.3, b_std=0.1, b_min_cutoff=0.01, temperature=300.0,
                               catalog_mode=abi.CATALOG_RANDOM, gamma0=0.1, ls_tol=1e-8, seed=5, instability_kmc=1,
                               fast_patching=1, patch_radius=3, sync_interval=4)
    o = po.OracleSim(c, E, nu)
    o.elastic_run()
    st = o.event_steps(12)
    ev = o.events()
    assert st.kmc_events == 12
    # events 4, 8, 12 end with a solve, the others are patched only
    assert [e.ls_iters > 0 for e in ev] == [(k + 1) % 4 == 0 for k in range(12)]
