"""Fast patching on the device (mgk_patch_init / mgk_patch_apply and
mgk_kmc_event's patch modes) against the oracle's restatement of
kmc_solver.hpp:276-332, :525-547, :866-913, :941-949 (3D :1256-1331,
:1535-1573, :1877-1972).

Bars: on a homogeneous grid whose window covers the whole grid, a patch
equals a full solve (1e-12); fed the oracle's state, one patch gives the
oracle's stress within 1e-10 (the kernels come from LS solves that agree to
~1e-12); whole runs with syncs every few flips give the oracle's counters,
event sequence and stress (1e-9)."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Simulation, abi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _dev_strain(D, seed):
    rng = np.random.default_rng(seed)
    C = 1e-2 * rng.standard_normal((D, D))
    C = 0.5 * (C + C.T)
    return C - np.trace(C) / D * np.eye(D)


@pytest.mark.parametrize("shape,R,voxel", [((16, 16), 8, 37), ((8, 8, 8), 4, 300), ((16, 16, 1), 8, 100),
                                           ((16, 16), 3, 5)])
def test_patch_matches_oracle_homogeneous(shape, R, voxel):
    # R = n/2 on an even grid: the window (n + 1 wide) wraps onto itself, so
    # the cells at offsets -R and +R coincide and take both contributions, in
    # the reference's loop order (the lane-per-cell kernel's multi-hit path)
    D = len(shape)
    N = int(np.prod(shape))
    kw = dict(dims=D, nx=shape[0], ny=shape[1], M=4, fast_patching=1, patch_radius=R, ls_tol=1e-12, ls_max_iter=500)
    if D == 3:
        kw["nz"] = shape[2]
    c = abi.default_sim_config(**kw)
    E, nu = np.full(N, 70e9), np.full(N, 0.3)
    d = Simulation(c, E, nu)
    o = po.OracleSim(c, E, nu)
    C = _dev_strain(D, 3)
    for s in (d, o):
        s.elastic_run()
        ep = s.get_field(abi.FIELD_EPS_PLASTIC)
        ep[voxel] = C
        s.set_field(abi.FIELD_EPS_PLASTIC, ep)
        s.patch(voxel, C)
    assert _rel(d.get_field(abi.FIELD_SIG), o.get_field(abi.FIELD_SIG)) < 1e-12


def test_patch_fed_oracle_state_heterogeneous():
    rng = np.random.default_rng(23)
    E = 85e9 * (1 + 0.15 * rng.standard_normal(512))
    nu = np.full(512, 0.35)
    c = abi.default_sim_config(dims=3, nx=8, ny=8, nz=8, M=4, fast_patching=1, patch_radius=2, ls_tol=1e-10,
                               catalog_mode=abi.CATALOG_RANDOM, temperature=300.0)
    d = Simulation(c, E, nu)
    o = po.OracleSim(c, E, nu)
    o.elastic_run()
    d.elastic_run()
    C = _dev_strain(3, 5)
    for s in (d, o):
        s.patch(77, C)
        s.patch(3, 2.0 * C)
    assert _rel(d.get_field(abi.FIELD_SIG), o.get_field(abi.FIELD_SIG)) < 1e-10


@pytest.mark.parametrize("kw", [
    dict(dims=2, nx=16, ny=16, M=4, patch_radius=3, sync_interval=4),
    dict(dims=2, nx=16, ny=16, M=4, patch_radius=5, sync_interval=3, instability_kmc=1),
    dict(dims=3, nx=8, ny=8, nz=8, M=6, patch_radius=2, sync_interval=5),
])
def test_patched_run_matches_oracle(kw):
    N = kw["nx"] * kw["ny"] * kw.get("nz", 1)
    rng = np.random.default_rng(23)
    E = 85e9 * (1 + 0.15 * rng.standard_normal(N))
    nu = np.full(N, 0.35)
    c = abi.default_sim_config(b_mean=1.3, b_std=0.1, b_min_cutoff=0.01, temperature=300.0,
                               catalog_mode=abi.CATALOG_RANDOM, gamma0=0.1, ls_tol=1e-8, seed=5, fast_patching=1,
                               **kw)
    run = abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6,
                                 max_events=60)
    o = po.OracleSim(c, E, nu)
    so = o.run(run)
    d = Simulation(c, E, nu)
    sd = d.run(run)
    print(kw, "oracle", so.kmc_events, so.cascades, so.ls_solves, "device", sd.kmc_events, sd.cascades, sd.ls_solves)
    assert so.kmc_events > 3
    assert (sd.kmc_events, sd.cascades, sd.cascade_flips, sd.elastic_steps) == \
           (so.kmc_events, so.cascades, so.cascade_flips, so.elastic_steps)
    eo, ed = o.events(), d.events()
    assert [e.flat_idx for e in ed] == [e.flat_idx for e in eo]
    assert [e.ls_iters > 0 for e in ed] == [e.ls_iters > 0 for e in eo]
    assert _rel(d.get_field(abi.FIELD_SIG), o.get_field(abi.FIELD_SIG)) < 1e-9
    assert _rel(d.get_field(abi.FIELD_EPS_PLASTIC), o.get_field(abi.FIELD_EPS_PLASTIC)) < 1e-14
