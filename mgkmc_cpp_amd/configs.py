"""The five benchmark/parity workloads, as simulation configurations.

c0  2D plane strain 32x32, homogeneous E = 85 GPa, nu = 0.35; 6 fixed shears
    of engineering magnitude 0.1 at k*30 deg; barriers N(1.0, 0.1) eV; T = 300 K,
    nu0 = 1e13 /s; uniaxial tension (sigma_yy = 0) in 1e-4 steps; 2000 events;
    LS tolerance 1e-8; FP64.
c1  as c0 with E a Gaussian random field (mean 85 GPa, CoV 0.15, correlation
    length 3 voxels); 5000 events.
c2  3D 64^3, E field with correlation length 4; 12 fixed deviatoric shears of
    magnitude 0.1; 20000 events; FP32 iterations with FP64 refinement to 1e-8.
c3  256 independent trajectories, grids drawn from 32^2, 64^2, 128^2, 32^3,
    48^3, 2000 events each, FP32+refine.
c4  Poisson arrivals (50 /s) of requests with grids 32^2 .. 128^3 and
    500-5000 events each.

The correlation length is the Gaussian-filter sigma in voxels
(material_generator.hpp:169); E values are generated in GPa and scaled by 1e9
as main.cpp:137-141 does for means below 1e6.
"""
import numpy as np

from . import abi
from .fields import correlated_field

E_MEAN_GPA = 85.0
E_COV = 0.15
NU = 0.35


def material(shape, corr, seed, heterogeneous=True, device=False):
    """E (Pa) and nu; device=True draws the field on the GPU
    (mgk_correlated_field: the same filter, Philox noise)."""
    N = int(np.prod(shape))
    if heterogeneous and device:
        from .solver import correlated_field_device
        E = correlated_field_device(shape, E_MEAN_GPA, E_COV * E_MEAN_GPA, corr, seed) * 1e9
    elif heterogeneous:
        E = correlated_field(shape, E_MEAN_GPA, E_COV * E_MEAN_GPA, corr, seed) * 1e9
    else:
        E = np.full(N, E_MEAN_GPA * 1e9)
    return E, np.full(N, NU)


def sim_config(shape, precision=abi.MGK_FP64, seed=42, **kw):
    dims = len(shape)
    base = dict(
        dims=dims, nx=shape[0], ny=shape[1], nz=shape[2] if dims == 3 else 1, pixel=1.0,
        plane_mode=abi.MGK_PLANE_STRAIN, M=6 if dims == 2 else 12, gamma0=0.1,
        barrier_type=abi.BARRIER_GAUSSIAN, b_mean=1.0, b_std=0.1, temperature=300.0, nu0=1e13,
        catalog_mode=abi.CATALOG_FIXED, catalog_magnitude=0.1, ls_tol=1e-8, ls_max_iter=200,
        precision=precision, seed=seed,
    )
    base.update(kw)
    return abi.default_sim_config(**base)


def c0(seed=42):
    shape = (32, 32)
    return sim_config(shape, seed=seed), material(shape, 0, seed, heterogeneous=False), 2000


def c1(seed=42):
    shape = (32, 32)
    return sim_config(shape, seed=seed), material(shape, 3.0, seed), 5000


def c2(seed=42, precision=abi.MGK_FP32_REFINE):
    shape = (64, 64, 64)
    return sim_config(shape, precision=precision, seed=seed), material(shape, 4.0, seed), 20000


C3_GRIDS = [(32, 32), (64, 64), (128, 128), (32, 32, 32), (48, 48, 48)]
C4_GRIDS = [(32, 32), (64, 64), (128, 128), (32, 32, 32), (48, 48, 48), (64, 64, 64), (128, 128, 128)]


def c3_grids(n=256, seed=7):
    rng = np.random.default_rng(seed)
    return [C3_GRIDS[k] for k in rng.integers(0, len(C3_GRIDS), n)]


def c3_requests(n=256, events=2000, seed=7, base_seed=1000, precision=abi.MGK_FP32_REFINE):
    """c3: n trajectories present from the start, grid k drawn with `seed`,
    trajectory i seeded base_seed + i.  The batch and serving workloads run
    instability_mode "kmc" (unstable modes are rated events), so the work of
    a request is its event count; "cascade" mode can flip for up to N rounds
    between two events.  Their inputs (E field, barriers, STZ strains) are
    drawn on the device."""
    from .serving import Request
    return [Request(rid=i, shape=g, events=events, seed=base_seed + i, arrival=0.0, precision=precision,
                    instability_kmc=1, device_inputs=True) for i, g in enumerate(c3_grids(n, seed))]


def c4_requests(n, rate=50.0, seed=11, events=(500, 5000), grids=C4_GRIDS, precision=abi.MGK_FP32_REFINE):
    """c4: n requests with Poisson arrivals at `rate` per second, a grid drawn
    from `grids` and an event count uniform in [events[0], events[1]], all
    fixed by `seed`; request i is seeded 100000 + seed * 1000 + i."""
    from .serving import Request
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    g = rng.integers(0, len(grids), n)
    ev = rng.integers(events[0], events[1] + 1, n)
    return [Request(rid=i, shape=tuple(grids[g[i]]), events=int(ev[i]), seed=100000 + seed * 1000 + i,
                    arrival=float(t[i]), precision=precision, instability_kmc=1, device_inputs=True)
            for i in range(n)]


def uniaxial_run(events, step_size=1e-4):
    """Strain-controlled uniaxial tension along xx with sigma_yy = 0 (2D) or
    sigma_yy = sigma_zz = 0 (3D), event cap `events`."""
    return dict(step_size=step_size, drv_i=0, drv_j=0, max_events=events, mixed_tol=1e6)
