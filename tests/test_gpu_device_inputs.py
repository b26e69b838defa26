"""Input generation on the device (mgk_kmc_draw, mgk_correlated_field): the
draws against the Philox restatement (oracle/philox_ref.py), the distributions
against their closed forms, and the field against the reference's FFT filter
(mgkmc_cpp_amd.fields.correlated_field, a restatement of
material_generator.hpp:130-258) given the same noise."""
import math

import numpy as np
import pytest

from mgkmc_cpp_amd import Context, Simulation, abi, fields
from mgkmc_cpp_amd.solver import correlated_field_device
from oracle import philox_ref

pytestmark = pytest.mark.gpu


def _ctx(shape, M):
    N = int(np.prod(shape))
    c = Context(shape, np.full(N, 70e9), np.full(N, 0.3))
    c.kmc_init(M=M)
    return c


def _params(**kw):
    p = abi.DrawParams()
    base = dict(barrier_type=abi.BARRIER_GAUSSIAN, b_mean=1.0, b_std=0.1, b_min_cutoff=-1.0, b_max_cutoff=-1.0,
                b_loc=1.0, b_scale=0.5, b_epsilon=0.1, b_ratio=0.8, gamma0=0.1, use_3d_barriers=0)
    base.update(kw)
    for k, v in base.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("btype", [abi.BARRIER_GAUSSIAN, abi.BARRIER_RAYLEIGH])
def test_barrier_draws_equal_philox_restatement(btype):
    ctx = _ctx((16, 16), 6)
    seed = 0x1234567890AB
    ctx.kmc_draw(_params(barrier_type=btype), seed)
    q0 = ctx.get_field(abi.FIELD_Q0)
    idx = list(range(0, 1536, 37)) + [1535]
    if btype == abi.BARRIER_GAUSSIAN:
        want = [philox_ref.barrier_gaussian(i, seed, 1.0, 0.1) for i in idx]
    else:
        want = [philox_ref.barrier_rayleigh(i, seed, 1.0, 0.5) for i in idx]
    assert np.abs(q0[idx] - np.array(want)).max() < 1e-13
    cat = ctx.get_field(abi.FIELD_CATALOG)                 # (N*M, 2, 2)
    for i in idx[:10]:
        gxx, gxy = philox_ref.planar_mode(i, seed, 0.1)
        assert abs(cat[i, 0, 0] - gxx) < 1e-15 and abs(cat[i, 1, 1] + gxx) < 1e-15 and abs(cat[i, 0, 1] - gxy) < 1e-15


def test_distributions_match_closed_forms():
    ctx = _ctx((64, 64), 25)                                # 102400 modes
    n = 64 * 64 * 25
    ctx.kmc_draw(_params(barrier_type=abi.BARRIER_RAYLEIGH, b_loc=0.5, b_scale=0.4), 7)
    q = ctx.get_field(abi.FIELD_Q0)
    mean_r, sd_r = 0.5 + 0.4 * math.sqrt(math.pi / 2), 0.4 * math.sqrt((4 - math.pi) / 2)
    assert abs(q.mean() - mean_r) < 5 * sd_r / math.sqrt(n) and abs(q.std() / sd_r - 1) < 0.02
    # modified Rayleigh: density x exp(-(x - mu)^2 / 2 s^2) on [max(0, mu - 6s), mu + 6s]
    ctx.kmc_draw(_params(barrier_type=abi.BARRIER_MODIFIED_RAYLEIGH, b_mean=2.0, b_std=0.6), 8)
    q = ctx.get_field(abi.FIELD_Q0)
    x = np.linspace(0.0, 5.6, 200001)
    w = x * np.exp(-(x - 2.0) ** 2 / (2 * 0.36))
    m1 = (x * w).sum() / w.sum()
    sd = math.sqrt((x * x * w).sum() / w.sum() - m1 * m1)
    assert abs(q.mean() - m1) < 5 * sd / math.sqrt(n) and abs(q.std() / sd - 1) < 0.02
    # planar STZ strains: gxx, gxy ~ N(0, (gamma0 / 2)^2)
    cat = ctx.get_field(abi.FIELD_CATALOG)
    for comp in (cat[:, 0, 0], cat[:, 0, 1]):
        assert abs(comp.mean()) < 5 * 0.05 / math.sqrt(n) and abs(comp.std() / 0.05 - 1) < 0.02
    assert np.array_equal(cat[:, 0, 0], -cat[:, 1, 1])


def test_glass_3d_modes():
    ctx = _ctx((16, 16, 16), 12)
    ctx.kmc_draw(_params(use_3d_barriers=1, gamma0=0.2), 9)
    c = ctx.get_field(abi.FIELD_CATALOG)                   # (N*M, 3, 3)
    assert np.abs(np.trace(c, axis1=1, axis2=2)).max() < 1e-15
    assert np.array_equal(c, c.transpose(0, 2, 1))
    # |c|_F sqrt(2) / gamma0 = sqrt(-2 ln u): Rayleigh(1), mean sqrt(pi / 2)
    r = np.sqrt((c * c).sum(axis=(1, 2))) * math.sqrt(2) / 0.2
    assert abs(r.mean() - math.sqrt(math.pi / 2)) < 0.015 and abs(r.std() - math.sqrt((4 - math.pi) / 2)) < 0.015


@pytest.mark.parametrize("shape,corr", [((32, 32), 3.0), ((48, 32), 2.5), ((32, 32, 32), 4.0), ((16, 16, 1), 2.0)])
def test_field_equals_reference_filter_given_the_noise(shape, corr):
    N = int(np.prod(shape))
    noise = fields.box_muller_normals(11, N)
    want = fields.correlated_field(shape, 85.0, 12.75, corr, 11, clip_min=60.0)
    got = correlated_field_device(shape, 85.0, 12.75, corr, 0, clip_min=60.0, noise=noise)
    assert np.abs(got - want).max() < 1e-12 * 85.0


def test_field_from_device_noise_statistics():
    shape, corr = (64, 64, 64), 4.0
    f = correlated_field_device(shape, 85.0, 12.75, corr, 3)
    assert abs(f.mean() - 85.0) < 1e-10 and abs(f.std() - 12.75) < 1e-9
    g = f.reshape(shape) - 85.0
    rho = (g * np.roll(g, 1, axis=0)).mean() / (g * g).mean()
    assert abs(rho - math.exp(-1.0 / (4 * corr * corr))) < 0.01   # Gaussian filter: exp(-r^2 / 4 sigma^2)
    assert not np.array_equal(f, correlated_field_device(shape, 85.0, 12.75, corr, 4))


def test_simulation_with_device_draws_runs_and_repeats():
    N = 32 * 32
    E = correlated_field_device((32, 32), 85e9, 12.75e9, 3.0, 5)
    nu = np.full(N, 0.35)
    cfg = abi.default_sim_config(dims=2, nx=32, ny=32, M=6, b_mean=1.0, b_std=0.1, temperature=300.0, ls_tol=1e-8,
                                 instability_kmc=1, device_draws=1, seed=77)
    a, b = Simulation(cfg, E, nu), Simulation(cfg, E, nu)
    for s in (a, b):
        s.elastic_run()
        s.event_steps(20)
    assert len(a.events()) == 20
    assert np.array_equal(a.get_field(abi.FIELD_Q0), b.get_field(abi.FIELD_Q0))
    assert [e.flat_idx for e in a.events()] == [e.flat_idx for e in b.events()]


@pytest.mark.parametrize("shape,at,corr", [((32, 32), (0, 0), 3.0), ((16, 12), (5, 7), 2.0),
                                           ((8, 10, 6), (3, 0, 5), 1.5)])
def test_field_filter_closed_form_impulse(shape, at, corr):
    # a unit impulse through the device filter (three circular convolutions)
    # equals the standardised Gaussian kernel summed mode by mode
    from test_oracle_pinning import _impulse_field
    noise = np.zeros(shape)
    noise[at] = 1.0
    got = correlated_field_device(shape, 0.0, 1.0, corr, 0, noise=noise)
    assert np.abs(got - _impulse_field(shape, at, corr)).max() < 1e-12
