"""KMC event primitives and trajectories on the device against the oracle.

* fed the oracle's state, update_barriers / rates / total rate agree to
  rounding and the selected event index is bit-exact for every u2 drawn;
* adversarial targets on exact prefix boundaries (T = 0, rates exactly
  V*nu0 or 0) go through the exact sequential fallback and still agree;
* apply_flip agrees bit-exactly;
* whole trajectories of the drop-in host driver reproduce the oracle's event
  sequence, with stress within 1e-10."""
import math

import numpy as np
import pytest

from mgkmc_cpp_amd import Context, Simulation, abi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _cfg(**kw):
    base = dict(dims=2, nx=16, ny=16, M=6, gamma0=0.14, b_mean=1.0, b_std=0.1, temperature=300.0, seed=42,
                ls_tol=1e-8, catalog_mode=abi.CATALOG_FIXED, catalog_magnitude=0.1)
    base.update(kw)
    return abi.default_sim_config(**base)


def _field(N, seed=3, cov=0.15):
    rng = np.random.default_rng(seed)
    return 85e9 * (1.0 + cov * rng.standard_normal(N)), np.full(N, 0.35)


def _device_from_oracle(osim, c, E, nu):
    shape = (c.nx, c.ny) + ((c.nz,) if c.dims == 3 else ())
    ctx = Context(shape, E, nu, pixel=c.pixel)
    tau = osim.tau()
    ctx.kmc_init(M=c.M, softening_directional=c.softening_directional, softening_cap=c.softening_cap, jp=c.jp,
                 jt=c.jt, neighbor_softening_fraction=c.neighbor_softening_fraction, tau=tau,
                 volume=c.pixel ** 3, temperature=c.temperature, nu0=c.nu0,
                 stability_threshold=c.stability_threshold, instability_kmc=c.instability_kmc,
                 scale_rate_by_volume=c.scale_rate_by_volume)
    for f in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_Q0, abi.FIELD_CATALOG,
              abi.FIELD_SOFT, abi.FIELD_LAST_EVENT_TIME, abi.FIELD_PREV_DIR):
        ctx.set_field(f, osim.get_field(f))
    return ctx


def _evolved_oracle(c, E, nu, events=25):
    osim = po.OracleSim(c, E, nu)
    osim.event_steps(events)
    return osim


@pytest.mark.parametrize("dims,directional", [(2, 0), (2, 1), (3, 1)])
def test_rates_and_selection_fed_oracle_state(dims, directional):
    kw = dict(dims=dims, softening_directional=directional, instability_kmc=1)
    if dims == 3:
        kw.update(nx=8, ny=8, nz=8, M=12)
    c = _cfg(**kw)
    N = c.nx * c.ny * (c.nz if dims == 3 else 1)
    E, nu = _field(N)
    osim = _evolved_oracle(c, E, nu)
    ctx = _device_from_oracle(osim, c, E, nu)
    _, _, t = osim.macro()
    osim.update_barriers()
    total_o, rates_o, idx_o = osim.rates()
    total_d, n_unst, n_incl = ctx.update_rates(t)
    assert n_incl == len(rates_o)
    assert _rel(ctx.get_field(abi.FIELD_Q), osim.get_field(abi.FIELD_Q)) < 1e-14
    assert abs(total_d - total_o) <= 1e-13 * total_o
    rng = np.random.default_rng(11)
    for u2 in rng.random(3000):
        want = idx_o[po.select_sequential(rates_o, u2 * total_o)]
        got, q = ctx.select(u2)
        assert got == want, (u2, got, want)


@pytest.mark.parametrize("catalog,directional", [(abi.CATALOG_FIXED, 0), (abi.CATALOG_RANDOM, 1)])
def test_rates_and_selection_fed_oracle_state_64cube(catalog, directional):
    # 3.1 M modes in 3072 rate blocks (shared catalog staged in LDS, or
    # per-voxel catalog rows)
    c = _cfg(dims=3, nx=64, ny=64, nz=64, M=12, softening_directional=directional, instability_kmc=1,
             catalog_mode=catalog)
    N = c.nx * c.ny * c.nz
    E, nu = _field(N)
    osim = _evolved_oracle(c, E, nu, events=2)
    ctx = _device_from_oracle(osim, c, E, nu)
    _, _, t = osim.macro()
    osim.update_barriers()
    total_o, rates_o, idx_o = osim.rates()
    total_d, n_unst, n_incl = ctx.update_rates(t)
    assert n_incl == len(rates_o) == N * c.M
    assert _rel(ctx.get_field(abi.FIELD_Q), osim.get_field(abi.FIELD_Q)) < 1e-14
    # The oracle's left-to-right total (the reference's order) carries up to
    # n u of rounding over 3.1 M terms: once the running sum is large, small
    # rates fall below its half ulp (1.3e-11 of the total here).  The
    # device's blocked sum stays within 1e-12 of the exactly rounded sum,
    # and selection allows for the n u difference (its ambiguity bound).
    exact = math.fsum(rates_o)
    assert abs(total_o - exact) <= N * c.M * 2.0 ** -53 * exact
    assert abs(total_d - exact) <= 1e-12 * exact
    rng = np.random.default_rng(12)
    for u2 in rng.random(300):
        want = idx_o[po.select_sequential(rates_o, u2 * total_o)]
        got, q = ctx.select(u2)
        assert got == want, (u2, got, want)
    ctx.close()


@pytest.mark.parametrize("kw", [
    dict(nx=48, ny=32, M=7, instability_kmc=1),                                  # 10752 modes: 10.5 chunks
    dict(nx=48, ny=32, M=7, instability_kmc=0, b_mean=0.3, b_std=0.3, b_min_cutoff=-1.0),   # excluded modes
    dict(dims=3, nx=16, ny=16, nz=16, M=12, instability_kmc=1),                  # 49152 modes: 48 chunks
])
def test_exact_selection_pass_matches_oracle(kw):
    # the reference's left-to-right selection run as written on the device
    # (one wave), against the oracle's, over several chunks of 1024 modes
    c = _cfg(**kw)
    N = c.nx * c.ny * (c.nz if c.dims == 3 else 1)
    E, nu = _field(N)
    osim = po.OracleSim(c, E, nu)
    osim.elastic_run()
    ctx = _device_from_oracle(osim, c, E, nu)
    _, _, t = osim.macro()
    osim.update_barriers()
    total_o, rates_o, idx_o = osim.rates()
    ctx.update_rates(t)
    assert kw.get("instability_kmc") or len(rates_o) < N * c.M
    rng = np.random.default_rng(5)
    us = [0.0, float(np.nextafter(1.0, 0.0))] + list(rng.random(40))
    for u2 in us:
        want = idx_o[po.select_sequential(rates_o, u2 * total_o)]
        got, q = ctx.select(u2, exact=True)
        assert got == want, (u2, got, want)
        assert got == ctx.select(u2)[0]


def test_selection_on_exact_boundaries():
    # T = 0 and instability_mode "kmc": rates are exactly V*nu0 (Q <= 0) or 0,
    # so prefix sums are exact and targets can sit on a boundary.
    c = _cfg(nx=8, ny=8, M=6, temperature=0.0, instability_kmc=1, b_mean=0.0, b_std=1.0, b_min_cutoff=-10.0)
    N = 64
    E, nu = _field(N)
    osim = po.OracleSim(c, E, nu)
    ctx = _device_from_oracle(osim, c, E, nu)
    osim.update_barriers()
    total_o, rates_o, idx_o = osim.rates()
    total_d, _, _ = ctx.update_rates(0.0)
    assert total_d == total_o
    nz = int((rates_o > 0).sum())
    assert 0 < nz < len(rates_o)
    us = [k / nz for k in range(nz + 1)] + [np.nextafter(k / nz, 1.0) for k in range(nz)] + [0.0]
    for u2 in us:
        if u2 >= 1.0:
            continue
        want = idx_o[po.select_sequential(rates_o, u2 * total_o)]
        got, _ = ctx.select(u2)
        assert got == want, (u2, got, want)


def test_flip_and_redraw_match_oracle():
    c = _cfg(nx=4, ny=4, M=3, neighbor_softening_fraction=0.3, catalog_mode=abi.CATALOG_RANDOM)
    E, nu = _field(16)
    osim = _evolved_oracle(c, E, nu, events=5)
    ctx = _device_from_oracle(osim, c, E, nu)
    for flat, t in ((7, 1.25), (0, 1.5), (47, 2.0)):
        osim.apply_flip(flat, t)
        ctx.apply_flip(flat, t)
        for f in (abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_LAST_EVENT_TIME, abi.FIELD_PREV_DIR):
            assert np.array_equal(ctx.get_field(f), osim.get_field(f)), f
    # redraw upload: catalog / Q0 of the flipped voxel replaced after the flip used the old entry
    cat = np.array([[[0.01, 0.02], [0.02, -0.01]]] * 3)
    q0 = np.array([1.1, 1.2, 1.3])
    ctx.apply_flip(5, 3.0, new_catalog=cat, new_Q0=q0, first_mode=0)
    got_cat = ctx.get_field(abi.FIELD_CATALOG)[3:6]
    assert np.array_equal(got_cat, cat) and np.array_equal(ctx.get_field(abi.FIELD_Q0)[3:6], q0)


@pytest.mark.parametrize("shape", [(8, 1, 8), (4, 4, 1), (4, 4, 4)])
def test_flip_on_thin_and_small_grids_matches_oracle(shape):
    # an axis of length 1 makes neighbours alias (the flipped voxel is its own
    # neighbour along that axis): the flip then runs on one lane in the
    # reference's order; 4^3 takes the one-lane-per-cell path
    c = _cfg(dims=3, nx=shape[0], ny=shape[1], nz=shape[2], M=3, neighbor_softening_fraction=0.3,
             catalog_mode=abi.CATALOG_RANDOM)
    N = int(np.prod(shape))
    E, nu = _field(N)
    osim = _evolved_oracle(c, E, nu, events=4)
    ctx = _device_from_oracle(osim, c, E, nu)
    for k, t in enumerate((1.25, 1.5, 2.0, 2.5)):
        flat = (k * 7 + 1) % (N * c.M)
        osim.apply_flip(flat, t)
        ctx.apply_flip(flat, t)
        for f in (abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_LAST_EVENT_TIME, abi.FIELD_PREV_DIR):
            assert np.array_equal(ctx.get_field(f), osim.get_field(f)), (shape, flat, f)


def test_select_flip_chain_equals_select_then_flip():
    c = _cfg(nx=8, ny=8, M=4, instability_kmc=1, catalog_mode=abi.CATALOG_RANDOM,
             neighbor_softening_fraction=0.2)
    E, nu = _field(64)
    osim = _evolved_oracle(c, E, nu, events=6)
    a = _device_from_oracle(osim, c, E, nu)
    b = _device_from_oracle(osim, c, E, nu)
    rng = np.random.default_rng(5)
    for t, u2 in ((1.0, 0.3), (1.5, 0.91), (2.0, 0.0003)):
        cat = rng.standard_normal((4, 2, 2)) * 0.05
        cat = 0.5 * (cat + np.transpose(cat, (0, 2, 1)))
        q0 = rng.uniform(0.8, 1.2, 4)
        a.update_rates(t)
        b.update_rates(t)
        flat, _ = a.select(u2)
        a.apply_flip(flat, t, new_catalog=cat, new_Q0=q0, first_mode=0)
        b.select_flip(u2, t, new_catalog=cat, new_Q0=q0, first_mode=0)
        assert b.last_selection()[0] == flat
        for f in (abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_LAST_EVENT_TIME, abi.FIELD_PREV_DIR,
                  abi.FIELD_CATALOG, abi.FIELD_Q0):
            assert np.array_equal(a.get_field(f), b.get_field(f)), f
    # redraw of the selected mode only (both redraw switches off)
    a.update_rates(3.0)
    b.update_rates(3.0)
    flat, _ = a.select(0.5)
    a.apply_flip(flat, 3.0, new_catalog=cat[:1], new_Q0=q0[:1], first_mode=flat % 4)
    b.select_flip(0.5, 3.0, new_catalog=cat[:1], new_Q0=q0[:1], first_mode=-1)
    assert b.last_selection()[0] == flat
    for f in (abi.FIELD_CATALOG, abi.FIELD_Q0, abi.FIELD_EPS_PLASTIC):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f


@pytest.mark.parametrize("kw,events", [
    (dict(), 40),                                                          # 2D fixed catalog, kmc-free cascade mode
    (dict(instability_kmc=1, catalog_mode=abi.CATALOG_RANDOM), 40),        # random catalog, host-drawn redraws
    (dict(softening_directional=1, neighbor_softening_fraction=0.2), 30),
    (dict(dims=3, nx=8, ny=8, nz=8, M=12), 15),
])
def test_trajectory_matches_oracle(kw, events):
    c = _cfg(**kw)
    N = c.nx * c.ny * (c.nz if c.dims == 3 else 1)
    E, nu = _field(N, seed=17)
    osim = po.OracleSim(c, E, nu)
    dsim = Simulation(c, E, nu)
    osim.elastic_run()
    dsim.elastic_run()
    so = osim.event_steps(events)
    sd = dsim.event_steps(events)
    eo, ed = osim.events(), dsim.events()
    assert len(eo) == len(ed) == so.kmc_events == sd.kmc_events
    assert [e.flat_idx for e in ed] == [e.flat_idx for e in eo]
    for a, b in zip(ed, eo):
        assert a.ls_iters == b.ls_iters
        assert abs(a.time - b.time) <= 1e-12 * abs(b.time)
        assert _rel(np.array(a.sig_macro), np.array(b.sig_macro)) < 1e-10
    assert _rel(dsim.get_field(abi.FIELD_SIG), osim.get_field(abi.FIELD_SIG)) < 1e-10
    assert _rel(dsim.get_field(abi.FIELD_EPS_PLASTIC), osim.get_field(abi.FIELD_EPS_PLASTIC)) < 1e-15


def test_run_simulation_with_mixed_bc_and_cascades():
    # large strain steps drive avalanches: events, cascades (unstable sort,
    # per-voxel dedupe, host RNG order) and the yy stress-free mixed BC loop.
    # Oracle: 150 events, ~96 cascades, 5 elastic steps, < 1 s on one core.
    c = _cfg(nx=16, ny=16, M=4, b_mean=1.3, b_std=0.1, b_min_cutoff=0.01, temperature=300.0,
             catalog_mode=abi.CATALOG_RANDOM, gamma0=0.1, ls_tol=1e-8)
    rng = np.random.default_rng(23)
    E = 85e9 * (1 + 0.15 * rng.standard_normal(256))
    nu = np.full(256, 0.35)
    run = abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6,
                                 max_events=150)
    osim = po.OracleSim(c, E, nu)
    dsim = Simulation(c, E, nu)
    so = osim.run(run)
    sd = dsim.run(run)
    assert so.cascades > 10 and so.elastic_steps > 0
    assert (sd.kmc_events, sd.cascades, sd.cascade_flips, sd.elastic_steps) == \
           (so.kmc_events, so.cascades, so.cascade_flips, so.elastic_steps)
    assert [e.flat_idx for e in dsim.events()] == [e.flat_idx for e in osim.events()]
    eo, so_, to = osim.macro()
    ed, sd_, td = dsim.macro()
    assert _rel(sd_, so_) < 1e-9 and abs(td - to) <= 1e-12 * abs(to)


_STATE_FIELDS = (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_LAST_EVENT_TIME,
                 abi.FIELD_PREV_DIR, abi.FIELD_CATALOG, abi.FIELD_Q0)


def test_device_event_equals_host_chain():
    # mgk_kmc_event (rates, t_wait, selection, flip and solve in one round
    # trip) against the host chain update_rates -> t_wait on the host ->
    # select_flip -> elastic_run: bit for bit.
    c = _cfg(nx=16, ny=16, M=4, instability_kmc=1, catalog_mode=abi.CATALOG_RANDOM,
             neighbor_softening_fraction=0.2)
    E, nu = _field(256)
    osim = _evolved_oracle(c, E, nu, events=6)
    em, _, t = osim.macro()
    a = _device_from_oracle(osim, c, E, nu)
    b = _device_from_oracle(osim, c, E, nu)
    a.elastic_run(em, tol=1e-8)
    b.elastic_run(em, tol=1e-8)
    em_a = em_b = em
    rng = np.random.default_rng(9)
    for _ in range(4):
        u1, u2 = rng.random(2)
        cat = rng.standard_normal((4, 2, 2)) * 0.05
        cat = 0.5 * (cat + np.transpose(cat, (0, 2, 1)))
        q0 = rng.uniform(0.8, 1.2, 4)
        L = -np.log(u1)
        R, _, _ = b.update_rates(t)
        tw = L / R
        eb = em_b.copy()
        eb[0, 1] += 1.0 * (2.0 * tw)
        b.select_flip(u2, t + tw, new_catalog=cat, new_Q0=q0, first_mode=0)
        rb = b.elastic_run(eb, tol=1e-8)
        out, em_a2, sm_a2 = a.kmc_event(em_a, t, L, u2, strain_rate=2.0, drv=(0, 1), new_catalog=cat, new_Q0=q0,
                                        first_mode=0, tol=1e-8)
        assert out.outcome == abi.EVENT_DONE
        assert out.total_rate == R and out.t_wait == tw
        assert out.flat_idx == b.last_selection()[0] and out.iters == rb["iters"]
        assert np.array_equal(em_a2, rb["eps_macro"]) and np.array_equal(sm_a2, rb["sig_macro"])
        for f in _STATE_FIELDS:
            assert np.array_equal(a.get_field(f), b.get_field(f)), f
        t = t + tw
        em_a, em_b = em_a2, rb["eps_macro"]
    # t_wait >= remaining: nothing changes
    before = {f: a.get_field(f) for f in _STATE_FIELDS}
    out, em2, _ = a.kmc_event(em_a, t, 50.0, 0.5, remaining=1e-300, tol=1e-8)
    assert out.outcome == abi.EVENT_REJECTED and out.flat_idx == -1 and out.total_rate > 0
    assert np.array_equal(em2, em_a)
    for f in _STATE_FIELDS:
        assert np.array_equal(a.get_field(f), before[f]), f
    # the next event still matches the host chain after a rejected attempt
    R, _, _ = b.update_rates(t)
    out, em_a2, _ = a.kmc_event(em_a, t, 0.7, 0.25, tol=1e-8)
    b.select_flip(0.25, t + 0.7 / R)
    eb = em_b.copy()
    eb[0, 0] += 0.7 / R
    rb = b.elastic_run(eb, tol=1e-8)
    assert out.outcome == abi.EVENT_DONE and np.array_equal(em_a2, rb["eps_macro"])
    for f in _STATE_FIELDS:
        assert np.array_equal(a.get_field(f), b.get_field(f)), f


def test_device_event_without_rate_or_with_unstable_modes():
    # T = 0 in cascade mode: every rated mode has rate 0 -> no event
    c = _cfg(nx=8, ny=8, M=3, temperature=0.0)
    E, nu = _field(64)
    osim = po.OracleSim(c, E, nu)
    a = _device_from_oracle(osim, c, E, nu)
    em = np.zeros((2, 2))
    em[0, 0] = 1e-3
    a.elastic_run(em, tol=1e-8)
    before = {f: a.get_field(f) for f in _STATE_FIELDS}
    out, em2, _ = a.kmc_event(em, 0.0, 1.0, 0.5, tol=1e-8)
    assert out.outcome == abi.EVENT_NO_RATE and out.total_rate == 0.0
    assert np.array_equal(em2, em)
    # a stability threshold above every barrier makes all modes unstable
    c2 = _cfg(nx=8, ny=8, M=3, temperature=300.0, stability_threshold=1e3)
    osim2 = po.OracleSim(c2, E, nu)
    b = _device_from_oracle(osim2, c2, E, nu)
    b.elastic_run(em, tol=1e-8)
    out, _, _ = b.kmc_event(em, 0.0, 1.0, 0.5, tol=1e-8, cascade_check=True)
    assert out.outcome == abi.EVENT_UNSTABLE and out.n_unstable == 64 * 3
    out, _, _ = b.kmc_event(em, 0.0, 1.0, 0.5, tol=1e-8, cascade_check=False)
    assert out.outcome == abi.EVENT_DONE
    for f in _STATE_FIELDS:
        assert np.array_equal(a.get_field(f), before[f]), f


@pytest.mark.parametrize("shape", [(16, 16, 16), (16, 16, 32), (32, 32)])
def test_batch_equals_separate_fp32_refine(shape):
    # FP32+refine: 16^3 runs its CG steps as plane passes with folded
    # reductions, 16x16x32 and 32^2 the line passes.  Each trajectory of a
    # batch keeps its own CG cycle length, and a pending CG iterate is absorbed
    # the same way wherever the refinement rounds end, so a batch equals the
    # same trajectories run one by one, bit for bit.
    from mgkmc_cpp_amd import SimulationBatch
    seeds, events = (42, 7, 1234), 8
    D = len(shape)
    geo = dict(dims=D, nx=shape[0], ny=shape[1]) | (dict(nz=shape[2], M=12) if D == 3 else dict(M=6))
    cfgs = [_cfg(seed=sd, precision=abi.MGK_FP32_REFINE, **geo) for sd in seeds]
    mats = [_field(int(np.prod(shape)), seed=17 + i) for i in range(len(seeds))]
    single = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    batched = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    for s in single + batched:
        s.elastic_run()
    st_single = [s.event_steps(events) for s in single]
    batch = SimulationBatch(batched)
    st_batch = batch.event_steps(events)
    for a, b, sa, sb in zip(single, batched, st_single, st_batch):
        assert sa.kmc_events == sb.kmc_events == events
        assert (sa.ls_iters, sa.ls_iters_fp32) == (sb.ls_iters, sb.ls_iters_fp32)
        ea, eb = a.events(), b.events()
        assert [e.flat_idx for e in ea] == [e.flat_idx for e in eb]
        assert [e.time for e in ea] == [e.time for e in eb]
        for f in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_Q0):
            assert np.array_equal(a.get_field(f), b.get_field(f)), f
    # against the FP64 oracle: the same events.  Each path stops at tol 1e-8
    # with its own truncation error, so the FP32+refine stress is compared
    # with the converged solution of the same event sequence (the oracle at
    # tol 1e-12): within 1e-6 of max |sigma|, or within twice the FP64 path's
    # own error at tol 1e-8 where that is larger.
    osim = po.OracleSim(cfgs[0], *mats[0])
    osim.elastic_run()
    osim.event_steps(events)
    conv = po.OracleSim(_cfg(seed=seeds[0], **(geo | dict(ls_tol=1e-12))), *mats[0])
    conv.elastic_run()
    conv.event_steps(events)
    flats = [e.flat_idx for e in batched[0].events()]
    assert [e.flat_idx for e in osim.events()] == flats == [e.flat_idx for e in conv.events()]
    s_conv = conv.get_field(abi.FIELD_SIG)
    err32 = _rel(batched[0].get_field(abi.FIELD_SIG), s_conv)
    err64 = _rel(osim.get_field(abi.FIELD_SIG), s_conv)
    assert err32 < max(1e-6, 2.0 * err64), (err32, err64)
    batch.close()


@pytest.mark.parametrize("kw,events,seeds", [
    (dict(), 25, (42, 7, 1234)),                                           # 2D fixed catalog, cascade mode
    (dict(instability_kmc=1, catalog_mode=abi.CATALOG_RANDOM, neighbor_softening_fraction=0.2), 25, (42, 7, 1234)),
    (dict(dims=3, nx=8, ny=8, nz=8, M=12), 10, (42, 7, 1234)),
    # FP64 iterations on the plane passes: the batched and single-context
    # instantiations of a pass must round alike
    (dict(dims=3, nx=16, ny=16, nz=16, M=12, instability_kmc=1), 6, (42, 7)),
    (dict(), 15, (42,)),                                                   # a batch of one
])
def test_batch_equals_separate_trajectories(kw, events, seeds):
    # trajectories advanced together (one launch sequence per step) equal the
    # same trajectories run one by one, bit for bit in FP64, and the oracle
    from mgkmc_cpp_amd import SimulationBatch
    cfgs = [_cfg(seed=sd, **kw) for sd in seeds]
    N = cfgs[0].nx * cfgs[0].ny * (cfgs[0].nz if cfgs[0].dims == 3 else 1)
    mats = [_field(N, seed=17 + i) for i in range(len(seeds))]
    single = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    batched = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    for s in single + batched:
        s.elastic_run()
    st_single = [s.event_steps(events) for s in single]
    batch = SimulationBatch(batched)
    st_batch = batch.event_steps(events)
    for a, b, sa, sb in zip(single, batched, st_single, st_batch):
        assert (sa.kmc_events, sa.cascades, sa.cascade_flips) == (sb.kmc_events, sb.cascades, sb.cascade_flips)
        ea, eb = a.events(), b.events()
        assert [e.flat_idx for e in ea] == [e.flat_idx for e in eb]
        assert [e.time for e in ea] == [e.time for e in eb]
        assert [list(e.sig_macro) for e in ea] == [list(e.sig_macro) for e in eb]
        for f in (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_SOFT, abi.FIELD_Q0):
            assert np.array_equal(a.get_field(f), b.get_field(f)), f
    # and the first trajectory against the oracle
    osim = po.OracleSim(cfgs[0], *mats[0])
    osim.elastic_run()
    so = osim.event_steps(events)
    assert so.kmc_events == st_batch[0].kmc_events
    assert [e.flat_idx for e in osim.events()] == [e.flat_idx for e in batched[0].events()]
    assert _rel(batched[0].get_field(abi.FIELD_SIG), osim.get_field(abi.FIELD_SIG)) < 1e-10
    batch.close()


@pytest.mark.parametrize("precision", [abi.MGK_FP64, abi.MGK_FP32_REFINE])
def test_batch_caps_leave_finished_trajectories_alone(precision):
    # per-trajectory event caps: a trajectory that reached its cap (or starts
    # at cap 0) sits in the batch as an inactive row while the others go on.
    # Every field of its state -- Q included, which an inactive row's rate pass
    # must not recompute -- equals the same trajectory run alone for its cap.
    from mgkmc_cpp_amd import SimulationBatch
    seeds, caps = (42, 7, 1234, 99), (3, 9, 0, 6)
    geo = dict(instability_kmc=1, catalog_mode=abi.CATALOG_RANDOM, neighbor_softening_fraction=0.2,
               precision=precision)
    cfgs = [_cfg(seed=sd, **geo) for sd in seeds]
    mats = [_field(256, seed=17 + i) for i in range(len(seeds))]
    single = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    batched = [Simulation(c, E, nu) for c, (E, nu) in zip(cfgs, mats)]
    for s in single + batched:
        s.elastic_run()
    st_single = [s.event_steps(cap) if cap > 0 else None for s, cap in zip(single, caps)]
    batch = SimulationBatch(batched)
    st_batch = batch.event_steps(max(caps), caps=caps)
    fields = _STATE_FIELDS + (abi.FIELD_Q,)
    for a, b, sa, sb, cap in zip(single, batched, st_single, st_batch, caps):
        assert sb.kmc_events == cap
        if sa is not None:
            assert sa.kmc_events == cap
            assert [e.flat_idx for e in a.events()] == [e.flat_idx for e in b.events()]
            assert [e.time for e in a.events()] == [e.time for e in b.events()]
        for f in fields:
            assert np.array_equal(a.get_field(f), b.get_field(f)), (cap, f)
    batch.close()


@pytest.mark.parametrize("precision", [abi.MGK_FP64, abi.MGK_FP32_REFINE])
def test_event_selection_edges_name_a_mode(precision):
    # the device-decided event always names a mode when the total rate is
    # positive.  Cascade mode with a threshold of -1 eV: the first rate block's
    # 1024 modes (Q0 = -0.5) are neither rated nor unstable.  u2 = 0 gives a
    # target of 0, which block 0 "reaches" with a prefix of 0 although it holds
    # no included mode: the answer is the first included mode after it (the
    # reference's first prefix >= 0).  u2 = 1 - 2^-53 puts the target at the
    # total, where the in-block sum can round below the block prefix.  Both
    # must equal the reference's left-to-right answer (mgk_select_exact), on
    # the FP64 path (exact pass) and on the FP32+refine chain (parallel pass).
    shape, M = (16, 16), 6
    N = 256
    E, nu = _field(N)
    q0 = np.full(N * M, 1.0)
    q0[:1024] = -0.5
    kw = dict(M=M, temperature=300.0, stability_threshold=-1.0)
    em = np.zeros((2, 2))
    for u2 in (0.0, 1.0 - 2.0 ** -53):
        ref = Context(shape, E, nu)
        ref.kmc_init(**kw)
        ref.set_field(abi.FIELD_Q0, q0)
        ref.elastic_run(em.copy(), tol=1e-8)
        R, n_unst, n_inc = ref.update_rates(0.0)
        assert R > 0 and n_unst == 0 and n_inc == N * M - 1024
        want, _ = ref.select(u2, exact=True)
        assert want == (1024 if u2 == 0.0 else N * M - 1)
        a = Context(shape, E, nu, precision=precision)
        a.kmc_init(**kw)
        a.set_field(abi.FIELD_Q0, q0)
        a.elastic_run(em.copy(), tol=1e-8)
        out, _, _ = a.kmc_event(em.copy(), 0.0, 1.0, u2, tol=1e-8, cascade_check=True)
        assert out.outcome == abi.EVENT_DONE, out.outcome
        assert out.flat_idx == want, (u2, out.flat_idx, want)
        ref.close()
        a.close()
