"""Checkpoint and resume (no counterpart in the reference, whose
output.checkpoint_interval is parsed and ignored, main.cpp:252-258).

A run continued from a checkpoint must equal the uninterrupted run bit for
bit: fields, macros, time, counters and the event sequence, for the FP64 and
FP32+refine paths (the latter also carries the CG tuning state), at strain-step
boundaries and inside a strain step (event cap)."""
import numpy as np
import pytest

from mgkmc_cpp_amd import Context, MgkError, Simulation, abi

pytestmark = pytest.mark.gpu

FIELDS = (abi.FIELD_EPS, abi.FIELD_SIG, abi.FIELD_EPS_PLASTIC, abi.FIELD_Q, abi.FIELD_Q0, abi.FIELD_CATALOG,
          abi.FIELD_SOFT, abi.FIELD_LAST_EVENT_TIME, abi.FIELD_PREV_DIR)


def _cfg(precision):
    return abi.default_sim_config(dims=2, nx=16, ny=16, M=4, b_mean=1.3, b_std=0.1, b_min_cutoff=0.01,
                                  temperature=300.0, catalog_mode=abi.CATALOG_RANDOM, gamma0=0.1, ls_tol=1e-8,
                                  seed=5, precision=precision)


def _material():
    rng = np.random.default_rng(23)
    return 85e9 * (1 + 0.15 * rng.standard_normal(256)), np.full(256, 0.35)


def _same(a, b):
    for f in FIELDS:
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
    ea, sa, ta = a.macro()
    eb, sb, tb = b.macro()
    assert np.array_equal(ea, eb) and np.array_equal(sa, sb) and ta == tb
    assert [(e.flat_idx, e.time, e.ls_iters) for e in a.events()] == [(e.flat_idx, e.time, e.ls_iters) for e in b.events()]


@pytest.mark.parametrize("precision", [abi.MGK_FP64, abi.MGK_FP32_REFINE])
def test_resume_at_step_boundary(tmp_path, precision):
    E, nu = _material()
    # oracle, FP64: 20 events and 6 cascades over the 6 strain steps
    run = abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6)
    a = Simulation(_cfg(precision), E, nu)
    a.set_checkpoint(2, tmp_path / "ck_{step}")
    sa = a.run(run)
    assert all((tmp_path / f"ck_{k}.mgkc").exists() for k in (2, 4, 6))
    assert sa.kmc_events > 5 and sa.elastic_steps == 6
    b = Simulation(_cfg(precision), E, nu)
    b.load(tmp_path / "ck_2.mgkc")
    sb = b.run(run)
    assert (sb.kmc_events, sb.cascades, sb.cascade_flips, sb.elastic_steps) == \
           (sa.kmc_events, sa.cascades, sa.cascade_flips, sa.elastic_steps)
    _same(a, b)


@pytest.mark.parametrize("precision", [abi.MGK_FP64, abi.MGK_FP32_REFINE])
def test_resume_inside_a_strain_step(tmp_path, precision):
    E, nu = _material()
    full = abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6,
                                  max_events=18)
    half = abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6,
                                  max_events=9)
    ref = Simulation(_cfg(precision), E, nu)
    sr = ref.run(full)
    assert sr.kmc_events == 18
    a = Simulation(_cfg(precision), E, nu)
    a.set_checkpoint("end", tmp_path / "stop")
    assert a.run(half).kmc_events == 9
    b = Simulation(_cfg(precision), E, nu)
    b.load(tmp_path / "stop.mgkc")
    sb = b.run(full)
    assert (sb.kmc_events, sb.cascades, sb.elastic_steps) == (sr.kmc_events, sr.cascades, sr.elastic_steps)
    _same(ref, b)


def test_state_roundtrip_and_checks():
    E, nu = _material()
    a = Simulation(_cfg(abi.MGK_FP64), E, nu)
    a.event_steps(10)
    blob = a.state_save()
    b = Simulation(_cfg(abi.MGK_FP64), E, nu)
    b.state_load(blob)
    for f in FIELDS:
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
    c = Simulation(_cfg(abi.MGK_FP64), E * 1.01, nu)
    with pytest.raises(MgkError, match="another material"):
        c.state_load(blob)
    with pytest.raises(MgkError, match="magic"):
        c.state_load(b"x" * len(blob))
    d = Context((16, 16), E, nu)   # no KMC state
    with pytest.raises(MgkError, match="mode count"):
        d.state_load(blob)
