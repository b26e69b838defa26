"""The drop-in C++ headers (mgkmc_cpp_amd/host) against a caller written the
way the reference's main.cpp calls its own (tests/dropin/reference_style_caller.cpp:
green_operator_3d + spectral_solver_3d, main.cpp:540-580; BarrierGenerator +
the KmcSimulation3D constructor + run_simulation, main.cpp:379-395), built on
the drop-in's own tensor type and on a second layout-identical one that
stands for Eigen's (tests/dropin/layout_type.hpp).

CPU: both builds compile and their host-only functions (fftfreq, Lame
constants, the Green operator's projector identity, fft3d/ifft3d, the stress
helper) pass.  GPU: both builds run the elastic branch (sigma against the
oracle's LS solve within 1e-10) and the KMC branch (macro stress, time and
plastic strain against the oracle's run_simulation), and write the same bytes."""
import os
import subprocess

import numpy as np
import pytest

from mgkmc_cpp_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "tests", "dropin")
HOST = os.path.join(ROOT, "mgkmc_cpp_amd", "host")
LIB = os.path.join(ROOT, "mgkmc_cpp_amd", "lib")


@pytest.mark.parametrize("variant", ["builtin", "layout"])
def test_reference_style_caller_compiles_and_runs_host_part(variant, tmp_path):
    exe = str(tmp_path / ("caller_" + variant))
    extra = ["-DCALLER_LAYOUT_TYPE"] if variant == "layout" else []
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror"] + extra +
                       ["-I" + HOST, "-I" + DROPIN, os.path.join(DROPIN, "reference_style_caller.cpp"), "-o", exe,
                        "-L" + LIB, "-lmgkmc_gpu", "-Wl,-rpath," + LIB], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    r = subprocess.run([exe, "host"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0 and "host-only checks passed" in r.stdout, r.stdout


def _material(shape):
    N = int(np.prod(shape))
    rng = np.random.default_rng(5)
    return 85e9 * (1.0 + 0.1 * rng.standard_normal(N)), np.full(N, 0.35)


@pytest.mark.gpu
def test_reference_style_caller_runs_on_the_device(tmp_path):
    from oracle import pyoracle as po
    shape = (16, 16, 16)
    N = int(np.prod(shape))
    E, nu = _material(shape)
    inp = tmp_path / "in.bin"
    np.concatenate([np.array(shape, dtype=np.float64), E]).tofile(inp)
    outs = {}
    for variant in ("builtin", "layout"):
        exe = os.path.join(DROPIN, "_build", "caller_" + variant)
        assert os.path.exists(exe), f"{exe} is not built (python -m mgkmc_cpp_amd.build)"
        out = tmp_path / (variant + ".bin")
        print(f"[dropin] running {variant} caller", flush=True)
        r = subprocess.run([exe, "device", str(inp), str(out)], cwd=tmp_path, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=300)
        assert r.returncode == 0, r.stdout
        outs[variant] = out.read_bytes()
    assert outs["builtin"] == outs["layout"]
    v = np.frombuffer(outs["builtin"], dtype=np.float64)
    sigM = v[:9].reshape(3, 3, order="F")
    sig = v[9:9 + 9 * N].reshape(N, 3, 3).transpose(0, 2, 1)
    kmc = v[9 + 9 * N:]
    # elastic branch against the oracle's LS solve
    bar = np.zeros((3, 3))
    bar[0, 0] = 1e-3
    ref = po.ls_solve(shape, E, nu, bar, tol=1e-8)
    assert np.abs(sig - ref["sig"]).max() < 1e-10 * np.abs(ref["sig"]).max()
    assert np.abs(sigM - ref["sig"].mean(axis=0)).max() < 1e-10 * np.abs(ref["sig"]).max()
    # KMC branch against the oracle's run_simulation with the same settings
    cfg = abi.default_sim_config(dims=3, nx=16, ny=16, nz=16, M=6, gamma0=0.02, b_mean=1.0, b_std=0.1,
                                 temperature=330.0, instability_kmc=1, seed=42)
    run = abi.default_run_config(n_global_steps=4, step_size=1e-4, drv_i=0, drv_j=0, mixed_tol=1e6,
                                 mixed_max_iter=50, targets={(1, 1): 0.0, (2, 2): 0.0})
    o = po.OracleSim(cfg, E, nu)
    st = o.run(run)
    assert st.kmc_events > 5
    _, s_o, t_o = o.macro()
    s_d = kmc[:9].reshape(3, 3, order="F")
    assert np.abs(s_d - s_o).max() <= 1e-9 * np.abs(s_o).max()
    assert abs(kmc[9] - t_o) <= 1e-12 * abs(t_o)
    ep = o.get_field(abi.FIELD_EPS_PLASTIC).sum(axis=0)
    assert np.abs(kmc[10:19].reshape(3, 3, order="F") - ep).max() <= 1e-12
