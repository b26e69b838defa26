"""run_simulation's outputs from the device path: the summary and global logs
(init_logs / do_logging, kmc_solver.hpp:566-719) and the VTU files
(vtk_interval, vtk_elastic_only; final file :998-1015).

Logs: the headers are the reference's text; every KMC line of the global log
equals, byte for byte, the reference's printf format applied to the recorded
event's macros and time.  VTU: the files the interval asks for exist, and the
last one equals the reference writer (oracle restatement) applied to the
final fields."""
import re

import numpy as np
import pytest

from mgkmc_cpp_amd import Simulation, abi
from oracle import vtu_ref

pytestmark = pytest.mark.gpu

SUMMARY_HEADER = ("Timestamp              Elapsed(s)   Step     Type            Eps_xx       Sig_xx(GPa)     KMC      "
                  "Cascade  Flips    SimTime(s)")
GLOBAL_FMT = ("%-10d %-12d %-10d %-15.6f %-15.6f %-15.6f %-15.6f %-15.6f %-15.6f %-15.3f %-15.3f %-15.3f %-15.3f "
              "%-15.3f %-15.3f %-14d %-17d %-15.6e\n")


def _global_header():
    cols = [("GlobalStep", 10), ("ElasticStep", 12), ("KMCStep", 10)] + \
           [(n, 15) for n in ("Eps_xx", "Eps_yy", "Eps_zz", "Eps_xy", "Eps_xz", "Eps_yz", "Sig_xx(GPa)", "Sig_yy(GPa)",
                              "Sig_zz(GPa)", "Sig_xy(GPa)", "Sig_xz(GPa)", "Sig_yz(GPa)")] + \
           [("CascadeSteps", 14), ("TotalCascadeFlips", 17)]
    return "".join(n.ljust(w) + " " for n, w in cols) + "SimTime(s)\n"


def test_logs_and_vtu_of_a_run(tmp_path):
    c = abi.default_sim_config(dims=2, nx=16, ny=16, M=4, b_mean=1.3, b_std=0.1, b_min_cutoff=0.01, temperature=300.0,
                               catalog_mode=abi.CATALOG_RANDOM, gamma0=0.1, ls_tol=1e-8, seed=5)
    rng = np.random.default_rng(23)
    E = 85e9 * (1 + 0.15 * rng.standard_normal(256))
    nu = np.full(256, 0.35)
    sim = Simulation(c, E, nu)
    sim.set_output(tmp_path, vtk_interval="4", vtk_elastic_only=True, summary_log=True, global_log=True)
    st = sim.run(abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6))
    assert st.kmc_events > 5 and st.elastic_steps == 6

    summary = (tmp_path / "summary_log.txt").read_text().splitlines()
    assert summary[0] == SUMMARY_HEADER and summary[1] == "-" * len(SUMMARY_HEADER)
    assert re.match(r"^Total Duration: \d+\.\d\d seconds \(\d+h \d\dm \d\ds\)$", summary[-1])
    assert summary[-2].startswith("Simulation Finish Time: ")

    glog = (tmp_path / "global_log.txt").read_text().splitlines(keepends=True)
    assert glog[0] == _global_header()
    body = glog[1:]
    # KMC lines: no cascade rounds, one flip (ELASTIC: 0 0, CASCADE: rounds >= 1)
    kmc_lines = [l for l in body if l.split()[15] == "0" and l.split()[16] == "1"]
    events = sim.events()
    assert len(kmc_lines) == len(events) == st.kmc_events
    nu_avg = nu.mean()
    for k, (line, e) in enumerate(zip(kmc_lines, events)):
        em, sm = np.array(e.eps_macro[:4]), np.array(e.sig_macro[:4])   # column-major xx, yx, xy, yy
        gstep, estep = int(line.split()[0]), int(line.split()[1])
        want = GLOBAL_FMT % (gstep, estep, k, em[0], em[3], 0.0, em[2], 0.0, 0.0, sm[0] / 1e9, sm[3] / 1e9,
                             nu_avg * (sm[0] + sm[3]) / 1e9, sm[2] / 1e9, 0.0, 0.0, 0, 1, e.time)
        assert line == want

    # VTU: step 0, every 4th elastic step, and a final file unless the last step was one
    steps = [(int(l.split()[0]), int(l.split()[1])) for l in body]
    assert (tmp_path / "vtk_step_000000.vtu").exists()
    last_global = steps[-1][0]
    for g, es in steps:
        if es == 4 and (tmp_path / f"vtk_step_{g:06d}.vtu").exists():
            break
    else:
        raise AssertionError("no VTU for elastic step 4")
    final = tmp_path / f"vtk_step_{last_global:06d}_final.vtu"
    assert final.exists() == (last_global % 4 != 0)

    # vtk_interval "last": one file, the final fields
    out2 = tmp_path / "last"
    sim2 = Simulation(c, E, nu)
    sim2.set_output(out2, vtk_interval="last")
    sim2.run(abi.default_run_config(n_global_steps=6, step_size=3e-3, targets={(1, 1): 0.0}, mixed_tol=1e6))
    files = sorted(p.name for p in out2.iterdir())
    assert files == [f"vtk_step_{last_global:06d}_final.vtu"]
    want = vtu_ref.export_to_vtu_2d(16, 16, sim2.get_field(abi.FIELD_EPS), sim2.get_field(abi.FIELD_SIG), E, nu, 1.0,
                                    sim2.get_field(abi.FIELD_EPS_PLASTIC), sim2.get_field(abi.FIELD_SOFT))
    assert (out2 / files[0]).read_text() == want
