"""Pins the oracle's KMC restatement (kmc_solver.hpp) against brute-force
Python written from the reference's formulas and against the random stream
reproduced from numpy's MT19937: constructor draws, update_barriers, rates,
selection, apply_flip (with periodic neighbour duplicates) and run-loop
termination.  CPU only."""
import math

import numpy as np

from mgkmc_cpp_amd import abi
from oracle import pyoracle as po


class Stream:
    """libstdc++ mt19937 + generate_canonical + normal_distribution (polar,
    cached per object), rebuilt from numpy's independent MT19937."""

    def __init__(self, seed):
        bg = np.random.MT19937()
        bg._legacy_seeding(seed)
        self.w = bg.random_raw(200000).astype(np.uint64)
        self.pos = 0

    def U(self):
        a, b = float(self.w[self.pos]), float(self.w[self.pos + 1])
        self.pos += 2
        r = (a + b * 2.0 ** 32) / 2.0 ** 64
        return r if r < 1.0 else math.nextafter(1.0, 0.0)

    class Normal:
        def __init__(self, s, mean=0.0, sd=1.0):
            self.s, self.mean, self.sd, self.saved = s, mean, sd, None

        def __call__(self):
            if self.saved is not None:
                r, self.saved = self.saved, None
            else:
                while True:
                    x = 2.0 * self.s.U() - 1.0
                    y = 2.0 * self.s.U() - 1.0
                    r2 = x * x + y * y
                    if not (r2 > 1.0 or r2 == 0.0):
                        break
                m = math.sqrt(-2 * math.log(r2) / r2)
                self.saved = x * m
                r = y * m
            return r * self.sd + self.mean


def _cfg(**kw):
    base = dict(dims=2, nx=4, ny=4, M=3, gamma0=0.14, b_mean=1.0, b_std=0.1, temperature=300.0,
                instability_kmc=1, seed=42)
    base.update(kw)
    return abi.default_sim_config(**base)


def test_constructor_draw_order_2d():
    c = _cfg()
    N = 16
    sim = po.OracleSim(c, np.full(N, 85e9), np.full(N, 0.35))
    s = Stream(42)
    std_normal = Stream.Normal(s)          # one object for all catalog draws (kmc_solver.hpp:223)
    Q0 = []
    cat = []
    for v in range(N):
        for m in range(c.M):
            Q0.append(max(Stream.Normal(s, 1.0, 0.1)(), 0.1))   # fresh object per barrier (:38), clipped (:88)
            gxx = 0.5 * 0.14 * std_normal()
            gxy = 0.5 * 0.14 * std_normal()
            cat.append([[gxx, gxy], [gxy, -gxx]])
    assert np.array_equal(sim.get_field(abi.FIELD_Q0), np.array(Q0))
    assert np.array_equal(sim.get_field(abi.FIELD_CATALOG), np.array(cat))


def test_constructor_draw_order_3d_glass():
    c = _cfg(dims=3, nx=2, ny=2, nz=2, M=2, use_3d_barriers=1, seed=9)
    N = 8
    sim = po.OracleSim(c, np.full(N, 85e9), np.full(N, 0.35))
    s = Stream(9)
    Q0, cat = [], []
    for v in range(N):
        for m in range(c.M):
            Q0.append(max(Stream.Normal(s, 1.0, 0.1)(), 0.1))
            g = Stream.Normal(s)             # fresh per mode (kmc_solver.hpp:1055)
            Mx = np.array([[g() for j in range(3)] for i in range(3)])
            S = 0.5 * (Mx + Mx.T)
            G = S - (1.0 / 3.0) * (S[0, 0] + S[1, 1] + S[2, 2]) * np.eye(3)
            n = math.sqrt(sum(G[i, j] ** 2 for j in range(3) for i in range(3)))
            G = G * (1.0 / (math.sqrt(2.0) * n))
            u = 1e-15 + (1.0 - 1e-15) * s.U()
            cat.append(G * (0.14 * math.sqrt(-2.0 * math.log(u))))
    assert np.array_equal(sim.get_field(abi.FIELD_Q0), np.array(Q0))
    assert np.abs(sim.get_field(abi.FIELD_CATALOG) - np.array(cat)).max() < 1e-17


def _brute_Q(sim, c, time, tau):
    N, M, D = sim.N, sim.M, sim.D
    sig = sim.get_field(abi.FIELD_SIG)
    cat = sim.get_field(abi.FIELD_CATALOG)
    Q0 = sim.get_field(abi.FIELD_Q0)
    soft = sim.get_field(abi.FIELD_SOFT)
    last = sim.get_field(abi.FIELD_LAST_EVENT_TIME)
    prev = sim.get_field(abi.FIELD_PREV_DIR)
    V = c.pixel ** 3
    Q = np.zeros(N * M)
    for v in range(N):
        g_t = soft[v, 1]
        gtc = 0.0
        if g_t > 0.0:
            if math.isinf(tau):
                gtc = g_t
            elif last[v] == -math.inf:
                gtc = 0.0
            else:
                gtc = g_t * math.exp(-max(time - last[v], 0.0) / tau)
        gb = soft[v, 0] + gtc
        if c.softening_cap > 0.0 and gb > c.softening_cap:
            gb = c.softening_cap
        for m in range(M):
            i = v * M + m
            mod = 1.0
            if c.softening_directional:
                dot = nm = npp = 0.0
                for a in range(D):
                    for b in range(D):
                        dot += cat[i, a, b] * prev[v, a, b]
                        nm += cat[i, a, b] * cat[i, a, b]
                        npp += prev[v, a, b] * prev[v, a, b]
                if not (math.sqrt(npp) < 1e-12 or math.sqrt(nm) < 1e-12):
                    ct = dot / (math.sqrt(nm) * math.sqrt(npp))
                    mod = (1.0 + ct) ** 2 / 4.0
            w = 0.0
            for a in range(D):
                for b in range(D):
                    w += sig[v, a, b] * cat[i, a, b]
            Q[i] = Q0[i] * math.exp(-(mod * gb)) - 0.5 * V * (w / 1e9) * 6.241509
    return Q


def _perturbed_state(sim, seed):
    rng = np.random.default_rng(seed)
    N, D = sim.N, sim.D
    s = rng.standard_normal((N, D, D)) * 1e9
    sim.set_field(abi.FIELD_SIG, 0.5 * (s + np.transpose(s, (0, 2, 1))))
    sim.set_field(abi.FIELD_SOFT, np.abs(rng.standard_normal((N, 2))) * 0.3)
    last = rng.uniform(0.0, 2.0, N)
    last[rng.random(N) < 0.3] = -np.inf
    sim.set_field(abi.FIELD_LAST_EVENT_TIME, last)
    p = rng.standard_normal((N, D, D)) * 0.05
    p[rng.random(N) < 0.3] = 0.0
    sim.set_field(abi.FIELD_PREV_DIR, 0.5 * (p + np.transpose(p, (0, 2, 1))))


def test_update_barriers_brute_force():
    for dims, directional in ((2, 0), (2, 1), (3, 1)):
        kw = dict(dims=dims, softening_directional=directional, q_act_temp=0.37)
        if dims == 3:
            kw.update(nx=2, ny=3, nz=2)
        c = _cfg(**kw)
        N = c.nx * c.ny * (c.nz if dims == 3 else 1)
        sim = po.OracleSim(c, np.full(N, 85e9), np.full(N, 0.35))
        _perturbed_state(sim, 3)
        sim.set_macro(np.zeros((dims, dims)), 1.5)
        sim.update_barriers()
        tau = sim.tau()
        assert abs(tau - 1.0 / (1e13 * math.exp(-0.37 / (8.617e-5 * 300.0)))) < 1e-15 * tau
        assert np.array_equal(sim.get_field(abi.FIELD_Q), _brute_Q(sim, c, 1.5, tau))


def test_rates_and_selection_brute_force():
    for imode in (0, 1):
        c = _cfg(instability_kmc=imode, b_mean=0.02, b_std=0.05, b_min_cutoff=-1.0)
        sim = po.OracleSim(c, np.full(16, 85e9), np.full(16, 0.35))
        _perturbed_state(sim, 5)
        sim.update_barriers()
        Q = sim.get_field(abi.FIELD_Q)
        beta = 1.0 / (8.617e-5 * 300.0)
        rates, idx, tot = [], [], 0.0
        for i, q in enumerate(Q):
            if imode == 1 or q > 0.0:
                r = 1e13 if q <= 0.0 else 1e13 * math.exp(-q * beta)
                rates.append(r)
                idx.append(i)
                tot += r
        total, r_o, i_o = sim.rates()
        assert total == tot and np.array_equal(r_o, rates) and np.array_equal(i_o, idx)
        # first prefix >= target, default last (kmc_solver.hpp:816-826)
        pref = np.cumsum(rates)   # numpy cumsum is sequential left to right
        for target in list(pref[:5]) + [0.0, tot, tot * 0.37, np.nextafter(pref[2], 0.0)]:
            want = next((k for k, s in enumerate(pref) if s >= target), len(rates) - 1)
            assert po.select_sequential(np.array(rates), target) == want


def test_apply_flip_neighbour_duplicates():
    # 2x2 periodic grid: the 8 neighbours of a voxel hit 3 distinct voxels,
    # 3 or 2 times each, and every hit accumulates (kmc_solver.hpp:444-461)
    c = _cfg(nx=2, ny=2, M=2, neighbor_softening_fraction=0.25, softening_cap=5.0)
    sim = po.OracleSim(c, np.full(4, 85e9), np.full(4, 0.35))
    cat = sim.get_field(abi.FIELD_CATALOG)
    flat = 3 * 2 + 1   # voxel 3 = (x=1, y=1), mode 1
    e = cat[flat]
    e11, e22, e12 = e[0, 0], e[1, 1], e[0, 1]
    ss = (e12 * e12) + (e22 * e22 + e11 * e11 + (e11 - e22) ** 2) / 6.0
    sim.apply_flip(flat, 0.75)
    soft = sim.get_field(abi.FIELD_SOFT)
    gp = np.zeros(4)
    gt = np.zeros(4)
    gp[3] = min(0.0 + 10.0 * ss, 5.0)
    gt[3] = 30.0 * ss
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            if dx == 0 and dy == 0:
                continue
            nb = ((1 + dx + 2) % 2) * 2 + (1 + dy + 2) % 2
            gp[nb] = min(gp[nb] + 0.25 * 10.0 * ss, 5.0)
            gt[nb] += 0.25 * 30.0 * ss
    assert np.array_equal(soft[:, 0], gp) and np.array_equal(soft[:, 1], gt)
    assert np.array_equal(sim.get_field(abi.FIELD_EPS_PLASTIC)[3], e)
    assert np.array_equal(sim.get_field(abi.FIELD_PREV_DIR)[3], e)
    assert sim.get_field(abi.FIELD_LAST_EVENT_TIME)[3] == 0.75


def test_fixed_catalog_magnitudes():
    for dims, M in ((2, 6), (3, 12)):
        for m in range(M):
            a = po.fixed_mode(dims, m, M, 0.1)
            assert abs(np.trace(a)) < 1e-17 and np.array_equal(a, a.T)
            assert abs(np.sqrt((a * a).sum()) - 0.1 / np.sqrt(2.0)) < 1e-16
    # 2D angles k*30 degrees: (g/2)[[cos 2t, sin 2t], [sin 2t, -cos 2t]]
    a = po.fixed_mode(2, 1, 6, 0.1)
    assert abs(a[0, 0] - 0.05 * math.cos(math.pi / 3)) < 1e-17 and abs(a[0, 1] - 0.05 * math.sin(math.pi / 3)) < 1e-17


def test_event_cap_and_instability_termination():
    N = 64
    c = _cfg(nx=8, ny=8, M=2, b_mean=1.0, b_std=0.1, catalog_mode=abi.CATALOG_FIXED, ls_tol=1e-6)
    sim = po.OracleSim(c, np.full(N, 85e9), np.full(N, 0.35))
    st = sim.run(abi.default_run_config(n_global_steps=10 ** 9, step_size=1e4, max_events=7))
    assert st.event_cap_reached == 1 and st.kmc_events == 7 and len(sim.events()) == 7
    # every mode "unstable" (Q <= thr) in kmc mode: more than pct*N consecutive
    # KMC_INSTAB events terminate the run (kmc_solver.hpp:918-929)
    c2 = _cfg(nx=8, ny=8, M=2, stability_threshold=100.0, temperature=0.0, ls_tol=1e-6)
    sim2 = po.OracleSim(c2, np.full(N, 85e9), np.full(N, 0.35))
    st2 = sim2.run(abi.default_run_config(n_global_steps=10, step_size=1e4, max_kmc_steps_pct=0.05))
    assert st2.terminated == 1
    assert st2.kmc_events == int(0.05 * N)   # the terminating event is not counted


# ---------------------------------------------------------------------------
# run_cascade (kmc_solver.hpp:464-564) restated in numpy on a 4x4 grid: the
# whole loop -- update_barriers, find_unstable_indices, the sort by Q, one
# flip per voxel per round (the first, most negative, hit), the flips and the
# redraw draws in sorted order, an elastic_run per round, the local_step > N
# stop -- against the oracle's C++ restatement.  The LS solve here is a
# numpy transcription of spectral_solver_2d (linear_elastic_solver.hpp:146-309)
# with numpy.fft, independent of the oracle's own FFT.
# ---------------------------------------------------------------------------
def _np_green_2d(nx, ny, pixel, lam0, mu0):
    xi = [2.0 * math.pi * np.fft.fftfreq(n, d=pixel) for n in (nx, ny)]   # Nyquist at -n/2, as fftfreq (:22-33)
    A = 1.0 / (4.0 * mu0)
    B = (lam0 + mu0) / (mu0 * (lam0 + 2.0 * mu0))
    G = np.zeros((nx, ny, 2, 2, 2, 2))
    d = np.eye(2)
    for x in range(nx):
        for y in range(ny):
            q = np.array([xi[0][x], xi[1][y]])
            n2 = q @ q
            if n2 == 0.0:
                continue
            for k in range(2):
                for h in range(2):
                    for i in range(2):
                        for j in range(2):
                            G[x, y, k, h, i, j] = (A * (d[k, i] * q[h] * q[j] + d[h, i] * q[k] * q[j] + d[k, j] * q[h] * q[i]
                                                        + d[h, j] * q[k] * q[i]) / n2 - B * q[k] * q[h] * q[i] * q[j] / n2 ** 2)
    return G


class _NpCascadeModel:
    """KmcSimulation2D's state and the parts of it run_cascade touches."""

    def __init__(self, osim, c, E, nu, stream):
        self.c, self.s = c, stream
        self.nx, self.ny, self.M = c.nx, c.ny, c.M
        self.N = c.nx * c.ny
        self.E, self.nu = np.asarray(E, float), np.asarray(nu, float)
        self.eps = osim.get_field(abi.FIELD_EPS)
        self.sig = osim.get_field(abi.FIELD_SIG)
        self.eps_p = osim.get_field(abi.FIELD_EPS_PLASTIC)
        self.Q0 = osim.get_field(abi.FIELD_Q0)
        self.cat = osim.get_field(abi.FIELD_CATALOG)
        self.soft = osim.get_field(abi.FIELD_SOFT)
        self.last = osim.get_field(abi.FIELD_LAST_EVENT_TIME)
        self.prev = osim.get_field(abi.FIELD_PREV_DIR)
        self.eps_macro, _, self.time = osim.macro()
        self.tau = osim.tau()
        ea = sum(self.E) / self.N
        na = sum(self.nu) / self.N
        mu0 = ea / (2.0 * (1.0 + na))
        lam0 = ea * na / ((1.0 + na) * (1.0 - 2.0 * na))
        self.lam0, self.mu0 = lam0, mu0
        self.mu = self.E / (2.0 * (1.0 + self.nu))
        self.lam = self.E * self.nu / ((1.0 + self.nu) * (1.0 - 2.0 * self.nu))
        self.G = _np_green_2d(c.nx, c.ny, c.pixel, lam0, mu0).reshape(self.N, 2, 2, 2, 2)

    def _stress(self, eps):
        e = eps - self.eps_p
        tr = e[:, 0, 0] + e[:, 1, 1]
        return 2.0 * self.mu[:, None, None] * e + (self.lam * tr)[:, None, None] * np.eye(2)

    def elastic_run(self, tol=1e-6, max_iter=200):
        # spectral_solver_2d with eps_bar = eps_macro, warm start (:180-191)
        bar = self.eps_macro.copy()
        eps = self.eps + (bar - self.eps.mean(axis=0))
        for _ in range(max_iter):
            sig = self._stress(eps)
            tr = eps[:, 0, 0] + eps[:, 1, 1]
            tau = sig - (2.0 * self.mu0 * eps + (self.lam0 * tr)[:, None, None] * np.eye(2))
            th = np.fft.fft2(tau.reshape(self.nx, self.ny, 2, 2), axes=(0, 1), norm="ortho").reshape(self.N, 2, 2)
            eh = -np.einsum("nkhrs,nrs->nkh", self.G, th)
            et = np.fft.ifft2(eh.reshape(self.nx, self.ny, 2, 2), axes=(0, 1), norm="ortho").real.reshape(self.N, 2, 2)
            new = bar + (et - et.mean(axis=0))
            diff = math.sqrt(((new - eps) ** 2).sum()) / (math.sqrt((eps ** 2).sum()) + 1e-20)
            eps = new
            if diff < tol:
                break
        self.eps = eps
        self.sig = self._stress(eps)
        self.eps_macro = eps.mean(axis=0)

    def Q(self):
        c, M = self.c, self.M
        Q = np.zeros(self.N * M)
        for v in range(self.N):
            g_t = self.soft[v, 1]
            gtc = 0.0
            if g_t > 0.0:
                if math.isinf(self.tau):
                    gtc = g_t
                elif self.last[v] != -math.inf:
                    gtc = g_t * math.exp(-max(self.time - self.last[v], 0.0) / self.tau)
            gb = self.soft[v, 0] + gtc
            if c.softening_cap > 0.0 and gb > c.softening_cap:
                gb = c.softening_cap
            for m in range(M):
                i = v * M + m
                w = sum(self.sig[v, a, b] * self.cat[i, a, b] for a in range(2) for b in range(2))
                Q[i] = self.Q0[i] * math.exp(-gb) - 0.5 * c.pixel ** 3 * (w / 1e9) * 6.241509
        return Q

    def flip(self, v, m):
        c, i = self.c, v * self.M + m
        e = self.cat[i]
        self.eps_p[v] = self.eps_p[v] + e
        e11, e22, e12 = e[0, 0], e[1, 1], e[0, 1]
        ss = (e12 * e12) + (e22 * e22 + e11 * e11 + (e11 - e22) ** 2) / 6.0
        gp = self.soft[v, 0] + c.jp * ss
        self.soft[v, 0] = min(gp, c.softening_cap) if c.softening_cap > 0.0 else gp
        self.soft[v, 1] = c.jt * ss
        self.last[v] = self.time
        if c.neighbor_softening_fraction > 0.0:
            x, y = divmod(v, self.ny)
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    if dx == 0 and dy == 0:
                        continue
                    nb = ((x + dx + self.nx) % self.nx) * self.ny + (y + dy + self.ny) % self.ny
                    g = self.soft[nb, 0] + c.neighbor_softening_fraction * c.jp * ss
                    self.soft[nb, 0] = min(g, c.softening_cap) if c.softening_cap > 0.0 else g
                    self.soft[nb, 1] += c.neighbor_softening_fraction * c.jt * ss

    def redraw(self, v):
        # redraw_directions and redraw_barriers (:494-512): a fresh normal
        # object per flip; per mode gxx, gxy, then a gaussian barrier
        # (BarrierGenerator: fresh object, clipped at min_cutoff)
        c = self.c
        sn = Stream.Normal(self.s)
        for mm in range(self.M):
            i = v * self.M + mm
            gxx = 0.5 * c.gamma0 * sn()
            gxy = 0.5 * c.gamma0 * sn()
            self.cat[i] = [[gxx, gxy], [gxy, -gxx]]
            self.Q0[i] = max(Stream.Normal(self.s, c.b_mean, c.b_std)(), c.b_min_cutoff)

    def run_cascade(self, rounds_log):
        total, local, limit = 0, 0, self.N
        while True:
            Q = self.Q()
            unstable = [i for i in range(self.N * self.M) if Q[i] < self.c.stability_threshold]   # ascending (:412-421)
            if not unstable:
                break
            assert len({Q[i] for i in unstable}) == len(unstable), "ties would leave std::sort's order open"
            unstable.sort(key=lambda i: Q[i])
            done, flips = set(), []
            for i in unstable:
                v, m = divmod(i, self.M)
                if v in done:
                    continue
                done.add(v)
                flips.append(i)
                C = self.cat[i].copy()
                self.flip(v, m)
                self.prev[v] = C
                self.redraw(v)
            rounds_log.append(dict(flips=flips, Q=Q, eps_p=self.eps_p.copy()))
            self.elastic_run()
            total += len(flips)
            local += 1
            if local > limit:
                break
        return local, total


def _cascade_setup(**kw):
    base = dict(nx=4, ny=4, M=3, instability_kmc=0, neighbor_softening_fraction=0.2, seed=5)
    base.update(kw)
    c = _cfg(**base)
    N = 16
    rng = np.random.default_rng(11)
    E = 85e9 * (1.0 + 0.15 * rng.standard_normal(N))
    nu = np.full(N, 0.35)
    osim = po.OracleSim(c, E, nu)
    s = Stream(c.seed)
    std_normal = Stream.Normal(s)   # the constructor's draws (:223-237) put the stream where the oracle's is
    for _ in range(N * c.M):
        Stream.Normal(s, c.b_mean, c.b_std)()
        std_normal()
        std_normal()
    return c, E, nu, osim, s


def test_run_cascade_brute_force_4x4():
    c, E, nu, osim, s = _cascade_setup()
    # unstable modes with distinct barriers; voxel 2 holds two (the more
    # negative, mode 1, must be the one flipped), voxel 11 two as well
    Q0 = osim.get_field(abi.FIELD_Q0)
    for i, q in ((2 * 3 + 0, -0.30), (2 * 3 + 1, -0.35), (5 * 3 + 2, -0.10), (11 * 3 + 0, -0.20), (11 * 3 + 2, -0.05)):
        Q0[i] = q
    osim.set_field(abi.FIELD_Q0, Q0)
    osim.set_macro(np.array([[1e-3, 0.0], [0.0, 0.0]]), 0.5)
    osim.elastic_run()
    model = _NpCascadeModel(osim, c, E, nu, s)
    cat_before = model.cat.copy()
    log = []
    want = model.run_cascade(log)
    got = osim.run_cascade()
    assert got == want, (got, want)
    assert want[1] >= 4 and want[0] >= 1
    # the first round flips in ascending Q, one mode per voxel: mode 1 of
    # voxel 2 (-0.35) before mode 0 (-0.30, skipped), mode 0 of voxel 11
    # (-0.20) and not mode 2 (-0.05)
    r0 = log[0]
    assert [r0["Q"][i] for i in r0["flips"]] == sorted(r0["Q"][i] for i in r0["flips"])
    assert r0["flips"][0] == 7 and 6 not in r0["flips"]
    assert 33 in r0["flips"] and 35 not in r0["flips"] and 17 in r0["flips"]
    assert np.array_equal(r0["eps_p"][2], cat_before[7])
    # every flip and redraw in the reference's order: the random draws land
    # on the same voxels, so Q0 and the catalog are equal bit for bit
    assert np.array_equal(osim.get_field(abi.FIELD_Q0), model.Q0)
    assert np.array_equal(osim.get_field(abi.FIELD_CATALOG), model.cat)
    assert np.array_equal(osim.get_field(abi.FIELD_EPS_PLASTIC), model.eps_p)
    assert np.array_equal(osim.get_field(abi.FIELD_SOFT), model.soft)
    assert np.array_equal(osim.get_field(abi.FIELD_PREV_DIR), model.prev)
    assert np.array_equal(osim.get_field(abi.FIELD_LAST_EVENT_TIME), model.last)
    sig = osim.get_field(abi.FIELD_SIG)
    assert np.abs(sig - model.sig).max() < 1e-11 * np.abs(model.sig).max()


def test_run_cascade_stops_after_n_plus_one_rounds():
    # a threshold above every barrier keeps all modes unstable: each round
    # flips every voxel once, and the loop stops when local_step > N
    # (kmc_solver.hpp:556-557): N + 1 rounds of N flips
    c, E, nu, osim, s = _cascade_setup(stability_threshold=1e3, neighbor_softening_fraction=0.0)
    osim.set_macro(np.array([[1e-3, 0.0], [0.0, 0.0]]), 0.0)
    osim.elastic_run()
    model = _NpCascadeModel(osim, c, E, nu, s)
    log = []
    want = model.run_cascade(log)
    assert want == (17, 17 * 16)
    assert all(len(r["flips"]) == 16 for r in log)
    assert osim.run_cascade() == want
    assert np.array_equal(osim.get_field(abi.FIELD_Q0), model.Q0)
    assert np.array_equal(osim.get_field(abi.FIELD_EPS_PLASTIC), model.eps_p)
