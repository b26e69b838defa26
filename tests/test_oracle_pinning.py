"""Pins the CPU oracle (oracle/mgkmc_oracle.cpp) against what is fixed
independently of it: numpy's FFT (pocketfft C), the reference's own vendored
pocketfft (oracle/_ref, built from the reference tree), brute-force DFTs,
closed forms of the Green operator and of the homogeneous-medium
Lippmann-Schwinger solution (Navier equation in Fourier space), equilibrium
of converged heterogeneous fields, and numpy's own MT19937 for the random
stream.  CPU only."""
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HAVE_REF = os.path.exists(po.REF_LIB)


# ---------------------------------------------------------------------------
# FFT (pocketfft c2c with ortho scaling, linear_elastic_solver.hpp:122-143)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(4, 4), (8, 8), (6, 10), (7, 9), (48,), (32, 32), (8, 8, 8), (12, 6, 5),
                                   (11, 13), (49, 3)])
def test_fft_matches_numpy_pocketfft(shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = po.fft_c2c(x, True)
    b = po.fft_c2c(x, False)
    assert np.abs(f - np.fft.fftn(x, norm="ortho")).max() < 1e-14 * np.abs(f).max()
    assert np.abs(b - np.fft.ifftn(x, norm="ortho")).max() < 1e-14 * np.abs(b).max()


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("shape", [(32, 32), (64, 64, 64), (48, 48), (15, 9)])
def test_fft_matches_reference_pocketfft(shape):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for fwd in (True, False):
        a = po.fft_c2c(x, fwd)
        r = po.fft_c2c(x, fwd, ref=True)
        assert np.abs(a - r).max() < 1e-14 * np.abs(r).max()


def test_fft_brute_force_dft():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 6)) + 1j * rng.standard_normal((4, 6))
    n0, n1 = x.shape
    out = np.zeros_like(x)
    for k0 in range(n0):
        for k1 in range(n1):
            s = 0j
            for j0 in range(n0):
                for j1 in range(n1):
                    s += x[j0, j1] * np.exp(-2j * np.pi * (k0 * j0 / n0 + k1 * j1 / n1))
            out[k0, k1] = s / np.sqrt(x.size)
    assert np.abs(po.fft_c2c(x, True) - out).max() < 1e-13


# ---------------------------------------------------------------------------
# Green operator (linear_elastic_solver.hpp:70-119, :344-396)
# ---------------------------------------------------------------------------
def _C0(D, lam, mu):
    I = np.eye(D)
    return (lam * np.einsum("ij,kl->ijkl", I, I) + mu * (np.einsum("ik,jl->ijkl", I, I)
                                                         + np.einsum("il,jk->ijkl", I, I)))


@pytest.mark.parametrize("shape", [(6, 8), (5, 4, 6)])
def test_green_projector_identities(shape):
    D = len(shape)
    lam0, mu0 = po.lame(D, 85e9, 0.35)
    G = po.green(shape, 0.7, lam0, mu0)
    C0 = _C0(D, lam0, mu0)
    assert np.all(G[0] == 0.0)
    rng = np.random.default_rng(4)
    for idx in rng.choice(np.arange(1, G.shape[0]), 12, replace=False):
        g = G[idx]
        # minor and major symmetries
        # (k, h) symmetry is bitwise: term1's summation order is unchanged under k <-> h
        assert np.array_equal(g, np.transpose(g, (1, 0, 2, 3)))
        scale = np.abs(g).max()
        assert np.abs(g - np.transpose(g, (0, 1, 3, 2))).max() < 1e-15 * scale
        assert np.abs(g - np.transpose(g, (2, 3, 0, 1))).max() < 1e-15 * scale
        # projector: Gamma0 : C0 : Gamma0 = Gamma0
        ggg = np.einsum("abij,ijkl,klcd->abcd", g, C0, g)
        assert np.abs(ggg - g).max() < 1e-12 * np.abs(g).max()
        # Gamma0 : C0 : sym(xi (x) a) = sym(xi (x) a) for the voxel's own xi
        coords = np.unravel_index(idx, shape)
        xi = np.array([(c if c < (n + 1) // 2 else c - n) / (n * 0.7) for c, n in zip(coords, shape)]) * 2 * np.pi
        a = rng.standard_normal(D)
        s = 0.5 * (np.outer(xi, a) + np.outer(a, xi))
        back = np.einsum("abij,ijkl,kl->ab", g, C0, s)
        assert np.abs(back - s).max() < 1e-12 * np.abs(s).max()


def test_green_degree_zero_in_xi():
    lam0, mu0 = po.lame(2, 70e9, 0.3)
    a = po.green((8, 6), 1.0, lam0, mu0)
    b = po.green((8, 6), 0.37, lam0, mu0)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


# ---------------------------------------------------------------------------
# LS solve: closed forms
# ---------------------------------------------------------------------------
def _hooke(D, lam, mu, e):
    return 2 * mu * e + lam * np.trace(e) * np.eye(D)


def _navier_solution(shape, lam, mu, ep, bar):
    """Exact periodic solution of div C0:(eps(u) - eps_p) = 0 with mean strain
    eps_bar: u_hat = -i K(xi)^-1 t(xi), K_ik = C0_ijkl xi_j xi_l,
    t_i = C0_ijkl xi_j eps_p_hat_kl (acoustic-tensor form, no Green tensor)."""
    D = len(shape)
    N = int(np.prod(shape))
    C0 = _C0(D, lam, mu)
    eph = np.fft.fftn(ep.reshape(shape + (D, D)), axes=tuple(range(D)))
    freqs = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    xi = np.stack(freqs, axis=-1).reshape(N, D) * 2 * np.pi
    eph = eph.reshape(N, D, D)
    out = np.zeros((N, D, D), dtype=complex)
    for k in range(1, N):
        x = xi[k]
        K = np.einsum("ijkl,j,l->ik", C0, x, x)
        t = np.einsum("ijkl,j,kl->i", C0, x, eph[k])
        w = np.linalg.solve(K, t)
        out[k] = 0.5 * (np.outer(w, x) + np.outer(x, w))
    et = np.fft.ifftn(out.reshape(shape + (D, D)), axes=tuple(range(D))).real.reshape(N, D, D)
    return bar[None] + et


@pytest.mark.parametrize("shape", [(9, 7), (15, 15), (5, 7, 3)])
def test_homogeneous_eigenstrain_closed_form(shape):
    D = len(shape)
    N = int(np.prod(shape))
    E = np.full(N, 85e9)
    nu = np.full(N, 0.35)
    lam, mu = po.lame(D, 85e9, 0.35)
    rng = np.random.default_rng(5)
    ep = np.zeros((N, D, D))
    for v in rng.choice(N, 3, replace=False):
        a = rng.standard_normal((D, D)) * 0.05
        ep[v] = 0.5 * (a + a.T)
    bar = np.diag([1e-3] + [0.0] * (D - 1))
    r = po.ls_solve(shape, E, nu, bar, eps_plastic=ep, tol=1e-12)
    exact = _navier_solution(shape, lam, mu, ep, bar)
    assert np.abs(r["eps"] - exact).max() < 1e-12 * np.abs(exact).max()
    # homogeneous medium: one iteration reaches the solution, the second measures diff ~ 0
    assert r["iters"] == 2
    assert r["diff"] < 1e-13


def test_uniform_strain_is_exact():
    shape = (8, 8)
    E = np.full(64, 70e9)
    nu = np.full(64, 0.3)
    bar = np.array([[1e-3, 2e-4], [2e-4, -3e-4]])
    for ps in (False, True):
        r = po.ls_solve(shape, E, nu, bar, tol=1e-10, plane_stress=ps)
        lam, mu = po.lame(2, 70e9, 0.3, plane_stress=ps)
        assert r["iters"] == 1
        assert np.abs(r["eps"] - bar[None]).max() < 1e-18
        assert np.abs(r["sig"] - _hooke(2, lam, mu, bar)[None]).max() < 1e-6


@pytest.mark.parametrize("shape", [(15, 13), (7, 9, 5)])
def test_heterogeneous_equilibrium(shape):
    D = len(shape)
    N = int(np.prod(shape))
    rng = np.random.default_rng(6)
    E = 85e9 * (1 + 0.2 * rng.standard_normal(N))
    nu = np.full(N, 0.35)
    bar = np.diag([1e-3] + [0.0] * (D - 1))
    r = po.ls_solve(shape, E, nu, bar, tol=1e-13, max_iter=500)
    sh = np.fft.fftn(r["sig"].reshape(shape + (D, D)), axes=tuple(range(D))).reshape(N, D, D)
    freqs = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    xi = np.stack(freqs, axis=-1).reshape(N, D)
    div = np.einsum("kij,kj->ki", sh, xi)
    assert np.abs(div[1:]).max() < 1e-9 * np.abs(sh).max()
    # compatibility: the fluctuation is sym(xi (x) w): its xi-orthogonal shear part vanishes in 2D
    assert abs(r["eps_macro"][0, 0] - 1e-3) < 1e-15


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference tree absent)")
def test_ls_solve_same_with_reference_pocketfft():
    shape = (32, 32)
    rng = np.random.default_rng(8)
    E = 85e9 * (1 + 0.15 * rng.standard_normal(1024))
    nu = np.full(1024, 0.35)
    bar = np.diag([1e-3, 0.0])
    a = po.ls_solve(shape, E, nu, bar, tol=1e-8)
    b = po.ls_solve(shape, E, nu, bar, tol=1e-8, ref=True)
    assert a["iters"] == b["iters"]
    assert np.abs(a["sig"] - b["sig"]).max() < 1e-13 * np.abs(b["sig"]).max()


# ---------------------------------------------------------------------------
# random stream (libstdc++ 11: mt19937, generate_canonical, Marsaglia polar)
# ---------------------------------------------------------------------------
def _mt_words(seed, k):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    return bg.random_raw(k).astype(np.uint64)


def test_mt19937_words_match_numpy():
    for seed in (42, 1, 12345):
        assert np.array_equal(po.rng_stream(seed, 0, 1000), _mt_words(seed, 1000).astype(float))


def test_uniform_is_generate_canonical():
    w = _mt_words(42, 200)
    u = (w[0::2].astype(float) + w[1::2].astype(float) * 2.0 ** 32) / 2.0 ** 64
    assert np.array_equal(po.rng_stream(42, 1, 100), u)


def test_normal_is_marsaglia_polar_with_cache():
    w = _mt_words(7, 4000)
    pos = 0

    def U():
        nonlocal pos
        a, b = float(w[pos]), float(w[pos + 1])
        pos += 2
        return (a + b * 2.0 ** 32) / 2.0 ** 64

    out = []
    saved = None
    while len(out) < 50:
        if saved is not None:
            out.append(saved)
            saved = None
            continue
        while True:
            x = 2.0 * U() - 1.0
            y = 2.0 * U() - 1.0
            r2 = x * x + y * y
            if 0.0 < r2 <= 1.0:
                break
        m = math.sqrt(-2.0 * math.log(r2) / r2)
        saved = x * m
        out.append(y * m)
    assert np.array_equal(po.rng_stream(7, 2, 50), np.array(out))


def test_correlated_field_statistics():
    f = po.correlated_field((32, 32), 85.0, 12.75, 3.0, 42)
    assert abs(f.mean() - 85.0) < 1e-12
    assert abs(f.std() - 12.75) < 1e-10
    g = po.correlated_field((32, 32), 85.0, 12.75, 3.0, 42, clip_min=60.0)
    assert g.min() == 60.0


def _impulse_field(shape, at, corr, coeff=None):
    """Closed form of the filter step of generate_correlated_field_cpp
    (material_generator.hpp:139-219) for a unit impulse at voxel `at`:
    g(x) = (1/N) sum_f exp(-2 pi^2 corr^2 |f|^2) cos(2 pi f.(x - at)) with
    f = k/n in [-1/2, 1/2) per axis, then standardised with the population
    deviation (:222-235).  Written as an explicit sum over modes (no FFT)."""
    shape = tuple(shape)
    N = int(np.prod(shape))
    coeff = -2.0 * math.pi ** 2 * corr ** 2 if coeff is None else coeff
    axes = []
    for ax, n in enumerate(shape):
        k = np.arange(n)
        f = np.where(k < (n + 1) // 2, k, k - n) / n
        x = (np.arange(n) - at[ax]) % n
        # per-axis factor: exp(coeff f^2) and the phase 2 pi f x
        axes.append((f, x))
    g = np.zeros(shape)
    grids = np.meshgrid(*[a[1] for a in axes], indexing="ij")
    for idx in np.ndindex(*shape):
        fs = [axes[a][0][idx[a]] for a in range(len(shape))]
        h = math.exp(coeff * sum(f * f for f in fs))
        ph = sum(2.0 * math.pi * fs[a] * grids[a] for a in range(len(shape)))
        g += h * np.cos(ph)
    g = (g / N).reshape(-1)
    m = g.mean()
    sd = math.sqrt(((g - m) ** 2).sum() / N)
    return (g - m) / sd


def test_correlated_field_filter_closed_form():
    # a unit impulse through the generator's filter gives the standardised
    # Gaussian kernel of the closed form; the standardisation (an affine map)
    # cannot hide a wrong filter, because the kernel's shape is compared voxel
    # by voxel.  A filter with the wrong constant (pi^2 instead of 2 pi^2) or
    # with frequencies not divided by n is far off.
    from mgkmc_cpp_amd import fields
    for shape, at, corr in (((32, 32), (0, 0), 3.0), ((16, 12), (5, 7), 2.0), ((8, 10, 6), (3, 0, 5), 1.5)):
        N = int(np.prod(shape))
        noise = np.zeros(shape)
        noise[at] = 1.0
        want = _impulse_field(shape, at, corr)
        got = po.correlated_field_noise(shape, 0.0, 1.0, corr, noise)
        assert np.abs(got - want).max() < 1e-12, shape
        host = fields.correlated_field(shape, 0.0, 1.0, corr, 0, noise=noise)
        assert np.abs(host - want).max() < 1e-12, shape
        # the mean and deviation targets are applied after standardising
        got2 = po.correlated_field_noise(shape, 85.0, 12.75, corr, noise)
        assert np.abs(got2 - (85.0 + 12.75 * want)).max() < 1e-10
        wrong_const = _impulse_field(shape, at, corr, coeff=-math.pi ** 2 * corr ** 2)
        assert np.abs(wrong_const - want).max() > 1e-2
        # corr is the real-space deviation of the kernel: the filter
        # exp(-2 pi^2 s^2 f^2) is the transform of exp(-x^2 / (2 s^2)), so the
        # kernel (above its periodic floor) falls to exp(-1/2) at x = s
        if len(shape) == 2 and shape[0] == 32:
            prof = got.reshape(shape)[:, 0]
            base = prof.min()
            r = (prof[3] - base) / (prof[0] - base)
            assert abs(r - math.exp(-0.5)) < 0.02, r
