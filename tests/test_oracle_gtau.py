"""Pins of oracle/gtau.py (G(tau), mu bisection, the SDE's second-order term) by
closed forms that do not use its code path."""
import math

import numpy as np
import pytest

from oracle import gtau as T
from oracle.lattice import Lattice, matsubara

BETA = 2.0


def test_fdfit_exact_for_quartics():
    rng = np.random.default_rng(0)
    c = rng.standard_normal(5)
    dt = 0.1
    L = 12
    t = dt * np.arange(L)
    p = sum(c[i] * t ** i for i in range(5))
    dp = lambda x: sum(i * c[i] * x ** (i - 1) for i in range(1, 5))
    fd1, fd2 = T.fdfit(p, dt)
    assert abs(fd1 - dp(0.0)) < 1e-10 * (1 + abs(dp(0.0)))
    assert abs(fd2 - dp(t[-1])) < 1e-10 * (1 + abs(dp(t[-1])))


def _int_cubic(c, om, beta):
    """int_0^beta (c0 + c1 t + c2 t^2 + c3 t^3) e^{i om t} dt, closed form."""
    if abs(om) < 1e-12:
        return sum(c[k] * beta ** (k + 1) / (k + 1) for k in range(4))
    io = 1j * om
    e = np.exp(io * beta)
    J = [(e - 1.0) / io]
    for k in range(1, 4):
        J.append(beta ** k * e / io - k / io * J[k - 1])
    return sum(c[k] * J[k] for k in range(4))


@pytest.mark.parametrize("fermionic", [True, False])
def test_nfourier_exact_for_cubics(fermionic):
    rng = np.random.default_rng(1)
    c = rng.standard_normal(4)
    L, iwmax = 15, 6
    t = BETA / L * np.arange(L + 1)
    p = c[0] + c[1] * t + c[2] * t ** 2 + c[3] * t ** 3
    d0 = c[1]
    d1 = c[1] + 2 * c[2] * BETA + 3 * c[3] * BETA ** 2
    got = T.nfourier(fermionic, L, iwmax, d0, d1, p, BETA)
    for i in range(1, iwmax + 1):
        om = (2 * i - 1) * math.pi / BETA if fermionic else 2 * (i - 1) * math.pi / BETA
        exp = _int_cubic(c, om, BETA)
        assert abs(got[i - 1] - exp) < 1e-11 * (1 + abs(exp)), (i, got[i - 1], exp)


def test_g_tau_bare_and_half_filling():
    lat = Lattice(4, 8, 1)
    sigma = np.zeros(lat.nx * lat.ny * lat.nf, dtype=np.complex128)
    mu = 0.3
    g = T.g_tau_k(lat, sigma, mu, BETA)
    tau = BETA / (lat.nf - 1) * np.arange(lat.nf)
    for kx in range(lat.nx):
        for ky in range(lat.ny):
            e = -2 * math.cos(2 * math.pi * kx / lat.nx) - 2 * math.cos(2 * math.pi * ky / lat.ny)
            exp = -np.exp((BETA - tau) * (e - mu)) / (1 + math.exp(BETA * (e - mu)))
            assert np.allclose(g[kx, ky], exp, rtol=1e-14, atol=1e-15)
    # half filling at mu = 0 (particle-hole symmetric lattice): n = 1, and the bisection stays at mu = 0
    n0 = T.particle_number(T.grt(lat, sigma, 0.0, BETA))
    assert abs(n0 - 1.0) < 1e-13
    mu_b, _, steps = T.bisect_mu(lat, sigma, 0.0, BETA, 2.0, 1.0)
    assert mu_b == 0.0 and steps == 1


@pytest.mark.parametrize("ngrain", [2, 3])
def test_g_tau_fine_grid_covers_every_momentum(ngrain):
    """With coarse graining each grain lands on its own fine-grid point
    2 pi (i - 1) / (Nx Ng): at Sigma = 0 the fine grid is the bare G of the fine lattice."""
    lat = Lattice(4, 8, 1)
    sigma = np.zeros(lat.nx * lat.ny * lat.nf, dtype=np.complex128)
    mu = -0.4
    g = T.g_tau_k(lat, sigma, mu, BETA, ngrain)
    n = lat.nx * ngrain
    tau = BETA / (lat.nf - 1) * np.arange(lat.nf)
    for i in range(n):
        for j in range(n):
            e = -2 * math.cos(2 * math.pi * i / n) - 2 * math.cos(2 * math.pi * j / n)
            exp = -np.exp((BETA - tau) * (e - mu)) / (1 + math.exp(BETA * (e - mu)))
            assert np.allclose(g[i, j], exp, rtol=1e-13, atol=1e-15), (i, j)
    assert abs(T.particle_number(T.grt(lat, sigma, 0.0, BETA, ngrain)) - 1.0) < 1e-13


def test_g_tau_constant_sigma_shifts_the_band():
    """Sigma = s (constant, real) is the bare G at e + s; the Matsubara part of G(tau)
    then carries only the truncation of the frequency sum: ~1e-4 inside (0, beta) and
    ~1e-3 at the two ends for Nf = 64, where a wrong phase sign or 1/beta is O(s)."""
    lat = Lattice(4, 64, 1)
    s, mu = 0.3, 0.1
    sigma = np.full(lat.nx * lat.ny * lat.nf, s + 0j)
    g = T.g_tau_k(lat, sigma, mu, BETA)
    tau = BETA / (lat.nf - 1) * np.arange(lat.nf)
    for kx in range(lat.nx):
        for ky in range(lat.ny):
            e = -2 * math.cos(2 * math.pi * kx / lat.nx) - 2 * math.cos(2 * math.pi * ky / lat.ny) + s
            exp = -np.exp((BETA - tau) * (e - mu)) / (1 + math.exp(BETA * (e - mu)))
            d = np.abs(g[kx, ky] - exp)
            assert d[1:-1].max() < 3e-4 and d[[0, -1]].max() < 3e-3, (kx, ky, d.max())


def _second_order_analytic(lat, mu, U):
    """Sigma2(k, i nu) = U^2/N^2 sum_{p,p'} (1-n_p)(1-n_p') n_p'' (1 + e^{-beta E}) / (i nu - E),
    E = e_p + e_p' - e_p'', p'' = p + p' - k (energies from mu): the tau integral of
    U^2 G0(r,tau)^2 G0(-r,beta-tau) done in closed form."""
    nx, nf = lat.nx, lat.nf
    N = nx * nx
    k = 2 * math.pi * np.arange(nx) / nx
    eps = -2 * np.cos(k)[:, None] - 2 * np.cos(k)[None, :] - mu
    n = 1.0 / (np.exp(BETA * eps) + 1.0)
    out = np.zeros(N * nf, dtype=np.complex128)
    for kx in range(nx):
        for ky in range(nx):
            acc = np.zeros(nf, dtype=np.complex128)
            nus = np.array([matsubara(v, nf, BETA) for v in range(1, nf + 1)])
            for px in range(nx):
                for py in range(nx):
                    for qx in range(nx):
                        for qy in range(nx):
                            ppx, ppy = (px + qx - kx) % nx, (py + qy - ky) % nx
                            E = eps[px, py] + eps[qx, qy] - eps[ppx, ppy]
                            w = (1 - n[px, py]) * (1 - n[qx, qy]) * n[ppx, ppy] * (1 + math.exp(-BETA * E))
                            acc += w / (nus - E)
            out[(kx * nx + ky) * nf:(kx * nx + ky + 1) * nf] = U * U / N ** 2 * acc
    return out


def test_order_2nd_bare_matches_closed_form():
    lat = Lattice(4, 64, 1)
    sigma = np.zeros(lat.nx * lat.ny * lat.nf, dtype=np.complex128)
    mu, U = 0.2, 2.0
    g = T.grt(lat, sigma, mu, BETA)
    got = T.order_2nd(lat, g, U, BETA)
    exp = _second_order_analytic(lat, mu, U)
    err = np.linalg.norm(got - exp) / np.linalg.norm(exp)
    assert err < 1e-5, err  # cubic-spline quadrature on Nf = 64 tau points (measured 1.7e-6)


def test_bisection_hits_the_bare_density():
    """At Sigma = 0 the density is 2/N sum_k f(e_k - mu): the bisection's mu solves it to
    the stopping rule |n - n_target| <= 1e-6."""
    from scipy.optimize import brentq
    lat = Lattice(4, 8, 1)
    sigma = np.zeros(lat.nx * lat.ny * lat.nf, dtype=np.complex128)
    k = 2 * math.pi * np.arange(lat.nx) / lat.nx
    e = (-2 * np.cos(k)[:, None] - 2 * np.cos(k)[None, :]).ravel()
    dens = lambda m: 2.0 / e.size * np.sum(1.0 / (np.exp(BETA * (e - m)) + 1.0))
    target = 0.8
    mu_exact = brentq(lambda m: dens(m) - target, -6, 6, xtol=1e-14)
    mu, g, steps = T.bisect_mu(lat, sigma, 0.5, BETA, 2.0, target)
    assert abs(T.particle_number(g) - target) <= 1e-6
    assert abs(dens(mu) - target) <= 1e-6
    assert abs(mu - mu_exact) < 1e-5 and steps > 5


def _lindhard(lat, mu, beta, q, w, channel, weight=None):
    """Bare bubbles at Sigma = 0 in closed form (the tau integrals done exactly):
    ph: 1/N sum_k FF(k)^2 (f_k - f_{k+q}) / (i w + e_k - e_{k+q})   (-beta f (1 - f) when degenerate)
    pp: 1/(2N) sum_k FF(k)^2 (1 - f_k - f_{q-k}) / (i w - e_k - e_{q-k})."""
    nx = lat.nx
    k = 2 * math.pi * np.arange(nx) / nx
    eps = (-2 * np.cos(k)[:, None] - 2 * np.cos(k)[None, :]) - mu
    f = 1.0 / (np.exp(beta * eps) + 1.0)
    om = 2j * math.pi * w / beta
    s = 0j
    for a in range(nx):
        for b in range(nx):
            wt = 1.0 if weight is None else weight[a, b]
            if channel == "ph":
                a2, b2 = (a + q[0]) % nx, (b + q[1]) % nx
                de = eps[a, b] - eps[a2, b2]
                if abs(de) < 1e-12 and w == 0:
                    x = -beta * f[a, b] * (1 - f[a, b])
                else:
                    x = (f[a, b] - f[a2, b2]) / (om + de)
            else:
                a2, b2 = (q[0] - a) % nx, (q[1] - b) % nx
                x = 0.5 * (1 - f[a, b] - f[a2, b2]) / (om - eps[a, b] - eps[a2, b2])
            s += wt * x
    return s / (nx * nx)


def test_bare_bubbles_are_the_lindhard_functions():
    """calc_Gkw_Grt's bubbles at Sigma = 0 against the closed forms, for the plain and the
    form-factor weighted versions (the weight sits on the momentum k of the first Green's
    function); the form factor 1 reproduces the plain bubbles exactly."""
    from oracle import sus as SU
    lat = Lattice(4, 64, 1)
    sigma = np.zeros(lat.nx * lat.ny * lat.nf, dtype=np.complex128)
    mu = 0.3
    g = T.grt(lat, sigma, mu, BETA)
    ffs = [T.ffg(lat, sigma, mu, BETA, 1, f) for f in range(4)]
    ph, pp, phf, ppf = T.chi0(lat, g, ffs, BETA)
    assert np.array_equal(phf[:, 0], ph) and np.array_equal(ppf[:, 0], pp)
    fs = SU.ff_sdwave(lat.nx)
    nw = lat.nf // 2
    for (i, j) in [(1, 1), (2, 1), (3, 2), (3, 3)]:
        for w in (0, 1, 5):
            idx = ((i * (i - 1)) // 2 + j - 1) * nw + w
            q = (i - 1, j - 1)
            for got, ch in ((ph[idx], "ph"), (pp[idx], "pp")):
                exp = _lindhard(lat, mu, BETA, q, w, ch)
                assert abs(got - exp) < 2e-6 * max(abs(exp), 0.05), (i, j, w, ch, got, exp)
            for f in (1, 2, 3):
                for got, ch in ((phf[idx, f], "ph"), (ppf[idx, f], "pp")):
                    exp = _lindhard(lat, mu, BETA, q, w, ch, weight=fs[:, :, f] ** 2)
                    assert abs(got - exp) < 2e-6 * max(abs(exp), 0.05), (i, j, w, ch, f, got, exp)
