"""Pins of oracle/bse.py: the ladder update and the bubble chi0."""
import math

import numpy as np
import pytest

from oracle import bse as B
from oracle.lattice import Lattice
from tests.helpers import crandn, random_chi_compact, rel_fro, random_vertices

# c_r of SRC/parquet_BSE.f90:73,89,105,120 (prefactor / f_damping)
C_REF = {0: 0.5, 1: 0.5, 2: -0.25, 3: 0.25}


@pytest.mark.parametrize("nl,nf,f", [(1, 8, 1.0), (1, 8, 0.5), (5, 4, 0.3), (9, 4, 0.7)])
def test_ladder_fixed_point(nl, nf, f):
    """At F = (I - kappa_r Gamma chi)^-1 Gamma (kappa_r = 2 c_r alpha) the damped update is F - Gamma.

    The solve is LAPACK (numpy.linalg.solve), independent of the oracle's ZGEMM chain;
    a dropped term, a wrong sign of c_r, a transposed chi or a wrong mixing weight fails.
    """
    rng = np.random.default_rng(7)
    lat = Lattice(4, nf, nl)
    nq, nz, beta = 3, lat.nz, 2.0
    alpha = 1.0 / (beta * lat.nx * lat.ny)
    gam = random_vertices(rng, nz, nq, diag=2.0, scale=0.3)
    chic = [random_chi_compact(rng, nl, nf, nq, 2.0), random_chi_compact(rng, nl, nf, nq, 2.0)]
    chi = [B.chi_compact_to_dense(lat, c) for c in chic]
    F = []
    for r in range(4):
        kappa = 2 * C_REF[r] * alpha
        fr = np.zeros_like(gam[r])
        ch = chi[0 if r < 2 else 1]
        for q in range(nq):
            a = np.eye(nz) - kappa * gam[r][:, :, q] @ ch[:, :, q]
            fr[:, :, q] = np.linalg.solve(a, gam[r][:, :, q])
        F.append(fr)
    phi = B.bse_ladder(lat, gam, F, chi, beta, f)
    for r in range(4):
        assert rel_fro(phi[r], F[r] - gam[r]) < 1e-12, r


@pytest.mark.parametrize("nl,nf", [(1, 8), (5, 4), (9, 4)])
def test_closed_form_is_the_ladder_fixed_point(nl, nf):
    """bse_closed gives Phi with ladder(Gamma, F = Gamma + Phi) = Phi for any damping:
    the closed form is pinned by the reference's own update, not by its formula."""
    rng = np.random.default_rng(11)
    lat = Lattice(4, nf, nl)
    nq, nz, beta = 2, lat.nz, 2.0
    gam = random_vertices(rng, nz, nq, diag=2.0, scale=0.3)
    chi = [B.chi_compact_to_dense(lat, random_chi_compact(rng, nl, nf, nq, 2.0)) for _ in range(2)]
    phi = B.bse_closed(lat, gam, chi, beta)
    F = [gam[r] + phi[r] for r in range(4)]
    for f in (1.0, 0.4):
        again = B.bse_ladder(lat, gam, F, chi, beta, f)
        for r in range(4):
            assert rel_fro(again[r], phi[r]) < 1e-12, (r, f)
    # and it is not the trivial solution
    assert min(np.linalg.norm(p) for p in phi) > 1e-3


def test_ladder_first_iteration():
    """F = Gamma (the reference's first iteration, SRC/preprocessing.f90:100-116):
    Phi = 2 f c_r alpha Gamma chi Gamma."""
    rng = np.random.default_rng(3)
    lat = Lattice(4, 8, 5)
    nq, beta, f = 2, 3.0, 0.4
    alpha = 1.0 / (beta * lat.nc)
    gam = random_vertices(rng, lat.nz, nq)
    chi = [B.chi_compact_to_dense(lat, random_chi_compact(rng, 5, 8, nq)) for _ in range(2)]
    phi = B.bse_ladder(lat, gam, gam, chi, beta, f)
    for r in range(4):
        ch = chi[r // 2]
        exp = np.stack([2 * f * C_REF[r] * alpha * gam[r][:, :, q] @ ch[:, :, q] @ gam[r][:, :, q]
                        for q in range(nq)], axis=2)
        assert rel_fro(phi[r], exp) < 1e-13


def _bubble_bruteforce(lat, q, beta, mu, ngrain):
    """chi_ph/pp from whole-array k-space products (no index maps)."""
    nx, nf = lat.nx, lat.nf
    qx, qy, qw = q
    half = (ngrain + 1) // 2
    k = 2 * math.pi * np.arange(nx) / nx
    g = (np.arange(1, ngrain + 1) - half) * (2 * math.pi / nx) / ngrain
    nu = np.arange(1, nf + 1)

    def G(kshift_x, kshift_y, nus, gx, gy):
        # dispersion on the shifted grid (momentum index shift, grain offsets)
        px = k[(np.arange(nx) + kshift_x) % nx][:, None] + gx
        py = k[(np.arange(nx) + kshift_y) % nx][None, :] + gy
        eps = -2 * np.cos(px) - 2 * np.cos(py)
        z = 1j * math.pi / beta * (2 * (nus - nf // 2 - 1) + 1)
        return 1.0 / (z[None, None, :] - eps[:, :, None] + mu)

    chi_k_ph = np.zeros((nx, nx, nf), complex)
    chi_k_pp = np.zeros((nx, nx, nf), complex)
    for ix in range(ngrain):
        for iy in range(ngrain):
            gk = G(0, 0, nu, g[ix], g[iy])
            gkq = G(qx - 1, qy - 1, nu + qw - 1, g[ix], g[iy])
            chi_k_ph += gk * gkq / ngrain**2
            # q - k: momentum index (q-1) - (k-1); grain reversed
            gq = np.zeros((nx, nx, nf), complex)
            rx, ry = g[ngrain - 1 - ix], g[ngrain - 1 - iy]
            for a in range(nx):
                for b in range(nx):
                    ma, mb = (qx - 1 - a) % nx, (qy - 1 - b) % nx
                    px, py = k[ma] + rx, k[mb] + ry
                    eps = -2 * np.cos(px) - 2 * np.cos(py)
                    nn = nf + qw - nu
                    z = 1j * math.pi / beta * (2 * (nn - nf // 2 - 1) + 1)
                    gq[a, b] = 1.0 / (z - eps + mu)
            chi_k_pp += gk * gq / ngrain**2
    out = []
    for ck in (chi_k_ph, chi_k_pp):
        d = np.zeros((lat.nz, lat.nz), complex)
        for l1 in range(lat.nl):
            for l2 in range(lat.nl):
                w = lat.ff[:, :, l1] * lat.ff[:, :, l2]
                d[l1 * nf + np.arange(nf), l2 * nf + np.arange(nf)] = np.einsum("xyn,xy->n", ck, w)
        out.append(d)
    return out


@pytest.mark.parametrize("nl,ngrain,q", [(1, 1, (2, 1, 3)), (5, 1, (3, 2, 2)), (5, 2, (2, 2, 1))])
def test_calc_chi_bruteforce(nl, ngrain, q):
    lat = Lattice(4, 8, nl)
    beta, mu = 2.0, 0.1
    sigma = np.zeros(lat.nc * lat.nf, complex)  # Sigma = 0: G is the bare lattice G
    ph, pp = B.calc_chi(lat, *q, sigma, mu, beta, ngrain)
    bph, bpp = _bubble_bruteforce(lat, q, beta, mu, ngrain)
    assert rel_fro(ph, bph) < 1e-13
    assert rel_fro(pp, bpp) < 1e-13


def test_calc_chi_nl1_closed_form():
    """Nl=1: chi(nu,nu) = (1/Nc) sum_k G(k,nu) G(k+q,nu+w) (FF_s = 1/Nx)."""
    lat = Lattice(4, 8, 1)
    beta, mu = 2.0, 0.0
    rng = np.random.default_rng(1)
    sigma = 0.1 * (rng.standard_normal(lat.nc * lat.nf) + 1j * rng.standard_normal(lat.nc * lat.nf))
    qx, qy, qw = 3, 1, 2
    ph, _ = B.calc_chi(lat, qx, qy, qw, sigma, mu, beta, 1)
    ekg = B.ek_grain(lat, 1)
    for nu in range(1, lat.nf + 1):
        acc = 0
        for kx in range(1, 5):
            for ky in range(1, 5):
                g1 = B.green_coarse(lat, ekg, sigma, mu, beta, 0, 0, kx, ky, nu)
                g2 = B.green_coarse(lat, ekg, sigma, mu, beta, 0, 0, (kx + qx - 2) % 4 + 1,
                                    (ky + qy - 2) % 4 + 1, nu + qw - 1)
                acc += g1 * g2
        assert abs(ph[nu - 1, nu - 1] - acc / lat.nc) < 1e-13


def test_calc_chi_hartree_outside_box_is_a_mu_shift():
    """A constant self-energy s inside the box with Sigma_H = s outside is the
    same as Sigma = 0 at mu - s: pins that frequencies outside 1..Nf (reached by
    nu + w - 1 and Nf - nu + w) take Sigma_H, as get_green_coarse does."""
    lat = Lattice(4, 8, 5)
    beta, mu, s = 2.0, 0.2, 0.3 - 0.1j
    sigma = np.full(lat.nc * lat.nf, s)
    sigma_h = np.full((lat.nx, lat.ny), s)
    for q in [(2, 1, 3), (3, 3, 4)]:
        ph, pp = B.calc_chi(lat, *q, sigma, mu, beta, 1, sigma_h)
        ph0, pp0 = B.calc_chi(lat, *q, np.zeros_like(sigma), mu - s, beta, 1)
        assert rel_fro(ph, ph0) < 1e-14 and rel_fro(pp, pp0) < 1e-14
        # without Sigma_H the out-of-box frequencies differ
        ph1, _ = B.calc_chi(lat, *q, sigma, mu, beta, 1)
        assert rel_fro(ph1, ph0) > 1e-6


def test_gamma_update():
    rng = np.random.default_rng(0)
    a = [crandn(rng, (3, 3, 2)) for _ in range(4)]
    b = [crandn(rng, (3, 3, 2)) for _ in range(4)]
    for x, y, z in zip(B.gamma_update(a, b), a, b):
        assert np.array_equal(x, y - z)
