"""Pins of oracle/sus.py (calc_sus) by quantities pinned elsewhere and by closed forms."""
import math

import numpy as np
import pytest

from oracle import bse as B
from oracle import sde as D
from oracle import sus as S
from oracle.lattice import Lattice

BETA, U = 2.0, 2.0


def _sigma(rng, lat, scale=0.2):
    n = lat.nx * lat.ny * lat.nf
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


@pytest.mark.parametrize("nx", [4, 8])
def test_ff_sdwave_check(nx):
    """FF_sdwave_check (SRC/parquet_check.f90:274-281) against the pinned form factors."""
    lat = Lattice(nx, 4, 5)
    fs = S.ff_sdwave(nx)
    err = (np.abs(lat.ff[:, :, 0] - fs[:, :, 0] / nx).sum() + np.abs(lat.ff[:, :, 2] - fs[:, :, 1] / nx).sum()
           + np.abs(lat.ff[:, :, 3] - math.sqrt(2.0) / nx * fs[:, :, 2]).sum())
    assert err < 1e-12


def test_gg_s_wave_is_the_auxiliary_bubble():
    rng = np.random.default_rng(0)
    lat = Lattice(4, 4, 5)
    sigma = _sigma(rng, lat)
    sh = 0.1 * (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    ph, pp = S.gg(lat, sigma, sh, 0.2, BETA, ngrain=2, f_range=1)
    aph, app = D.chi_aux(lat, sigma, sh, 0.2, BETA, ngrain=2, f_range=1)
    assert np.abs(ph[:, :, 0] - aph).max() < 1e-14 * np.abs(aph).max()
    assert np.abs(pp[:, :, 0] - app).max() < 1e-14 * np.abs(app).max()


def test_gg_d_and_p_wave_are_bubble_diagonals():
    rng = np.random.default_rng(1)
    lat = Lattice(4, 4, 5)
    nf = lat.nf
    sigma = _sigma(rng, lat)
    ph, pp = S.gg(lat, sigma, None, 0.1, BETA)
    for iq in (0, 5, 11):
        cph, cpp = B.calc_chi(lat, int(lat.qx[iq]), int(lat.qy[iq]), int(lat.qw[iq]), sigma, 0.1, BETA)
        for nu in range(nf):
            for g, c in ((ph, cph), (pp, cpp)):
                assert abs(g[2 * nf + nu, iq, 1] - lat.nx * c[2 * nf + nu, 2 * nf + nu]) < 1e-13
                assert abs(g[3 * nf + nu, iq, 2] - lat.nx / math.sqrt(2.0) * c[3 * nf + nu, 3 * nf + nu]) < 1e-13


def _gg(rng, lat, f_range):
    return S.gg(lat, _sigma(rng, lat), None, 0.0, BETA, f_range=f_range)


def test_contraction_identity_vertex():
    rng = np.random.default_rng(2)
    lat = Lattice(4, 4, 1)
    gph, gpp = _gg(rng, lat, 0)
    c = 0.7 - 0.2j
    eye = np.repeat(np.eye(lat.nz, dtype=np.complex128)[:, :, None], lat.nq, axis=2)
    zero = np.zeros_like(eye)
    sus = S.calc_sus(lat, [c * eye] * 4, [zero] * 4, gph, gpp, U, BETA)
    norm = 1.0 / (lat.nx ** 4 * BETA ** 2)
    for r in range(4):
        g = gph if r < 2 else gpp
        assert np.allclose(sus[r], c * np.sum(g * g, axis=0) * norm, rtol=1e-13, atol=0)


def test_contraction_bare_term_outside_the_box():
    rng = np.random.default_rng(3)
    lat = Lattice(4, 4, 1)
    nf = lat.nf
    gph, gpp = _gg(rng, lat, 1)
    zero = np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128)
    sus = S.calc_sus(lat, [zero] * 4, [zero] * 4, gph, gpp, U, BETA, f_range=1)
    norm = 1.0 / (lat.nx ** 4 * BETA ** 2) / lat.ff[0, 0, 0] ** 2
    for r, bare in enumerate((U, -U, 2 * U, 0.0)):
        g = gph if r < 2 else gpp
        full, box = g.sum(axis=0), g[nf:2 * nf].sum(axis=0)
        exp = bare * (full * full - box * box) * norm
        assert np.abs(sus[r] - exp).max() <= 1e-13 * max(np.abs(exp).max(), 1e-300)


def test_contraction_clamped_phi_entry():
    rng = np.random.default_rng(4)
    lat = Lattice(4, 4, 1)
    nf = lat.nf
    gph, gpp = _gg(rng, lat, 1)
    zero = np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128)
    phi = zero.copy()
    phi[0, 0, :] = 1.5  # nu = 1 in both arguments: every frequency at or below the box bottom reads it
    sus = S.calc_sus(lat, [zero] * 4, [phi] * 4, gph, gpp, 0.0, BETA, f_range=1)
    norm = 1.0 / (lat.nx ** 4 * BETA ** 2)
    for r in range(4):
        g = gph if r < 2 else gpp
        low = g[:nf + 1].sum(axis=0)  # nu = -Nf + 1 .. 1
        exp = 1.5 * (low * low - g[nf] * g[nf]) * norm
        assert np.abs(sus[r] - exp).max() <= 1e-13 * np.abs(exp).max()
