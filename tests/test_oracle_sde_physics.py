"""The SDE oracle (oracle/sde.py, a restatement of SRC/parquet_SDE.f90:27-258)
against the Schwinger-Dyson equation evaluated directly in momentum and
frequency space (tests/parquet_bruteforce.py: sde_kspace), for vertices in the
span of the Nl form factors, so that the form-factor projection is exact.

This pins what the second-order diagram (zero vertices) leaves open: the
contraction of G_r_LL with the auxiliary bubbles in PhChi / PPChi for Nl > 1
(which index is the external one), FF_inv per operation and form factor, the
form-factor signs of the time-reversed half, the pp term's G(Q - k), and with
f_range > 0 the enlarged frequency sums (vertices clamped to the box edge,
Lambda replaced by its bare limit outside the box).  The
test vertices are C4v-symmetric but not symmetric under k <-> k', so a
transposed contraction fails.  Planted mistakes in the oracle (each undone after
its check) must fail the comparison.
"""
import numpy as np
import pytest

from oracle import sde as S
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro
from tests.parquet_bruteforce import KSpace, sde_kspace, symmetric_phi_full

BETA, U, MU = 2.0, 2.0, 0.1
TOL = 1e-12


def _inputs(nx, nf, nl, seed):
    lat = Lattice(nx, nf, nl)
    rng = np.random.default_rng(seed)
    fulls, refs = symmetric_phi_full(lat, rng)
    scale = 0.05
    fulls = [f * scale for f in fulls]
    refs = [r * scale for r in refs]
    lam_d = crandn(rng, (nf, nf, lat.nw), 0.3)
    lam_m = crandn(rng, (nf, nf, lat.nw), 0.3)
    # Sigma(k, nu) = a(k) + i b(k) nu / (1 + nu^2): C4v and inversion symmetric,
    # Sigma(k, -nu) = conj Sigma(k, nu); Sigma_H = a(k)
    ks = KSpace(lat)
    kx = 2 * np.pi * ks.kvec[:, 0] / nx
    ky = 2 * np.pi * ks.kvec[:, 1] / nx
    a = 0.3 * (-2 * np.cos(kx) - 2 * np.cos(ky)) + 0.1 * np.cos(kx) * np.cos(ky)
    b = -0.5 - 0.2 * (np.cos(kx) + np.cos(ky))
    nu = (2.0 * (np.arange(nf) - nf // 2) + 1.0) * np.pi / BETA
    sig_k = a[:, None] + 1j * b[:, None] * (nu / (1.0 + nu * nu))[None, :]
    sig_h_k = a.astype(np.complex128)
    # the oracle's layouts: Sigma index ((kx-1) Ny + ky-1) Nf + nu-1, Sigma_H (Nx, Ny)
    sigma = np.zeros(nx * nx * nf, dtype=np.complex128)
    sigma_h = np.zeros((nx, nx), dtype=np.complex128)
    for k in range(ks.nc):
        x, y = ks.kvec[k]
        sigma[(x * nx + y) * nf:(x * nx + y + 1) * nf] = sig_k[k]
        sigma_h[x, y] = sig_h_k[k]
    return lat, ks, fulls, refs, lam_d, lam_m, sig_k, sig_h_k, sigma, sigma_h


def _oracle(lat, ks, refs, lam_d, lam_m, sigma, sigma_h, f_range=0):
    got = S.calc_sde(lat, refs, (lam_d, lam_m), sigma, sigma_h, MU, BETA, U, f_range=f_range)
    nx, nf = lat.nx, lat.nf
    out = np.zeros((ks.nc, nf), dtype=np.complex128)
    for k in range(ks.nc):
        x, y = ks.kvec[k]
        out[k] = got[(x * nx + y) * nf:(x * nx + y + 1) * nf]
    return out


@pytest.mark.parametrize("nx,nf,nl,f_range", [(4, 8, 1, 0), (4, 6, 5, 0), (4, 4, 9, 0), (4, 6, 5, 1), (4, 4, 1, 2)])
def test_sde_oracle_equals_kspace_sde(nx, nf, nl, f_range):
    lat, ks, fulls, refs, lam_d, lam_m, sig_k, sig_h_k, sigma, sigma_h = _inputs(nx, nf, nl, 10 + nl)
    exp = sde_kspace(lat, fulls, lam_d, lam_m, sig_k, sig_h_k, MU, BETA, U, f_range=f_range)
    got = _oracle(lat, ks, refs, lam_d, lam_m, sigma, sigma_h, f_range)
    err = rel_fro(got, exp)
    assert err < TOL, err
    # the vertices carry weight: without them the result moves far outside the bar
    zero = [np.zeros_like(r) for r in refs]
    assert rel_fro(_oracle(lat, ks, zero, lam_d, lam_m, sigma, sigma_h, f_range), exp) > 1e-3


def _planted(monkeypatch, kind):
    if kind == "transposed_contraction":
        orig = S._v_chi
        monkeypatch.setattr(S, "_v_chi", lambda lat, caux, g0, g1, c1, fr: orig(
            lat, caux, np.ascontiguousarray(g0.transpose(1, 0, 2)), np.ascontiguousarray(g1.transpose(1, 0, 2)),
            c1, fr))
    elif kind == "ph_weight":
        monkeypatch.setattr(S, "ph_chi", lambda lat, cph, g_d, g_m, f_range=0: S._v_chi(lat, cph, g_d, g_m, -1.0,
                                                                                           f_range))
    elif kind == "pp_sign":
        monkeypatch.setattr(S, "pp_chi", lambda lat, cpp, g_s, g_t, f_range=0: S._v_chi(lat, cpp, g_s, g_t, -1.0,
                                                                                           f_range))
    elif kind == "no_ff_signs":
        pass  # applied to the lattice below
    else:
        raise ValueError(kind)


@pytest.mark.parametrize("kind", ["transposed_contraction", "ph_weight", "pp_sign", "no_ff_signs"])
def test_planted_sde_mistakes_are_caught(monkeypatch, kind):
    nx, nf, nl = 4, 6, 5
    lat, ks, fulls, refs, lam_d, lam_m, sig_k, sig_h_k, sigma, sigma_h = _inputs(nx, nf, nl, 15)
    exp = sde_kspace(lat, fulls, lam_d, lam_m, sig_k, sig_h_k, MU, BETA, U)
    _planted(monkeypatch, kind)
    lat2 = Lattice(nx, nf, nl)
    if kind == "no_ff_signs":
        assert (lat2.signs == -1).any()
        lat2.signs = np.ones_like(lat2.signs)
    got = _oracle(lat2, ks, refs, lam_d, lam_m, sigma, sigma_h)
    assert rel_fro(got, exp) > 1e-6, kind
