"""GPU parity of calc_chi (SRC/parquet_chi.f90:438-523) against the oracle.

The oracle's calc_chi is pinned by a brute-force bubble, the Nl = 1 closed form
and the Hartree mu-shift invariant (tests/test_oracle_bse.py).  Here the
library's bubbles for every q slot (compact Nl x Nl blocks per frequency) are
compared with it on sampled slots, with and without coarse graining and with
an explicit or derived Sigma_H.
"""
import sys

import numpy as np
import pytest

from oracle import bse as B
from oracle.lattice import Lattice
from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _compact(lat, dense):
    nf, nl = lat.nf, lat.nl
    out = np.zeros((nl, nl, nf), complex)
    nu = np.arange(nf)
    for l1 in range(nl):
        for l2 in range(nl):
            out[l1, l2] = dense[l1 * nf + nu, l2 * nf + nu]
    return out


@pytest.mark.parametrize("nx,nf,nl,ngrain,explicit_h",
                         [(4, 8, 1, 1, False), (4, 8, 5, 2, True), (6, 4, 9, 1, True), (8, 16, 5, 3, False)])
def test_calc_chi_parity(nx, nf, nl, ngrain, explicit_h):
    from tups_amd import CHI, TupsContext

    _say(f"calc_chi parity nx={nx} nf={nf} nl={nl} ngrain={ngrain}")
    rng = np.random.default_rng(nx + nf + nl + ngrain)
    lat = Lattice(nx, nf, nl)
    beta, mu = 2.0, 0.15
    sigma = 0.2 * (rng.standard_normal(lat.nc * nf) + 1j * rng.standard_normal(lat.nc * nf))
    if explicit_h:
        sigma_h = 0.1 * (rng.standard_normal((nx, nx)) + 1j * rng.standard_normal((nx, nx)))
    else:  # the reference's rule: Sigma_H(kx, ky) = Sigma(k, Nf)
        sigma_h = np.array([[sigma[((kx * nx) + ky) * nf + nf - 1] for ky in range(nx)] for kx in range(nx)])
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=beta, f_damping=0.5) as c:
        c.calc_chi(sigma, mu, ngrain, sigma_h if explicit_h else None)
        ph = c.download(CHI, 0)
        pp = c.download(CHI, 1)
        qx, qy, qw = c.q_map()
    slots = sorted(set(np.linspace(0, len(qx) - 1, 6).astype(int).tolist()))
    for q in slots:
        eph, epp = B.calc_chi(lat, int(qx[q]), int(qy[q]), int(qw[q]), sigma, mu, beta, ngrain, sigma_h)
        e1, e2 = rel_fro(ph[:, :, :, q], _compact(lat, eph)), rel_fro(pp[:, :, :, q], _compact(lat, epp))
        _say(f"  q={q} ({qx[q]},{qy[q]},{qw[q]}): ph {e1:.1e} pp {e2:.1e}")
        assert e1 < TOL and e2 < TOL
