"""GPU parity of the susceptibilities (tups_calc_sus) against the oracle (oracle/sus.py).

calc_sus of SRC/parquet_sus.f90:18-362: the Green's-function pairs GG_ph / GG_pp
with the four external form factors over the enlarged box, contracted with F
inside the box and with the bare vertex and the clamped Phi outside it; random
F, Phi and Sigma (with Sigma_H), Nl 1 and 5, f_range 0 and 1, coarse graining 1
and 2; 1e-12 relative Frobenius error per channel.
"""
import sys

import numpy as np
import pytest

from oracle import sus as S
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro

pytestmark = pytest.mark.gpu
BETA, U = 2.0, 2.0


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize("nx,nf,nl,f_range,ngrain", [(4, 8, 1, 0, 1), (4, 8, 5, 0, 2), (4, 8, 1, 1, 1),
                                                     (4, 6, 5, 1, 1)])
def test_sus_parity(nx, nf, nl, f_range, ngrain):
    from tups_amd import F, GAMMA, TupsContext

    _say(f"sus parity nx={nx} nf={nf} nl={nl} f_range={f_range} ngrain={ngrain}")
    rng = np.random.default_rng(nx * 100 + nf * 10 + nl + f_range)
    lat = Lattice(nx, nf, nl)
    fv = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.5) for _ in range(4)]
    phi = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.3) for _ in range(4)]
    sigma = crandn(rng, (nx * nx * nf,), 0.2).ravel()
    sh = crandn(rng, (nx, nx), 0.1)
    mu = 0.15
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=BETA, f_damping=0.5) as c:
        for r in range(4):
            c.upload(F, r, fv[r])
            c.upload(GAMMA, r, phi[r])
        c.set_sigma(sigma, mu, ngrain=ngrain, sigma_h=sh)
        got = c.calc_sus(U, f_range=f_range)
        for r in range(4):  # the same with the vertex arrays holding Gamma = F - Phi
            c.upload(GAMMA, r, fv[r] - phi[r])
        upd = c.calc_sus(U, f_range=f_range, after_update=True)
    ph, pp = S.gg(lat, sigma, sh, mu, BETA, ngrain, f_range)
    exp = S.calc_sus(lat, fv, phi, ph, pp, U, BETA, f_range)
    for r in range(4):
        e = exp[r].T  # (4 ff, Nq)
        err, err_u = rel_fro(got[r], e), rel_fro(upd[r], e)
        _say(f"  channel {r}: rel err {err:.2e}, after the update {err_u:.2e}")
        assert err < 1e-12 and err_u < 1e-12


def test_sus_needs_sigma():
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=4, nf=8, nl=1, beta=BETA, f_damping=0.5) as c:
        with pytest.raises(TupsError):
            c.calc_sus(U)
