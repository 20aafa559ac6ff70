"""GPU parity of the self-energy (tups_sde) against the oracle.

calc_SDE (SRC/parquet_SDE.f90:27-258; with use_U2 also calc_order_2nd from the
device Grt, whose own parity is in test_gpu_gtau.py): the oracle
(oracle/sde.py) is pinned by the brute-force second-order diagram and the
full-zone star sum; here the GPU kernels must reproduce it on random vertices,
Lambda and Sigma at 1e-12 relative Frobenius error, across Nl, the enlarged
frequency box (f_range), coarse graining and the use_U2 form of LChi.
"""
import sys

import numpy as np
import pytest

from oracle import gtau as T
from oracle import sde as S
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12
BETA, U = 2.0, 2.0


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize(
    "nx,nf,nl,f_range,ngrain,use_u2,with_sh",
    [(4, 8, 1, 0, 1, False, False),
     (4, 6, 5, 1, 2, True, True),
     (8, 8, 5, 0, 1, False, False),
     (8, 4, 9, 0, 1, False, True)],
)
def test_sde_parity(nx, nf, nl, f_range, ngrain, use_u2, with_sh):
    from tups_amd import GAMMA, LAMBDA, TupsContext

    _say(f"sde parity nx={nx} nf={nf} nl={nl} f_range={f_range} ngrain={ngrain} use_u2={use_u2}")
    rng = np.random.default_rng(nx * 100 + nf * 10 + nl)
    lat = Lattice(nx, nf, nl)
    gv = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.3) for _ in range(4)]
    lam = [crandn(rng, (nf, nf, lat.nw), 1.0) for _ in range(4)]
    sigma = crandn(rng, (nx * nx * nf,), 0.2).ravel()
    sigma_h = crandn(rng, (nx, nx), 0.1) if with_sh else None
    mu = 0.15
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=BETA, f_damping=0.5) as c:
        for r in range(4):
            c.upload(GAMMA, r, gv[r])
            c.upload(LAMBDA, r, lam[r])
        u2 = 0.0
        if use_u2:  # Grt of calc_Gkw_Grt at this mu: one evaluation at 0.5 ((mu + U) + (mu - U))
            c.set_sigma(sigma, mu, ngrain=ngrain)
            mu_g, _, _ = c.gkw_grt(U, 1.0, mu, cal_chi=True)
            u2 = T.order_2nd(lat, T.grt(lat, sigma, mu_g, BETA, ngrain), U, BETA, ngrain)
        got = c.sde(sigma, mu, U, ngrain=ngrain, f_range=f_range, use_u2=use_u2, sigma_h=sigma_h)
        again = c.sde(sigma, mu, U, ngrain=ngrain, f_range=f_range, use_u2=use_u2, sigma_h=sigma_h)
    sh = sigma_h
    if sh is None:  # the default rule Sigma_H(k) = Sigma(k, Nf)
        sh = sigma.reshape(nx, nx, nf)[:, :, nf - 1].copy()
    exp = S.calc_sde(lat, gv, (lam[0], lam[1]), sigma, sh, mu, BETA, U, ngrain=ngrain, f_range=f_range,
                     use_u2=use_u2) + u2
    err = rel_fro(got, exp)
    _say(f"  rel err {err:.2e}")
    assert err < TOL
    assert np.array_equal(got, again)  # deterministic: no atomics in the reductions


def test_sde_rejects_partial_context():
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=4, nf=8, nl=1, nq=3, beta=BETA, f_damping=0.5) as c:
        with pytest.raises(TupsError):
            c.sde(np.zeros(4 * 4 * 8, dtype=np.complex128), 0.0, U)
