"""GPU parity of calc_Gkw_Grt's mu bisection and G(r, tau) (tups_gkw_grt) and of
the SDE's second-order term, against the oracle (oracle/gtau.py).

The device search evaluates six levels of the bisection tree per pass and walks
them as the sequential loop would, so mu and the number of evaluations must be
the oracle's exactly (the one-workgroup search of small grids and the
multi-pass search of larger ones, 8x8 at Ngrain 4); Grt (direct 2D DFTs on the
device, numpy FFTs in the oracle) agrees to 1e-12 relative Frobenius error.
"""
import sys

import numpy as np
import pytest

from oracle import gtau as T
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12
BETA, U = 2.0, 2.0


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize("nx,nf,ngrain,n_target",
                         [(4, 16, 1, 0.9), (4, 16, 3, 1.1), (8, 32, 2, 0.75), (8, 32, 4, 0.8), (6, 8, 1, 1.0)])
def test_gkw_grt_bisection_parity(nx, nf, ngrain, n_target):
    from tups_amd import TupsContext

    _say(f"gkw_grt nx={nx} nf={nf} ngrain={ngrain} n={n_target}")
    rng = np.random.default_rng(nx * 1000 + nf * 10 + ngrain)
    lat = Lattice(nx, nf, 1)
    sigma = crandn(rng, (nx * nx * nf,), 0.2).ravel()
    mu0 = 0.3
    with TupsContext(nx=nx, nf=nf, nl=1, beta=BETA, f_damping=0.5) as c:
        c.set_sigma(sigma, mu0, ngrain=ngrain)
        mu, steps, grt = c.gkw_grt(U, n_target, mu0)
        mu_dev, steps_dev = c.get_mu()
    mu_o, grt_o, steps_o = T.bisect_mu(lat, sigma, mu0, BETA, U, n_target, ngrain)
    err = rel_fro(grt, grt_o)
    _say(f"  mu {mu!r} (oracle {mu_o!r}) steps {steps} (oracle {steps_o}) Grt rel err {err:.2e}")
    assert steps == steps_o and mu == mu_o
    assert mu_dev == mu and steps_dev == steps
    assert err < TOL
    assert abs(T.particle_number(grt) - n_target) <= 1e-6


def test_gkw_grt_cal_chi_evaluates_once():
    """cal_chi: one evaluation at 0.5 ((mu + U) + (mu - U)) and no search (SRC/parquet_chi.f90:186-189)."""
    from tups_amd import TupsContext

    nx, nf = 4, 16
    rng = np.random.default_rng(5)
    lat = Lattice(nx, nf, 1)
    sigma = crandn(rng, (nx * nx * nf,), 0.2).ravel()
    mu0 = 0.15
    with TupsContext(nx=nx, nf=nf, nl=1, beta=BETA, f_damping=0.5) as c:
        c.set_sigma(sigma, mu0)
        mu, steps, grt = c.gkw_grt(U, 0.5, mu0, cal_chi=True)
    assert steps == 1 and mu == 0.5 * ((mu0 + U) + (mu0 - U))
    assert rel_fro(grt, T.grt(lat, sigma, mu, BETA)) < TOL


@pytest.mark.parametrize("nx,nf,ngrain", [(4, 8, 1),     # one-workgroup search
                                          (8, 32, 4)])   # multi-workgroup passes + tree walk
def test_gkw_grt_reports_no_convergence(nx, nf, ngrain):
    """A density the bracket cannot reach (n = 5 > 2) ends after 64 evaluations with
    TUPS_ENOCONV, where the reference would loop on; every launch exits normally."""
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=nx, nf=nf, nl=1, beta=BETA, f_damping=0.5) as c:
        c.set_sigma(np.zeros(nx * nx * nf, dtype=np.complex128), 0.0, ngrain=ngrain)
        with pytest.raises(TupsError) as ei:
            c.gkw_grt(U, 5.0, 0.0)
        assert "ENOCONV" in str(ei.value)
        with pytest.raises(TupsError):
            c.get_mu()
        mu, steps, _ = c.gkw_grt(U, 1.0, 0.0)  # the context stays usable
        assert steps == 1 and mu == 0.0


def test_gkw_grt_needs_sigma():
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=4, nf=8, nl=1, beta=BETA, f_damping=0.5) as c:
        with pytest.raises(TupsError):
            c.gkw_grt(U, 1.0, 0.0)


@pytest.mark.parametrize("nx,nf,ngrain", [(4, 16, 1), (4, 8, 2), (8, 16, 1)])
def test_order_2nd_parity(nx, nf, ngrain):
    """The SDE with use_u2 = the oracle's SDE (use_U2 form of LChi) + calc_order_2nd of
    the device Grt, at 1e-12."""
    from oracle import sde as S
    from tups_amd import GAMMA, LAMBDA, TupsContext

    _say(f"order_2nd nx={nx} nf={nf} ngrain={ngrain}")
    rng = np.random.default_rng(nx * 31 + nf + ngrain)
    lat = Lattice(nx, nf, 1)
    gv = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.3) for _ in range(4)]
    lam = [crandn(rng, (nf, nf, lat.nw), 1.0) for _ in range(4)]
    sigma = crandn(rng, (nx * nx * nf,), 0.2).ravel()
    mu0 = 0.1
    with TupsContext(nx=nx, nf=nf, nl=1, beta=BETA, f_damping=0.5) as c:
        for r in range(4):
            c.upload(GAMMA, r, gv[r])
            c.upload(LAMBDA, r, lam[r])
        c.set_sigma(sigma, mu0, ngrain=ngrain)
        mu, _, grt = c.gkw_grt(U, 1.0, mu0)
        got = c.sde(sigma, mu, U, ngrain=ngrain, use_u2=True)
    sh = sigma.reshape(nx, nx, nf)[:, :, nf - 1].copy()
    grt_o = T.grt(lat, sigma, mu, BETA, ngrain)
    assert rel_fro(grt, grt_o) < TOL
    u2 = T.order_2nd(lat, grt_o, U, BETA, ngrain)
    exp = S.calc_sde(lat, gv, (lam[0], lam[1]), sigma, sh, mu, BETA, U, ngrain=ngrain, use_u2=True) + u2
    err = rel_fro(got, exp)
    _say(f"  rel err {err:.2e} (second-order share {np.linalg.norm(u2) / np.linalg.norm(exp):.2f})")
    assert err < TOL


def test_sde_use_u2_needs_grt():
    from tups_amd import TupsContext, TupsError

    nx, nf = 4, 8
    with TupsContext(nx=nx, nf=nf, nl=1, beta=BETA, f_damping=0.5) as c:
        with pytest.raises(TupsError):
            c.sde(np.zeros(nx * nx * nf, dtype=np.complex128), 0.0, U, use_u2=True)


@pytest.mark.parametrize("nx,nf,ngrain", [(4, 16, 1), (6, 8, 2), (8, 16, 1)])
def test_chi0_parity(nx, nf, ngrain):
    """The bare bubbles of calc_Gkw_Grt with cal_chi (plain and form-factor weighted,
    ph and pp) from the device Grt against the oracle at the same mu, 1e-12."""
    from tups_amd import TupsContext

    _say(f"chi0 nx={nx} nf={nf} ngrain={ngrain}")
    rng = np.random.default_rng(nx * 7 + nf + ngrain)
    lat = Lattice(nx, nf, 1)
    sigma = crandn(rng, (nx * nx * nf,), 0.2).ravel()
    mu0 = 0.2
    with TupsContext(nx=nx, nf=nf, nl=1, beta=BETA, f_damping=0.5) as c:
        c.set_sigma(sigma, mu0, ngrain=ngrain)
        mu, _, _ = c.gkw_grt(U, 1.0, mu0, cal_chi=True, want_grt=False)
        got = c.chi0()
    g = T.grt(lat, sigma, mu, BETA, ngrain)
    ffs = [T.ffg(lat, sigma, mu, BETA, ngrain, f) for f in range(4)]
    exp = T.chi0(lat, g, ffs, BETA, ngrain)
    for name, a, b in zip(("ph", "pp", "ph_ff", "pp_ff"), got, exp):
        err = rel_fro(a, b)
        _say(f"  {name}: rel err {err:.2e}")
        assert err < TOL, name


def test_chi0_needs_grt():
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=4, nf=8, nl=1, beta=BETA, f_damping=0.5) as c:
        c.set_sigma(np.zeros(4 * 4 * 8, dtype=np.complex128), 0.0)
        with pytest.raises(TupsError):
            c.chi0()
