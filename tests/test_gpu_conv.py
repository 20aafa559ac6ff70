"""GPU parity of the convergence check (tups_conv) against the oracle (oracle/conv.py).

conv() of SRC/parquet_check.f90:39-150: the Sigma change (average, maximum and its
location) and the per-channel vertex maxima with their change since the last
call.  The maxima pick the same element (max is order-independent) and agree to
the last bits of hypot (device and host libm may differ by an ulp: 1e-15); the
average is a sum in another order (1e-14).
"""
import sys

import numpy as np
import pytest

from oracle import conv as O
from tests.helpers import crandn

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _check(got, exp):
    assert (got["kx"], got["ky"], got["nu"]) == (exp["kx"], exp["ky"], exp["nu"])
    assert abs(got["sigma_max"] - exp["sigma_max"]) <= 1e-15 * exp["sigma_max"]
    assert abs(got["sigma_avg"] - exp["sigma_avg"]) <= 1e-14 * exp["sigma_avg"]
    for g, e in zip(got["f_max"], exp["f_max"]):
        assert abs(g - e) <= 1e-15 * e
    for g, e in zip(got["f_rel_err"], exp["f_rel_err"]):
        assert g == e or abs(g - e) <= 1e-6 * abs(e) + 1e-15
    assert got["converged"] == exp["converged"]


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 1), (8, 16, 5)])
def test_conv_host_pair_parity(nx, nf, nl):
    from tups_amd import F, TupsContext

    _say(f"conv nx={nx} nf={nf} nl={nl}")
    rng = np.random.default_rng(nx + nf + nl)
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as c:
        fv = [crandn(rng, (c.nz, c.nz, c.nq), 0.3) for _ in range(4)]
        fv[1][2, 3, 1] = 9.0 + 4.0j  # a planted largest element
        for r in range(4):
            c.upload(F, r, fv[r])
        sigma = crandn(rng, (nx * nx * nf,), 0.5).ravel()
        old = sigma * (1.0 + 1e-3 * crandn(rng, (nx * nx * nf,), 1.0).ravel())
        prev = [0.0] * 4
        for call in range(3):
            got = c.conv(sigma_old=old, sigma=sigma)
            exp = O.conv(old, sigma, fv, prev, nx, nx, nf)
            _check(got, exp)
            prev = exp["f_max"]
            fv = [f * (1.0 + 1e-9 * (call + 1)) for f in fv]
            for r in range(4):
                c.upload(F, r, fv[r])
            old = sigma * (1.0 + 1e-10)
        assert got["converged"]  # 1e-10 Sigma change and 2e-9 vertex change < 1e-8


def test_conv_device_pair_after_iteration():
    """After tups_iteration_full the device holds SigmaOld (the Sigma the iteration
    started from) and the new Sigma: conv on that pair equals conv on the host copies."""
    from tups_amd import F, GAMMA, LAMBDA, TupsContext

    nx, nf, nl, U, mu = 4, 8, 1, 2.0, 0.1
    rng = np.random.default_rng(3)
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as c:
        for r in range(4):
            c.upload(GAMMA, r, crandn(rng, (c.nz, c.nz, c.nq), 0.05) + 0.5 * np.eye(c.nz)[:, :, None])
            c.upload(F, r, crandn(rng, (c.nz, c.nz, c.nq), 0.05) + 0.5 * np.eye(c.nz)[:, :, None])
            c.upload(LAMBDA, r, crandn(rng, (nf, nf, c.nw), 0.1))
        c.calc_M()
        sigma0 = crandn(rng, (nx * nx * nf,), 0.1).ravel()
        c.set_sigma(sigma0, mu)
        with pytest.raises(Exception):
            c.conv()  # no pair before an iteration
        c.iteration_full(U, graph=True)
        sigma1 = c.get_sigma()
        got = c.conv()
        fv = [c.download(F, r) for r in range(4)]
        exp = O.conv(sigma0, sigma1, fv, [0.0] * 4, nx, nx, nf)
        _check(got, exp)
        assert not got["converged"]
