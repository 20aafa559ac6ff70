"""tups_conv_noalloc (the convergence check with its partials and staged Sigma
pair in the library's fixed device array) against tups_conv, through the C-ABI.

Two contexts hold the same vertices, so each entry point keeps its own
previous-maximum history and every field must agree exactly: the two run the
same kernels on the same data in the same order.  Two calls in a row, the
second on a smaller context, so stale scratch contents of the first call would
show; a third call on the device Sigma pair of an iteration; and a Sigma beyond
the fixed scratch is refused with an error, not run.
"""
import sys

import numpy as np
import pytest

from tests.helpers import crandn

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _same(a, b):
    for k in ("sigma_avg", "sigma_max", "kx", "ky", "nu", "converged"):
        assert a[k] == b[k], (k, a[k], b[k])
    for k in ("f_max", "f_rel_err"):
        assert a[k] == b[k], (k, a[k], b[k])


def _pair(nx, nf, nl, seed):
    from tups_amd import F, TupsContext

    rng = np.random.default_rng(seed)
    ca = TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5)
    cb = TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5)
    fv = [crandn(rng, (ca.nz, ca.nz, ca.nq), 0.3) for _ in range(4)]
    for r in range(4):
        ca.upload(F, r, fv[r])
        cb.upload(F, r, fv[r])
    sigma = crandn(rng, (nx * nx * nf,), 0.5).ravel()
    old = sigma * (1.0 + 1e-3 * crandn(rng, (nx * nx * nf,), 1.0).ravel())
    return ca, cb, fv, old, sigma


@pytest.mark.parametrize("sizes", [((8, 32, 5), (4, 8, 1))])
def test_noalloc_conv_equals_conv_on_two_calls(sizes):
    for call, (nx, nf, nl) in enumerate(sizes):
        _say(f"noalloc conv call {call}: nx={nx} nf={nf} nl={nl}")
        ca, cb, fv, old, sigma = _pair(nx, nf, nl, 7 + call)
        try:
            for rep in range(2):  # the second pair of calls also checks the history (f_rel_err)
                a = ca.conv(sigma_old=old, sigma=sigma)
                b = cb.conv(sigma_old=old, sigma=sigma, noalloc=True)
                _same(a, b)
                assert np.isfinite(b["f_max"]).all() and b["sigma_max"] > 0
                sigma = sigma * (1.0 + 1e-9)
        finally:
            ca.close()
            cb.close()


def test_noalloc_conv_device_pair_after_iteration():
    from tups_amd import F, GAMMA, LAMBDA, TupsContext

    nx, nf, nl, U, mu = 4, 8, 1, 2.0, 0.1
    _say(f"noalloc conv device pair: nx={nx} nf={nf} nl={nl}")
    ctxs = [TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) for _ in range(2)]
    try:
        for c in ctxs:
            rng = np.random.default_rng(3)
            for r in range(4):
                c.upload(GAMMA, r, crandn(rng, (c.nz, c.nz, c.nq), 0.05) + 0.5 * np.eye(c.nz)[:, :, None])
                c.upload(F, r, crandn(rng, (c.nz, c.nz, c.nq), 0.05) + 0.5 * np.eye(c.nz)[:, :, None])
                c.upload(LAMBDA, r, crandn(rng, (nf, nf, c.nw), 0.1))
            c.calc_M()
            c.set_sigma(crandn(rng, (nx * nx * nf,), 0.1).ravel(), mu)
            c.iteration_full(U, graph=False)
        _same(ctxs[0].conv(), ctxs[1].conv(noalloc=True))
    finally:
        for c in ctxs:
            c.close()


def test_noalloc_conv_rejects_sigma_beyond_scratch():
    from tups_amd import TupsContext
    from tups_amd.runtime import TupsError

    nx, nf = 32, 128  # Nx Ny Nf = 131072 > 65536
    _say(f"noalloc conv refusal: nx={nx} nf={nf}")
    with TupsContext(nx=nx, nf=nf, nl=1, beta=2.0, f_damping=0.5, nq=4) as c:
        s = np.ones(nx * nx * nf, dtype=np.complex128)
        with pytest.raises(TupsError, match="fixed scratch"):
            c.conv(sigma_old=s, sigma=s, noalloc=True)
