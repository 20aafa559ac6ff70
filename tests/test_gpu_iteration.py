"""One parquet iteration of the path as a replayed HIP graph (tups_iteration).

The graph runs the same kernels as the step-by-step calls BSE(ite) ->
calc_PhiR -> calculate_PAE -> G = F - G (SRC/loop.f90:78-139,219-222), so every
replay must equal the eager sequence bit for bit.  Replays alternate between
the two vertex-buffer parities, and host work with synchronisation runs
between them.
"""
import sys

import numpy as np
import pytest

from tests.helpers import crandn, random_chi_compact, random_vertices

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _setup(c, seed):
    from tups_amd import CHI, F, GAMMA, LAMBDA

    rng = np.random.default_rng(seed)
    gam = random_vertices(rng, c.nz, c.nq, diag=0.5, scale=0.05)
    fv = random_vertices(rng, c.nz, c.nq, diag=0.5, scale=0.05)
    for r in range(4):
        c.upload(GAMMA, r, gam[r])
        c.upload(F, r, fv[r])
        c.upload(LAMBDA, r, crandn(rng, (c.nf, c.nf, c.nw), 0.1))
    for p in range(2):
        c.upload(CHI, p, random_chi_compact(rng, c.nl, c.nf, c.nq, 0.5))
    c.calc_M()
    c.sync()


@pytest.mark.parametrize("nx,nf,nl,split", [(4, 8, 1, False), (4, 8, 5, False), (6, 8, 1, True)])
def test_graph_iteration_equals_eager(nx, nf, nl, split):
    from tups_amd import BSE_LADDER_FP64, BSE_LADDER_SPLIT, F, GAMMA, PHIR, TupsContext

    mode = BSE_LADDER_SPLIT if split else BSE_LADDER_FP64
    _say(f"graph iteration nx={nx} nf={nf} nl={nl} mode={mode}")
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as a, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as b, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as e:
        for c in (a, b, e):
            _setup(c, 17)
        for it in range(4):
            a.BSE(it, mode=mode)
            a.calc_PhiR(it)
            a.calculate_PAE()
            a.gamma_update()
            b.iteration(mode, graph=True)
            e.iteration(mode, graph=False)
            for which in (GAMMA, F, PHIR):
                for r in range(4):
                    x = a.download(which, r)
                    assert np.array_equal(x, b.download(which, r)), (it, which, r)
                    assert np.array_equal(x, e.download(which, r)), (it, which, r)
            assert np.isfinite(a.download(GAMMA, 0)).all()


def test_iteration_needs_m_tensors():
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=4, nf=8, nl=1, beta=2.0, f_damping=0.5) as c:
        with pytest.raises(TupsError):
            c.iteration(graph=True)


@pytest.mark.parametrize("nx,nf,nl,split,f_range", [(4, 8, 1, False, 0), (4, 8, 5, False, 1), (6, 8, 1, True, 0)])
def test_full_iteration_graph_equals_steps(nx, nf, nl, split, f_range):
    """calc_chi -> BSE -> PhiR -> PAE -> SDE -> Sigma -> G = F - G with Sigma on the
    device, as a graph and eagerly, against the same library calls made one by one
    with Sigma passed through the host (SRC/loop.f90:67-222 without the mu bisection)."""
    from tups_amd import BSE_LADDER_FP64, BSE_LADDER_SPLIT, F, GAMMA, PHIR, TupsContext

    mode = BSE_LADDER_SPLIT if split else BSE_LADDER_FP64
    U, mu = 2.0, 0.1
    _say(f"full iteration nx={nx} nf={nf} nl={nl} mode={mode} f_range={f_range}")
    rng = np.random.default_rng(23)
    sigma0 = crandn(rng, (nx * nx * nf,), 0.1).ravel()
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as a, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as b, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as e:
        for c in (a, b, e):
            _setup(c, 29)
        b.set_sigma(sigma0, mu)
        e.set_sigma(sigma0, mu)
        sigma = sigma0.copy()
        for it in range(3):
            a.calc_chi(sigma, mu)
            a.BSE(it, mode=mode)
            a.calc_PhiR(it)
            a.calculate_PAE()
            new_sigma = a.sde(sigma, mu, U, f_range=f_range)
            a.gamma_update()
            sigma = new_sigma
            b.iteration_full(U, mode, graph=True, f_range=f_range)
            e.iteration_full(U, mode, graph=False, f_range=f_range)
            assert np.array_equal(sigma, b.get_sigma()), it
            assert np.array_equal(sigma, e.get_sigma()), it
            for which in (GAMMA, F, PHIR):
                for r in range(4):
                    x = a.download(which, r)
                    assert np.array_equal(x, b.download(which, r)), (it, which, r)
                    assert np.array_equal(x, e.download(which, r)), (it, which, r)
            assert np.isfinite(sigma).all()


def test_full_iteration_needs_sigma():
    from tups_amd import TupsContext, TupsError

    with TupsContext(nx=4, nf=8, nl=1, beta=2.0, f_damping=0.5) as c:
        _setup(c, 3)
        with pytest.raises(TupsError):
            c.iteration_full(2.0, graph=True)


@pytest.mark.parametrize("nx,nf,nl,split,ngrain", [(4, 8, 1, False, 1), (4, 8, 5, False, 2), (6, 8, 1, True, 1)])
def test_full_iteration_with_mu_search_equals_steps(nx, nf, nl, split, ngrain):
    """The whole loop body of SRC/loop.f90:67-222: calc_Gkw_Grt's mu bisection ->
    calc_chi -> BSE -> PhiR -> PAE -> SDE with use_U2 (its second-order term from
    this iteration's Grt) -> Sigma -> G = F - G, as a graph and eagerly, against the
    library calls made one by one with Sigma and mu passed through the host."""
    from tups_amd import BSE_LADDER_FP64, BSE_LADDER_SPLIT, F, GAMMA, PHIR, TupsContext

    mode = BSE_LADDER_SPLIT if split else BSE_LADDER_FP64
    U, mu0, n_target = 2.0, 0.4, 0.95
    _say(f"full iteration with mu search nx={nx} nf={nf} nl={nl} mode={mode} ngrain={ngrain}")
    rng = np.random.default_rng(41)
    sigma0 = crandn(rng, (nx * nx * nf,), 0.1).ravel()
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as a, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as b, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as e:
        for c in (a, b, e):
            _setup(c, 31)
        b.set_sigma(sigma0, mu0, ngrain=ngrain)
        e.set_sigma(sigma0, mu0, ngrain=ngrain)
        sigma, mu = sigma0.copy(), mu0
        for it in range(3):
            a.set_sigma(sigma, mu, ngrain=ngrain)
            mu, steps, _ = a.gkw_grt(U, n_target, mu)
            a.calc_chi(sigma, mu, ngrain=ngrain)
            a.BSE(it, mode=mode)
            a.calc_PhiR(it)
            a.calculate_PAE()
            new_sigma = a.sde(sigma, mu, U, ngrain=ngrain, use_u2=True)
            a.gamma_update()
            sigma = new_sigma
            b.iteration_full(U, mode, graph=True, use_u2=True, n_particle=n_target)
            e.iteration_full(U, mode, graph=False, use_u2=True, n_particle=n_target)
            for x in (b, e):
                assert x.get_mu() == (mu, steps), it
                assert np.array_equal(sigma, x.get_sigma()), it
            for which in (GAMMA, F, PHIR):
                for r in range(4):
                    x = a.download(which, r)
                    assert np.array_equal(x, b.download(which, r)), (it, which, r)
                    assert np.array_equal(x, e.download(which, r)), (it, which, r)
            assert np.isfinite(sigma).all()
            _say(f"  it {it}: mu {mu:.6f} after {steps} evaluations")


def test_full_iteration_graph_follows_sigma_h_rule():
    """The Sigma commit inside the full-iteration graph applies the Sigma_H rule in force at
    capture (Sigma_H = Sigma(k, Nf) when set_sigma got no Sigma_H, else the explicit one).
    Switching the rule between calls must re-capture: graph and eager stay bit-identical."""
    from tups_amd import BSE_LADDER_FP64, F, GAMMA, TupsContext

    nx, nf, nl, U, mu = 4, 8, 1, 2.0, 0.1
    rng = np.random.default_rng(61)
    sigma0 = crandn(rng, (nx * nx * nf,), 0.1).ravel()
    sigh = crandn(rng, (nx, nx), 0.2)
    with TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as b, \
            TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as e:
        for c in (b, e):
            _setup(c, 37)
        it = 0
        for explicit in (True, False, True):
            s = b.get_sigma() if it else sigma0
            for c in (b, e):
                c.set_sigma(s, mu, sigma_h=sigh if explicit else None)
            # two iterations per rule: the second one's bubbles read the Sigma_H that the
            # first one's commit left
            for _ in range(2):
                _say(f"  sigma_h rule switch, iteration {it}: explicit Sigma_H = {explicit}")
                b.iteration_full(U, BSE_LADDER_FP64, graph=True)
                e.iteration_full(U, BSE_LADDER_FP64, graph=False)
                assert np.array_equal(b.get_sigma(), e.get_sigma()), it
                for which in (GAMMA, F):
                    for r in range(4):
                        assert np.array_equal(b.download(which, r), e.download(which, r)), (it, which, r)
                it += 1
