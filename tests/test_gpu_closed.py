"""GPU parity of the closed-form BSE (TUPS_BSE_CLOSED_LU) against the oracle.

The closed form returns the fixed point of BSE(ite) (SRC/parquet_BSE.f90:62-132)
for fixed Gamma: Phi = (I - kappa_r Gamma chi)^-1 Gamma - Gamma.  It is pinned
three ways: against the oracle's LAPACK solve, by the library's own ladder
returning Phi unchanged when fed F = Gamma + Phi, and by the singular case
(kappa Gamma chi = I) being reported instead of returned.
"""
import sys

import numpy as np
import pytest

from oracle import bse as B
from oracle.lattice import Lattice
from tests.helpers import random_chi_compact, random_vertices, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _ctx(**kw):
    from tups_amd import TupsContext

    return TupsContext(**kw)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize("variant", [0, 1, 2])  # panels of 16, blocks of 64, blocks with register sub-panels
@pytest.mark.parametrize(
    "nx,nf,nl,nq",
    [(4, 16, 1, 0),     # Nz=16: one panel
     (4, 8, 5, 0),      # Nz=40: partial last panel (8 columns)
     (4, 6, 5, 3),      # Nz=30: panel of 14, odd tile edges
     (4, 8, 9, 2),      # Nz=72: two update tiles, a block of 64 and one of 8
     (8, 64, 5, 2),     # Nz=320: the batched-BSE shape, five blocks
     (8, 64, 9, 1)],    # Nz=576: above the register sub-panel's 512 rows
)
def test_closed_form_parity(nx, nf, nl, nq, variant):
    from tups_amd import BSE_CLOSED_LU, CHI, F, GAMMA
    from tups_amd.runtime import K_BSE_CLOSED

    _say(f"closed parity nx={nx} nf={nf} nl={nl} nq={nq} variant={variant}")
    rng = np.random.default_rng(nx * 31 + nf + nl)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=0.5) as c:
        c.set_variant(variant, K_BSE_CLOSED)
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        for r in range(4):
            c.upload(GAMMA, r, gam[r])
            c.upload(F, r, gam[r])
        for p in range(2):
            c.upload(CHI, p, chic[p])
        c.BSE(0, mode=BSE_CLOSED_LU)
        got = [c.download(GAMMA, r) for r in range(4)]
    exp = B.bse_closed(lat, gam, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0)
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: rel err {err:.2e}")
        assert err < TOL


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("nx,nf,nl,nq", [(4, 8, 9, 2), (8, 64, 5, 1)])
@pytest.mark.parametrize("refined", [False, True])
def test_closed_form_parity_with_pivoting(nx, nf, nl, nq, variant, refined):
    """Gamma without its diagonal and ten times larger: I - kappa Gamma chi is
    far from diagonally dominant, so partial pivoting swaps rows at most
    columns (226-295 of the 320 at D = 320, 65-67 of 72 at D = 72, counted with
    LAPACK's getrf on the same matrices).  The condition numbers reach 3e3, so
    the bar is 1e-10 against the oracle's LAPACK solve; a misplaced swap gives
    errors of order one."""
    from tups_amd import BSE_CLOSED_IR, BSE_CLOSED_LU, CHI, F, GAMMA
    from tups_amd.runtime import K_BSE_CLOSED, K_BSE_IR

    _say(f"closed parity, pivoting: nx={nx} nf={nf} nl={nl} nq={nq} variant={variant} refined={refined}")
    rng = np.random.default_rng(nx * 31 + nf + nl)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=0.5) as c:
        c.set_variant(variant, K_BSE_IR if refined else K_BSE_CLOSED)
        gam = random_vertices(rng, c.nz, c.nq, diag=0.0, scale=3.0)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        for r in range(4):
            c.upload(GAMMA, r, gam[r])
            c.upload(F, r, gam[r])
        for p in range(2):
            c.upload(CHI, p, chic[p])
        c.BSE(0, mode=BSE_CLOSED_IR if refined else BSE_CLOSED_LU)
        got = [c.download(GAMMA, r) for r in range(4)]
    exp = B.bse_closed(lat, gam, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0)
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: rel err {err:.2e}")
        assert err < 1e-10


def test_closed_form_is_fixed_point_of_gpu_ladder():
    from tups_amd import BSE_CLOSED_LU, CHI, F, GAMMA

    rng = np.random.default_rng(3)
    nx, nf, nl = 4, 8, 5
    with _ctx(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.3) as c:
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        for p in range(2):
            c.upload(CHI, p, chic[p])
        for r in range(4):
            c.upload(GAMMA, r, gam[r])
        c.BSE(0, mode=BSE_CLOSED_LU)
        phi = [c.download(GAMMA, r) for r in range(4)]
        for r in range(4):
            c.upload(GAMMA, r, gam[r])
            c.upload(F, r, gam[r] + phi[r])
        c.BSE(0)  # the damped FP64 ladder
        again = [c.download(GAMMA, r) for r in range(4)]
    for r in range(4):
        assert rel_fro(again[r], phi[r]) < TOL, r


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_closed_form_reports_singular_system(variant):
    """Gamma_d = I and chi_ph = 1/kappa_d make I - kappa_d Gamma_d chi_ph = 0."""
    from tups_amd import BSE_CLOSED_LU, CHI, GAMMA, TupsError
    from tups_amd.runtime import K_BSE_CLOSED

    nx, nf, nl = 4, 8, 1
    with _ctx(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as c:
        c.set_variant(variant, K_BSE_CLOSED)
        D, nq = c.nz, c.nq
        eye = np.asfortranarray(np.repeat(np.eye(D, dtype=complex)[:, :, None], nq, axis=2))
        for r in range(4):
            c.upload(GAMMA, r, eye)
        kappa_d = 2 * 0.5 * c.alpha
        c.upload(CHI, 0, np.asfortranarray(np.full((1, 1, nf, nq), 1.0 / kappa_d, complex)))
        c.upload(CHI, 1, np.asfortranarray(np.zeros((1, 1, nf, nq), complex)))
        with pytest.raises(TupsError):
            c.BSE(0, mode=BSE_CLOSED_LU)


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize(
    "nx,nf,nl,nq",
    [(4, 16, 1, 0), (4, 8, 5, 0), (4, 6, 5, 3), (4, 8, 9, 2), (8, 64, 5, 2), (8, 64, 9, 1)],
)
def test_refined_closed_form_parity(nx, nf, nl, nq, variant):
    """FP32 LU (Gauss-Jordan inverse) + FP64 iterative refinement: the residual
    |Gamma - A X|_F / |Gamma|_F reaches 1e-12 for every (q, r) and X matches the
    oracle's FP64 LAPACK solve to cond(A) times that; the FP32 inverse alone
    (no refinement step allowed) does not reach the tolerance and is reported."""
    from tups_amd import BSE_CLOSED_IR, CHI, F, GAMMA, TupsError

    from tups_amd.runtime import K_BSE_IR

    _say(f"refined closed parity nx={nx} nf={nf} nl={nl} nq={nq} variant={variant}")
    rng = np.random.default_rng(nx * 31 + nf + nl)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=0.5) as c:
        c.set_variant(variant, K_BSE_IR)
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        for p in range(2):
            c.upload(CHI, p, chic[p])

        def load():
            for r in range(4):
                c.upload(GAMMA, r, gam[r])
                c.upload(F, r, gam[r])

        load()
        c.BSE(0, mode=BSE_CLOSED_IR)
        steps, res = c.solve_info()
        got = [c.download(GAMMA, r) for r in range(4)]
        _say(f"  refinement steps {steps}, residual {res:.2e}")
        assert 1 <= steps <= 10 and res <= 1e-12
        load()
        c.set_refinement(1e-12, 0)
        with pytest.raises(TupsError, match="did not reach"):
            c.BSE(0, mode=BSE_CLOSED_IR)
        steps0, res0 = c.solve_info()
        assert steps0 == 0 and 1e-12 < res0 < 1e-4
    exp = B.bse_closed(lat, gam, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0)
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: rel err {err:.2e}")
        assert err < 1e-11


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_refined_closed_form_reports_singular_system(variant):
    from tups_amd import BSE_CLOSED_IR, CHI, GAMMA, TupsError
    from tups_amd.runtime import K_BSE_IR

    nx, nf, nl = 4, 8, 1
    with _ctx(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5) as c:
        c.set_variant(variant, K_BSE_IR)
        D, nq = c.nz, c.nq
        eye = np.asfortranarray(np.repeat(np.eye(D, dtype=complex)[:, :, None], nq, axis=2))
        for r in range(4):
            c.upload(GAMMA, r, eye)
        kappa_d = 2 * 0.5 * c.alpha
        c.upload(CHI, 0, np.asfortranarray(np.full((1, 1, nf, nq), 1.0 / kappa_d, complex)))
        c.upload(CHI, 1, np.asfortranarray(np.zeros((1, 1, nf, nq), complex)))
        with pytest.raises(TupsError):
            c.BSE(0, mode=BSE_CLOSED_IR)
