"""GPU parity of the split (INT8 MFMA) ladder against the FP64 oracle.

The split mode is the precision form of BSE(ite) (SRC/parquet_BSE.f90:62-132):
the same damped update with its complex GEMM evaluated exactly on integers
(operands rounded to ta / tb bits per row / column) through residues modulo
Gaussian moduli.  Its bar is a relative Frobenius error <= 1e-10 per channel
at the default 12 moduli; each added modulus must lower the error, down to
FP64 rounding at 16, and per-row / per-column scaling must keep rows of very
different magnitude accurate on their own.
"""
import math
import sys

import numpy as np
import pytest

from oracle import bse as B
from oracle.lattice import Lattice
from tests.helpers import crandn, random_chi_compact, random_vertices, rel_fro

pytestmark = pytest.mark.gpu
SPLIT_TOL = 1e-10


def _ctx(**kw):
    from tups_amd import TupsContext

    return TupsContext(**kw)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _run(c, gam, fv, chic, moduli, qbatch=0):
    from tups_amd import BSE_LADDER_SPLIT, CHI, F, GAMMA

    for r in range(4):
        c.upload(GAMMA, r, gam[r])
        c.upload(F, r, fv[r])
    for p in range(2):
        c.upload(CHI, p, chic[p])
    c.set_split(moduli, qbatch)
    c.BSE(0, mode=BSE_LADDER_SPLIT)
    return [c.download(GAMMA, r) for r in range(4)]


@pytest.mark.parametrize(
    "nx,nf,nl,nq,f,qbatch",
    [(4, 16, 1, 0, 0.5, 0),     # Nz=16: one partial tile
     (4, 12, 5, 0, 0.5, 0),     # Nz=60: inner dimension 120 padded to 128
     (4, 8, 9, 0, 0.3, 7),      # Nz=72: two tiles, several passes
     (8, 64, 5, 3, 0.5, 0),     # Nz=320: the batched-BSE shape
     (16, 64, 9, 2, 1.0, 1),    # Nz=576, one q per pass
     (8, 64, 1, 5, 0.5, 2),     # Nz=64: exact tile, ragged last pass
     (4, 112, 9, 1, 0.5, 0)],   # Nz=1008: bubble blocks too large for LDS beside the columns
)
def test_split_ladder_parity(nx, nf, nl, nq, f, qbatch):
    _say(f"split parity nx={nx} nf={nf} nl={nl} nq={nq} qbatch={qbatch}")
    rng = np.random.default_rng(nx * 100 + nf + nl)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=f) as c:
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        fv = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        got = _run(c, gam, fv, chic, 12, qbatch)
    exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, f)
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: rel err {err:.2e}")
        assert err < SPLIT_TOL


def test_split_error_falls_with_moduli():
    """Each modulus adds ~7.6 bits to M and so ~3.8 bits to each operand's
    integer width: the error must fall by >= 4x per modulus until FP64
    rounding, which 16 moduli reach."""
    rng = np.random.default_rng(5)
    nx, nf, nl = 4, 16, 5
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=4, beta=2.0, f_damping=1.0) as c:
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        fv = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, 1.0)
        errs = []
        for n in range(8, 17):
            got = _run(c, gam, fv, chic, n)
            errs.append(max(rel_fro(got[r], exp[r]) for r in range(4)))
            _say(f"  moduli {n}: max rel err {errs[-1]:.2e}")
    assert errs[0] < 1e-5
    for a, b in zip(errs, errs[1:]):
        assert b < a / 4 or b < 1e-14
    assert errs[-1] < 1e-14


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 5), (4, 64, 1)])
def test_split_rows_and_columns_of_very_different_scale(nx, nf, nl):
    """Gamma and F with rows scaled by 1e-9 and columns by 1e6: every element of
    Phi, measured in units of its own row and column scale, stays accurate.
    A single scale per matrix would leave the small rows and columns with
    errors of order 2^-35 * 1e15."""
    rng = np.random.default_rng(9)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=1.0) as c:
        D, nq = c.nz, c.nq
        rs = np.where(np.arange(D) % 3 == 0, 1e-9, 1.0)
        cs = np.where(np.arange(D) % 5 == 0, 1e6, 1.0)
        w = (rs[:, None] * cs[None, :])[:, :, None]
        gam = [np.asfortranarray(crandn(rng, (D, D, nq)) * w) for _ in range(4)]
        fv = [np.asfortranarray(crandn(rng, (D, D, nq)) * w) for _ in range(4)]
        gam[1][:, :, 0] = 0.0  # a zero matrix in one operand
        chic = [random_chi_compact(rng, nl, nf, nq) for _ in range(2)]
        got = _run(c, gam, fv, chic, 12)
    exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, 1.0)
    for r in range(4):
        for q in range(nq):
            e = np.abs(exp[r][:, :, q]) / w[:, :, 0]
            d = np.abs(got[r][:, :, q] - exp[r][:, :, q]) / w[:, :, 0]
            assert d.max() <= 1e-9 * e.max(), (r, q, d.max(), e.max())
    # Gamma_m(q=0) = 0 with f = 1: Phi is exactly zero (zero exponents, zero residues)
    assert not np.any(got[1][:, :, 0])


def test_split_settings_rejected():
    from tups_amd import TupsError

    with _ctx(nx=4, nf=8, nl=1) as c:
        for n, qb in ((7, 0), (17, 0), (12, -1)):
            with pytest.raises(TupsError):
                c.set_split(n, qb)
        c.set_split(16, 3)


def test_split_matches_fp64_on_batched_workload():
    """The batched-BSE inputs (Gamma = U I + X, sigma^2 = (0.3U)^2/D, scaled bare
    bubble) at Nz=320: split vs the FP64 path of the same library, per channel."""
    from tups_amd import BSE_LADDER_SPLIT, GAMMA
    from tups_amd.synth import batched_bse_inputs

    errs = {}
    with _ctx(nx=8, nf=64, nl=5, nq=8, beta=2.0, f_damping=0.5) as c:
        batched_bse_inputs(c, seed=3)
        c.BSE(0)
        ref = [c.download(GAMMA, r) for r in range(4)]
        for n in (12, 16):
            batched_bse_inputs(c, seed=3)
            c.set_split(n)
            c.BSE(0, mode=BSE_LADDER_SPLIT)
            got = [c.download(GAMMA, r) for r in range(4)]
            errs[n] = [rel_fro(got[r], ref[r]) for r in range(4)]
            _say(f"  moduli {n}: " + " ".join(f"{e:.2e}" for e in errs[n]))
    assert max(errs[12]) < SPLIT_TOL
    assert max(errs[16]) < 1e-13


@pytest.mark.parametrize("nx,nf,nl,nq", [(8, 64, 5, 2), (4, 8, 9, 0), (16, 64, 9, 1)])
def test_every_split_gemm_tile_and_autotune(nx, nf, nl, nq):
    """The 64 x 64 (register- and DMA-staged, 4 and 5 LDS stages), 160 x 160 and
    192 x 192 GEMM tiles (each with its own padding of D) and the autotuner's
    pick all meet the split bar."""
    from tups_amd.runtime import K_SPLIT_GEMM

    rng = np.random.default_rng(nx + nf + nl + 5)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=0.5) as c:
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        fv = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, 0.5)
        variants = c.variants(K_SPLIT_GEMM)
        assert len(variants) == 5
        for v, name in variants + [(None, "autotuned")]:
            _say(f"split tile {name} nx={nx} nf={nf} nl={nl} nq={c.nq}")
            if v is None:
                for r in range(4):
                    from tups_amd import F, GAMMA
                    c.upload(GAMMA, r, gam[r])
                    c.upload(F, r, fv[r])
                name, ms = c.autotune(K_SPLIT_GEMM)
                assert ms > 0
            else:
                c.set_variant(v, K_SPLIT_GEMM)
            got = _run(c, gam, fv, chic, 12)
            for r in range(4):
                assert rel_fro(got[r], exp[r]) < SPLIT_TOL, (name, r)
