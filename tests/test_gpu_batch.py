"""GPU parity of the batching form (tups_batch_*): q points with their own Nf_q, Nl_q.

Each q of the batch is one BSE(ite) update (SRC/parquet_BSE.f90:62-132) at its
own (Nf_q, Nl_q); the oracle runs the loop-faithful four-ZGEMM update per q at
that size.  A uniform batch must equal the uniform context path bit for bit
when both run the same tile code, and two chained calls must equal two oracle
iterations (the batch keeps Phi as the next Gamma, as G_r_LL is overwritten).
"""
import sys

import numpy as np
import pytest

from oracle import bse as B
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12
BETA, FD = 2.0, 0.5


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _inputs(rng, nf, nl):
    gam, fv, chi = [[] for _ in range(4)], [[] for _ in range(4)], [[], []]
    for f, l in zip(nf, nl):
        d = f * l
        for r in range(4):
            g = crandn(rng, (d, d), 0.3 / np.sqrt(d)) + 2.0 * np.eye(d)
            gam[r].append(np.asfortranarray(g))
            fv[r].append(np.asfortranarray(g + crandn(rng, (d, d), 0.05)))
        for p in range(2):
            chi[p].append(crandn(rng, (l, l, f), 1.0))
    return gam, fv, chi


def _oracle(nx, nf, nl, gam, fv, chi):
    out = [[] for _ in range(4)]
    lats = {}
    for q, (f, l) in enumerate(zip(nf, nl)):
        lat = lats.setdefault((f, l), Lattice(nx, f, l))
        dense = [B.chi_compact_to_dense(lat, chi[p][q][:, :, :, None]) for p in range(2)]
        phi = B.bse_ladder(lat, [gam[r][q][:, :, None] for r in range(4)],
                           [fv[r][q][:, :, None] for r in range(4)], dense, BETA, FD)
        for r in range(4):
            out[r].append(phi[r][:, :, 0])
    return out


def _run(nx, nf, nl, gam, fv, chi, calls=1, variant=None, autotune=False):
    from tups_amd import CHI, F, GAMMA, TupsBatch

    with TupsBatch(nx, nf, nl, beta=BETA, f_damping=FD) as b:
        if variant is not None:
            b.set_variant(variant)
        for r in range(4):
            b.upload(GAMMA, r, gam[r])
            b.upload(F, r, fv[r])
        for p in range(2):
            b.upload(CHI, p, chi[p])
        if autotune:
            name, ms = b.autotune()
            _say(f"  autotune picked {name} ({ms:.3f} ms)")
            assert name and ms > 0
        for _ in range(calls):
            b.BSE()
        b.sync()
        return [b.download(GAMMA, r) for r in range(4)]


@pytest.mark.parametrize(
    "nx,nf,nl",
    [(4, [16, 8, 6, 12, 64, 2, 10, 4], [1, 5, 9, 5, 1, 5, 9, 1]),   # D 16, 40, 54, 60, 64, 10, 90, 4
     (8, [64, 40, 16, 64, 30], [9, 5, 1, 5, 9]),                     # D 576, 200, 16, 320, 270
     (4, [2], [1])],                                                 # a single 2 x 2 block
)
def test_batch_parity(nx, nf, nl):
    _say(f"batch parity nx={nx} nf={nf} nl={nl}")
    rng = np.random.default_rng(sum(nf) * 7 + sum(nl))
    gam, fv, chi = _inputs(rng, nf, nl)
    got = _run(nx, nf, nl, gam, fv, chi)
    exp = _oracle(nx, nf, nl, gam, fv, chi)
    for r in range(4):
        for q in range(len(nf)):
            err = rel_fro(got[r][q], exp[r][q])
            assert err < TOL, (r, q, nf[q], nl[q], err)
    _say("  max rel err", max(rel_fro(got[r][q], exp[r][q]) for r in range(4) for q in range(len(nf))))


def test_every_batch_variant_matches_oracle():
    """Each tile choice and item order of the batch kernel against the oracle; the
    two item orders of one tile choice agree bit for bit (same per-item code)."""
    from tups_amd import TupsBatch

    nx, nf, nl = 4, [8, 12, 6, 16, 4], [9, 5, 9, 1, 5]
    rng = np.random.default_rng(23)
    gam, fv, chi = _inputs(rng, nf, nl)
    exp = _oracle(nx, nf, nl, gam, fv, chi)
    with TupsBatch(nx, [2], [1]) as probe:
        variants = probe.variants()
    assert len(variants) >= 2
    first = None
    for v, name in variants:
        _say(f"  variant {v} {name}")
        got = _run(nx, nf, nl, gam, fv, chi, variant=v)
        for r in range(4):
            for q in range(len(nf)):
                assert rel_fro(got[r][q], exp[r][q]) < TOL, (name, r, q)
        if v % 3 == 0:
            first = got
        else:  # another item order of the same tile code
            for r in range(4):
                for q in range(len(nf)):
                    assert np.array_equal(got[r][q], first[r][q]), (name, r, q)


def test_autotune_leaves_the_batch_unchanged():
    nx, nf, nl = 4, [8, 10, 4], [9, 5, 1]
    rng = np.random.default_rng(5)
    gam, fv, chi = _inputs(rng, nf, nl)
    got = _run(nx, nf, nl, gam, fv, chi, autotune=True)
    exp = _oracle(nx, nf, nl, gam, fv, chi)
    for r in range(4):
        for q in range(len(nf)):
            assert rel_fro(got[r][q], exp[r][q]) < TOL, (r, q)


def test_batch_two_calls_chain():
    nx, nf, nl = 4, [8, 12, 6], [5, 1, 9]
    _say(f"batch chain nx={nx} nf={nf} nl={nl}")
    rng = np.random.default_rng(11)
    gam, fv, chi = _inputs(rng, nf, nl)
    got = _run(nx, nf, nl, gam, fv, chi, calls=2)
    phi1 = _oracle(nx, nf, nl, gam, fv, chi)
    exp = _oracle(nx, nf, nl, phi1, fv, chi)
    for r in range(4):
        for q in range(len(nf)):
            assert rel_fro(got[r][q], exp[r][q]) < TOL, (r, q)


@pytest.mark.parametrize("nl,variant", [(5, "nl5_g4"), (1, "nl1_g16"), (9, "nl9_g4")])
def test_uniform_batch_equals_context_path(nl, variant):
    """Same tile code, same inputs: the batch and the uniform launch agree exactly."""
    from tups_amd import CHI, F, GAMMA, TupsContext

    nx, nf, nq = 4, 8, 3
    _say(f"batch vs context nl={nl} variant={variant}")
    rng = np.random.default_rng(nl)
    gam, fv, chi = _inputs(rng, [nf] * nq, [nl] * nq)
    got = _run(nx, [nf] * nq, [nl] * nq, gam, fv, chi)
    with TupsContext(nx=nx, nf=nf, nl=nl, nq=nq, beta=BETA, f_damping=FD) as c:
        ids = {name: v for v, name in c.variants()}
        c.set_variant(ids[variant])
        for r in range(4):
            c.upload(GAMMA, r, np.stack(gam[r], axis=2))
            c.upload(F, r, np.stack(fv[r], axis=2))
        for p in range(2):
            c.upload(CHI, p, np.stack(chi[p], axis=3))
        c.BSE(0)
        ref = [c.download(GAMMA, r) for r in range(4)]
    for r in range(4):
        for q in range(nq):
            assert np.array_equal(got[r][q], ref[r][:, :, q]), (r, q)


def test_batch_rejects_bad_shapes():
    from tups_amd import TupsBatch, TupsError

    with pytest.raises(TupsError):
        TupsBatch(4, [7], [5])      # odd Nf_q
    with pytest.raises(TupsError):
        TupsBatch(4, [8], [3])      # Nl_q outside {1, 5, 9}
    with TupsBatch(4, [8, 4], [5, 1]) as b:
        with pytest.raises(ValueError):
            b.upload_packed(0, 0, np.zeros(3, dtype=np.complex128))


@pytest.mark.parametrize(
    "nx,nf,nl",
    [(4, [16, 8, 6, 12, 64, 2, 10, 4, 8, 16], [1, 5, 9, 5, 1, 5, 9, 1, 5, 1]),  # classes with one and two q
     (8, [64, 40, 16, 64, 30, 64], [9, 5, 1, 5, 9, 5])],                         # D 576 (192 tile) .. 16
)
def test_batch_split_parity(nx, nf, nl):
    """tups_batch_bse_split: every q of every (Nf_q, Nl_q) class within the split
    bar (relative Frobenius 1e-10 at 12 moduli) of the oracle's FP64 update."""
    from tups_amd import CHI, F, GAMMA, TupsBatch

    _say(f"batch split nx={nx} nf={nf} nl={nl}")
    rng = np.random.default_rng(sum(nf) + 7)
    gam, fv, chi = _inputs(rng, nf, nl)
    exp = _oracle(nx, nf, nl, gam, fv, chi)
    with TupsBatch(nx, nf, nl, beta=BETA, f_damping=FD) as b:
        for r in range(4):
            b.upload(GAMMA, r, gam[r])
            b.upload(F, r, fv[r])
        for p in range(2):
            b.upload(CHI, p, chi[p])
        b.BSE_split(12)
        b.sync()
        got = [b.download(GAMMA, r) for r in range(4)]
    for r in range(4):
        for q in range(len(nf)):
            assert rel_fro(got[r][q], exp[r][q]) < 1e-10, (r, q, nf[q], nl[q])
