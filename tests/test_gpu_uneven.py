"""GPU parity of the uneven-box iteration (tups_uneven_*): IBZ q = (m, w) with the
box Nf_w per bosonic frequency and Nl_m form factors per momentum.

The BSE of each q is the reference's update (SRC/parquet_BSE.f90:62-132) on its
own box; the crossing follows the library's definition for uneven boxes (the
reference's calc_PhiR + calculate_PAE on the largest box of the embedded
vertices, restricted to each box: oracle.parquet.crossing_uneven, parity
unpinned as no reference output exists; on uniform boxes it reduces exactly to
calculate_pae); then G = F - G (SRC/loop.f90:219-222).  The graph-replayed
iteration must equal the eager one bit for bit over several iterations.
Bar: relative Frobenius error <= 1e-12 per channel.
"""
import sys

import numpy as np
import pytest

from oracle import bse as B
from oracle import parquet as P
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12
BETA, FD = 2.0, 0.5
NX, NF_MAX = 4, 8
NF_W = [8, 6, 4, 4]          # per bosonic frequency, largest box first
NL_M = [1, 5, 9, 5, 1, 9]    # per IBZ momentum (NIBZ = 6 at Nx = 4)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _boxes():
    nw = NF_MAX // 2
    nq = len(NL_M) * nw
    return [NF_W[q % nw] for q in range(nq)], [NL_M[q // nw] for q in range(nq)]


def _inputs(seed):
    rng = np.random.default_rng(seed)
    nf_q, nl_q = _boxes()
    gam, fv, chi = [[] for _ in range(4)], [[] for _ in range(4)], [[], []]
    for f, l in zip(nf_q, nl_q):
        d = f * l
        for r in range(4):
            g = crandn(rng, (d, d), 0.3 / np.sqrt(d)) + 2.0 * np.eye(d)
            gam[r].append(np.asfortranarray(g))
            fv[r].append(np.asfortranarray(g + crandn(rng, (d, d), 0.05)))
        for p in range(2):
            chi[p].append(crandn(rng, (l, l, f), 1.0))
    lam = [crandn(rng, (NF_MAX, NF_MAX, NF_MAX // 2), 0.2) for _ in range(4)]
    return gam, fv, chi, lam


def _oracle_iteration(gam, fv, chi, lam, lat_max, mt):
    nf_q, nl_q = _boxes()
    phi = [[] for _ in range(4)]
    lats = {}
    for q, (f, l) in enumerate(zip(nf_q, nl_q)):
        lat = lats.setdefault((f, l), Lattice(NX, f, l))
        dense = [B.chi_compact_to_dense(lat, chi[p][q][:, :, :, None]) for p in range(2)]
        out = B.bse_ladder(lat, [gam[r][q][:, :, None] for r in range(4)],
                           [fv[r][q][:, :, None] for r in range(4)], dense, BETA, FD)
        for r in range(4):
            phi[r].append(out[r][:, :, 0])
    f_new = P.crossing_uneven(lat_max, phi, lam, nf_q, nl_q, mt)
    g_new = [[f_new[r][q] - phi[r][q] for q in range(len(nf_q))] for r in range(4)]
    return g_new, f_new


def _upload(u, gam, fv, chi, lam):
    from tups_amd import CHI, F, GAMMA, LAMBDA

    for r in range(4):
        u.batch.upload(GAMMA, r, gam[r])
        u.batch.upload(F, r, fv[r])
        u.ctx.upload(LAMBDA, r, lam[r])
    for p in range(2):
        u.batch.upload(CHI, p, chi[p])


def _cat(blocks):
    return np.concatenate([np.ravel(b, order="F") for b in blocks])


def test_uneven_iteration_parity():
    from tups_amd import F, GAMMA, TupsUneven

    _say(f"uneven iteration nx={NX} nf_w={NF_W} nl_m={NL_M}")
    gam, fv, chi, lam = _inputs(11)
    lat_max = Lattice(NX, NF_MAX, max(NL_M))
    mt = P.calc_m_tensors(lat_max)
    with TupsUneven(NX, NF_MAX, NF_W, NL_M, beta=BETA, f_damping=FD) as u:
        _upload(u, gam, fv, chi, lam)
        u.iteration(graph=False)
        u.sync()
        got_g = [u.batch.download(GAMMA, r) for r in range(4)]
        got_f = [u.batch.download(F, r) for r in range(4)]
    exp_g, exp_f = _oracle_iteration(gam, fv, chi, lam, lat_max, mt)
    for r in range(4):
        ef, eg = rel_fro(_cat(got_f[r]), _cat(exp_f[r])), rel_fro(_cat(got_g[r]), _cat(exp_g[r]))
        _say(f"  channel {r}: F {ef:.2e}, G = F - Phi {eg:.2e}")
        assert ef < TOL and eg < TOL


def test_uneven_graph_replays_equal_eager():
    from tups_amd import F, GAMMA, TupsUneven

    gam, fv, chi, lam = _inputs(12)
    res = {}
    for graph in (False, True):
        _say(f"uneven iteration x3, graph={graph}")
        with TupsUneven(NX, NF_MAX, NF_W, NL_M, beta=BETA, f_damping=FD) as u:
            _upload(u, gam, fv, chi, lam)
            for _ in range(3):
                u.iteration(graph=graph)
                u.sync()
            res[graph] = ([_cat(u.batch.download(GAMMA, r)) for r in range(4)],
                          [_cat(u.batch.download(F, r)) for r in range(4)])
    for a, b in zip(res[False][0] + res[False][1], res[True][0] + res[True][1]):
        assert np.array_equal(a, b)
    assert all(np.isfinite(x).all() for x in res[True][0])


def test_uneven_rejects_bad_boxes():
    from tups_amd import TupsError, TupsUneven

    with pytest.raises(TupsError):
        TupsUneven(NX, NF_MAX, [6, 6, 4, 4], NL_M)   # no frequency with the largest box
    with pytest.raises(TupsError):
        TupsUneven(NX, NF_MAX, NF_W, [1, 5, 9, 5, 1, 7])
