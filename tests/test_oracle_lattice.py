"""Pins of oracle/lattice.py: index maps, C4v operations, form factors, exp tables.

Every check compares the oracle with something it does not compute itself: a
closed form, a group/counting identity, or a direct evaluation.
"""
import math

import numpy as np
import pytest

from oracle import lattice as L

SYM_MATS = {
    1: ((1, 0), (0, 1)), 2: ((1, 0), (0, -1)), 3: ((-1, 0), (0, 1)), 4: ((-1, 0), (0, -1)),
    5: ((0, 1), (1, 0)), 6: ((0, 1), (-1, 0)), 7: ((0, -1), (-1, 0)), 8: ((0, -1), (1, 0)),
}


@pytest.mark.parametrize("nx,nf", [(2, 4), (4, 8), (6, 4), (8, 16), (16, 8)])
def test_ibz_map_closed_form(nx, nf):
    lat = L.Lattice(nx, nf, 1, nr=1)
    nxi = (nx + 2) // 2
    seen = set()
    for q in range(lat.nq):
        kx, ky, w = int(lat.qx[q]), int(lat.qy[q]), int(lat.qw[q])
        assert 1 <= ky <= kx <= nxi and 1 <= w <= nf // 2
        assert (((kx - 1) * kx) // 2 + ky - 1) * (nf // 2) + w == q + 1
        seen.add((kx, ky, w))
    assert len(seen) == lat.nq == nxi * (nxi + 1) // 2 * nf // 2


@pytest.mark.parametrize("nx", [2, 4, 6, 8, 10, 16])
def test_stars_cover_bz_exactly_once(nx):
    nxi = (nx + 2) // 2
    count = np.zeros((nx, nx), dtype=int)
    for kx in range(1, nxi + 1):
        for ky in range(1, kx + 1):
            sl = L.list_symmetries(kx, ky, nxi)
            for s in range(1, 9):
                if sl[s - 1]:
                    a, b = L.sym_index(s, kx, ky, nx, nx)
                    count[a - 1, b - 1] += 1
    assert (count == 1).all()


def test_group_closure_and_inverses():
    mats = {s: np.array(m) for s, m in SYM_MATS.items()}
    for s, m in mats.items():
        x, y = L.sym_vec(s, 3, 7)
        assert (x, y) == tuple(m @ np.array([3, 7]))
        assert (mats[L.SYM_INV[s]] @ m == np.eye(2, dtype=int)).all()
        for t, n in mats.items():
            assert any((m @ n == k).all() for k in mats.values())


@pytest.mark.parametrize("nx", [4, 6, 8, 16])
def test_form_factor_closed_forms(nx):
    """The reference's FF_sdwave_check (SRC/parquet_check.f90:277-281) extended to all nine."""
    lat = L.Lattice(nx, 4, 9)
    k = 2 * math.pi * np.arange(nx) / nx
    cx, cy = np.cos(k)[:, None], np.cos(k)[None, :]
    sx, sy = np.sin(k)[:, None], np.sin(k)[None, :]
    one = np.ones((nx, nx))
    expect = [
        one / nx,
        (cx + cy) / nx,
        (cx - cy) / nx,
        math.sqrt(2) * sx * one / nx,
        math.sqrt(2) * sy * one / nx,
        2 * cx * cy / nx,
        -2 * sx * sy / nx,
        math.sqrt(2) * np.sin(k[:, None] + k[None, :]) / nx,
        math.sqrt(2) * np.sin(k[None, :] - k[:, None]) / nx,
    ]
    for l in range(9):
        assert np.abs(lat.ff[:, :, l] - expect[l]).max() < 1e-13, l
    gram = np.einsum("xyl,xym->lm", lat.ff, lat.ff)
    assert np.abs(gram - np.eye(9)).max() < 1e-13
    assert list(lat.signs) == [1, 1, 1, -1, -1, 1, 1, -1, -1]


@pytest.mark.parametrize("nx,nl", [(4, 5), (4, 9), (6, 9), (8, 9)])
def test_symmetry_law_ties_ff_reps_and_sym_op_L(nx, nl):
    """FF(S k, l) = fac(S, l) FF(k, l_sym(S, l)) for every operation and form factor."""
    lat = L.Lattice(nx, 4, nl)
    for s in range(1, 9):
        fwd = np.zeros_like(lat.ff)
        for ix in range(1, nx + 1):
            for iy in range(1, nx + 1):
                a, b = L.sym_index(s, ix, iy, nx, nx)
                fwd[ix - 1, iy - 1] = lat.ff[a - 1, b - 1]
        # FF_inv of the inverse operation is FF(S k)
        assert np.array_equal(lat.ff_inv[L.SYM_INV[s] - 1], fwd)
        for l in range(nl):
            ls, f = lat.lsym[s - 1, l], lat.lfac[s - 1, l]
            assert np.abs(fwd[:, :, l] - f * lat.ff[:, :, ls]).max() < 1e-13, (s, l)


@pytest.mark.parametrize("nx", [4, 8])
def test_exp_tables_against_full_zone_sums(nx):
    lat = L.Lattice(nx, 4, 9, nr=None)
    lat_all = L.Lattice(nx, 4, 1, nr=lat.nibz)
    # sum over stars of e^{iQ.R} == sum over the BZ == Nc delta_{R,0}
    s = lat_all.expSqR.sum(axis=(0, 2))
    expect = np.zeros(lat_all.nr, dtype=complex)
    expect[0] = nx * nx
    assert np.abs(s - expect).max() < 1e-10
    # sum over all shells R and their stars of e^{-ik.R} == Nc delta_{k,0}
    t = lat_all.expqSR.sum(axis=(1, 2))
    expect = np.zeros(nx * nx, dtype=complex)
    expect[0] = nx * nx
    assert np.abs(t - expect).max() < 1e-10
    # individual entries against direct evaluation
    for m, (qx, qy) in enumerate(lat.ibz_points):
        sl = L.list_symmetries(qx, qy, lat.nx_ibz)
        for sym in range(1, 9):
            for r in range(lat.nr):
                rx, ry = lat.ibz_points[r]
                if not sl[sym - 1]:
                    assert lat.expSqR[m, r, sym - 1] == 0
                    continue
                v = np.array(SYM_MATS[sym]) @ np.array([qx - 1, qy - 1])
                ph = 2 * math.pi / nx * (v[0] * (rx - 1) + v[1] * (ry - 1))
                assert abs(lat.expSqR[m, r, sym - 1] - np.exp(1j * ph)) < 1e-12


def test_nl_from_neighbours():
    for nx in (4, 6, 8, 16):
        assert [L.count_nl(nx, n) for n in (1, 2, 3)] == [1, 5, 9]
    assert [L.nr_for_nl(n) for n in (1, 5, 9)] == [1, 4, 6]


def test_frequency_helpers_have_their_physical_meaning():
    beta, nf = 2.0, 16
    nu_ = lambda v: L.matsubara(v, nf, beta)
    om = lambda w: 2j * math.pi * (w - 1) / beta
    for v in range(1, nf + 1):
        assert abs(nu_(L.mF(v, nf)) + nu_(v)) < 1e-12
        for w in range(1, nf // 2 + 1):
            assert abs(nu_(L.FpB(v, w)) - (nu_(v) + om(w))) < 1e-12
            assert abs(nu_(L.FmB(v, w)) - (nu_(v) - om(w))) < 1e-12
            for w2 in range(1, nf // 2 + 1):
                assert abs(nu_(L.BmBmF(w, w2, v, nf)) - (om(w) - nu_(v) - om(w2))) < 1e-12
                assert abs(nu_(L.BpBmF(w, w2, v, nf)) - (om(w) + om(w2) - nu_(v))) < 1e-12
    for x in range(-2 * nf + 1, 3 * nf):
        assert L.pick_nu(x, nf) == min(max(x, 1), nf)
