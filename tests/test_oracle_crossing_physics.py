"""Pin of the crossing oracle by the parquet relations in momentum/frequency space.

calc_phir + calc_m_tensors + calculate_pae (oracle/parquet.py, restating
SRC/parquet_PhiR.f90:35-258, SRC/parquet_EQMatrix.f90:21-360,
SRC/parquet_EQOuter.f90:16-118 and SRC/parquet_EQContribs.f90:44-524) must equal
the SU(2) parquet crossing evaluated directly on the k-space vertex
(tests/parquet_bruteforce.py) when the real-space sum covers every lattice
vector (NR = number of IBZ points).  One comparison covers the eight
frequency maps with their conj / signs / mF, the pickNu clamps, the factor
table, the Nl>1 M contraction and the PhiR transform together.

The second half plants mistakes in the oracle (on the CPU) and checks that the
comparison catches each of them.
"""
import numpy as np
import pytest

from oracle import parquet as P
from oracle.lattice import FmB, FpB, Lattice, mF, pick_nu, BmBmF
from tests.helpers import rel_fro
from tests.parquet_bruteforce import crossing_kspace, symmetric_phi


def _full_r_lattice(nx, nf, nl):
    return Lattice(nx, nf, nl, nr=Lattice(nx, nf, nl).nibz)


def _oracle_crossing(lat, phi):
    lam0 = [np.zeros((lat.nf, lat.nf, lat.nw), complex) for _ in range(4)]
    f = P.calculate_pae(lat, phi, lam0, [P.calc_phir(lat, p) for p in phi], P.calc_m_tensors(lat))
    return [a - b for a, b in zip(f, phi)]


def _max_err(lat, phi, exp=None):
    exp = crossing_kspace(lat, phi) if exp is None else exp
    got = _oracle_crossing(lat, phi)
    return max(rel_fro(g, e) for g, e in zip(got, exp))


@pytest.mark.parametrize("nx,nf,nl", [(4, 4, 1), (4, 8, 1), (4, 6, 5), (4, 4, 9), (6, 4, 5), (6, 6, 9)])
def test_pae_equals_kspace_crossing(nx, nf, nl):
    lat = _full_r_lattice(nx, nf, nl)
    phi = symmetric_phi(lat, np.random.default_rng(100 + nx + nf + nl))
    assert _max_err(lat, phi) < 1e-13


def test_kspace_crossing_exchange_symmetry():
    """Independent of the oracle: at Q = 0 the singlet crossing is even and the triplet one
    odd under (k', nu') -> (-k', omega - nu') (pp exchange of the outgoing pair), which in the
    l-basis reads F(l1 nu, l2 nu') = +/- signs(l2) F(l1 nu, l2 nu*), nu* = omega - nu'.  Where
    nu* is in the box the ph and ph-bar terms of the two sides map onto each other with no
    clamped argument, so the relation is exact there."""
    lat = _full_r_lattice(4, 8, 5)
    phi = symmetric_phi(lat, np.random.default_rng(7))
    phi[2][:] = 0
    phi[3][:] = 0
    out = crossing_kspace(lat, phi)
    nf, nl, nw = lat.nf, lat.nl, lat.nw
    checked = 0
    for w in range(1, nw + 1):  # momentum slot 0 is Q = 0
        fs = out[2][:, :, w - 1].reshape(nf, nl, nf, nl, order="F")
        ft = out[3][:, :, w - 1].reshape(nf, nl, nf, nl, order="F")
        for n1 in range(nf):
            for n2 in range(nf):
                # nu* = omega - nu': index Nf - nu' + w (1-based), SRC/parquet_util.f90:1332-1339
                n2s = nf - (n2 + 1) + w - 1
                if not (0 <= n2s < nf):
                    continue
                sg = lat.signs.astype(float)
                assert np.allclose(fs[n1, :, n2, :], fs[n1, :, n2s, :] * sg[None, :], atol=1e-12)
                assert np.allclose(ft[n1, :, n2, :], -ft[n1, :, n2s, :] * sg[None, :], atol=1e-12)
                checked += 1
    assert checked > 10
    assert np.abs(out[3]).max() > 1e-3


# --- planted mistakes -------------------------------------------------------------------

def _mut_maps(routine, tr_flag, fn):
    orig = P.routine_maps

    def mutated(r, tr, nf, qw, w2):
        res = orig(r, tr, nf, qw, w2)
        if r == routine and tr == tr_flag:
            return [x for x in (fn(m, nf, qw, w2) for m in res) if x is not None]
        return res
    return mutated


def _in_box(v, nf):
    return 1 <= v <= nf


MAP_MUTATIONS = {
    # target column of ph-bar -> ph written as nu - omega2 instead of nu + omega2
    "phbar_ph_target_FmB": _mut_maps("phbar_to_ph", False, lambda m, nf, qw, w2:
                                     (m[0], FmB(m[0], w2), m[2], m[3]) if _in_box(FmB(m[0], w2), nf) else None),
    # source of pp -> ph transposed
    "pp_ph_source_swapped": _mut_maps("pp_to_ph", False, lambda m, nf, qw, w2: (m[0], m[1], m[3], m[2])),
    # time-reversed pp -> ph without the mirror of the target column
    "pp_ph_tr_no_mirror": _mut_maps("pp_to_ph", True, lambda m, nf, qw, w2: (m[0], mF(m[1], nf), m[2], m[3])),
    # ph -> pp time reversal with the unmirrored first source frequency
    "ph_pp_tr_no_mF": _mut_maps("ph_to_pp", True, lambda m, nf, qw, w2: (m[0], m[1], m[0], m[3])),
    # clamp of ph-bar -> pp with the bosonic arguments exchanged
    "phbar_pp_bmbmf_order": _mut_maps("phbar_to_pp", False, lambda m, nf, qw, w2:
                                      (m[0], m[1], m[2], pick_nu(BmBmF(w2, qw, m[0], nf), nf))),
    # ph-bar -> ph second source frequency nu + omega2 instead of nu + omega
    "phbar_ph_source_w2": _mut_maps("phbar_to_ph", False, lambda m, nf, qw, w2:
                                    (m[0], m[1], m[2], pick_nu(FpB(m[0], w2), nf))),
}


def _table_mutation(channel, idx, which, value):
    t = {k: list(v) for k, v in P.PAE_TABLE.items()}
    r = list(t[channel][idx])
    r[which] = value
    t[channel][idx] = tuple(r)
    return t


TABLE_MUTATIONS = {
    "m_phbar_pp_fac2_sign": _table_mutation(1, 2, 2, +0.5),
    "t_pp_ph_fac1": _table_mutation(3, 0, 1, 0.5),
    "d_ph_pp_fac2_sign": _table_mutation(0, 1, 2, -0.5),
    "s_pp_ph_swapped": _table_mutation(2, 0, 0, "phbar_to_ph"),
}


@pytest.fixture(scope="module")
def planted_case():
    lat = _full_r_lattice(4, 6, 5)
    phi = symmetric_phi(lat, np.random.default_rng(31))
    return lat, phi, crossing_kspace(lat, phi)


def test_planted_case_is_clean(planted_case):
    lat, phi, exp = planted_case
    assert _max_err(lat, phi, exp) < 1e-13


@pytest.mark.parametrize("name", sorted(MAP_MUTATIONS))
def test_planted_map_mistake_is_caught(planted_case, monkeypatch, name):
    lat, phi, exp = planted_case
    monkeypatch.setattr(P, "routine_maps", MAP_MUTATIONS[name])
    assert _max_err(lat, phi, exp) > 1e-3


@pytest.mark.parametrize("name", sorted(TABLE_MUTATIONS))
def test_planted_factor_mistake_is_caught(planted_case, monkeypatch, name):
    lat, phi, exp = planted_case
    monkeypatch.setattr(P, "PAE_TABLE", TABLE_MUTATIONS[name])
    assert _max_err(lat, phi, exp) > 1e-3


def test_planted_conj_and_sign_mistakes_are_caught(planted_case, monkeypatch):
    lat, phi, exp = planted_case
    orig = P.add_contribution

    def no_conj(lat_, F, work, M, routine, tr, w2, f1, f2, qs=None):
        return orig(lat_, F, np.conj(work) if tr else work, M, routine, tr, w2, f1, f2, qs)

    monkeypatch.setattr(P, "add_contribution", no_conj)
    assert _max_err(lat, phi, exp) > 1e-3
    monkeypatch.setattr(P, "add_contribution", orig)
    monkeypatch.setattr(lat, "signs", np.ones_like(lat.signs))
    assert _max_err(lat, phi, exp) > 1e-3


@pytest.mark.parametrize("tensor,op", [("phbar_to_ph", "swap34"), ("pp_to_ph", "swap34"),
                                       ("ph_to_pp", "conj"), ("phbar_to_pp", "swap12")])
def test_planted_m_mistake_is_caught(planted_case, monkeypatch, tensor, op):
    lat, phi, exp = planted_case
    orig = P.calc_m_tensors

    def mutated(lat_):
        mt = orig(lat_)
        m = mt[tensor]
        if op == "swap34":
            mt[tensor] = np.ascontiguousarray(m.transpose(0, 1, 3, 2, 4, 5))
        elif op == "swap12":
            mt[tensor] = np.ascontiguousarray(m.transpose(1, 0, 2, 3, 4, 5))
        else:
            mt[tensor] = np.conj(m)
        return mt

    monkeypatch.setattr(P, "calc_m_tensors", mutated)
    assert _max_err(lat, phi, exp) > 1e-3
