"""Pins of oracle/parquet.py: q->omega transpose, PhiR, M tensors, crossing."""
import itertools

import numpy as np
import pytest

from oracle import parquet as P
from oracle.lattice import NS, SYM_INV, Lattice, list_symmetries, sym_index
from tests.helpers import crandn, rel_fro


# --- q -> omega distribution ---------------------------------------------------------

@pytest.mark.parametrize("nx,nf", [(4, 8), (8, 16)])
def test_q_to_omega_transpose(nx, nf):
    lat = Lattice(nx, nf, 1)
    # one rank: identity
    assert all(P.index_r_from_index_q(lat, q, 1) == q for q in range(1, lat.nq + 1))
    # several ranks: each rank's slots are filled exactly once by the q it owns
    for ntasks in (2, 4):
        wpt = P.w_per_task(nf, ntasks)
        slots = {}
        for q in range(1, lat.nq + 1):
            owner = P.get_w_task(int(lat.qw[q - 1]), nf, ntasks)
            key = (owner, P.index_r_from_index_q(lat, q, ntasks))
            assert key not in slots
            slots[key] = q
            assert 1 <= key[1] <= lat.nibz * wpt


# --- PhiR ----------------------------------------------------------------------------

def _full_zone_phir(lat, phi):
    """(1/Nx) sum over the full BZ with Phi(S Q) built from the IBZ value."""
    nz2 = lat.nz * lat.nz
    out = np.zeros((nz2, lat.nw, lat.nr), dtype=complex)
    for m, (qx, qy) in enumerate(lat.ibz_points):
        sl = list_symmetries(qx, qy, lat.nx_ibz)
        for s in range(1, NS + 1):
            if not sl[s - 1]:
                continue
            sx, sy = sym_index(s, qx, qy, lat.nx, lat.ny)
            for w in range(lat.nw):
                v = P.sym_permute(lat, phi[:, :, m * lat.nw + w], s).reshape(-1, order="F")
                for r in range(lat.nr):
                    rx, ry = lat.ibz_points[r]
                    ph = 2 * np.pi / lat.nx * ((sx - 1) * (rx - 1) + (sy - 1) * (ry - 1))
                    out[:, w, r] += v * np.exp(1j * ph) / lat.nx
    return out


@pytest.mark.parametrize("nx,nf,nl", [(4, 4, 1), (4, 4, 5), (6, 4, 9)])
def test_phir_full_zone_fourier(nx, nf, nl):
    lat = Lattice(nx, nf, nl)
    rng = np.random.default_rng(11)
    phi = crandn(rng, (lat.nz, lat.nz, lat.nq))
    assert rel_fro(P.calc_phir(lat, phi), _full_zone_phir(lat, phi)) < 1e-13


def test_phir_nl1_closed_form():
    """Nl = NR = 1: PhiR = (1/Nx) sum_{Q in BZ} Phi(Q), Phi constant on stars."""
    lat = Lattice(4, 8, 1)
    rng = np.random.default_rng(2)
    phi = crandn(rng, (lat.nz, lat.nz, lat.nq))
    out = P.calc_phir(lat, phi).reshape(lat.nz, lat.nz, lat.nw, order="F")
    for w in range(lat.nw):
        acc = 0
        for m, (qx, qy) in enumerate(lat.ibz_points):
            acc = acc + sum(list_symmetries(qx, qy, lat.nx_ibz)) * phi[:, :, m * lat.nw + w]
        assert rel_fro(out[:, :, w], acc / lat.nx) < 1e-14


# --- M tensors -----------------------------------------------------------------------

@pytest.mark.parametrize("nx", [4, 8])
def test_m_nl1_is_one(nx):
    lat = Lattice(nx, 4, 1)
    mt = P.calc_m_tensors(lat)
    for name, m in mt.items():
        assert np.abs(m - 1.0).max() < 1e-13, name


def _m_bruteforce(lat):
    """Unfactorised double k-sums of the four tensors, straight from the definitions."""
    nx, nl, nr = lat.nx, lat.nl, lat.nr
    ks = [(kx, ky) for ky in range(nx) for kx in range(nx)]
    FF = lambda k, l: lat.ff[k[0] % nx, k[1] % nx, l]
    # FF_inv(S,k) = FF(S^-1 k)
    FFi = lambda s, k, l: lat.ff[tuple(np.array(sym_index(SYM_INV[s], k[0] % nx + 1, k[1] % nx + 1, nx, nx)) - 1) + (l,)]

    def E(k, r, s):
        return lat.expqSR[k[1] % nx * nx + k[0] % nx, r, s - 1]

    out = {n: np.zeros((nl,) * 4 + (lat.nibz, nr), complex) for n in P.TARGETS}
    for m, (qx, qy) in enumerate(lat.ibz_points):
        Q = (qx - 1, qy - 1)
        for r in range(nr):
            for s in range(1, NS + 1):
                for l1, l2, l3, l4 in itertools.product(range(nl), repeat=4):
                    a = b = c = d = e = f = g = h = 0
                    for k in ks:
                        kq = (k[0] + Q[0], k[1] + Q[1])
                        qk = (Q[0] - k[0], Q[1] - k[1])
                        a += np.conj(E(k, r, s)) * FF(k, l1) * FFi(s, k, l3) * FFi(s, kq, l4)
                        b += E(k, r, s) * FF(k, l2)
                        c += FF(k, l1) * FFi(s, k, l3) * E(k, r, s)
                        d += FF(k, l2) * FFi(s, k, l4) * E(k, r, s)
                        e += FF(k, l1) * FFi(s, k, l3) * np.conj(E(k, r, s))
                        f += FF(k, l2) * FFi(s, k, l4) * np.conj(E(k, r, s))
                        g += FFi(s, qk, l4) * FF(k, l2) * E(k, r, s)
                    eQ = E(Q, r, s)
                    out["phbar_to_ph"][l1, l2, l3, l4, m, r] += a * b
                    out["pp_to_ph"][l1, l2, l3, l4, m, r] += eQ * c * d
                    out["ph_to_pp"][l1, l2, l3, l4, m, r] += eQ * e * f
                    out["phbar_to_pp"][l1, l2, l3, l4, m, r] += e * g
    return out


def test_m_bruteforce_small():
    lat = Lattice(4, 4, 5, nr=2)
    mt = P.calc_m_tensors(lat)
    bf = _m_bruteforce(lat)
    for name in P.TARGETS:
        assert rel_fro(mt[name], bf[name]) < 1e-12, name


# --- crossing ------------------------------------------------------------------------

def _inverse_w2(routine, tr, nf, nw, qw, nu1, nu2):
    """Gather form: the w2 and source (nu_a, nu_b) that reach target (nu1, nu2), or None."""
    clamp = lambda x: min(max(x, 1), nf)
    if routine in ("phbar_to_ph", "phbar_to_pp"):
        w2 = (nu2 - nu1 + 1) if not tr else (nu1 - nu2 + 1)
        if not (1 <= w2 <= nw) or (tr and w2 < 2):
            return None
        if routine == "phbar_to_ph":
            nb = clamp(nu1 + qw - 1) if not tr else clamp(nf + 1 - (nu1 + qw - 1))
        else:
            nb = clamp(qw - nu1 - w2 + nf + 1) if not tr else clamp(nu1 - qw - w2 + 2)
        return w2, (nu1 if not tr else nf + 1 - nu1), nb
    if routine == "pp_to_ph":
        w2 = nu1 + nu2 + qw - nf - 1 if not tr else nf + 3 - qw - nu1 - nu2
        if not (1 <= w2 <= nw) or (tr and w2 < 2):
            return None
        return (w2, nu1, nu2) if not tr else (w2, nf + 1 - nu1, nu1 + qw + w2 - 2)
    w2 = qw - nu1 - nu2 + nf + 1 if not tr else nu1 + nu2 - qw - nf + 1
    if not (1 <= w2 <= nw) or (tr and w2 < 2):
        return None
    return (w2, nu1, nu2) if not tr else (w2, nf + 1 - nu1, nu1 - qw - w2 + 2)


@pytest.mark.parametrize("nf", [4, 8, 16])
def test_crossing_maps_invert(nf):
    """Forward scatter maps of the eight routines == the gather form used on the GPU;
    each target entry is reached at most once per routine pair."""
    nw = nf // 2
    for routine in P.TARGETS:
        for qw in range(1, nw + 1):
            hits = {}
            for tr in (False, True):
                for w2 in range(1 if not tr else 2, nw + 1):
                    for (nu1, nu2, na, nb) in P.routine_maps(routine, tr, nf, qw, w2):
                        assert 1 <= nu2 <= nf and 1 <= na <= nf and 1 <= nb <= nf
                        assert (nu1, nu2) not in hits
                        hits[(nu1, nu2)] = (w2, na, nb)
                        assert _inverse_w2(routine, tr, nf, nw, qw, nu1, nu2) == (w2, na, nb)
            for nu1 in range(1, nf + 1):
                for nu2 in range(1, nf + 1):
                    inv = [_inverse_w2(routine, tr, nf, nw, qw, nu1, nu2) for tr in (False, True)]
                    inv = [x for x in inv if x is not None]
                    assert len(inv) <= 1
                    assert (len(inv) == 1) == ((nu1, nu2) in hits)


def test_pae_linear_in_phir():
    lat = Lattice(4, 4, 5)
    rng = np.random.default_rng(5)
    mt = P.calc_m_tensors(lat)
    phi = [crandn(rng, (lat.nz, lat.nz, lat.nq)) for _ in range(4)]
    lam = P.lambda_pa(lat, 2.0)
    zero = [np.zeros((lat.nz ** 2, lat.nw, lat.nr), complex, order="F") for _ in range(4)]
    x = [crandn(rng, (lat.nz ** 2, lat.nw, lat.nr)) for _ in range(4)]
    y = [crandn(rng, (lat.nz ** 2, lat.nw, lat.nr)) for _ in range(4)]
    f0 = P.calculate_pae(lat, phi, lam, zero, mt)
    fx = P.calculate_pae(lat, phi, lam, x, mt)
    fy = P.calculate_pae(lat, phi, lam, y, mt)
    fxy = P.calculate_pae(lat, phi, lam, [a + b for a, b in zip(x, y)], mt)
    for r in range(4):
        assert rel_fro(fxy[r] - f0[r], (fx[r] - f0[r]) + (fy[r] - f0[r])) < 1e-12
        # PhiR = 0 leaves Phi + Lambda*Nc on the l = 1 block
        exp = phi[r].copy()
        for q in range(lat.nq):
            exp[:4, :4, q] += lam[r][:, :, lat.qw[q] - 1] * lat.nc
        assert rel_fro(f0[r], exp) < 1e-14


def test_pae_nl1_gather_form():
    """Nl = NR = 1 (M == 1): the oracle's scatter equals a gather written from the inverse maps."""
    lat = Lattice(4, 8, 1)
    rng = np.random.default_rng(9)
    mt = P.calc_m_tensors(lat)
    phi = [crandn(rng, (lat.nz, lat.nz, lat.nq)) for _ in range(4)]
    lam = P.lambda_pa(lat, 1.5)
    phir = [crandn(rng, (lat.nz ** 2, lat.nw, 1)) for _ in range(4)]
    got = P.calculate_pae(lat, phi, lam, phir, mt)
    nf, nw = lat.nf, lat.nw
    X = [p.reshape(nf, nf, nw, order="F") for p in phir]
    tab = {  # routine -> ((src0, src1), coefficients for (target a, target b))
        "phbar_to_ph": ((0, 1), (-0.5, -1.5), (-0.5, 0.5)),
        "pp_to_ph": ((2, 3), (0.5, 1.5), (-0.5, 0.5)),
        "ph_to_pp": ((0, 1), (0.5, -1.5), (0.5, 0.5)),
        "phbar_to_pp": ((0, 1), (0.5, -1.5), (-0.5, -0.5)),
    }
    exp = [p.copy() for p in phi]
    for q in range(lat.nq):
        qw = int(lat.qw[q])
        for r in range(4):
            exp[r][:, :, q] += lam[r][:, :, qw - 1] * lat.nc
        for routine, (srcs, ca, cb) in tab.items():
            ta, tb = P.TARGETS[routine]
            for nu1 in range(1, nf + 1):
                for nu2 in range(1, nf + 1):
                    for tr in (False, True):
                        hit = _inverse_w2(routine, tr, nf, nw, qw, nu1, nu2)
                        if hit is None:
                            continue
                        w2, na, nb = hit
                        x0 = X[srcs[0]][na - 1, nb - 1, w2 - 1]
                        x1 = X[srcs[1]][na - 1, nb - 1, w2 - 1]
                        if tr:
                            x0, x1 = np.conj(x0), np.conj(x1)
                        exp[ta][nu1 - 1, nu2 - 1, q] += (ca[0] * x0 + ca[1] * x1) / lat.nx
                        exp[tb][nu1 - 1, nu2 - 1, q] += (cb[0] * x0 + cb[1] * x1) / lat.nx
    for r in range(4):
        assert rel_fro(got[r], exp[r]) < 1e-13, r


def test_uneven_crossing_reduces_to_calculate_pae_on_uniform_boxes():
    """The uneven-box crossing (embed in the largest box, calculate_pae, extract) is
    calculate_pae itself when every q has the largest box."""
    lat = Lattice(4, 6, 5)
    rng = np.random.default_rng(5)
    phi = [crandn(rng, (lat.nz, lat.nz, lat.nq)) for _ in range(4)]
    lam = [crandn(rng, (lat.nf, lat.nf, lat.nw)) for _ in range(4)]
    nf_q, nl_q = [lat.nf] * lat.nq, [lat.nl] * lat.nq
    blocks = [[phi[r][:, :, q] for q in range(lat.nq)] for r in range(4)]
    for r in range(4):
        assert np.array_equal(P.uneven_embed(lat, blocks[r], nf_q, nl_q), phi[r])
    got = P.crossing_uneven(lat, blocks, lam, nf_q, nl_q)
    exp = P.calculate_pae(lat, phi, lam, [P.calc_phir(lat, p) for p in phi], P.calc_m_tensors(lat))
    for r in range(4):
        assert all(np.array_equal(got[r][q], exp[r][:, :, q]) for q in range(lat.nq))


def test_uneven_embed_clamps_and_truncates():
    """A q with a smaller box: frequencies outside it take the edge value (pickNu),
    form factors beyond its Nl are zero, and extract recovers the block exactly."""
    lat = Lattice(4, 8, 5)
    rng = np.random.default_rng(6)
    b = crandn(rng, (1 * 4, 1 * 4))           # Nl_q = 1, Nf_q = 4
    full = P.uneven_embed(lat, [b], [4], [1])[:, :, 0]
    # box-centred n = -4..3 on the largest box -> clamp to -2..1 on q's box
    c = np.clip(np.arange(8) - 4, -2, 1) + 2
    assert np.array_equal(full[:8, :8], b[np.ix_(c, c)])
    assert not full[8:, :].any() and not full[:, 8:].any()
    assert np.array_equal(P.uneven_extract(lat, full[:, :, None], [4], [1])[0], b)
