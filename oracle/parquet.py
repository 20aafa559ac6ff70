"""Oracle: Phi(q) -> Phi(R) transform, transformation tensors M, crossing (PAE).

Restates
* calc_PhiR / FTPhis / symop_arr   SRC/parquet_PhiR.f90:35-109, 196-258, 311-356
* getwTask / indexR_from_indexQ    SRC/parquet_PhiR.f90:277-307 (q -> omega transpose)
* calc_M_phbar_to_ph               SRC/parquet_EQMatrix.f90:21-123
* calc_M_pp_to_ph                  SRC/parquet_EQMatrix.f90:126-194
* calc_M_ph_to_pp                  SRC/parquet_EQMatrix.f90:198-271
* calc_M_phbar_to_pp               SRC/parquet_EQMatrix.f90:274-360
* calculate_PAE                    SRC/parquet_EQOuter.f90:16-118
* the eight contribution routines  SRC/parquet_EQContribs.f90:44-524

Pinned by (tests/test_oracle_parquet.py):
* getwTask/indexR_from_indexQ: on one rank the q->omega transpose is the
  identity; for several ranks it is a bijection onto the owner's slots;
* PhiR against a plain Fourier sum over the FULL Brillouin zone, with the
  full-zone Phi built from the IBZ one through the form-factor law
  FF(S k) = fac FF(k, l_sym) (each BZ point counted once);
* M: the Nl=1, NR=1 closed form M == 1; an unfactorised brute-force k,k'
  double sum on Nx=4 for all four tensors; the conjugation relation
  M_ph_to_pp(Q) = conj(M_pp_to_ph(-Q)) that ties two routines together;
* crossing: Nl=1, NR=1 against a brute-force frequency crossing written from
  the index maps of each routine (every target entry reached at most once per
  routine pair), and the linearity of calculate_PAE in PhiR.
* calc_phir + calc_m_tensors + calculate_pae together, for Nl in {1, 5, 9}
  (tests/test_oracle_crossing_physics.py): with NR = every IBZ shell the
  truncated-unity sum is exact, and the result must equal the SU(2) parquet
  crossing relations (d/m/s/t, Bickers 2004; Rohringer et al., PRB 86, 125114
  (2012)) evaluated directly on the momentum/frequency-space vertex built
  from C4v rotations and the conjugation symmetry
  V(k nu, k' nu'; Q, -omega) = conj V(k -nu, k' -nu'; Q, omega)
  (tests/parquet_bruteforce.py, which uses none of the maps, tables or
  tensors above).  Planted mistakes in the maps, the factor table, the
  time-reversal conj/signs and the M tensors are each caught.
"""
from __future__ import annotations

import numpy as np

from .lattice import NS, BmBmF, BpBmF, FmB, FpB, Lattice, mF, pick_nu


# --- q -> omega distribution (collapses to the identity on one rank) -------------

def w_per_task(nf: int, ntasks: int) -> int:
    """wperTask, SRC/parquet_ini.f90:346."""
    return int((nf - 1) // (2 * ntasks) + 1)


def get_w_task(w: int, nf: int, ntasks: int) -> int:
    """getwTask, SRC/parquet_PhiR.f90:277-292."""
    if ntasks <= nf // 2:
        return int((w - 1) // w_per_task(nf, ntasks))
    step = ntasks // (nf // 2)
    return (w - 1) * step


def index_r_from_index_q(lat: Lattice, idx_q: int, ntasks: int) -> int:
    """indexR_from_indexQ for a 1-based global q, SRC/parquet_PhiR.f90:295-307."""
    iw = int(lat.qw[idx_q - 1])
    idx_r = (idx_q - iw) // lat.nw
    wpt = w_per_task(lat.nf, ntasks)
    return idx_r * wpt + (iw - 1) % wpt + 1


# --- PhiR ----------------------------------------------------------------------------

def _nz_view(lat: Lattice, a: np.ndarray) -> np.ndarray:
    """(Nz, Nz, ...) -> (Nf, Nl, Nf, Nl, ...) view: row (l,nu) = l*Nf + nu."""
    return a.reshape((lat.nf, lat.nl, lat.nf, lat.nl) + a.shape[2:], order="F")


def sym_permute(lat: Lattice, phi: np.ndarray, s: int) -> np.ndarray:
    """symop_arr for one operation s (1-based) on (Nz, Nz, ...), SRC/parquet_PhiR.f90:311-356."""
    v = _nz_view(lat, phi)
    ls = lat.lsym[s - 1]
    lf = lat.lfac[s - 1].astype(np.float64)
    out = v[:, ls][:, :, :, ls]
    fac = lf[None, :, None, None] * lf[None, None, None, :]
    out = out * fac.reshape(fac.shape + (1,) * (phi.ndim - 2))
    return out.reshape(phi.shape, order="F")


def calc_phir(lat: Lattice, phi: np.ndarray, ws=None) -> np.ndarray:
    """PhiR_r(Nz*Nz, Nw, NR) from Phi_r(Nz, Nz, Nq); SRC/parquet_PhiR.f90:35-109, 196-258.

    One rank: PhiQ slot == global q (indexR_from_indexQ is the identity).
    The l-block permutation of symop_arr does not depend on Q, so it is applied
    after the sum over Q (to the Nz^2 x NR product) instead of to every PhiQ.
    ws: optional bosonic frequencies (0-based) to form; the others stay zero
    (for timing a sample of a large call).
    """
    nz2 = lat.nz * lat.nz
    out = np.zeros((nz2, lat.nw, lat.nr), dtype=np.complex128, order="F")
    for w1 in (range(lat.nw) if ws is None else ws):
        qs = np.arange(lat.nibz) * lat.nw + w1
        phiq = phi[:, :, qs].reshape((nz2, lat.nibz), order="F")
        acc = np.zeros((lat.nz, lat.nz, lat.nr), dtype=np.complex128, order="F")
        for s in range(1, NS + 1):
            t = (phiq @ lat.expSqR[:, :, s - 1]).reshape((lat.nz, lat.nz, lat.nr), order="F")
            acc += (1.0 / lat.nx) * sym_permute(lat, t, s)
        out[:, w1, :] = acc.reshape((nz2, lat.nr), order="F")
    return out


# --- transformation tensors -------------------------------------------------------

def _ffk(lat: Lattice):
    """FF and FF_inv re-indexed by k = (ky-1)*Nx + (kx-1)."""
    ffk = lat.ff.transpose(1, 0, 2).reshape(lat.nc, lat.nl)
    ffik = lat.ff_inv.transpose(0, 2, 1, 3).reshape(NS, lat.nc, lat.nl)
    return ffk, ffik


def _shift_index(lat: Lattice, qx: int, qy: int, sign: int):
    """k index of k+Q (sign=+1, SRC/parquet_EQMatrix.f90:79) or Q-k (sign=-1, :323)."""
    kx = np.arange(lat.nx)[None, :]
    ky = np.arange(lat.ny)[:, None]
    if sign > 0:
        sx = (kx + qx - 1) % lat.nx
        sy = (ky + qy - 1) % lat.ny
    else:
        sx = (qx - 1 - kx) % lat.nx
        sy = (qy - 1 - ky) % lat.ny
    return (sy * lat.nx + sx).reshape(-1)


def calc_m_tensors(lat: Lattice) -> dict:
    """The four M(l1,l2,l3,l4, Q, R) tensors, indexed by IBZ momentum Q (not by q).

    The reference stores them per local q (Nb); they depend on q only through
    its momentum, so M_q = M[..., momentum(q), R].  Shapes (Nl,Nl,Nl,Nl,NIBZ,NR).
    """
    ffk, ffik = _ffk(lat)
    e = lat.expqSR  # (Nc, NR, NS)
    nl, nr, nm = lat.nl, lat.nr, lat.nibz
    out = {}

    # phbar -> ph, SRC/parquet_EQMatrix.f90:21-123
    m2 = np.einsum("krs,kb->brs", e, ffk)
    m = np.zeros((nl, nl, nl, nl, nm, nr), dtype=np.complex128)
    for iq, (qx, qy) in enumerate(lat.ibz_points):
        kq = _shift_index(lat, qx, qy, +1)
        m1 = np.einsum("krs,ka,skc,skd->acdrs", np.conj(e), ffk, ffik, ffik[:, kq, :])
        m[:, :, :, :, iq, :] = np.einsum("acdrs,brs->abcdr", m1, m2)
    out["phbar_to_ph"] = m

    # pp -> ph, SRC/parquet_EQMatrix.f90:126-194
    m1 = np.einsum("ka,skc,krs->acrs", ffk, ffik, e)
    m = np.zeros((nl, nl, nl, nl, nm, nr), dtype=np.complex128)
    for iq, (qx, qy) in enumerate(lat.ibz_points):
        kq = (qy - 1) * lat.nx + qx - 1
        m[:, :, :, :, iq, :] = np.einsum("rs,acrs,bdrs->abcdr", e[kq], m1, m1)
    out["pp_to_ph"] = m

    # ph -> pp, SRC/parquet_EQMatrix.f90:198-271
    m1 = np.einsum("ka,skc,krs->acrs", ffk, ffik, np.conj(e))
    m = np.zeros((nl, nl, nl, nl, nm, nr), dtype=np.complex128)
    for iq, (qx, qy) in enumerate(lat.ibz_points):
        kq = (qy - 1) * lat.nx + qx - 1
        m[:, :, :, :, iq, :] = np.einsum("rs,acrs,bdrs->abcdr", e[kq], m1, m1)
    out["ph_to_pp"] = m

    # phbar -> pp, SRC/parquet_EQMatrix.f90:274-360
    m1 = np.einsum("ka,skc,krs->acrs", ffk, ffik, np.conj(e))
    m = np.zeros((nl, nl, nl, nl, nm, nr), dtype=np.complex128)
    for iq, (qx, qy) in enumerate(lat.ibz_points):
        qk = _shift_index(lat, qx, qy, -1)
        m2 = np.einsum("skd,kb,krs->bdrs", ffik[:, qk, :], ffk, e)
        m[:, :, :, :, iq, :] = np.einsum("acrs,bdrs->abcdr", m1, m2)
    out["phbar_to_pp"] = m
    return out


# --- crossing -------------------------------------------------------------------------

# (routine, factor for first target, factor for second target) per source channel,
# SRC/parquet_EQOuter.f90:85-112.  Targets: ph routines -> (d, m); pp routines -> (s, t).
PAE_TABLE = {
    0: [("phbar_to_ph", -0.5, -0.5), ("ph_to_pp", 0.5, 0.5), ("phbar_to_pp", 0.5, -0.5)],
    1: [("phbar_to_ph", -1.5, 0.5), ("ph_to_pp", -1.5, 0.5), ("phbar_to_pp", -1.5, -0.5)],
    2: [("pp_to_ph", 0.5, -0.5)],
    3: [("pp_to_ph", 1.5, 0.5)],
}
TARGETS = {"phbar_to_ph": (0, 1), "pp_to_ph": (0, 1), "ph_to_pp": (2, 3), "phbar_to_pp": (2, 3)}


def routine_maps(routine: str, tr: bool, nf: int, qw: int, w2: int):
    """Valid nu1 (1-based) and (nu2, nu_a, nu_b) for one contribution routine.

    SRC/parquet_EQContribs.f90:44-524 (line-by-line index maps and cycles).
    Source entry is PhiRWork((l3-1)*Nf + nu_a, (l4-1)*Nf + nu_b, w2, R2); TR
    routines additionally take conj and signs(l3)*signs(l4).
    """
    res = []
    for nu1 in range(1, nf + 1):
        if routine == "phbar_to_ph" and not tr:
            if FpB(nu1, w2) > nf:
                continue
            res.append((nu1, FpB(nu1, w2), nu1, pick_nu(FpB(nu1, qw), nf)))
        elif routine == "phbar_to_ph" and tr:
            if FmB(nu1, w2) > nf or FmB(nu1, w2) < 1:
                continue
            res.append((nu1, FmB(nu1, w2), mF(nu1, nf), pick_nu(mF(FpB(nu1, qw), nf), nf)))
        elif routine == "pp_to_ph" and not tr:
            v = BmBmF(w2, qw, nu1, nf)
            if v > nf or v < 1:
                continue
            res.append((nu1, v, nu1, v))
        elif routine == "pp_to_ph" and tr:
            v = FpB(nu1, qw + w2 - 1)
            if v > nf or v < 1:
                continue
            res.append((nu1, mF(v, nf), mF(nu1, nf), v))
        elif routine == "ph_to_pp" and not tr:
            v = BmBmF(qw, w2, nu1, nf)
            if v > nf or v < 1:
                continue
            res.append((nu1, v, nu1, v))
        elif routine == "ph_to_pp" and tr:
            v = BpBmF(qw, w2, nu1, nf)
            if v > nf or v < 1:
                continue
            res.append((nu1, v, mF(nu1, nf), FmB(nu1, qw + w2 - 1)))
        elif routine == "phbar_to_pp" and not tr:
            if FpB(w2, nu1) > nf:
                continue
            res.append((nu1, FpB(nu1, w2), nu1, pick_nu(BmBmF(qw, w2, nu1, nf), nf)))
        elif routine == "phbar_to_pp" and tr:
            if FmB(nu1, w2) > nf or FmB(nu1, w2) < 1:
                continue
            res.append((nu1, FmB(nu1, w2), mF(nu1, nf), pick_nu(FmB(nu1, qw + w2 - 1), nf)))
        else:
            raise ValueError(routine)
    return res


def add_contribution(lat: Lattice, F: list, phir_work: np.ndarray, M: np.ndarray, routine: str,
                     tr: bool, w2: int, fac1: float, fac2: float, qs=None):
    """One calc_* routine (vectorised over l1..l4 and nu1); F updated in place.

    phir_work: (Nz, Nz, Nw, NR) view of one channel's PhiR.
    M: (Nl,Nl,Nl,Nl,NIBZ,NR) tensor of this routine.
    qs: target q (0-based) to update, default all (each target q is independent).
    """
    t1, t2 = TARGETS[routine]
    nf, nl = lat.nf, lat.nl
    pw = phir_work.reshape((nf, nl, nf, nl, lat.nw, lat.nr), order="F")
    sgn = np.outer(lat.signs, lat.signs).astype(np.float64)  # (l3, l4)
    for q in (range(lat.nq) if qs is None else qs):
        qw = int(lat.qw[q])
        mq = lat.q_momentum_index(q)
        maps = routine_maps(routine, tr, nf, qw, w2)
        if not maps:
            continue
        nu1 = np.array([m[0] for m in maps]) - 1
        nu2 = np.array([m[1] for m in maps]) - 1
        nua = np.array([m[2] for m in maps]) - 1
        nub = np.array([m[3] for m in maps]) - 1
        for r2 in range(lat.nr):
            src = pw[nua, :, nub, :, w2 - 1, r2]  # (n, l3, l4)
            if tr:
                src = np.conj(src) * sgn[None, :, :]
            contrib = np.einsum("abcd,ncd->nab", M[:, :, :, :, mq, r2], src) * (1.0 / lat.nx)
            fv1 = F[t1][:, :, q].reshape((nf, nl, nf, nl), order="F")
            fv2 = F[t2][:, :, q].reshape((nf, nl, nf, nl), order="F")
            for i in range(len(nu1)):
                fv1[nu1[i], :, nu2[i], :] += contrib[i] * fac1
                fv2[nu1[i], :, nu2[i], :] += contrib[i] * fac2


def calculate_pae(lat: Lattice, phi: list, lam: list, phir: list, mt: dict, qs=None):
    """calculate_PAE: F_r = Phi_r + Lambda_r*Nc (l=1 block) + crossing; SRC/parquet_EQOuter.f90:16-118.

    phi: 4 arrays (Nz, Nz, Nq) (G_r_LL after BSE); lam: 4 arrays (Nf, Nf, Nw);
    phir: 4 arrays (Nz*Nz, Nw, NR); mt: output of calc_m_tensors.
    qs: optional list of target q (0-based); only those slices get the crossing
    (the others keep Phi + Lambda*Nc), for checks at large shapes on a sample of q.
    Returns the 4 new F_r arrays.
    """
    F = [np.array(p, dtype=np.complex128, order="F", copy=True) for p in phi]
    aux = abs(1.0 / lat.ff[0, 0, 0] * 1.0 / lat.ff[0, 0, 0])
    nf = lat.nf
    for q in range(lat.nq):
        qw = int(lat.qw[q])
        for r in range(4):
            F[r][:nf, :nf, q] += lam[r][:, :, qw - 1] * aux
    for channel in range(4):
        work = phir[channel].reshape((lat.nz, lat.nz, lat.nw, lat.nr), order="F")
        for w2 in range(1, lat.nw + 1):
            for (routine, f1, f2) in PAE_TABLE[channel]:
                add_contribution(lat, F, work, mt[routine], routine, False, w2, f1, f2, qs)
            if w2 > 1:
                for (routine, f1, f2) in PAE_TABLE[channel]:
                    add_contribution(lat, F, work, mt[routine], routine, True, w2, f1, f2, qs)
    return F


def lambda_pa(lat: Lattice, U: float):
    """Parquet-approximation Lambda: L_d=U, L_m=-U, L_s=2U, L_t=0, SRC/parquet_util.f90:311-314."""
    shape = (lat.nf, lat.nf, lat.nw)
    return [np.full(shape, U, dtype=np.complex128, order="F"),
            np.full(shape, -U, dtype=np.complex128, order="F"),
            np.full(shape, 2.0 * U, dtype=np.complex128, order="F"),
            np.zeros(shape, dtype=np.complex128, order="F")]


# ---- uneven boxes ------------------------------------------------------------------
#
# The reference fixes one Nf and one Nl for every q (SRC/parquet_ini.f90:289-348),
# so a crossing between q points with different boxes has no reference semantics.
# The library defines it as the uniform crossing on the largest box of the
# vertices embedded in it (form factors beyond Nl_q zero, frequencies outside
# q's box clamped to its edge, as pickNu does, SRC/parquet_EQContribs.f90:17-39),
# restricted to each q's box.  These restate that definition; on uniform boxes
# both are the identity.  Parity unpinned: no reference output exists.

def uneven_embed(lat_max: Lattice, blocks: list, nf_q, nl_q) -> np.ndarray:
    """Per-q blocks (D_q, D_q) -> (Nz, Nz, Nq) on the largest box lat_max."""
    nz, nf = lat_max.nz, lat_max.nf
    out = np.zeros((nz, nz, len(blocks)), dtype=np.complex128, order="F")
    n = np.arange(nf) - nf // 2
    for q, b in enumerate(blocks):
        f, l = int(nf_q[q]), int(nl_q[q])
        c = np.clip(n, -(f // 2), f // 2 - 1) + f // 2
        src = np.concatenate([l1 * f + c for l1 in range(l)])
        dst = np.concatenate([l1 * nf + np.arange(nf) for l1 in range(l)])
        out[np.ix_(dst, dst, [q])] = np.asarray(b)[np.ix_(src, src)][:, :, None]
    return out


def uneven_extract(lat_max: Lattice, full: np.ndarray, nf_q, nl_q) -> list:
    """(Nz, Nz, Nq) on the largest box -> per-q blocks (D_q, D_q)."""
    nf = lat_max.nf
    out = []
    for q in range(full.shape[2]):
        f, l = int(nf_q[q]), int(nl_q[q])
        sh = nf // 2 - f // 2
        rows = np.concatenate([l1 * nf + sh + np.arange(f) for l1 in range(l)])
        out.append(np.asfortranarray(full[np.ix_(rows, rows, [q])][:, :, 0]))
    return out


def crossing_uneven(lat_max: Lattice, phi_blocks: list, lam: list, nf_q, nl_q, mt=None) -> list:
    """F_r per q box from the per-q Phi_r blocks (4 lists): embed, calc_phir,
    calculate_pae on the largest box, extract."""
    emb = [uneven_embed(lat_max, phi_blocks[r], nf_q, nl_q) for r in range(4)]
    phir = [calc_phir(lat_max, e) for e in emb]
    mt = calc_m_tensors(lat_max) if mt is None else mt
    f = calculate_pae(lat_max, emb, lam, phir, mt)
    return [uneven_extract(lat_max, f[r], nf_q, nl_q) for r in range(4)]
