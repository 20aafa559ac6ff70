"""The parquet crossing evaluated directly in momentum and frequency space.

Test infrastructure: an independent check of oracle/parquet.py (calc_phir,
calc_m_tensors, calculate_pae), written from the physics and not from the
reference's index maps.  Nothing here uses sym_op_L, FF_inv, expSqR, expqSR,
the M tensors, the routine maps or the factor table of oracle/parquet.py.

Conventions (fixed by the reference's bubbles and ladder, not by the crossing):
* fermionic nu_n = (2n+1) pi/beta, n in [-Nf/2, Nf/2-1] <-> box index n + Nf/2 + 1
  (get_freq, SRC/parquet_util.f90:1332-1339); bosonic omega_m = 2 pi m/beta,
  stored for m = 0 .. Nf/2-1 (w = m + 1);
* a ph vertex F_ph(k, k'; q) pairs (k, k+q) with (k', k'+q) — the ph bubble is
  G(k) G(k+q) (SRC/parquet_chi.f90:470-486); a pp vertex F_pp(k, k'; q) pairs
  (k, q-k) with (k', q-k') — the pp bubble is G(k) G(q-k) (:488-498); k = (k, nu);
* the l-basis vertex is V(l nu, l' nu'; q) = sum_{k,k'} f_l(k) f_l'(k') V(k nu, k' nu'; q)
  with sum_k f_l(k) f_l'(k) = delta (SRC/parquet_formfactors.f90:11-90), so the bare
  U is U*Nc on the s-wave block (SRC/parquet_EQOuter.f90:50-59).

Symmetries used to extend the stored vertex (IBZ momenta, omega >= 0):
* C4v: V(S k, S k'; S Q) = V(k, k'; Q) for the eight point-group operations
  (rotations of momentum vectors, SRC/parquet_util.f90:896-978);
* complex conjugation of a real Hamiltonian with inversion symmetry:
  V(k nu, k' nu'; Q, -omega) = conj V(k -nu, k' -nu'; Q, omega).
Fermionic arguments outside the box are clamped to its edge, the reference's
treatment of the asymptotics (pickNu, SRC/parquet_EQContribs.f90:17-39).

Crossing relations of the SU(2) parquet equations in the density/magnetic/
singlet/triplet channels (Bickers, "Self-consistent many-body theory for
condensed matter systems", 2004; Rohringer, Valli, Toschi, PRB 86, 125114
(2012)), in the conventions above, with 4-vectors k = (k, nu), q = (Q, omega):

  F_d += -1/2 Phi_d(k, k+q; k'-k) - 3/2 Phi_m(k, k+q; k'-k)
         +1/2 Phi_s(k, k'; k+k'+q) + 3/2 Phi_t(k, k'; k+k'+q)
  F_m += -1/2 Phi_d(k, k+q; k'-k) + 1/2 Phi_m(k, k+q; k'-k)
         -1/2 Phi_s(k, k'; k+k'+q) + 1/2 Phi_t(k, k'; k+k'+q)
  F_s += +1/2 Phi_d(k, k'; q-k-k') - 3/2 Phi_m(k, k'; q-k-k')
         +1/2 Phi_d(k, q-k'; k'-k) - 3/2 Phi_m(k, q-k'; k'-k)
  F_t += +1/2 Phi_d(k, k'; q-k-k') + 1/2 Phi_m(k, k'; q-k-k')
         -1/2 Phi_d(k, q-k'; k'-k) - 1/2 Phi_m(k, q-k'; k'-k)

A term enters when its transfer frequency is a stored bosonic frequency or the
negative of one (|omega'| <= omega_{Nf/2-1}).  The projection onto the Nl form
factors is exact when the real-space sum of the truncated-unity form runs over
every lattice vector (NR = number of IBZ points), which is the setting the
tests use.
"""
from __future__ import annotations

import numpy as np

from oracle.lattice import NS, Lattice, sym_vec

SOURCES = ("d", "m", "s", "t")
# target -> list of (relation, source channel, coefficient)
RELATIONS = {
    0: [("phbar", 0, -0.5), ("phbar", 1, -1.5), ("pp_ph", 2, 0.5), ("pp_ph", 3, 1.5)],
    1: [("phbar", 0, -0.5), ("phbar", 1, 0.5), ("pp_ph", 2, -0.5), ("pp_ph", 3, 0.5)],
    2: [("ph_pp", 0, 0.5), ("ph_pp", 1, -1.5), ("phbar_pp", 0, 0.5), ("phbar_pp", 1, -1.5)],
    3: [("ph_pp", 0, 0.5), ("ph_pp", 1, 0.5), ("phbar_pp", 0, -0.5), ("phbar_pp", 1, -0.5)],
}


class KSpace:
    """Full-zone, full-frequency k-space vertices built from IBZ l-basis data."""

    def __init__(self, lat: Lattice):
        self.lat = lat
        nx = lat.nx
        self.nc = nx * nx
        # f[l, k] with k = kx + nx*ky (0-based components)
        self.f = np.stack([lat.ff[:, :, l].T.reshape(-1) for l in range(lat.nl)])
        kx, ky = np.meshgrid(np.arange(nx), np.arange(nx), indexing="xy")
        self.kvec = np.stack([kx.reshape(-1), ky.reshape(-1)], axis=1)  # (Nc, 2)
        # operation tables on k indices: rot[s, k] = index of S_s k
        self.rot = np.zeros((NS, self.nc), dtype=np.int64)
        for s in range(1, NS + 1):
            for k, (x, y) in enumerate(self.kvec):
                sx, sy = sym_vec(s, int(x), int(y))
                self.rot[s - 1, k] = (sx % nx) + nx * (sy % nx)
        self.inv_rot = np.argsort(self.rot, axis=1)  # S^-1
        self.ibz_k = [((qx - 1) % nx) + nx * ((qy - 1) % nx) for (qx, qy) in lat.ibz_points]

    def add(self, a, b, sign=1):
        """k index of k_a + sign*k_b."""
        nx = self.lat.nx
        v = self.kvec[a] + sign * self.kvec[b]
        return (v[..., 0] % nx) + nx * (v[..., 1] % nx)

    def neg(self, a):
        nx = self.lat.nx
        v = -self.kvec[a]
        return (v[..., 0] % nx) + nx * (v[..., 1] % nx)

    def to_k(self, vl):
        """(Nf, Nl, Nf, Nl) l-basis block -> (Nc, Nf, Nc, Nf) k-space block."""
        return np.einsum("iajb,ak,bp->kipj", vl, self.f, self.f, optimize=True)

    def to_l(self, vk):
        """(Nc, Nf, Nc, Nf) -> (Nf, Nl, Nf, Nl)."""
        return np.einsum("kipj,ak,bp->iajb", vk, self.f, self.f, optimize=True)

    def symmetric_full_zone(self, phi_l):
        """phi_l: (Nf, Nl, Nf, Nl, NIBZ, Nw) at the IBZ momenta.  Returns the C4v-symmetric
        k-space vertex on the whole zone, (Nc[Q], Nw, Nc, Nf, Nc, Nf), and its projection back
        on the IBZ momenta in the reference layout (Nz, Nz, Nq)."""
        lat = self.lat
        nc, nf, nw = self.nc, lat.nf, lat.nw
        out = np.zeros((nc, nw, nc, nf, nc, nf), dtype=np.complex128)
        cnt = np.zeros(nc)
        for m, qk in enumerate(self.ibz_k):
            for w in range(nw):
                vk = self.to_k(phi_l[:, :, :, :, m, w])
                for s in range(NS):
                    Q = self.rot[s, qk]
                    # V(S k, S k'; S Q) = V(k, k'; Q)  ->  V_SQ(k, k') = V_Q(S^-1 k, S^-1 k')
                    ik = self.inv_rot[s]
                    out[Q, w] += vk[ik][:, :, ik, :]
                    if w == 0:
                        cnt[Q] += 1
        out /= cnt[:, None, None, None, None, None]
        # the reference's input: projection at the IBZ momenta
        ref = np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128, order="F")
        for m, qk in enumerate(self.ibz_k):
            for w in range(nw):
                vl = self.to_l(out[qk, w])  # (Nf, Nl, Nf, Nl): row (l, nu) = l*Nf + nu
                ref[:, :, m * nw + w] = vl.reshape(lat.nz, lat.nz, order="F")
        return out, ref


def _source(ks: KSpace, full, Q, m2, a_k, a_n, b_k, b_n):
    """Phi(k_a n_a, k_b n_b; Q, omega_{m2}) for signed m2 with |m2| < Nw, vectorised.

    Negative m2 through V(k nu, k' nu'; Q, -omega) = conj V(k -nu, k' -nu'; Q, omega);
    n_b (and n_a) are clamped to the box first.  Returns (values, valid mask)."""
    nf, nw = ks.lat.nf, ks.lat.nw
    lo, hi = -nf // 2, nf // 2 - 1
    a_n = np.clip(a_n, lo, hi)
    b_n = np.clip(b_n, lo, hi)
    valid = np.abs(m2) <= nw - 1
    m2c = np.where(valid, m2, 0)
    neg = m2c < 0
    ia = np.where(neg, -a_n - 1, a_n) - lo
    ib = np.where(neg, -b_n - 1, b_n) - lo
    v = full[Q, np.abs(m2c), a_k, ia, b_k, ib]
    v = np.where(neg, np.conj(v), v)
    return np.where(valid, v, 0.0), valid


def crossing_kspace(lat: Lattice, phi: list, seed_full=None):
    """Crossing contributions to F_d, F_m, F_s, F_t at the IBZ q (reference layout
    (Nz, Nz, Nq)), from the reducible vertices phi (reference layout), by the
    relations in the module docstring.

    phi must be the projection of C4v-symmetric vertices (symmetric_full_zone's
    second output); the full-zone k-space vertices are rebuilt from it here.
    """
    ks = KSpace(lat)
    nf, nw, nc, nl = lat.nf, lat.nw, ks.nc, lat.nl
    fulls = []
    for r in range(4):
        # q = m*Nw + w (IBZ momentum slowest), SRC/parquet_util.f90:110-117
        pl = phi[r].reshape(nf, nl, nf, nl, nw, lat.nibz, order="F").transpose(0, 1, 2, 3, 5, 4)
        full, _ = ks.symmetric_full_zone(pl)
        fulls.append(full)
    # grids over the target entry (k, n, k', n')
    k = np.arange(nc)[:, None, None, None]
    kp = np.arange(nc)[None, None, :, None]
    n = (np.arange(nf) - nf // 2)[None, :, None, None]
    npr = (np.arange(nf) - nf // 2)[None, None, None, :]
    shape = (nc, nf, nc, nf)
    k, kp, n, npr = (np.broadcast_to(x, shape) for x in (k, kp, n, npr))
    out = [np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128, order="F") for _ in range(4)]
    for mq, Q in enumerate(ks.ibz_k):
        Qa = np.full(shape, Q)
        for w in range(nw):
            om = w  # omega index m of the target transfer
            terms = {}
            # ph-bar into ph: Phi(k n, k+Q n+om; k'-k, n'-n)
            terms["phbar"] = (ks.add(kp, k, -1), npr - n, k, n, ks.add(k, Qa), n + om)
            # pp into ph: Phi(k n, k' n'; k+k'+Q, n+n'+1+om)
            terms["pp_ph"] = (ks.add(ks.add(k, kp), Qa), n + npr + 1 + om, k, n, kp, npr)
            # ph into pp: Phi(k n, k' n'; Q-k-k', om-n-n'-1)
            terms["ph_pp"] = (ks.add(Qa, ks.add(k, kp), -1), om - n - npr - 1, k, n, kp, npr)
            # ph-bar into pp: Phi(k n, Q-k' om-n'-1; k'-k, n'-n)
            terms["phbar_pp"] = (ks.add(kp, k, -1), npr - n, k, n, ks.add(Qa, kp, -1), om - npr - 1)
            cache = {}
            for tgt in range(4):
                acc = np.zeros(shape, dtype=np.complex128)
                for rel, src, coef in RELATIONS[tgt]:
                    key = (rel, src)
                    if key not in cache:
                        Qs, m2, ak, an, bk, bn = terms[rel]
                        cache[key] = _source(ks, fulls[src], Qs, m2, ak, an, bk, bn)[0]
                    acc += coef * cache[key]
                vl = ks.to_l(acc)
                out[tgt][:, :, mq * nw + w] = vl.reshape(lat.nz, lat.nz, order="F")
    return out


def symmetric_phi(lat: Lattice, rng) -> list:
    """Four random reducible vertices made C4v-symmetric (little-group averaged), in the
    reference layout (Nz, Nz, Nq)."""
    ks = KSpace(lat)
    out = []
    for _ in range(4):
        pl = (rng.standard_normal((lat.nf, lat.nl, lat.nf, lat.nl, lat.nibz, lat.nw))
              + 1j * rng.standard_normal((lat.nf, lat.nl, lat.nf, lat.nl, lat.nibz, lat.nw)))
        _, ref = ks.symmetric_full_zone(pl)
        out.append(ref)
    return out


def symmetric_phi_full(lat: Lattice, rng):
    """As symmetric_phi, also returning the full-zone k-space vertices
    ((Nc[Q], Nw, Nc, Nf, Nc, Nf) each)."""
    ks = KSpace(lat)
    fulls, refs = [], []
    for _ in range(4):
        pl = (rng.standard_normal((lat.nf, lat.nl, lat.nf, lat.nl, lat.nibz, lat.nw))
              + 1j * rng.standard_normal((lat.nf, lat.nl, lat.nf, lat.nl, lat.nibz, lat.nw)))
        full, ref = ks.symmetric_full_zone(pl)
        fulls.append(full)
        refs.append(ref)
    return fulls, refs


def sde_kspace(lat: Lattice, fulls, lam_d, lam_m, sigma_k, sigma_h_k, mu, beta, U, f_range=0):
    """The Schwinger-Dyson self-energy evaluated directly in momentum/frequency space.

    Sigma(k, nu) = -U/(2 beta^2 Nc^2) sum_{Q in BZ} sum_{|omega| < Nw}
        [ G(k+Q, nu+omega) sum_{k', nu' in box} ((L_d - L_m)(nu, nu'; omega)
                 + Phi_d(k nu, k' nu'; Q omega) - 3 Phi_m(k nu, k' nu'; Q omega)) G(k', nu') G(k'+Q, nu'+omega)
        + G(Q-k, omega-nu) sum_{k', nu' in box} (Phi_s + Phi_t)(k nu, k' nu'; Q omega) G(k', nu') G(Q-k', omega-nu') ]

    The SU(2) SDE with one bare U vertex, F_updown = (Lambda + ph + ph-bar + pp)
    parts, the ph-bar part relabelled into the ph form (hence Phi_d - 3 Phi_m,
    Eckhardt, Honerkamp, Held, Kauch, arXiv:1912.07469, the paper the
    reference's README cites), the
    local parquet Lambda entering as a k-independent vertex.  The external line k
    attaches to the vertex's first pair; the bubble to its second.  Negative omega
    through V(k nu, k' nu'; Q, -omega) = conj V(k -nu, k' -nu'; Q, omega) for the
    vertices and Lambda.  G(k, nu) = 1 / (i nu - eps_k + mu - Sigma(k, nu)) with
    Sigma_H(k) outside the box.  fulls: symmetric_phi_full's k-space vertices;
    lam_*: (Nf, Nf, Nw) as stored; sigma_k: (Nc, Nf) with k = kx + Nx ky,
    sigma_h_k: (Nc,).  f_range > 0 extends the nu' sums to (2 f_range + 1) Nf
    frequencies: there the vertices take their value at the nearest box frequency
    (the reference's asymptotics, pickNu) and Lambda its bare limit U_d - U_m = 2U.
    Returns Sigma (Nc, Nf)."""
    ks = KSpace(lat)
    nx, nf, nw, nc = lat.nx, lat.nf, lat.nw, ks.nc
    kx = 2.0 * np.pi * ks.kvec[:, 0] / nx
    ky = 2.0 * np.pi * ks.kvec[:, 1] / nx
    eps = -2.0 * np.cos(kx) - 2.0 * np.cos(ky)

    def G(k, n):
        """k: k index array, n: box index array (any integer; outside the box Sigma_H)."""
        k, n = np.broadcast_arrays(k, n)
        inside = (n >= 0) & (n < nf)
        nu = (2.0 * (n - nf // 2) + 1.0) * np.pi / beta
        sig = np.where(inside, sigma_k[k, np.clip(n, 0, nf - 1)], sigma_h_k[k])
        return 1.0 / (1j * nu - eps[k] + mu - sig)

    def vertex(full, Q, om):
        if om >= 0:
            return full[Q, om]
        return np.conj(full[Q, -om][:, ::-1, :, ::-1])

    def lam(L, om):
        return L[:, :, om] if om >= 0 else np.conj(L[::-1, ::-1, -om])

    kk = np.arange(nc)[:, None]
    nn = np.arange(nf)[None, :]
    ne = np.arange(-f_range * nf, (f_range + 1) * nf)[None, :]   # the nu' range
    ce = np.clip(ne[0], 0, nf - 1)                                # its nearest box frequency
    inside = (ne[0] >= 0) & (ne[0] < nf)
    out = np.zeros((nc, nf), dtype=np.complex128)
    for Q in range(nc):
        for om in range(-(nw - 1), nw):
            b_ph = G(kk, ne) * G(ks.add(kk, Q), ne + om)                     # (k', n')
            b_pp = G(kk, ne) * G(ks.add(np.full_like(kk, Q), kk, -1), om + nf - 1 - ne)
            v_ph = vertex(fulls[0], Q, om) - 3.0 * vertex(fulls[1], Q, om)    # (k, n, k', n' in box)
            v_pp = vertex(fulls[2], Q, om) + vertex(fulls[3], Q, om)
            x_ph = np.einsum("anbm,bm->an", v_ph[:, :, :, ce], b_ph)
            lam_e = np.where(inside[None, :], (lam(lam_d, om) - lam(lam_m, om))[:, ce], 2.0 * U)
            x_ph += np.einsum("nm,bm->n", lam_e, b_ph)[None, :]
            x_pp = np.einsum("anbm,bm->an", v_pp[:, :, :, ce], b_pp)
            out += G(ks.add(kk, Q), nn + om) * x_ph
            out += G(ks.add(np.full_like(kk, Q), kk, -1), om + nf - 1 - nn) * x_pp
    return -U / (2.0 * beta * beta * nc * nc) * out
