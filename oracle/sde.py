"""Oracle: the l-space Schwinger-Dyson equation (self-energy from the vertices).

Test infrastructure only: the checker for the GPU path, never shipped or timed.

Restates calc_SDE (SRC/parquet_SDE.f90:27-258) without the use_U2
second-order term calc_order_2nd (:263-340, FFTPACK + spline Matsubara
transform), with its helpers
* calc_chi_aux        SRC/parquet_SDE.f90:346-432
* calc_LChi           SRC/parquet_SDE.f90:436-497
* calc_PhChi          SRC/parquet_SDE.f90:501-537
* calc_PPChi          SRC/parquet_SDE.f90:540-575
* fillpickNu_range    SRC/parquet_SDE.f90:650-678
* get_green           SRC/parquet_util.f90:1278-1299 (Gkw, SRC/parquet_chi.f90:208)
* symmetrize_array    SRC/parquet_util.f90:1215-1272
* Ek                  SRC/parquet_util.f90:128-140
and the index maps index_FaddB / index_minusF (SRC/parquet_util.f90:1345-1381).

Pinned by (tests/test_oracle_sde.py):
* the second-order diagram: with every reducible vertex zero and the
  parquet-approximation Lambda (L_d = U, L_m = -U) the l-space SDE reduces to
  Sigma(k, nu) = -U^2/(beta^2 Nc^2) sum_{Q in BZ} sum_s G(k+Q, nu+s)
                 sum_{k', nu' in box} G(k', nu') G(k'+Q, nu'+s)
  over bosonic shifts |s| < Nf/2, evaluated by brute force over the whole zone
  (no IBZ, no stars, no form factors) -- this pins LChi, the star sum, FF_inv,
  the time-reversed half and the symmetrisation;
* at Nl = 1 (s-wave, invariant under every operation) the star sum over the
  IBZ equals a plain sum over every Q of the zone with the vertex of Q's IBZ
  representative, for random vertices -- this pins PhChi and PPChi's place in
  the sum;
* linearity in the vertices;
* the Schwinger-Dyson equation evaluated directly in momentum and frequency
  space (tests/test_oracle_sde_physics.py) for C4v-symmetric vertices in the
  span of the Nl form factors at Nl = 1, 5, 9 and f_range = 0, 1, 2: this pins
  the (row, column) orientation of G_r_LL in PhChi / PPChi (the external line on
  the row), FF_inv per operation, the form-factor signs of the time-reversed
  half, the pp term and the enlarged frequency sums with the pickNu clamp.
"""
from __future__ import annotations

import math

import numpy as np

from .bse import ek_grain, green_coarse
from .lattice import NS, Lattice, list_symmetries, matsubara, sym_index


def ek(lat: Lattice) -> np.ndarray:
    """Ek(kx, ky) (0-based), SRC/parquet_util.f90:128-140."""
    k = 2.0 * math.pi * np.arange(lat.nx) / lat.nx
    return -2.0 * np.cos(k)[:, None] - 2.0 * np.cos(k)[None, :]


def green(lat: Lattice, sigma, sigma_h, mu, beta, kx, ky, nu, e=None):
    """get_green for 1-based (kx, ky, nu): Gkw inside the box, Sigma_H outside."""
    e = ek(lat) if e is None else e
    z = matsubara(nu, lat.nf, beta)
    if nu < 1 or nu > lat.nf:
        s = sigma_h[kx - 1, ky - 1]
    else:
        s = sigma[((kx - 1) * lat.ny + ky - 1) * lat.nf + nu - 1]
    return 1.0 / (z - e[kx - 1, ky - 1] + mu - s)


def faddb(kx, ky, nu, qx, qy, w, nx, ny):
    """index_FaddB: (k + q, nu + w - 1)."""
    return (kx - 1 + qx - 1) % nx + 1, (ky - 1 + qy - 1) % ny + 1, nu + w - 1


def minusf(kx, ky, nu, nx, ny, nf):
    """index_minusF: (-k, Nf + 1 - nu)."""
    return (-(kx - 1) + nx) % nx + 1, (-(ky - 1) + ny) % ny + 1, -nu + nf + 1


def pick_range(x, nf, f_range):
    """pickNu_range(x) for x in 1..(2 f_range + 1) Nf: clamp of x - f_range Nf to [1, Nf]."""
    return min(max(x - f_range * nf, 1), nf)


def chi_aux(lat: Lattice, sigma, sigma_h, mu, beta, ngrain=1, f_range=0):
    """Chi_aux_ph, Chi_aux_pp ((2 f_range + 1) Nf Nl, Nq): SRC/parquet_SDE.f90:346-432."""
    nx, ny, nf, nl = lat.nx, lat.ny, lat.nf, lat.nl
    ext = (2 * f_range + 1) * nf
    ekg = ek_grain(lat, ngrain)
    ph = np.zeros((ext * nl, lat.nq), dtype=np.complex128)
    pp = np.zeros((ext * nl, lat.nq), dtype=np.complex128)
    for iq in range(lat.nq):
        qx, qy, qw = int(lat.qx[iq]), int(lat.qy[iq]), int(lat.qw[iq])
        for l4 in range(1, nl + 1):
            for l40 in range(-f_range * nf + 1, (f_range + 1) * nf + 1):
                idx = (l4 - 1) * ext + f_range * nf + l40 - 1
                for kx in range(1, nx + 1):
                    for ky in range(1, ny + 1):
                        ff = lat.ff[kx - 1, ky - 1, l4 - 1]
                        kqx, kqy, kqw = faddb(kx, ky, l40, qx, qy, qw, nx, ny)
                        for gx in range(ngrain):
                            for gy in range(ngrain):
                                gk = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kx, ky, l40, sigma_h)
                                gkq = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kqx, kqy, kqw, sigma_h)
                                ph[idx, iq] += gk * gkq * ff * (1.0 / (ngrain * ngrain))
                        mkx, mky, mkw = minusf(kx, ky, l40, nx, ny, nf)
                        qkx, qky, qkw = faddb(mkx, mky, mkw, qx, qy, qw, nx, ny)
                        for gx in range(ngrain):
                            for gy in range(ngrain):
                                gk = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kx, ky, l40, sigma_h)
                                gqk = green_coarse(lat, ekg, sigma, mu, beta, ngrain - gx - 1, ngrain - gy - 1,
                                                  qkx, qky, qkw, sigma_h)
                                pp[idx, iq] += gk * gqk * ff * (1.0 / (ngrain * ngrain))
    return ph, pp


def l_chi(lat: Lattice, cph, lam_d, lam_m, U, use_u2=False, f_range=0):
    """LChi (Nz, Nq): SRC/parquet_SDE.f90:436-497 (only the l = 1 block is nonzero)."""
    nf, nz = lat.nf, lat.nz
    ext = (2 * f_range + 1) * nf
    ff11 = lat.ff[0, 0, 0]
    out = np.zeros((nz, lat.nq), dtype=np.complex128)
    for iq in range(lat.nq):
        w = int(lat.qw[iq])
        for i3 in range(1, nf + 1):  # idxL3 <= Nf, l4 = 1
            acc = 0j
            for l40 in range(-f_range * nf + 1, (f_range + 1) * nf + 1):
                c = cph[f_range * nf + l40 - 1, iq]
                if 0 < l40 < nf + 1:
                    acc += lam_d[i3 - 1, l40 - 1, w - 1] * (1.0 / (ff11 * ff11)) * c
                    acc -= lam_m[i3 - 1, l40 - 1, w - 1] * (1.0 / (ff11 * ff11)) * c
                    if use_u2:
                        acc -= 2.0 * U * (1.0 / (ff11 * ff11)) * c
                elif not use_u2:
                    acc += 2.0 * U * (1.0 / (ff11 * ff11)) * c
            out[i3 - 1, iq] = acc
    return out


def ph_chi(lat: Lattice, cph, g_d, g_m, f_range=0):
    """PhChi (Nz, Nq) = sum (G_d - 3 G_m)(idxL3, (l4, pick(l4_0)), q) Chi_aux_ph: SRC/parquet_SDE.f90:501-537."""
    return _v_chi(lat, cph, g_d, g_m, -3.0, f_range)


def pp_chi(lat: Lattice, cpp, g_s, g_t, f_range=0):
    """PPChi (Nz, Nq) = sum (G_s + G_t)(idxL3, (l4, pick(l4_0)), q) Chi_aux_pp: SRC/parquet_SDE.f90:540-575."""
    return _v_chi(lat, cpp, g_s, g_t, 1.0, f_range)


def _v_chi(lat, caux, g0, g1, c1, f_range):
    nf, nl = lat.nf, lat.nl
    ext = (2 * f_range + 1) * nf
    out = np.zeros((lat.nz, lat.nq), dtype=np.complex128)
    for iq in range(lat.nq):
        for l4 in range(1, nl + 1):
            for l40 in range(-f_range * nf + 1, (f_range + 1) * nf + 1):
                idx4 = (l4 - 1) * ext + f_range * nf + l40 - 1
                col = (l4 - 1) * nf + pick_range(l40 + f_range * nf, nf, f_range) - 1
                out[:, iq] += (g0[:, col, iq] + c1 * g1[:, col, iq]) * caux[idx4, iq]
    return out


def calc_sde(lat: Lattice, gv, lam, sigma, sigma_h, mu, beta, U, ngrain=1, f_range=0, use_u2=False):
    """Sigma over the whole zone (Nx Ny Nf), SRC/parquet_SDE.f90:27-258 without calc_order_2nd.

    gv: the four reducible vertices G_r_LL (Nz, Nz, Nq) (the arrays the
    reference passes at SRC/loop.f90:154, after calculate_PAE); lam: (L_d, L_m)."""
    nx, ny, nf, nl = lat.nx, lat.ny, lat.nf, lat.nl
    cph, cpp = chi_aux(lat, sigma, sigma_h, mu, beta, ngrain, f_range)
    lc = l_chi(lat, cph, lam[0], lam[1], U, use_u2, f_range)
    pc = ph_chi(lat, cph, gv[0], gv[1], f_range)
    sc = pp_chi(lat, cpp, gv[2], gv[3], f_range)
    e = ek(lat)
    nred2 = lat.nibz * nf
    sig_v = np.zeros(nred2, dtype=np.complex128)
    for k0 in range(1, nf + 1):
        for kx in range(1, lat.nx_ibz + 1):
            for ky in range(1, kx + 1):
                mkx, mky, mkw = minusf(kx, ky, k0, nx, ny, nf)
                acc = 0j
                for iq in range(lat.nq):
                    qx, qy, qw = int(lat.qx[iq]), int(lat.qy[iq]), int(lat.qw[iq])
                    syms = list_symmetries(qx, qy, lat.nx_ibz)
                    for s in range(1, NS + 1):
                        if not syms[s - 1]:
                            continue
                        sx, sy = sym_index(s, qx, qy, nx, ny)
                        g_kq = green(lat, sigma, sigma_h, mu, beta, *faddb(kx, ky, k0, sx, sy, qw, nx, ny), e=e)
                        qmk = faddb(mkx, mky, mkw, sx, sy, qw, nx, ny)
                        g_qmk = green(lat, sigma, sigma_h, mu, beta, *qmk, e=e)
                        g_kmq = green(lat, sigma, sigma_h, mu, beta, *minusf(*qmk, nx, ny, nf), e=e)
                        for l3 in range(1, nl + 1):
                            fi = lat.ff_inv[s - 1, kx - 1, ky - 1, l3 - 1]
                            i3 = (l3 - 1) * nf + k0 - 1
                            acc += g_kq * fi * lc[i3, iq]
                            acc += g_kq * fi * pc[i3, iq]
                            acc += g_qmk * fi * sc[i3, iq]
                            if qw == 1:
                                continue
                            i3m = (l3 - 1) * nf + (-k0 + nf + 1) - 1
                            sg = lat.signs[l3 - 1]
                            acc += g_kmq * fi * sg * np.conj(lc[i3m, iq])
                            acc += g_kmq * fi * sg * np.conj(pc[i3m, iq])
                            acc += np.conj(g_kq) * fi * sg * np.conj(sc[i3m, iq])
                idx_k = (((kx - 1) * kx) // 2 + ky - 1) * nf + k0 - 1
                sig_v[idx_k] -= U / beta / beta / lat.nc / lat.nc / 2.0 * acc
    return symmetrize(lat, sig_v, nf)


def symmetrize(lat: Lattice, arr_in, num_freqs):
    """symmetrize_array: IBZ array -> whole zone, SRC/parquet_util.f90:1215-1272."""
    out = np.zeros(lat.nx * lat.ny * num_freqs, dtype=np.complex128)
    for qx in range(1, lat.nx_ibz + 1):
        for qy in range(1, qx + 1):
            syms = list_symmetries(qx, qy, lat.nx_ibz)
            for nu in range(1, num_freqs + 1):
                i = (((qx - 1) * qx) // 2 + qy - 1) * num_freqs + nu - 1
                for s in range(1, NS + 1):
                    if not syms[s - 1]:
                        continue
                    sx, sy = sym_index(s, qx, qy, lat.nx, lat.ny)
                    out[((sx - 1) * lat.ny + sy - 1) * num_freqs + nu - 1] += arr_in[i]
    return out
