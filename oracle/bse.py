"""Oracle: Bethe-Salpeter ladder update and the bubble chi0 in l-space.

Restates
* BSE(ite)           SRC/parquet_BSE.f90:19-135  (4 ZGEMMs + damped mixing per channel)
* calc_chi           SRC/parquet_chi.f90:438-523
* get_green_coarse   SRC/parquet_util.f90:1303-1327
* Ek_grain           SRC/parquet_util.f90:143-164
* index_FaddB/minusF SRC/parquet_util.f90:1345-1381

Pinned by (tests/test_oracle_bse.py):
* the fixed-point identity of the ladder: with F = (I - kappa_r Gamma chi)^-1 Gamma
  (LAPACK solve via numpy, kappa_r = 2 c_r alpha with the reference's c_r) the
  damped update returns exactly F - Gamma for every damping f;
* the ite-0 special case F = Gamma: Phi = 2 f c_r alpha Gamma chi Gamma;
* calc_chi against a brute-force triple loop on a tiny grid, and the Nl=1
  closed form chi(nu,nu) = (1/Nc) sum_k G G;
* bse_closed (the closed form of the ladder, no reference counterpart: the
  reference only iterates the damped update) by the same fixed-point identity
  run the other way: the ladder applied to F = Gamma + Phi_closed returns
  Phi_closed for every damping f.
"""
from __future__ import annotations

import math

import numpy as np

from .lattice import Lattice, matsubara

# channel order d, m, s, t; prefactor of the second ZGEMM pair divided by f_damping
# SRC/parquet_BSE.f90:73,89,105,120
C_CHANNEL = (0.5, 0.5, -0.25, 0.25)
# which bubble each channel uses: ph for d,m; pp for s,t (SRC/parquet_BSE.f90:72-124)
CHI_OF_CHANNEL = (0, 0, 1, 1)


def bse_ladder(lat: Lattice, gamma, F, chi, beta: float, f_damping: float):
    """One BSE(ite) call: returns Phi_r (the new G_r_LL) for r in d,m,s,t.

    gamma, F: lists of 4 complex arrays (Nz, Nz, Nq), Fortran order.
    chi: tuple (chi_ph, chi_pp) of dense (Nz, Nz, Nq) arrays.
    Loop-faithful: per q and channel the same four ZGEMMs as the reference.
    """
    alpha = 1.0 / (beta * lat.nx * lat.ny)
    out = []
    for r in range(4):
        g = gamma[r]
        fr = F[r]
        ch = chi[CHI_OF_CHANNEL[r]]
        c = C_CHANNEL[r] * f_damping
        phi = np.empty_like(g, order="F")
        for q in range(g.shape[2]):
            gq = g[:, :, q]
            fq = fr[:, :, q]
            cq = ch[:, :, q]
            chiG = alpha * (gq @ cq)
            tmp = c * (chiG @ fq)
            chiG = alpha * (cq @ gq)
            tmp += c * (fq @ chiG)
            phi[:, :, q] = tmp + (1.0 - f_damping) * (fq - gq)
        out.append(phi)
    return out


def bse_closed(lat: Lattice, gamma, chi, beta: float):
    """Fixed point of BSE(ite) for fixed Gamma: Phi_r = (I - kappa_r Gamma_r chi)^-1 Gamma_r - Gamma_r,
    kappa_r = 2 c_r alpha (c_r of SRC/parquet_BSE.f90:73,89,105,120), by LAPACK (numpy.linalg.solve).

    gamma: 4 arrays (Nz, Nz, Nq); chi: (chi_ph, chi_pp) dense (Nz, Nz, Nq).
    """
    alpha = 1.0 / (beta * lat.nx * lat.ny)
    out = []
    for r in range(4):
        kappa = 2.0 * C_CHANNEL[r] * alpha
        g = gamma[r]
        ch = chi[CHI_OF_CHANNEL[r]]
        phi = np.empty_like(g, order="F")
        eye = np.eye(g.shape[0])
        for q in range(g.shape[2]):
            a = eye - kappa * (g[:, :, q] @ ch[:, :, q])
            phi[:, :, q] = np.linalg.solve(a, g[:, :, q]) - g[:, :, q]
        out.append(phi)
    return out


def chi_compact_to_dense(lat: Lattice, chi_c: np.ndarray) -> np.ndarray:
    """(Nl, Nl, Nf, Nq) block-diagonal bubble -> dense (Nz, Nz, Nq) (Fortran order)."""
    nl, nf, nq = lat.nl, lat.nf, chi_c.shape[3]
    dense = np.zeros((lat.nz, lat.nz, nq), dtype=np.complex128, order="F")
    nu = np.arange(nf)
    for l1 in range(nl):
        for l2 in range(nl):
            dense[l1 * nf + nu, l2 * nf + nu, :] = chi_c[l1, l2, :, :]
    return dense


def ek_grain(lat: Lattice, ngrain: int) -> np.ndarray:
    """Ek_grain(ix2, iy2, ix, iy), SRC/parquet_util.f90:143-164 (0-based indices)."""
    dkx = 2.0 * math.pi / lat.nx
    dky = 2.0 * math.pi / lat.ny
    out = np.zeros((ngrain, ngrain, lat.nx, lat.ny))
    half = (ngrain + 1) // 2  # Fortran integer division
    for iy in range(1, lat.ny + 1):
        for ix in range(1, lat.nx + 1):
            for iy2 in range(1, ngrain + 1):
                for ix2 in range(1, ngrain + 1):
                    px = dkx * (ix - 1) + (ix2 - half) * dkx / float(ngrain)
                    py = dky * (iy - 1) + (iy2 - half) * dky / float(ngrain)
                    out[ix2 - 1, iy2 - 1, ix - 1, iy - 1] = -2.0 * math.cos(px) - 2.0 * math.cos(py)
    return out


def green_coarse(lat: Lattice, ekg, sigma, mu, beta, gx, gy, kx, ky, nu, sigma_h=None):
    """get_green_coarse for 1-based (kx, ky, nu), 0-based grain (gx, gy).

    sigma_h: Sigma_H(Nx, Ny) used outside the box (None: zero)."""
    z = matsubara(nu, lat.nf, beta)
    e = ekg[gx, gy, kx - 1, ky - 1]
    if nu < 1 or nu > lat.nf:
        sh = 0.0 if sigma_h is None else sigma_h[kx - 1, ky - 1]
        return 1.0 / (z - e + mu - sh)
    s = sigma[((kx - 1) * lat.ny + ky - 1) * lat.nf + nu - 1]
    return 1.0 / (z - e + mu - s)


def calc_chi(lat: Lattice, qx: int, qy: int, qw: int, sigma, mu: float, beta: float, ngrain: int = 1,
             sigma_h=None):
    """Bubbles chi_ph_LL, chi_pp_LL (Nz, Nz) for one 1-based bosonic (qx, qy, qw).

    SRC/parquet_chi.f90:438-523; loops vectorised over the fermionic index,
    order of the Nl x Nl x k sums kept.
    Returns dense (Nz, Nz) arrays (block-diagonal in nu).
    """
    nx, ny, nf = lat.nx, lat.ny, lat.nf
    ekg = ek_grain(lat, ngrain)
    chi_k_ph = np.zeros((nx, ny, nf), dtype=np.complex128)
    chi_k_pp = np.zeros((nx, ny, nf), dtype=np.complex128)
    for kx in range(1, nx + 1):
        for ky in range(1, ny + 1):
            kqx = (kx - 1 + qx - 1) % nx + 1
            kqy = (ky - 1 + qy - 1) % ny + 1
            mkx = (-(kx - 1) + nx) % nx + 1
            mky = (-(ky - 1) + ny) % ny + 1
            qmkx = (mkx - 1 + qx - 1) % nx + 1
            qmky = (mky - 1 + qy - 1) % ny + 1
            for nu in range(1, nf + 1):
                nu_kq = nu + qw - 1
                nu_qmk = (-nu + nf + 1) + qw - 1
                acc_ph = 0j
                acc_pp = 0j
                for gx in range(ngrain):
                    for gy in range(ngrain):
                        gk = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kx, ky, nu, sigma_h)
                        gkq = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kqx, kqy, nu_kq, sigma_h)
                        acc_ph += gk * gkq * 1.0 / ngrain / ngrain
                        gqmk = green_coarse(lat, ekg, sigma, mu, beta, ngrain - gx - 1, ngrain - gy - 1, qmkx, qmky,
                                            nu_qmk, sigma_h)
                        acc_pp += gk * gqmk * 1.0 / ngrain / ngrain
                chi_k_ph[kx - 1, ky - 1, nu - 1] = acc_ph
                chi_k_pp[kx - 1, ky - 1, nu - 1] = acc_pp
    ff = lat.ff
    chi_ph = np.zeros((lat.nz, lat.nz), dtype=np.complex128, order="F")
    chi_pp = np.zeros((lat.nz, lat.nz), dtype=np.complex128, order="F")
    nu = np.arange(nf)
    for l2 in range(lat.nl):
        for l1 in range(lat.nl):
            w = ff[:, :, l1] * ff[:, :, l2]
            chi_ph[l1 * nf + nu, l2 * nf + nu] = np.einsum("xyn,xy->n", chi_k_ph, w)
            chi_pp[l1 * nf + nu, l2 * nf + nu] = np.einsum("xyn,xy->n", chi_k_pp, w)
    return chi_ph, chi_pp


def gamma_update(F, phi):
    """G_r = F_r - G_r after the parquet step, SRC/loop.f90:219-222."""
    return [f - p for f, p in zip(F, phi)]
