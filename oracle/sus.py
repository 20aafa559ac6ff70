"""Oracle: the susceptibilities of calc_sus.

Test infrastructure only: the checker for the GPU path, never shipped or timed.

Restates
* fill_FF_sdwave                  SRC/parquet_formfactors.f90:792-815
  (the four external form factors 1, cos px - cos py, sin px, sin px sin py)
* calc_GG_ph / calc_GG_pp         SRC/parquet_sus.f90:528-643
  (GG(L, q, ff) = sum_k FF_sdwave(k, ff) FF(k, l) G(k, nu) G(k + q, nu + w - 1) / Ng^2,
  resp. G(Ng + 1 - g; q - k, Nf - nu + w), over the enlarged box of f_range)
* calc_sus_ph / calc_sus_pp       SRC/parquet_sus.f90:45-362
  (sus_r(q, ff) = sum_{L1, L2 in the box} F_r(L1, L2, q) GG(L1) GG(L2) / (Nx^2 Ny^2 beta^2)
  + over the pairs not both in the box: (bare_r / FF11^2 on the l = 1 block, bare =
  U, -U, 2U, 0 for d, m, s, t) + G_r(pickNu(l10), pickNu(l20), q), with G_r the
  reducible vertex Phi the loop holds at that point, SRC/loop.f90:193-196)

Pinned by (tests/test_oracle_sus.py):
* FF_sdwave against the form factors through the reference's own check
  FF_sdwave_check (SRC/parquet_check.f90:274-281);
* GG(ff = 1) equals the SDE's auxiliary bubble chi_aux (oracle/sde.py, pinned by
  the second-order diagram) on the enlarged box with coarse graining, and
  GG(l = 3, ff = 2) = Nx chi(3, 3), GG(l = 4, ff = 3) = Nx / sqrt(2) chi(4, 4) with
  chi from calc_chi (oracle/bse.py, pinned by brute force) at Nl = 5;
* the contraction by F = c I (the sum of GG^2), by the bare term alone at
  f_range = 1 (the difference of the squared sums over the enlarged box and the box,
  with the sign of each channel), and by a single Phi entry at nu = 1 (which the
  pickNu clamp repeats over every frequency below the box).
"""
from __future__ import annotations

import math

import numpy as np

from .bse import ek_grain, green_coarse
from .lattice import Lattice
from .sde import faddb, minusf


def ff_sdwave(nx: int) -> np.ndarray:
    """FF_sdwave(kx, ky, 1..4) as (Nx, Nx, 4)."""
    out = np.zeros((nx, nx, 4))
    for ky in range(1, nx + 1):
        for kx in range(1, nx + 1):
            px = 2.0 / nx * math.pi * (kx - 1)
            py = 2.0 / nx * math.pi * (ky - 1)
            out[kx - 1, ky - 1] = (1.0, math.cos(px) - math.cos(py), math.sin(px), math.sin(px) * math.sin(py))
    return out


def gg(lat: Lattice, sigma, sigma_h, mu, beta, ngrain=1, f_range=0):
    """GG_ph, GG_pp as (ext Nl, Nq, 4), index (l-1) ext + f_range Nf + nu - 1."""
    nx, ny, nf, nl = lat.nx, lat.ny, lat.nf, lat.nl
    ext = (2 * f_range + 1) * nf
    ekg = ek_grain(lat, ngrain)
    fs = ff_sdwave(nx)
    ph = np.zeros((ext * nl, lat.nq, 4), dtype=np.complex128)
    pp = np.zeros((ext * nl, lat.nq, 4), dtype=np.complex128)
    for iq in range(lat.nq):
        qx, qy, qw = int(lat.qx[iq]), int(lat.qy[iq]), int(lat.qw[iq])
        for l in range(1, nl + 1):
            for l0 in range(-f_range * nf + 1, (f_range + 1) * nf + 1):
                idx = (l - 1) * ext + f_range * nf + l0 - 1
                for ky in range(1, ny + 1):
                    for kx in range(1, nx + 1):
                        kqx, kqy, kqw = faddb(kx, ky, l0, qx, qy, qw, nx, ny)
                        mkx, mky, mkw = minusf(kx, ky, l0, nx, ny, nf)
                        qkx, qky, qkw = faddb(mkx, mky, mkw, qx, qy, qw, nx, ny)
                        sph = spp = 0j
                        for gy in range(ngrain):
                            for gx in range(ngrain):
                                gk = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kx, ky, l0, sigma_h)
                                gkq = green_coarse(lat, ekg, sigma, mu, beta, gx, gy, kqx, kqy, kqw, sigma_h)
                                gqk = green_coarse(lat, ekg, sigma, mu, beta, ngrain - gx - 1, ngrain - gy - 1,
                                                  qkx, qky, qkw, sigma_h)
                                sph += gk * gkq / (ngrain * ngrain)
                                spp += gk * gqk / (ngrain * ngrain)
                        w = fs[kx - 1, ky - 1, :] * lat.ff[kx - 1, ky - 1, l - 1]
                        ph[idx, iq, :] += w * sph
                        pp[idx, iq, :] += w * spp
    return ph, pp


def pick_nu(x, nf, f_range):
    """pickNu of fillpickNu_range (SRC/parquet_sus.f90:708-731) at argument x + f_range Nf."""
    return min(max(x, 1), nf)


def calc_sus(lat: Lattice, fv, phi, gg_ph, gg_pp, U, beta, f_range=0):
    """sus_r (Nq, 4) for r = d, m, s, t (the IBZ arrays sus_X_IBZ(Nred, 4))."""
    nf, nl = lat.nf, lat.nl
    ext = (2 * f_range + 1) * nf
    norm = 1.0 / (lat.nx * lat.nx * lat.ny * lat.ny * beta * beta)
    ff11 = lat.ff[0, 0, 0]
    bare = (U, -U, 2.0 * U, 0.0)
    out = []
    for r in range(4):
        g = gg_ph if r < 2 else gg_pp
        s = np.zeros((lat.nq, 4), dtype=np.complex128)
        for iq in range(lat.nq):
            # the box: F
            for l2 in range(nl):
                for n2 in range(1, nf + 1):
                    i2 = l2 * ext + f_range * nf + n2 - 1
                    for l1 in range(nl):
                        for n1 in range(1, nf + 1):
                            i1 = l1 * ext + f_range * nf + n1 - 1
                            s[iq] += fv[r][l1 * nf + n1 - 1, l2 * nf + n2 - 1, iq] * g[i1, iq] * g[i2, iq] * norm
            # the rest of the enlarged box: bare vertex on the l = 1 block and Phi at pickNu
            for l2 in range(nl):
                for n2 in range(-f_range * nf + 1, (f_range + 1) * nf + 1):
                    i2 = l2 * ext + f_range * nf + n2 - 1
                    for l1 in range(nl):
                        for n1 in range(-f_range * nf + 1, (f_range + 1) * nf + 1):
                            if 0 < n1 < nf + 1 and 0 < n2 < nf + 1:
                                continue
                            i1 = l1 * ext + f_range * nf + n1 - 1
                            gg2 = g[i1, iq] * g[i2, iq] * norm
                            if l1 == 0 and l2 == 0:
                                s[iq] += bare[r] / (ff11 * ff11) * gg2
                            p1, p2 = pick_nu(n1, nf, f_range), pick_nu(n2, nf, f_range)
                            s[iq] += phi[r][l1 * nf + p1 - 1, l2 * nf + p2 - 1, iq] * gg2
        out.append(s)
    return out
