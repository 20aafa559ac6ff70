"""Oracle: G in imaginary time, the chemical-potential bisection and the SDE's
second-order term.

Test infrastructure only: the checker for the GPU path, never shipped or timed.

Restates
* calc_Gkw_Grt, the G(r, tau) part   SRC/parquet_chi.f90:113-198
  (G on the fine grid, Matsubara -> tau with the bare G subtracted and added
  back in closed form, forward 2D FFT, particle number -2 G(r=0, tau=beta),
  bisection of mu between mu -/+ U to |n - nParticle| <= 1e-6)
* calc_order_2nd                     SRC/parquet_SDE.f90:263-340
* the bare bubbles of calc_Gkw_Grt with cal_chi, SRC/parquet_chi.f90:216-430
  (Chi0_ph = FT of -Grt(r, tau) Grt(-r, beta - tau), Chi0_pp of -1/2 Grt^2, and the
  form-factor versions with FFG = fftf2d of G(k, tau) FF_sdwave(k, ff)^2 in place of
  the second Grt; bosonic nfourier at the IBZ momenta)
* FDfit, nfourier                    SRC/MATH_mod.f90:95-212
* fftf2d / fftb2d                    SRC/MATH_mod.f90:215-259 (FFTPACK zfftf /
  zfftb: forward e^{-i}, divided by n1 n2; backward e^{+i}, unscaled) -- here
  numpy.fft.fft2 / ifft2, an independent implementation of the same transforms.
Gkw_c itself (the fine-grid G) is the coarse-grained Green's function of
oracle/bse.py.

Pinned by (tests/test_oracle_gtau.py):
* FDfit against the derivatives of quartic polynomials (its 5-point one-sided
  formulas are exact to degree 4);
* nfourier against the closed-form integral of a cubic polynomial times
  e^{i w tau} (the spline with exact end slopes reproduces a cubic exactly);
* G(k, tau) at Sigma = 0 against -e^{(beta-tau)(e-mu)} / (1 + e^{beta(e-mu)}),
  and the particle number at mu = 0 against half filling (the lattice
  dispersion is particle-hole symmetric), the bisection returning mu = 0;
* the second-order term at Sigma = 0 against the box-free brute-force
  Matsubara sum of the second-order diagram, to the spline's accuracy (the
  tau grid has Nf points; tolerance in the test);
* the bare bubbles at Sigma = 0 against the Lindhard functions (ph and pp, with
  and without the form-factor weight), and form factor 1 against the plain
  bubbles exactly.
"""
from __future__ import annotations

import math

import numpy as np

from .bse import ek_grain
from .lattice import Lattice, matsubara


def fine_index(lat: Lattice, ngrain: int, kx: int, ky: int, gx: int, gy: int):
    """(idx1, idx2), 1-based, of grain (gx, gy) of momentum (kx, ky) on the fine grid
    (SRC/parquet_chi.f90:135-139)."""
    i1 = (kx - 1) * ngrain + (gx - (ngrain + 1) // 2) + 1
    if i1 < 1:
        i1 += ngrain * lat.nx
    i2 = (ky - 1) * ngrain + (gy - (ngrain + 1) // 2) + 1
    if i2 < 1:
        i2 += ngrain * lat.ny
    return i1, i2


def g_tau_k(lat: Lattice, sigma, mu, beta, ngrain=1, weight=None):
    """G(k_fine, tau) before the FFT: (Nx Ng, Ny Ng, Nf) real, SRC/parquet_chi.f90:124-170
    (the nu and tau loops as numpy vectors); weight (Nx, Ny) multiplies every grain of a
    coarse momentum (the FF_sdwave^2 of FFG, SRC/parquet_chi.f90:341-343)."""
    nf = lat.nf
    ekg = ek_grain(lat, ngrain)
    out = np.zeros((lat.nx * ngrain, lat.ny * ngrain, nf))
    tau = beta / (nf - 1) * np.arange(nf)
    nus = np.arange(1, nf + 1)
    z = np.array([matsubara(int(nu), nf, beta) for nu in nus])
    # e^{-i pi/beta (2 (nu - Nf/2) - 1) tau}: rows tau, columns nu
    ph = -math.pi / beta * (2.0 * (nus[None, :] - nf // 2) - 1.0) * tau[:, None]
    phase = np.cos(ph) + 1j * np.sin(ph)
    for ky in range(1, lat.ny + 1):
        for kx in range(1, lat.nx + 1):
            s = sigma[((kx - 1) * lat.ny + ky - 1) * nf:((kx - 1) * lat.ny + ky) * nf]
            for gy in range(1, ngrain + 1):
                for gx in range(1, ngrain + 1):
                    i1, i2 = fine_index(lat, ngrain, kx, ky, gx, gy)
                    e = ekg[gx - 1, gy - 1, kx - 1, ky - 1]
                    d = 1.0 / (z + mu - e - s) - 1.0 / (z + mu - e)
                    acc = (phase * d[None, :]).real.sum(axis=1) / beta
                    out[i1 - 1, i2 - 1, :] = acc - np.exp((beta - tau) * (e - mu)) / (1.0 + math.exp(beta * (e - mu)))
                    if weight is not None:
                        out[i1 - 1, i2 - 1, :] *= weight[kx - 1, ky - 1]
    return out


def fftf2d(a):
    """fftf2d per tau slice: forward transform divided by n1 n2."""
    return np.fft.fft2(a, axes=(0, 1)) / (a.shape[0] * a.shape[1])


def fftb2d(a):
    """fftb2d per tau slice: backward transform, unscaled."""
    return np.fft.ifft2(a, axes=(0, 1)) * (a.shape[0] * a.shape[1])


def grt(lat: Lattice, sigma, mu, beta, ngrain=1):
    """Grt(r, tau) = fftf2d of G(k, tau) (complex (Nx Ng, Ny Ng, Nf))."""
    return fftf2d(g_tau_k(lat, sigma, mu, beta, ngrain).astype(np.complex128))


def particle_number(grt_rt):
    """t1 = -2 Grt(1, 1, Nf) (SRC/parquet_chi.f90:182)."""
    return float((-2.0 * grt_rt[0, 0, -1]).real)


def bisect_mu(lat: Lattice, sigma, mu, beta, U, n_target, ngrain=1, tol=1e-6):
    """The bisection of SRC/parquet_chi.f90:113-198: returns (mu, Grt, steps)."""
    hi, lo = mu + U, mu - U
    t1 = 3.0
    steps = 0
    g = None
    while abs(t1 - n_target) > tol:
        mu = 0.5 * (hi + lo)
        g = grt(lat, sigma, mu, beta, ngrain)
        t1 = particle_number(g)
        steps += 1
        if t1 > n_target:
            hi = mu
        else:
            lo = mu
    return mu, g, steps


def fdfit(gt, dt):
    """FDfit: one-sided 5-point first derivatives at both ends (SRC/MATH_mod.f90:95-107)."""
    L = len(gt)
    fd1 = (-25.0 * gt[0] + 48.0 * gt[1] - 36.0 * gt[2] + 16.0 * gt[3] - 3.0 * gt[4]) / 12.0 / dt
    fd2 = (25.0 * gt[L - 1] - 48.0 * gt[L - 2] + 36.0 * gt[L - 3] - 16.0 * gt[L - 4] + 3.0 * gt[L - 5]) / 12.0 / dt
    return fd1, fd2


def nfourier(fermionic: bool, L: int, iwmax: int, fd1, fd2, rin, beta):
    """nfourier (SRC/MATH_mod.f90:110-212): cubic spline through rin(0..L) on the
    grid tau_j = j beta / L with end slopes fd1, fd2, integrated exactly against
    e^{i w tau}, w = (2i-1) pi/beta (fermionic) or 2(i-1) pi/beta."""
    delta = beta / L
    r = np.asarray(rin, dtype=float)
    q = np.zeros(L + 1)
    u = np.zeros(L + 1)
    xm = np.zeros(L + 1)
    if fd1 > 0.99e30:
        q[0] = 0.0
        u[0] = 0.0
    else:
        q[0] = -0.5
        u[0] = (3.0 / delta) * ((r[1] - r[0]) / delta - fd1)
    for k in range(1, L):
        p = q[k - 1] / 2.0 + 2.0
        q[k] = -1.0 / 2.0 / p
        u[k] = 3.0 / delta ** 2 * (r[k + 1] + r[k - 1] - 2.0 * r[k])
        u[k] = (u[k] - u[k - 1] / 2.0) / p
    if fd2 > 0.99e30:
        q[L] = 0.0
        u[L] = 0.0
    else:
        q[L] = 0.5
        u[L] = (3.0 / delta) * (fd2 - (r[L] - r[L - 1]) / delta)
    xm[L] = (u[L] - q[L] * u[L - 1]) / (q[L] * q[L - 1] + 1.0)
    for k in range(L - 1, -1, -1):
        xm[k] = q[k] * xm[k + 1] + u[k]
    a = r[:L].copy()
    c = xm[:L] / 2.0
    b = (r[1:L + 1] - r[:L]) / delta - (2.0 * xm[:L] + xm[1:L + 1]) * delta / 6.0
    d = (xm[1:L + 1] - xm[:L]) / (6.0 * delta)
    out = np.zeros(iwmax, dtype=np.complex128)
    for i in range(1, iwmax + 1):
        om = (2.0 * i - 1.0) * math.pi / beta if fermionic else 2.0 * (i - 1) * math.pi / beta
        acc = 0j
        for j in range(1, L + 1):
            explus = np.exp(1j * om * delta * j)
            ex = np.exp(1j * om * delta * (j - 1))
            aj, bj, cj, dj = a[j - 1], b[j - 1], c[j - 1], d[j - 1]
            if abs(om) < 1e-6:
                acc += aj * delta + 0.5 * bj * delta ** 2 + cj * delta ** 3 / 3.0 + dj * delta ** 4 / 4.0
            else:
                acc += explus * ((-6.0 * dj) / om ** 4 + (2.0j * cj + 6.0 * delta * 1j * dj) / om ** 3
                                 + (bj + 2.0 * delta * cj + 3.0 * delta ** 2 * dj) / om ** 2
                                 + (-1j * aj - delta * 1j * bj - delta ** 2 * 1j * cj - delta ** 3 * 1j * dj) / om)
                acc += ex * (6.0 * dj / om ** 4 - 2.0j * cj / om ** 3 - bj / om ** 2 + 1j * aj / om)
        out[i - 1] = acc
    return out


def order_2nd(lat: Lattice, grt_rt, U, beta, ngrain=1):
    """calc_order_2nd (SRC/parquet_SDE.f90:263-340): (Nx Ny Nf) complex."""
    nf = lat.nf
    n1, n2 = lat.nx * ngrain, lat.ny * ngrain
    ggg = np.zeros((n1, n2, nf), dtype=np.complex128)
    for it in range(nf):
        for ky in range(1, n2 + 1):
            kym = 1 if ky == 1 else n2 - ky + 2
            for kx in range(1, n1 + 1):
                kxm = 1 if kx == 1 else n1 - kx + 2
                ggg[kx - 1, ky - 1, it] = (grt_rt[kx - 1, ky - 1, it] ** 2
                                           * grt_rt[kxm - 1, kym - 1, nf - it - 1] * U * U)
    ggg = fftb2d(ggg)
    out = np.zeros(lat.nx * lat.ny * nf, dtype=np.complex128)
    for kx in range(1, lat.nx + 1):
        kxc = (kx - 1) * ngrain + 1
        for ky in range(1, lat.ny + 1):
            kyc = (ky - 1) * ngrain + 1
            data = ggg[kxc - 1, kyc - 1, :].real
            fd1, fd2 = fdfit(data, beta / (nf - 1))
            ret = nfourier(True, nf - 1, nf // 2, fd1, fd2, data, beta)
            for nu in range(1, nf // 2 + 1):
                ik = ((kx - 1) * lat.ny + ky - 1) * nf + nu - 1
                imk = ((kx - 1) * lat.ny + ky - 1) * nf + nf - nu
                out[ik] += np.conj(ret[nf // 2 - nu])
                out[imk] += ret[nf // 2 - nu]
    return out


def ffg(lat: Lattice, sigma, mu, beta, ngrain, ff):
    """FFG for external form factor ff (0-based): fftf2d of G(k, tau) FF_sdwave(k_coarse, ff)^2
    (SRC/parquet_chi.f90:302-351)."""
    from .sus import ff_sdwave
    w = ff_sdwave(lat.nx)[:, :, ff] ** 2
    return fftf2d(g_tau_k(lat, sigma, mu, beta, ngrain, weight=w).astype(np.complex128))


def _bubble(lat: Lattice, prod, beta, ngrain):
    """fftb2d of prod(r, tau), read at the IBZ momenta (i, j), j <= i, then FDfit + nfourier
    (bosonic) per momentum: (NIBZ Nf/2,) in the order ((i (i-1))/2 + j - 1) Nf/2 + k."""
    nf = lat.nf
    t = fftb2d(prod)
    nxi = (lat.nx + 2) // 2
    out = np.zeros(nxi * (nxi + 1) // 2 * (nf // 2), dtype=np.complex128)
    for i in range(1, nxi + 1):
        for j in range(1, i + 1):
            data = t[(i - 1) * ngrain, (j - 1) * ngrain, :].real
            fd1, fd2 = fdfit(data, beta / (nf - 1))
            ret = nfourier(False, nf - 1, nf // 2, fd1, fd2, data, beta)
            base = ((i * (i - 1)) // 2 + j - 1) * (nf // 2)
            out[base:base + nf // 2] = ret
    return out


def _mirror(a):
    """a(-r, beta - tau): index 1 -> 1, else n - idx + 2 on both axes, Nf - tau + 1 in time."""
    return np.roll(a[::-1, ::-1, ::-1], shift=(1, 1), axis=(0, 1))


def chi0(lat: Lattice, grt_rt, ffgs, beta, ngrain=1):
    """The bare bubbles of calc_Gkw_Grt with cal_chi (SRC/parquet_chi.f90:216-430):
    Chi0_ph, Chi0_pp (Nred,) and Chi0_ph_FF, Chi0_pp_FF (Nred, 4), ffgs the four FFG."""
    ph = _bubble(lat, -grt_rt * _mirror(grt_rt), beta, ngrain)
    pp = _bubble(lat, -0.5 * grt_rt ** 2, beta, ngrain)
    ph_ff = np.stack([_bubble(lat, -grt_rt * _mirror(f), beta, ngrain) for f in ffgs], axis=1)
    pp_ff = np.stack([_bubble(lat, -0.5 * grt_rt * f, beta, ngrain) for f in ffgs], axis=1)
    return ph, pp, ph_ff, pp_ff
