"""Synthetic inputs of the batched-BSE workload (no reference counterpart).

The reference builds Gamma, F and chi inside its iteration; a throughput
measurement needs them at full size without running the whole solver.  This
module produces them the way the batched-BSE workload is specified:

* Gamma_r = U*I + X_r, X_r complex Gaussian with Re/Im iid N(0, sigma^2/2),
  sigma^2 = (0.3 U)^2 / D, generated on the device (``tups_synth_vertex``);
* F_r = Gamma_r (the reference's first iteration, SRC/preprocessing.f90:100-116),
  held in its own buffer;
* chi: the bare tight-binding bubble at mu = 0 (half filling),
  chi_ph(l,nu; Q,w) = sum_k G0(k,nu) G0(k+Q, nu+w) f_l(k)^2 and
  chi_pp(l,nu; Q,w) = sum_k G0(k,nu) G0(Q-k, w-nu) f_l(k)^2, diagonal in l,
  scaled per q so that |kappa_r| (1 + 2*0.3) U max|chi| = 0.85 — the norm bound
  of the spectral radius of kappa_r Gamma_r chi (||X||_2 ~ 2 sigma sqrt(D)).
  kappa = alpha for d,m and alpha/2 for s,t (the ladder kernels c_r = 1/2, 1/4).
The q slots are the 64 momenta of the 8x8 zone times 16 bosonic frequencies.
"""
from __future__ import annotations

import math

import numpy as np

from .runtime import CHI, F, GAMMA, TupsContext, build_tables


def bare_bubbles(nx: int, nf: int, nl: int, beta: float, qs):
    """Diagonal-in-l bubbles for q slots qs = [(qx0, qy0, w)] (0-based momenta, 1-based w)."""
    ff = build_tables(nx, nf, nl)[0]  # (nx, ny, nl)
    k = 2 * math.pi * np.arange(nx) / nx
    eps = -2 * np.cos(k)[:, None] - 2 * np.cos(k)[None, :]
    nu = np.arange(1, nf + 1)
    z = lambda n: 1j * math.pi / beta * (2 * (n - nf // 2 - 1) + 1)
    g0 = lambda e, n: 1.0 / (z(n)[None, None, :] - e[:, :, None])
    w2 = ff * ff  # (nx, ny, nl)
    ph = np.zeros((nl, nf, len(qs)), complex)
    pp = np.zeros((nl, nf, len(qs)), complex)
    gk = g0(eps, nu)
    for i, (qx, qy, w) in enumerate(qs):
        e_kq = np.roll(np.roll(eps, -qx, axis=0), -qy, axis=1)
        idx = (qx - np.arange(nx)) % nx
        idy = (qy - np.arange(nx)) % nx
        e_qk = eps[idx][:, idy]
        a = gk * g0(e_kq, nu + w - 1)
        b = gk * g0(e_qk, nf + w - nu)
        ph[:, :, i] = np.einsum("xyn,xyl->ln", a, w2)
        pp[:, :, i] = np.einsum("xyn,xyl->ln", b, w2)
    return ph, pp


def batched_bse_inputs(ctx: TupsContext, U: float = 2.0, seed: int = 0, nmom: int = 64, nomega: int = 16):
    """Fill a context with the batched-BSE workload (see module docstring)."""
    nx, nf, nl, nq, D = ctx.nx, ctx.nf, ctx.nl, ctx.nq, ctx.nz
    sigma = 0.3 * U / math.sqrt(D)
    for r in range(4):
        ctx.synth_vertex(GAMMA, r, seed, U, sigma)
        ctx.copy(F, GAMMA, r)
    qs = [((i // nomega) % nmom // nx, (i // nomega) % nmom % nx, i % nomega + 1) for i in range(nq)]
    ph, pp = bare_bubbles(nx, nf, nl, ctx.beta, qs)
    bound = (1.0 + 2 * 0.3) * U
    for p, (arr, kap) in enumerate(((ph, ctx.alpha), (pp, ctx.alpha / 2))):
        scale = 0.85 / (kap * bound * np.abs(arr).max(axis=(0, 1)))
        comp = np.zeros((nl, nl, nf, nq), complex, order="F")
        for l in range(nl):
            comp[l, l] = arr[l] * scale[None, :]
        ctx.upload(CHI, p, comp)
    ctx.sync()
    return qs
