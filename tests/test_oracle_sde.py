"""Pins of the SDE oracle (oracle/sde.py) by results that do not use its code path.

* the second-order diagram by brute force over the whole zone;
* at Nl = 1 the IBZ star sum against a plain sum over every Q of the zone;
* linearity in the reducible vertices.
"""
import math

import numpy as np
import pytest

from oracle import sde as S
from oracle.lattice import Lattice, matsubara, sym_index, list_symmetries, NS
from oracle.parquet import lambda_pa
from tests.helpers import crandn, rel_fro

BETA, U = 2.0, 2.0


def _g0(lat, mu=0.0):
    """Sigma = 0, Sigma_H = 0: the bare Green's function (lattice- and mirror-symmetric)."""
    return np.zeros(lat.nx * lat.ny * lat.nf, dtype=np.complex128), np.zeros((lat.nx, lat.ny), dtype=np.complex128)


def _second_order_bruteforce(lat, sigma, sigma_h, mu):
    """-U^2/(beta^2 Nc^2) sum_{Q in BZ} sum_{|s| < Nw} G(k+Q, nu+s) sum_{k', nu' in box} G(k',nu') G(k'+Q, nu'+s)."""
    nx, ny, nf, nw = lat.nx, lat.ny, lat.nf, lat.nw
    e = S.ek(lat)
    G = lambda kx, ky, nu: S.green(lat, sigma, sigma_h, mu, BETA, kx, ky, nu, e=e)
    shifts = range(-(nw - 1), nw)
    X = {}
    for qx in range(1, nx + 1):
        for qy in range(1, ny + 1):
            for s in shifts:
                acc = 0j
                for kx in range(1, nx + 1):
                    for ky in range(1, ny + 1):
                        kqx, kqy = (kx + qx - 2) % nx + 1, (ky + qy - 2) % ny + 1
                        for nu in range(1, nf + 1):
                            acc += G(kx, ky, nu) * G(kqx, kqy, nu + s)
                X[qx, qy, s] = acc
    out = np.zeros(nx * ny * nf, dtype=np.complex128)
    for kx in range(1, nx + 1):
        for ky in range(1, ny + 1):
            for nu in range(1, nf + 1):
                acc = 0j
                for (qx, qy, s), x in X.items():
                    acc += G((kx + qx - 2) % nx + 1, (ky + qy - 2) % ny + 1, nu + s) * x
                out[((kx - 1) * ny + ky - 1) * nf + nu - 1] = -U * U / BETA ** 2 / lat.nc ** 2 * acc
    return out


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 1), (4, 6, 5)])
def test_second_order_diagram(nx, nf, nl):
    lat = Lattice(nx, nf, nl)
    sigma, sigma_h = _g0(lat)
    zero = [np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128) for _ in range(4)]
    lam = lambda_pa(lat, U)
    got = S.calc_sde(lat, zero, (lam[0], lam[1]), sigma, sigma_h, 0.0, BETA, U)
    exp = _second_order_bruteforce(lat, sigma, sigma_h, 0.0)
    assert rel_fro(got, exp) < 1e-12


def test_star_sum_equals_zone_sum_nl1():
    """Nl = 1: replace the IBZ star sum by a sum over every Q with the vertex of its representative."""
    lat = Lattice(4, 8, 1)
    rng = np.random.default_rng(3)
    sigma, sigma_h = _g0(lat)
    gv = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.3) for _ in range(4)]
    lam = lambda_pa(lat, U)
    got = S.calc_sde(lat, gv, (lam[0], lam[1]), sigma, sigma_h, 0.0, BETA, U)
    # zone sum: every (Qx, Qy) with the IBZ slot whose star contains it
    cph, cpp = S.chi_aux(lat, sigma, sigma_h, 0.0, BETA)
    lc = S.l_chi(lat, cph, lam[0], lam[1], U)
    pc = S.ph_chi(lat, cph, gv[0], gv[1])
    sc = S.pp_chi(lat, cpp, gv[2], gv[3])
    rep = {}
    for m, (qx, qy) in enumerate(lat.ibz_points):
        syms = list_symmetries(qx, qy, lat.nx_ibz)
        for s in range(1, NS + 1):
            if syms[s - 1]:
                rep[sym_index(s, qx, qy, lat.nx, lat.ny)] = m
    assert len(rep) == lat.nc
    e = S.ek(lat)
    nx, ny, nf = lat.nx, lat.ny, lat.nf
    G = lambda kx, ky, nu: S.green(lat, sigma, sigma_h, 0.0, BETA, kx, ky, nu, e=e)
    f = lat.ff[0, 0, 0]
    exp = np.zeros(nx * ny * nf, dtype=np.complex128)
    for kx in range(1, nx + 1):
        for ky in range(1, ny + 1):
            for nu in range(1, nf + 1):
                acc = 0j
                for (Qx, Qy), m in rep.items():
                    for w in range(1, lat.nw + 1):
                        iq = m * lat.nw + w - 1
                        kq = ((kx + Qx - 2) % nx + 1, (ky + Qy - 2) % ny + 1)
                        qmk = ((Qx - kx) % nx + 1, (Qy - ky) % ny + 1)
                        acc += G(*kq, nu + w - 1) * f * (lc[nu - 1, iq] + pc[nu - 1, iq])
                        acc += G(*qmk, nf - nu + w) * f * sc[nu - 1, iq]
                        if w > 1:
                            kmq = ((kx - Qx) % nx + 1, (ky - Qy) % ny + 1)
                            acc += G(*kmq, nu - w + 1) * f * np.conj(lc[nf - nu, iq] + pc[nf - nu, iq])
                            acc += np.conj(G(*kq, nu + w - 1)) * f * np.conj(sc[nf - nu, iq])
                exp[((kx - 1) * ny + ky - 1) * nf + nu - 1] = -U / BETA ** 2 / lat.nc ** 2 / 2.0 * acc
    assert rel_fro(got, exp) < 1e-12


def test_linear_in_vertices():
    lat = Lattice(4, 6, 5)
    rng = np.random.default_rng(7)
    sigma = crandn(rng, (lat.nx * lat.ny * lat.nf,), 0.2).ravel()
    sigma_h = crandn(rng, (lat.nx, lat.ny), 0.1)
    lam = lambda_pa(lat, U)
    A = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.3) for _ in range(4)]
    B = [crandn(rng, (lat.nz, lat.nz, lat.nq), 0.3) for _ in range(4)]
    zero = [np.zeros_like(A[0]) for _ in range(4)]
    run = lambda gv: S.calc_sde(lat, gv, (lam[0], lam[1]), sigma, sigma_h, 0.1, BETA, U, f_range=1)
    s0, sa, sb = run(zero), run(A), run(B)
    sab = run([2.0 * a - 0.5 * b for a, b in zip(A, B)])
    assert rel_fro(sab - s0, 2.0 * (sa - s0) - 0.5 * (sb - s0)) < 1e-12
