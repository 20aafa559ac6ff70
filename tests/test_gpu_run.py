"""A whole run from the reference's parameter file (tups_amd/tups_run, Fortran, on the
device) against the oracle's chain from the same starting point.

The config-1 parameters (tests/golden/prms_config1: Nx=4, Nf=16, Nl=1, beta=2,
U=2, nParticle=1, use_U2) with ite_max raised to 2: Initialization's Sigma = 0,
mu = 0, G = F = 0 and the parquet Lambda (SRC/parquet_util.f90:170-205,296-314,
359-367), then two iterations of SRC/loop.f90 (mu bisection, calc_chi, BSE,
calc_PhiR, calculate_PAE, SDE + second-order term, G = F - G).  The Sigma file the
program writes (the Sigma.init record layout) must equal the oracle's Sigma to
1e-12 relative Frobenius error.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def _oracle_run(nx, nf, beta, U, n_target, iterations):
    from oracle import bse as B
    from oracle import gtau as T
    from oracle import parquet as P
    from oracle import sde as S
    from oracle.lattice import Lattice

    lat = Lattice(nx, nf, 1)
    zero = np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128, order="F")
    gam = [zero.copy() for _ in range(4)]
    fv = [zero.copy() for _ in range(4)]
    lam = [np.full((nf, nf, lat.nw), v, dtype=np.complex128, order="F") for v in (U, -U, 2 * U, 0.0)]
    mt = P.calc_m_tensors(lat)
    sigma = np.zeros(nx * nx * nf, dtype=np.complex128)
    mu = 0.0
    mus = []
    for _ in range(iterations):
        mu, grt, _ = T.bisect_mu(lat, sigma, mu, beta, U, n_target)
        mus.append(mu)
        sh = sigma.reshape(nx, nx, nf)[:, :, nf - 1].copy()
        ph = np.zeros_like(zero)
        pp = np.zeros_like(zero)
        for q in range(lat.nq):
            ph[:, :, q], pp[:, :, q] = B.calc_chi(lat, int(lat.qx[q]), int(lat.qy[q]), int(lat.qw[q]), sigma, mu,
                                                  beta, sigma_h=sh)
        phi = B.bse_ladder(lat, gam, fv, (ph, pp), beta, 0.5)
        phir = [P.calc_phir(lat, p) for p in phi]
        fv = P.calculate_pae(lat, phi, lam, phir, mt)
        sigma = (S.calc_sde(lat, phi, (lam[0], lam[1]), sigma, sh, mu, beta, U, use_u2=True)
                 + T.order_2nd(lat, grt, U, beta))
        gam = B.gamma_update(fv, phi)
    return sigma, mus


def test_run_from_parameter_file_matches_oracle(tmp_path):
    exe = ROOT / "tups_amd" / "tups_run"
    assert exe.exists(), "build the Fortran programs first (__graft_entry__.build())"
    rec = (ROOT / "tests" / "golden" / "prms_config1").read_text().splitlines()
    rec[11] = "  2        !maximum number of iterations"
    prm = tmp_path / "prms"
    prm.write_text("\n".join(rec) + "\n")
    out = tmp_path / "sigma.txt"
    print("tups_run on the config-1 parameters, 2 iterations", flush=True, file=sys.stderr)
    r = subprocess.run([str(exe), str(prm), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True, file=sys.stderr)
    assert r.returncode == 0
    lines = [ln for ln in r.stdout.splitlines() if ln.strip().startswith("iteration")]
    assert len(lines) == 2
    data = np.loadtxt(out)
    nx, nf = 4, 16
    assert data.shape == (nx * nx * nf, 5)
    idx = ((data[:, 0].astype(int) - 1) * nx + data[:, 1].astype(int) - 1) * nf + data[:, 2].astype(int) - 1
    assert np.array_equal(idx, np.arange(nx * nx * nf))
    got = data[:, 3] + 1j * data[:, 4]
    exp, mus = _oracle_run(nx, nf, 2.0, 2.0, 1.0, 2)
    err = rel_fro(got, exp)
    print(f"Sigma after 2 iterations: rel err {err:.2e}; oracle mu {mus}", flush=True, file=sys.stderr)
    assert err < 1e-12
    assert f"{mus[-1]:14.10f}".strip() in lines[-1]
