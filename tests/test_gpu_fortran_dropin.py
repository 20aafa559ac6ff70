"""The Fortran drop-in (tups_amd/fortran) driven in the reference's call order.

tups_fortran_driver runs BSE -> calculate_expSqR/expqSR -> calc_PhiR ->
calc_M_* -> calculate_PAE -> G = F - G (SRC/loop.f90:78-139, 219-222) through
ISO_C_BINDING, then BSE on the resident vertices (the drop-in skips their
upload) and BSE after the host halves F and declares it; every array it hands
back is compared with the oracle.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import bse as B
from oracle import parquet as P
from oracle.lattice import Lattice
from tests.helpers import crandn, random_chi_compact, random_vertices, rel_fro

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tups_amd" / "tups_fortran_driver"


@pytest.mark.parametrize("nx,nf,nl,f", [(4, 16, 1, 0.5), (4, 4, 5, 0.3)])
def test_fortran_dropin_iteration(tmp_path, nx, nf, nl, f):
    assert DRIVER.exists(), "build the Fortran driver with python -m tups_amd.build --fortran"
    print(f"fortran drop-in nx={nx} nf={nf} nl={nl}", file=sys.stderr, flush=True)
    beta = 2.0
    lat = Lattice(nx, nf, nl)
    rng = np.random.default_rng(17)
    gam = random_vertices(rng, lat.nz, lat.nq)
    fv = random_vertices(rng, lat.nz, lat.nq)
    chi = [B.chi_compact_to_dense(lat, random_chi_compact(rng, nl, nf, lat.nq)) for _ in range(2)]
    lam = [crandn(rng, (nf, nf, lat.nw)) for _ in range(4)]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array([nx, nf, nl, lat.nr], dtype=np.int32).tofile(fh)
        np.array([beta, f], dtype=np.float64).tofile(fh)
        for a in gam + fv + chi + lam:
            np.asfortranarray(a).ravel(order="F").tofile(fh)
    r = subprocess.run([str(DRIVER), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype=np.complex128)
    sizes = [lat.nz ** 2 * lat.nq] * 4 + [lat.nz ** 2 * lat.nw * lat.nr] * 4 \
        + [lat.nl ** 4 * lat.nq * lat.nr] * 4 + [lat.nz ** 2 * lat.nq] * 16
    assert out.size == sum(sizes)
    parts, o = [], 0
    for s in sizes:
        parts.append(out[o:o + s])
        o += s

    phi = B.bse_ladder(lat, gam, fv, chi, beta, f)
    phir = [P.calc_phir(lat, p) for p in phi]
    mt = P.calc_m_tensors(lat)
    f_ref = P.calculate_pae(lat, phi, lam, phir, mt)
    g_ref = B.gamma_update(f_ref, phi)
    # second BSE on the resident G = F - Phi and F; third with the host's F / 2
    phi2 = B.bse_ladder(lat, g_ref, f_ref, chi, beta, f)
    phi3 = B.bse_ladder(lat, phi2, [0.5 * x for x in f_ref], chi, beta, f)
    for r_ in range(4):
        assert rel_fro(parts[r_], phi[r_].ravel(order="F")) < 1e-12
        assert rel_fro(parts[4 + r_], phir[r_].ravel(order="F")) < 1e-12
        assert rel_fro(parts[12 + r_], f_ref[r_].ravel(order="F")) < 1e-12
        assert rel_fro(parts[16 + r_], g_ref[r_].ravel(order="F")) < 1e-12
        assert rel_fro(parts[20 + r_], phi2[r_].ravel(order="F")) < 1e-12
        assert rel_fro(parts[24 + r_], phi3[r_].ravel(order="F")) < 1e-12
    # M in the reference's per-q layout M(l1,l2,l3,l4,q,R)
    for i, name in enumerate(P.TARGETS):
        per_q = mt[name][:, :, :, :, np.arange(lat.nq) // lat.nw, :]
        assert rel_fro(parts[8 + i], per_q.ravel(order="F")) < 1e-12
