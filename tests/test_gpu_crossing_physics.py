"""GPU crossing against the momentum/frequency-space parquet relations, and at the
benchmarked shapes against the oracle.

* With NR = every IBZ shell the truncated-unity crossing is exact, so the HIP
  path (calc_PhiR -> M tensors -> calculate_PAE through the C ABI) must equal
  the SU(2) parquet crossing evaluated directly on the k-space vertex
  (tests/parquet_bruteforce.py) — a check that does not go through the
  oracle's restatement of SRC/parquet_EQContribs.f90:44-524 at all.
* At the shapes bench.py times (Nx=16, Nf=64, Nl=1, Nq=1440 and Nx=8, Nf=64,
  Nl=5, Nq=480) every q is computed on the GPU and a sample of target q is
  compared with oracle.parquet.calculate_pae (SRC/parquet_EQOuter.f90:16-118).

Bar: relative Frobenius error <= 1e-12 per channel.
"""
import sys

import numpy as np
import pytest

from oracle import parquet as P
from oracle.lattice import Lattice
from tests.helpers import crandn, rel_fro
from tests.parquet_bruteforce import crossing_kspace, symmetric_phi

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 1), (4, 6, 5), (4, 4, 9), (6, 4, 5)])
def test_gpu_crossing_equals_kspace_relations(nx, nf, nl):
    from tups_amd import F, GAMMA, LAMBDA, TupsContext

    lat = Lattice(nx, nf, nl, nr=Lattice(nx, nf, nl).nibz)
    _say(f"crossing vs k-space relations nx={nx} nf={nf} nl={nl} nr={lat.nr}")
    phi = symmetric_phi(lat, np.random.default_rng(50 + nx + nf + nl))
    exp = crossing_kspace(lat, phi)
    zero = np.zeros((nf, nf, lat.nw), complex, order="F")
    with TupsContext(nx=nx, nf=nf, nl=nl, nr=lat.nr) as c:
        for r in range(4):
            c.upload(GAMMA, r, phi[r])
            c.upload(LAMBDA, r, zero)
        c.calc_PhiR(0)
        c.calc_M()
        c.calculate_PAE()
        got = [c.download(F, r) - phi[r] for r in range(4)]
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: {err:.2e}")
        assert err < TOL


def _sample_q(lat, n, seed):
    """A spread of target q: the first and last slot, both zone-boundary frequencies and random ones."""
    rng = np.random.default_rng(seed)
    qs = {0, lat.nq - 1, lat.nw - 1, lat.nq - lat.nw}
    qs.update(int(x) for x in rng.choice(lat.nq, size=n, replace=False))
    return sorted(qs)


@pytest.mark.parametrize("nx,nf,nl,nsample", [(16, 64, 1, 8), (8, 64, 5, 4)])
def test_pae_parity_at_bench_shape(nx, nf, nl, nsample):
    from tups_amd import F, GAMMA, LAMBDA, TupsContext
    from tups_amd.runtime import K_PAE

    lat = Lattice(nx, nf, nl)
    qs = _sample_q(lat, nsample, nx + nl)
    _say(f"pae parity at bench shape nx={nx} nf={nf} nl={nl} nq={lat.nq}, sampled q {qs}")
    rng = np.random.default_rng(71)
    phi = [crandn(rng, (lat.nz, lat.nz, lat.nq)) for _ in range(4)]
    lam = [crandn(rng, (nf, nf, lat.nw)) for _ in range(4)]
    phir = [P.calc_phir(lat, p) for p in phi]
    exp = P.calculate_pae(lat, phi, lam, phir, P.calc_m_tensors(lat), qs=qs)
    with TupsContext(nx=nx, nf=nf, nl=nl) as c:
        for r in range(4):
            c.upload(GAMMA, r, phi[r])
            c.upload(LAMBDA, r, lam[r])
        c.calc_PhiR(0)
        c.calc_M()
        variants = c.variants(K_PAE) if nl == 1 else []
        for v, name in variants + [(None, "autotuned")]:
            if v is None:
                if nl == 1:
                    c.autotune(K_PAE)
            else:
                c.set_variant(v, K_PAE)
            _say(f"  variant {name}")
            c.calculate_PAE()
            for r in range(4):
                got = c.download(F, r)
                err = rel_fro(got[:, :, qs], exp[r][:, :, qs])
                assert err < TOL, (name, r, err)
