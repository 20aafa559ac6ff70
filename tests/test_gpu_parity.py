"""GPU parity: every hot-path routine through the C ABI against the oracle.

Bar: relative Frobenius error <= 1e-12 per channel for the
FP64 path; integer index maps enter through the crossing and are exact.
"""
import sys

import numpy as np
import pytest

from oracle import bse as B
from oracle import parquet as P
from oracle.lattice import Lattice
from tests.helpers import crandn, random_chi_compact, random_vertices, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _ctx(**kw):
    from tups_amd import TupsContext

    return TupsContext(**kw)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize(
    "nx,nf,nl,nq,f",
    [(4, 16, 1, 0, 0.5),      # config 1 shape (Nz=16, Nq=48)
     (4, 8, 5, 0, 0.5),       # Nz=40: partial tile
     (4, 8, 9, 0, 0.3),       # Nz=72: two tiles, masked
     (8, 64, 1, 6, 1.0),      # Nz=64: exact tile
     (8, 64, 5, 3, 0.5),      # Nz=320: the batched-BSE shape
     (16, 64, 9, 2, 0.5)],    # Nz=576: the large single-GPU shape
)
def test_bse_ladder_parity(nx, nf, nl, nq, f):
    from tups_amd import CHI, F, GAMMA

    _say(f"bse parity nx={nx} nf={nf} nl={nl} nq={nq}")
    rng = np.random.default_rng(nx * 1000 + nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=f) as c:
        lat = Lattice(nx, nf, nl)
        gam = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        fv = random_vertices(rng, c.nz, c.nq, diag=2.0, scale=0.3)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        for r in range(4):
            c.upload(GAMMA, r, gam[r])
            c.upload(F, r, fv[r])
        for p in range(2):
            c.upload(CHI, p, chic[p])
        c.BSE(0)
        got = [c.download(GAMMA, r) for r in range(4)]
        # F untouched
        assert np.array_equal(c.download(F, 0), fv[0])
    exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, f)
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: rel err {err:.2e}")
        assert err < TOL


def test_chi_dense_upload_packs_blocks():
    from tups_amd import CHI

    rng = np.random.default_rng(4)
    with _ctx(nx=4, nf=8, nl=5) as c:
        lat = Lattice(4, 8, 5)
        chic = random_chi_compact(rng, 5, 8, c.nq)
        c.upload_chi_dense(1, B.chi_compact_to_dense(lat, chic))
        assert np.array_equal(c.download(CHI, 1), chic)


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 1), (4, 4, 5), (6, 4, 9), (8, 8, 5)])
def test_phir_parity(nx, nf, nl):
    from tups_amd import GAMMA, PHIR

    _say(f"phir parity nx={nx} nf={nf} nl={nl}")
    rng = np.random.default_rng(21)
    lat = Lattice(nx, nf, nl)
    with _ctx(nx=nx, nf=nf, nl=nl) as c:
        phi = [crandn(rng, (c.nz, c.nz, c.nq)) for _ in range(4)]
        for r in range(4):
            c.upload(GAMMA, r, phi[r])
        c.calc_PhiR(0)
        got = [c.download(PHIR, r) for r in range(4)]
    for r in range(4):
        assert rel_fro(got[r], P.calc_phir(lat, phi[r])) < TOL


@pytest.mark.parametrize("nx,nl,nr", [(4, 1, 1), (4, 5, 4), (4, 9, 6), (8, 5, 4), (8, 9, 6), (16, 9, 6)])
def test_m_tensor_parity(nx, nl, nr):
    from tups_amd import M, M_NAMES

    _say(f"M parity nx={nx} nl={nl} nr={nr}")
    lat = Lattice(nx, 4, nl, nr=nr)
    with _ctx(nx=nx, nf=4, nl=nl, nr=nr) as c:
        c.calc_M()
        got = [c.download(M, i) for i in range(4)]
    ref = P.calc_m_tensors(lat)
    for i, name in enumerate(M_NAMES):
        err = rel_fro(got[i], ref[name])
        _say(f"  {name}: {err:.2e}")
        assert err < TOL


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 1), (4, 4, 5), (4, 4, 9), (6, 8, 1),
                                      (4, 8, 5),    # Nf^2 = 64: one exact column tile of the GEMM path
                                      (4, 12, 5),   # Nf^2 = 144: partial column tile
                                      (6, 4, 9)])   # Nl^2 NIBZ = 810 rows: many row tiles
def test_pae_parity(nx, nf, nl):
    from tups_amd import F, GAMMA, LAMBDA, PHIR

    _say(f"pae parity nx={nx} nf={nf} nl={nl}")
    rng = np.random.default_rng(33)
    lat = Lattice(nx, nf, nl)
    mt = P.calc_m_tensors(lat)
    lam = [crandn(rng, (nf, nf, nf // 2)) for _ in range(4)]
    with _ctx(nx=nx, nf=nf, nl=nl) as c:
        phi = [crandn(rng, (c.nz, c.nz, c.nq)) for _ in range(4)]
        phir = [crandn(rng, (c.nz ** 2, c.nw, c.nr)) for _ in range(4)]
        for r in range(4):
            c.upload(GAMMA, r, phi[r])
            c.upload(PHIR, r, phir[r])
            c.upload(LAMBDA, r, lam[r])
        c.calc_M()
        c.calculate_PAE()
        got = [c.download(F, r) for r in range(4)]
    exp = P.calculate_pae(lat, phi, lam, phir, mt)
    for r in range(4):
        err = rel_fro(got[r], exp[r])
        _say(f"  channel {r}: {err:.2e}")
        assert err < TOL


def test_pae_requires_m_and_phir():
    from tups_amd import TupsError

    with _ctx(nx=4, nf=4, nl=1) as c:
        with pytest.raises(TupsError):
            c.calculate_PAE()


def test_full_iteration_config1():
    """One parquet iteration of config 1 (Nx=4, Nf=16, Nl=1, U=2, beta=2) on the GPU
    against the oracle chain BSE -> PhiR -> M -> PAE -> Gamma update."""
    from tups_amd import CHI, F, GAMMA, LAMBDA

    nx, nf, nl, U, beta, fd = 4, 16, 1, 2.0, 2.0, 0.5
    lat = Lattice(nx, nf, nl)
    # reference initial condition: F = G = Lambda*Nc on the l=1 block (SRC/preprocessing.f90:100-116)
    lam = P.lambda_pa(lat, U)
    g0 = []
    for r in range(4):
        v = np.zeros((lat.nz, lat.nz, lat.nq), complex, order="F")
        for q in range(lat.nq):
            v[:nf, :nf, q] = lam[r][:, :, lat.qw[q] - 1] * lat.nc
        g0.append(v)
    sigma = np.zeros(lat.nc * nf, complex)
    chi_ph = np.zeros((lat.nz, lat.nz, lat.nq), complex, order="F")
    chi_pp = np.zeros_like(chi_ph)
    for q in range(lat.nq):
        chi_ph[:, :, q], chi_pp[:, :, q] = B.calc_chi(lat, lat.qx[q], lat.qy[q], lat.qw[q], sigma, 0.0, beta, 1)
    with _ctx(nx=nx, nf=nf, nl=nl, beta=beta, f_damping=fd) as c:
        for r in range(4):
            c.upload(GAMMA, r, g0[r])
            c.upload(F, r, g0[r])
            c.upload(LAMBDA, r, lam[r])
        c.upload_chi_dense(0, chi_ph)
        c.upload_chi_dense(1, chi_pp)
        c.BSE(0)
        c.calc_PhiR(0)
        c.calc_M()
        c.calculate_PAE()
        f_gpu = [c.download(F, r) for r in range(4)]
        c.gamma_update()
        g_gpu = [c.download(GAMMA, r) for r in range(4)]
        c.check_finite(GAMMA, 0)
    phi = B.bse_ladder(lat, g0, g0, [chi_ph, chi_pp], beta, fd)
    phir = [P.calc_phir(lat, p) for p in phi]
    f_ref = P.calculate_pae(lat, phi, lam, phir, P.calc_m_tensors(lat))
    g_ref = B.gamma_update(f_ref, phi)
    for r in range(4):
        assert rel_fro(f_gpu[r], f_ref[r]) < TOL
        assert rel_fro(g_gpu[r], g_ref[r]) < TOL


def test_synth_vertex_statistics():
    """Device generator: diagonal offset exact, Re/Im ~ N(0, sigma^2/2)."""
    from tups_amd import GAMMA

    with _ctx(nx=8, nf=64, nl=5, nq=4) as c:
        c.synth_vertex(GAMMA, 1, seed=5, diag=2.0, sigma=0.6)
        g = c.download(GAMMA, 1)
    off = g.copy()
    for q in range(g.shape[2]):
        off[:, :, q] -= 2.0 * np.eye(g.shape[0])
    v = off.ravel()
    assert abs(np.var(v.real) - 0.18) < 0.01 and abs(np.var(v.imag) - 0.18) < 0.01
    assert abs(np.mean(v)) < 0.01


@pytest.mark.parametrize("nx,nf,nl,nq", [(8, 64, 1, 4), (4, 8, 1, 0), (8, 64, 5, 2), (4, 6, 5, 0),
                                         (16, 64, 9, 2), (4, 8, 9, 0)])
def test_every_bse_variant_and_autotune(nx, nf, nl, nq):
    """Each tile variant (frequency chunk, LDS double buffering) and the autotuner's
    pick reproduce the oracle; Nf=6 exercises a partial frequency chunk."""
    from tups_amd import CHI, F, GAMMA

    rng = np.random.default_rng(nf * 100 + nl)
    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=0.5) as c:
        lat = Lattice(nx, nf, nl)
        gam = random_vertices(rng, c.nz, c.nq)
        fv = random_vertices(rng, c.nz, c.nq)
        chic = [random_chi_compact(rng, nl, nf, c.nq, 3.0) for _ in range(2)]
        exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, 0.5)
        for p in range(2):
            c.upload(CHI, p, chic[p])
        variants = c.variants()
        assert variants
        for v, name in variants + [(-1, "default"), (None, "autotuned")]:
            _say(f"variant {name} nx={nx} nf={nf} nl={nl}")
            for r in range(4):
                c.upload(GAMMA, r, gam[r])
                c.upload(F, r, fv[r])
            if v is None:
                name, ms = c.autotune()
                assert ms > 0
            else:
                c.set_variant(v)
            c.BSE(0)
            for r in range(4):
                assert rel_fro(c.download(GAMMA, r), exp[r]) < TOL, (name, r)


@pytest.mark.parametrize("nx,nf,nr", [(4, 8, 1), (6, 8, 2), (8, 16, 1)])
def test_every_crossing_variant_nl1(nx, nf, nr):
    """Nl=1 crossing (PhiR re-index pass + streaming kernel): every column
    variant and the autotuner's pick reproduce the oracle, also for NR > 1."""
    from tups_amd import F, GAMMA, LAMBDA, PHIR
    from tups_amd.runtime import K_PAE

    rng = np.random.default_rng(nx + nf + nr)
    lat = Lattice(nx, nf, 1, nr=nr)
    mt = P.calc_m_tensors(lat)
    lam = [crandn(rng, (nf, nf, nf // 2)) for _ in range(4)]
    with _ctx(nx=nx, nf=nf, nl=1, nr=nr) as c:
        phi = [crandn(rng, (c.nz, c.nz, c.nq)) for _ in range(4)]
        phir = [crandn(rng, (c.nz ** 2, c.nw, c.nr)) for _ in range(4)]
        for r in range(4):
            c.upload(GAMMA, r, phi[r])
            c.upload(PHIR, r, phir[r])
            c.upload(LAMBDA, r, lam[r])
        c.calc_M()
        exp = P.calculate_pae(lat, phi, lam, phir, mt)
        variants = c.variants(K_PAE)
        assert variants
        for v, name in variants + [(None, "autotuned")]:
            _say(f"crossing variant {name} nx={nx} nf={nf} nr={nr}")
            if v is None:
                c.autotune(K_PAE)
            else:
                c.set_variant(v, K_PAE)
            c.calculate_PAE()
            for r in range(4):
                assert rel_fro(c.download(F, r), exp[r]) < TOL, (name, r)


def test_rejected_tables_leave_context_unchanged():
    """tups_set_tables with an invalid irrep list returns an error and changes no table,
    even when the same call also carries new form factors and signs."""
    import ctypes as C

    from tups_amd import TupsError

    with _ctx(nx=4, nf=4, nl=5) as c:
        lib = c.lib
        lat = Lattice(4, 4, 5)
        ff0 = np.zeros_like(lat.ff, order="F")
        signs0 = np.zeros(5, np.int32)
        reps0 = np.zeros(5, np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        assert lib.tups_get_tables(c._h, dp(ff0), None, ip(signs0), ip(reps0), None, None) == 0
        bad_ff = np.asfortranarray(ff0 * 2.0)
        bad_signs = -signs0
        bad_reps = reps0.copy()
        bad_reps[1] = 9
        assert lib.tups_set_tables(c._h, dp(bad_ff), None, ip(bad_signs), ip(bad_reps), None, None) != 0
        ff1 = np.zeros_like(ff0)
        signs1 = np.zeros(5, np.int32)
        reps1 = np.zeros(5, np.int32)
        assert lib.tups_get_tables(c._h, dp(ff1), None, ip(signs1), ip(reps1), None, None) == 0
        assert np.array_equal(ff1, ff0) and np.array_equal(signs1, signs0) and np.array_equal(reps1, reps0)
        assert np.abs(ff0 - lat.ff).max() < 1e-15
        # the context still works after the rejected call
        c.calc_M()
        with pytest.raises(TupsError):
            c.calculate_PAE()  # PhiR was never formed


@pytest.mark.parametrize("nx,nf,nl,nq", [(8, 64, 5, 2), (4, 16, 1, 0), (4, 8, 9, 0)])
def test_ladder_product_forms_under_cancellation(nx, nf, nl, nq):
    """Gamma, F and chi nearly real (imaginary parts 1e-7 of the real ones) and
    large: Im(Gamma chi F) is then a small difference for the three-multiplication
    forms (Im = P3 - P1 - P2).  Every variant must still meet the parity bar on the
    complex result (relative Frobenius <= 1e-12); the four-product forms (the
    reference's ZGEMM arithmetic) must also give the small imaginary part alone to
    1e-12, while the three-product forms' error on it is reported."""
    from tups_amd import CHI, F, GAMMA

    rng = np.random.default_rng(77 + nl)
    eps = 1e-7

    def near_real(shape, scale, diag):
        re = rng.standard_normal(shape) * scale
        im = rng.standard_normal(shape) * scale * eps
        v = np.asfortranarray(re + 1j * im)
        if diag:
            for q in range(shape[2]):
                v[:, :, q] += diag * np.eye(shape[0])
        return v

    with _ctx(nx=nx, nf=nf, nl=nl, nq=nq, beta=2.0, f_damping=0.5) as c:
        lat = Lattice(nx, nf, nl)
        gam = [near_real((c.nz, c.nz, c.nq), 300.0, 2000.0) for _ in range(4)]
        fv = [near_real((c.nz, c.nz, c.nq), 300.0, 2000.0) for _ in range(4)]
        chic = [np.asfortranarray(near_real((nl, nl, nf, c.nq), 3.0, 0.0)) for _ in range(2)]
        exp = B.bse_ladder(lat, gam, fv, [B.chi_compact_to_dense(lat, x) for x in chic], 2.0, 0.5)
        for p in range(2):
            c.upload(CHI, p, chic[p])
        for v, name in c.variants():
            for r in range(4):
                c.upload(GAMMA, r, gam[r])
                c.upload(F, r, fv[r])
            c.set_variant(v)
            c.BSE(0)
            worst_c = worst_im = 0.0
            for r in range(4):
                got = c.download(GAMMA, r)
                worst_c = max(worst_c, rel_fro(got, exp[r]))
                worst_im = max(worst_im, rel_fro(got.imag, exp[r].imag))
            _say(f"  {name}: complex rel err {worst_c:.2e}, imaginary part alone {worst_im:.2e}")
            assert worst_c < TOL, name
            if not name.endswith(("_3m", "_3r")):
                assert worst_im < TOL, name
