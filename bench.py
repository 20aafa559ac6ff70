#!/usr/bin/env python3
"""Batched Bethe-Salpeter ladder throughput on one MI355X.

Workload (the batched-BSE configuration): 8x8 lattice, Nl = 5 form factors,
Nf = 64 fermionic frequencies -> matrix dimension D = 320 per transfer q;
Nq = 64 momenta x 16 bosonic frequencies = 1024 q; all four channels
(d, m, s, t).  One step = one BSE(ite) over the whole batch (the reference's
SRC/parquet_BSE.f90:62-132 loop), inputs resident in HBM.  One q-point = the
update of all four channels at one q.

Two forms of the same step are timed: the split (precision) form, whose GEMM
runs as exact INT8 MFMA products of residues and whose error against the
FP64 path is measured per channel, and the plain FP64 MFMA form.  `value` is
the split form's rate; the FP64 form's rate and roofline sit beside it.

Prints one JSON line with the measured throughput, live rooflines of the
dominant kernels (HIP events on the library's own stream) and a bounded CPU
baseline (the numpy oracle restatement with host BLAS, rank 0, N = 1 only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 MFMA, AMD published
INT8_PEAK_TOPS = 5000.0  # MI355X dense INT8 MFMA, AMD published
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E, AMD published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nq", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="CPU baseline budget of the ladder sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the D=576, Nq=4096 (261 GB) share")
    ap.add_argument("--moduli", type=int, default=12, help="CRT moduli of the split form")
    ap.add_argument("--qbatch", type=int, default=0, help="q points per split pass (0: library default)")
    return ap.parse_args()


def profile_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 counter summary
    (FETCH_SIZE x 2 KiB + WRITE_SIZE x 1 KiB), or None when the summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def _ladder_rate(lat, budget_s):
    """q-points/s of the oracle ladder (4 ZGEMMs per channel, host BLAS) over about budget_s."""
    import numpy as np

    from oracle import bse as B

    rng = np.random.default_rng(1)
    D, nl, nf = lat.nz, lat.nl, lat.nf
    nq = 2
    mk = lambda: np.asfortranarray(rng.standard_normal((D, D, nq)) + 1j * rng.standard_normal((D, D, nq)))
    gam = [mk() for _ in range(4)]
    fv = [mk() for _ in range(4)]
    chi = [B.chi_compact_to_dense(lat, np.asfortranarray(rng.standard_normal((nl, nl, nf, nq)) + 0j))
           for _ in range(2)]
    B.bse_ladder(lat, gam, fv, chi, 2.0, 0.5)  # warm-up
    done, t0 = 0, time.perf_counter()
    while True:
        B.bse_ladder(lat, gam, fv, chi, 2.0, 0.5)
        done += nq
        el = time.perf_counter() - t0
        if el >= budget_s:
            return done, el


def _crossing_cpu(nx, nf, nl, ws_sample, q_sample):
    """Oracle calc_PhiR (whole call, or extrapolated from a sample of bosonic frequencies)
    and calculate_PAE per target q (a sample of q) at one shape, one host process."""
    import numpy as np

    from oracle import parquet as P
    from oracle.lattice import Lattice

    lat = Lattice(nx, nf, nl)
    # the work does not depend on the values: constant inputs of the right shape
    phi = [np.full((lat.nz, lat.nz, lat.nq), 0.5 + 0.25j * r, dtype=np.complex128, order="F") for r in range(4)]
    lam = [np.full((nf, nf, lat.nw), 0.1 * r, dtype=np.complex128, order="F") for r in range(4)]
    ws = list(range(min(ws_sample, lat.nw)))
    t0 = time.perf_counter()
    phir = [P.calc_phir(lat, p, ws=ws) for p in phi]
    t_phir = (time.perf_counter() - t0) * lat.nw / len(ws)
    mt = P.calc_m_tensors(lat)
    qs = list(range(q_sample))
    t0 = time.perf_counter()
    P.calculate_pae(lat, phi, lam, phir, mt, qs=qs)
    t_pae = time.perf_counter() - t0
    return {"shape": f"Nx={nx} Nf={nf} Nl={nl} NR={lat.nr} Nq={lat.nq}",
            "phir_ms_per_call": round(t_phir * 1e3, 1),
            "phir_sample": f"{len(ws)} of {lat.nw} bosonic frequencies, scaled to the whole call",
            "pae_q_per_s": round(q_sample / t_pae, 3), "pae_ms_per_call": round(t_pae / q_sample * lat.nq * 1e3, 1),
            "pae_sample": f"{q_sample} target q of {lat.nq}, scaled to the whole call"}


def cpu_baseline(nx, nf, nl, budget_s):
    """The oracle restatement on the host cores: the ladder on a bounded sample of q with
    every BLAS thread and with one, calc_PhiR and calculate_PAE at the crossing shapes, and
    the config-1 iteration.  One process (the reference's MPI split collapses to one rank)."""
    from oracle.lattice import Lattice

    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threadpool_limits = None
        threads = os.cpu_count() or 1
    lat = Lattice(nx, nf, nl)
    done, el = _ladder_rate(lat, budget_s)
    out = {"value": done / el, "unit": "q-points/s", "cores": int(threads), "mpi_ranks": 1, "kind": "port",
           "sample": f"{done} q-points of the D={lat.nz} ladder (4 channels x 4 ZGEMM + mixing each, "
                     f"oracle restatement of SRC/parquet_BSE.f90:62-132 on host BLAS), {el:.1f} s"}
    if threadpool_limits is not None:
        with threadpool_limits(limits=1, user_api="blas"):
            d1, e1 = _ladder_rate(lat, budget_s / 2)
        out["one_thread"] = {"value": d1 / e1, "unit": "q-points/s", "cores": 1,
                             "sample": f"{d1} q-points, {e1:.1f} s"}
    out["crossing_nl1"] = _crossing_cpu(16, 64, 1, 32, 2)
    out["crossing_nl5"] = _crossing_cpu(8, 64, 5, 1, 1)
    out["config1_iteration"] = config1_cpu()
    return out


def config1_inputs(lat_or_ctx_nq, nz, nf, nw, seed=7):
    """Vertices, Lambda (parquet approximation, U = 2) and Sigma = 0 of the 4x4 / Nf=16 / Nl=1 case."""
    import numpy as np

    rng = np.random.default_rng(seed)
    shape = (nz, nz, lat_or_ctx_nq)
    mk = lambda d: np.asfortranarray(0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
                                     + d * np.eye(nz)[:, :, None])
    gam = [mk(0.5) for _ in range(4)]
    fv = [mk(0.5) for _ in range(4)]
    U = 2.0
    lam = [np.full((nf, nf, nw), v, dtype=np.complex128, order="F") for v in (U, -U, 2.0 * U, 0.0)]
    return gam, fv, lam, np.zeros(4 * 4 * nf, dtype=np.complex128)


CONFIG1_WORKLOAD = ("one full iteration, Nx=4 Nf=16 Nl=1 Nq=48, beta=2, U=2, nParticle=1, use_U2: mu bisection "
                    "and G(r, tau), calc_chi, BSE, calc_PhiR, calculate_PAE, SDE + second-order term, Sigma "
                    "update, G = F - G")


def config1_cpu():
    """One full iteration of the 4x4 / Nf=16 / Nl=1 / beta=2 / U=2 / nParticle=1 / use_U2 case
    on the oracle: calc_Gkw_Grt (mu bisection, G(r, tau)) -> calc_chi (every q) -> BSE ->
    calc_PhiR -> calculate_PAE -> SDE + calc_order_2nd -> G = F - G (SRC/loop.f90:67-222;
    M tensors built beforehand, not timed)."""
    import numpy as np

    from oracle import bse as B
    from oracle import gtau as T
    from oracle import parquet as P
    from oracle import sde as S
    from oracle.lattice import Lattice

    lat = Lattice(4, 16, 1)
    gam, fv, lam, sigma = config1_inputs(lat.nq, lat.nz, lat.nf, lat.nw)
    mt = P.calc_m_tensors(lat)
    t0 = time.perf_counter()
    mu, grt, _ = T.bisect_mu(lat, sigma, 0.0, 2.0, 2.0, 1.0)
    sh = sigma.reshape(4, 4, lat.nf)[:, :, lat.nf - 1].copy()
    ph = np.zeros((lat.nz, lat.nz, lat.nq), dtype=np.complex128, order="F")
    pp = np.zeros_like(ph)
    for q in range(lat.nq):
        ph[:, :, q], pp[:, :, q] = B.calc_chi(lat, int(lat.qx[q]), int(lat.qy[q]), int(lat.qw[q]), sigma, mu, 2.0,
                                              sigma_h=sh)
    phi = B.bse_ladder(lat, gam, fv, (ph, pp), 2.0, 0.5)
    phir = [P.calc_phir(lat, p) for p in phi]
    F = P.calculate_pae(lat, phi, lam, phir, mt)
    S.calc_sde(lat, phi, (lam[0], lam[1]), sigma, sh, mu, 2.0, 2.0, use_u2=True) + T.order_2nd(lat, grt, 2.0, 2.0)
    B.gamma_update(F, phi)
    el = time.perf_counter() - t0
    return {"ms_per_iteration": round(el * 1e3, 1), "cores": 1, "kind": "port",
            "workload": CONFIG1_WORKLOAD + "; oracle restatement in numpy"}


def reference_prms_run():
    """tups_amd/tups_run on the parameter file the reference ships (tests/golden/prms_reference):
    Initialization's start, then up to ite_max = 100 device iterations with the mu search,
    the SDE with use_U2 and conv(); wall time of the whole program (a child process)."""
    import subprocess

    root = os.path.dirname(os.path.abspath(__file__))
    exe = os.path.join(root, "tups_amd", "tups_run")
    prm = os.path.join(root, "tests", "golden", "prms_reference")
    out = {"workload": "the reference's prms: Nx=16, Nf=64, Nl=1 (Nq=1440), beta=5, U=2, nParticle=1, use_U2, "
                       "ite_max=100; whole program incl. set-up"}
    if not os.path.exists(exe):
        out["skipped"] = "tups_run not built"
        return out
    t0 = time.perf_counter()
    r = subprocess.run([exe, prm], capture_output=True, text=True, timeout=300)
    wall = time.perf_counter() - t0
    its = [ln for ln in r.stdout.splitlines() if ln.strip().startswith("iteration")]
    out.update({"returncode": r.returncode, "wall_s": round(wall, 3), "iterations": len(its),
                "converged": "converged" in r.stdout, "last": its[-1].strip() if its else None})
    print(f"[bench] reference prms run: {len(its)} iterations in {wall:.2f} s", file=sys.stderr, flush=True)
    return out


def large_share(moduli):
    """Nx=16, Nl=9, Nf=64 (D=576), Nq=4096 = 256 momenta x 16 bosonic frequencies:
    12 vertex arrays of 21.7 GB resident.  One FP64 and one split step (after a
    warm-up each), and the split error on 16 q of the same inputs."""
    import numpy as np
    from tups_amd import BSE_LADDER_FP64, BSE_LADDER_SPLIT, GAMMA, TupsContext, TupsError
    from tups_amd.synth import batched_bse_inputs

    out = {"workload": "batched BSE ladder, Nx=16, D=576 (Nl=9 x Nf=64), Nq=4096 (256 momenta x 16 bosonic "
                       "freq), 4 channels; Gamma, F, Phi of every channel resident (12 x 21.7 GB)"}
    try:
        c = TupsContext(nx=16, nf=64, nl=9, nq=4096, beta=2.0, f_damping=0.5)
    except TupsError as e:
        out["skipped"] = str(e)
        return out
    try:
        batched_bse_inputs(c, seed=0, nmom=256, nomega=16)
        ids = {n: v for v, n in c.variants()}
        c.set_variant(ids["nl9_g4_3r"])
        c.set_split(moduli, 0)
        fl = 16.0 * c.nz ** 3 * 4 * c.nq
        res = {}
        for mode, key in ((BSE_LADDER_FP64, "fp64"), (BSE_LADDER_SPLIT, "split")):
            c.BSE(0, mode=mode)
            c.sync()
            c.record(0)
            c.BSE(0, mode=mode)
            c.record(1)
            c.sync()
            ms = c.elapsed_ms(0, 1)
            res[key] = {"q_points_per_s": round(c.nq / (ms / 1e3), 1), "ms_per_step": round(ms, 1),
                        "algorithmic_tflops": round(fl / (ms / 1e3) / 1e12, 2)}
            print(f"[bench] large share {key}: {ms:.0f} ms per step", file=sys.stderr, flush=True)
        out.update(res)
        out["fp64"]["frac_fp64_peak"] = round(res["fp64"]["algorithmic_tflops"] / FP64_PEAK_TFLOPS, 4)
        out["split_vs_fp64"] = round(res["split"]["q_points_per_s"] / res["fp64"]["q_points_per_s"], 3)
    except TupsError as e:
        out["failed"] = str(e)
        return out
    finally:
        c.close()
    # split error on 16 q of the same inputs (the device generator is counter-based per element)
    e = TupsContext(nx=16, nf=64, nl=9, nq=16, beta=2.0, f_damping=0.5)
    batched_bse_inputs(e, seed=0, nmom=256, nomega=16)
    e.BSE(0)
    ref = [e.download(GAMMA, r) for r in range(4)]
    batched_bse_inputs(e, seed=0, nmom=256, nomega=16)
    e.set_split(moduli, 0)
    e.BSE(0, mode=BSE_LADDER_SPLIT)
    errs = [float(np.linalg.norm(e.download(GAMMA, r) - ref[r]) / np.linalg.norm(ref[r])) for r in range(4)]
    e.close()
    out["split_error"] = {"max": max(errs), "per_channel": errs, "vs": "FP64 form, 16 q of the same inputs"}
    return out


def phir_leg(c):
    """calc_PhiR on a full-IBZ context: HBM bytes (Phi read once per eight shells,
    PhiR written once) per call against 8 TB/s."""
    from tups_amd.runtime import K_PHIR

    c.calc_PhiR(0)
    c.set_timing(True)
    for _ in range(5):
        c.calc_PhiR(0)
    ms, n = c.kernel_time(K_PHIR)
    c.set_timing(False)
    nz2 = float(c.nz) ** 2
    nbytes = 16.0 * (4 * nz2 * c.nq * ((c.nr + 7) // 8) + 4 * nz2 * c.nw * c.nr)
    gbs = nbytes / (ms / n / 1e3) / 1e9
    print(f"[bench] PhiR Nx={c.nx} Nl={c.nl}: {ms / n:.3f} ms per call, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "avg_launch_ms": round(ms / n, 4), "bytes_per_call": nbytes,
            "kernel": "phir_block (one launch per symmetry-block shape)",
            "workload": f"calc_PhiR Nx={c.nx} Nf={c.nf} Nl={c.nl} NR={c.nr}, Nq={c.nq}, 4 channels"}


def dropin_leg():
    """One parquet iteration through the Fortran drop-in's transfers (tups_dropin.f90),
    emulated from Python on the same C entry points at Nx=8, Nf=64, Nl=5 (all 480
    IBZ q, D=320): per q a dense bubble pair packed and uploaded (BSE's calc_chi),
    G_r_LL and F_r_LL uploaded only when not resident, BSE -> download Phi,
    calc_PhiR -> download PhiR, calculate_PAE -> download F, G = F - G on the
    device -> download G; the host mirrors are pageable arrays filled in place,
    as the Fortran module arrays are.  Wall clock per iteration with the vertices
    resident and with them re-uploaded, beside the device time of the same
    kernels without transfers."""
    import numpy as np
    from tups_amd import GAMMA, LAMBDA, PHIR, TupsContext
    from tups_amd import F as FV

    c = TupsContext(nx=8, nf=64, nl=5, beta=2.0, f_damping=0.5)
    nz, nq = c.nz, c.nq
    for r in range(4):
        c.synth_vertex(GAMMA, r, 60 + r, 2.0, 0.3 / np.sqrt(nz))
        c.synth_vertex(FV, r, 70 + r, 2.0, 0.3 / np.sqrt(nz))
        c.upload(LAMBDA, r, np.full((c.nf, c.nf, c.nw), (2.0, -2.0, 4.0, 0.0)[r], dtype=np.complex128))
    g_host = [c.download(GAMMA, r) for r in range(4)]
    f_host = [c.download(FV, r) for r in range(4)]
    phir_host = [np.zeros(c.shape(PHIR), dtype=np.complex128, order="F") for _ in range(4)]
    rng = np.random.default_rng(11)
    chi_dense = np.zeros((nz, nz, 1), dtype=np.complex128, order="F")
    blocks = 0.1 * (rng.standard_normal((c.nl, c.nl, c.nf)) + 1j * rng.standard_normal((c.nl, c.nl, c.nf)))
    for nu in range(c.nf):
        idx = np.arange(c.nl) * c.nf + nu
        chi_dense[np.ix_(idx, idx, [0])] = blocks[:, :, nu][:, :, None]
    c.calc_M()
    vbytes = 16.0 * nz * nz * nq

    def iteration(resident):
        up = down = 0.0
        c.sync()
        t0 = time.perf_counter()
        for q in range(nq):
            c.upload_chi_dense(0, chi_dense, q)
            c.upload_chi_dense(1, chi_dense, q)
        up += 2 * 16.0 * c.nl * c.nl * c.nf * nq
        if not resident:
            for r in range(4):
                c.upload(GAMMA, r, g_host[r])
                c.upload(FV, r, f_host[r])
            up += 8 * vbytes
        c.BSE(0)
        for r in range(4):
            c.download(GAMMA, r, out=g_host[r])
        c.calc_PhiR(0)
        for r in range(4):
            c.download(PHIR, r, out=phir_host[r])
        c.calculate_PAE()
        for r in range(4):
            c.download(FV, r, out=f_host[r])
        c.gamma_update()
        for r in range(4):
            c.download(GAMMA, r, out=g_host[r])
        c.sync()
        down += 12 * vbytes + 4 * 16.0 * phir_host[0].size
        return (time.perf_counter() - t0) * 1e3, up, down

    iteration(False)
    t_res, up_res, down_res = min(iteration(True) for _ in range(2))
    t_re, up_re, down_re = iteration(False)
    c.sync()
    c.record(0)
    c.BSE(0)
    c.calc_PhiR(0)
    c.calculate_PAE()
    c.gamma_update()
    c.record(1)
    c.sync()
    gpu_ms = c.elapsed_ms(0, 1)
    c.close()
    print(f"[bench] drop-in iteration: {t_res:.0f} ms resident, {t_re:.0f} ms re-uploading, device {gpu_ms:.1f} ms",
          file=sys.stderr, flush=True)
    return {"ms_resident": round(t_res, 1), "ms_reupload": round(t_re, 1), "gpu_kernels_ms": round(gpu_ms, 2),
            "pcie_bytes_resident": {"host_to_device": up_res, "device_to_host": down_res},
            "pcie_bytes_reupload": {"host_to_device": up_re, "device_to_host": down_re},
            "effective_gbs_resident": round((up_res + down_res) / (t_res / 1e3) / 1e9, 1),
            "workload": f"drop-in iteration Nx=8 Nf=64 Nl=5, Nq={nq} (all IBZ q), D={nz}: BSE (FP64), calc_PhiR, "
                        "calculate_PAE, G = F - G with the transfers of tups_dropin.f90 from pageable host arrays"}


def main():
    import numpy as np

    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU q-sharding "
              "of this path is a design note, not a run", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from tups_amd import BSE_CLOSED_IR, BSE_CLOSED_LU, BSE_LADDER_FP64, BSE_LADDER_SPLIT, GAMMA, TupsContext
    from tups_amd.runtime import K_BSE_CLOSED, K_BSE_IR, K_BSE_LADDER, K_SPLIT_GEMM, K_SPLIT_PREP
    from tups_amd.synth import batched_bse_inputs

    nx, nf, nl, U = 8, 64, 5, 2.0
    t_setup = time.perf_counter()
    ctx = TupsContext(nx=nx, nf=nf, nl=nl, nq=args.nq, beta=2.0, f_damping=0.5)
    batched_bse_inputs(ctx, U=U, seed=0)
    D, nq = ctx.nz, ctx.nq
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s D={D} nq={nq}", file=sys.stderr, flush=True)

    def timed(mode):
        for _ in range(args.warmup):
            ctx.BSE(0, mode=mode)
        ctx.sync()
        ctx.record(0)
        for _ in range(args.steps):
            ctx.BSE(0, mode=mode)
        ctx.record(1)
        ctx.sync()
        return ctx.elapsed_ms(0, 1)

    # --- FP64 form ---------------------------------------------------------------------
    # the three-multiplication complex products first (reported beside), then the
    # four-product variants (the reference's ZGEMM arithmetic): the plain path
    ctx.set_product_forms(True)
    var3, _ = ctx.autotune()
    ms3 = timed(BSE_LADDER_FP64)
    qps3 = nq * args.steps / (ms3 / 1e3)
    print(f"[bench] fp64 with three-multiplication products ({var3}): {qps3:.1f} q-points/s", file=sys.stderr,
          flush=True)
    ctx.set_product_forms(False)
    variant, tune_ms = ctx.autotune()
    print(f"[bench] autotuner picked {variant} ({tune_ms:.2f} ms/launch)", file=sys.stderr, flush=True)
    ms64 = timed(BSE_LADDER_FP64)
    qps64 = nq * args.steps / (ms64 / 1e3)
    print(f"[bench] fp64: {args.steps} steps {ms64:.1f} ms -> {qps64:.1f} q-points/s", file=sys.stderr, flush=True)
    nprof = max(2, min(args.steps, 5))
    ctx.set_timing(True)
    for _ in range(nprof):
        ctx.BSE(0)
    lad_ms, lad_n = ctx.kernel_time(K_BSE_LADDER)
    ctx.set_timing(False)
    q_per_launch = nq * nprof / lad_n
    flops_per_launch = 16.0 * D ** 3 * 4 * q_per_launch  # 2 complex DxD GEMMs per (q, channel)
    achieved = flops_per_launch / (lad_ms / lad_n / 1e3) / 1e12
    roof64 = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
              "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
              "traffic": profile_traffic("bse_ladder_fp64_D320_nq1024.md"),
              "kernel": f"bse_ladder_kernel variant {variant}", "avg_launch_ms": round(lad_ms / lad_n, 4),
              "flops_per_launch": flops_per_launch}

    # --- split form ----------------------------------------------------------------------
    ctx.set_split(args.moduli, args.qbatch)
    batched_bse_inputs(ctx, U=U, seed=0)
    svar, _ = ctx.autotune(K_SPLIT_GEMM)
    print(f"[bench] split GEMM tile: {svar}", file=sys.stderr, flush=True)
    batched_bse_inputs(ctx, U=U, seed=0)
    ms_sp = timed(BSE_LADDER_SPLIT)
    qps = nq * args.steps / (ms_sp / 1e3)
    print(f"[bench] split: {args.steps} steps {ms_sp:.1f} ms -> {qps:.1f} q-points/s", file=sys.stderr, flush=True)
    ctx.set_timing(True)
    for _ in range(nprof):
        ctx.BSE(0, mode=BSE_LADDER_SPLIT)
    gemm_ms, gemm_n = ctx.kernel_time(K_SPLIT_GEMM)
    prep_ms, prep_n = ctx.kernel_time(K_SPLIT_PREP)
    ctx.set_timing(False)
    q_per_gemm = nq * nprof / gemm_n
    # algorithmic INT8 ops: 2 real GEMMs (D x D x 2D) per modulus per (q, channel)
    ops_per_launch = 2.0 * args.moduli * 2.0 * D * D * (2 * D) * 4 * q_per_gemm
    ach_sp = ops_per_launch / (gemm_ms / gemm_n / 1e3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(ach_sp, 1), "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                "frac": round(ach_sp / INT8_PEAK_TOPS, 4),
                "traffic": profile_traffic("bse_split_crt_D320_nq1024.md"),
                "kernel": f"split GEMM {svar} ({args.moduli} moduli)", "avg_launch_ms": round(gemm_ms / gemm_n, 4),
                "ops_per_launch": ops_per_launch, "prep_ms_per_launch": round(prep_ms / prep_n, 4),
                "launches_per_step": gemm_n // nprof}
    # --- closed form (fixed point by pivoted Gauss-Jordan) --------------------------------
    batched_bse_inputs(ctx, U=U, seed=0)
    ctx.BSE(0, mode=BSE_CLOSED_LU)  # warm-up (allocates its workspace)
    ctx.sync()
    ncl = max(1, min(args.steps, 2))
    ctx.set_timing(True)
    for _ in range(ncl):
        ctx.BSE(0, mode=BSE_CLOSED_LU)
    cl_ms, cl_n = ctx.kernel_time(K_BSE_CLOSED)
    ctx.set_timing(False)
    cl_step_ms = cl_ms / ncl
    cl_flops = 12.0 * D ** 3 * 4 * nq  # Gauss-Jordan on D right-hand sides, 4 channels
    closed = {"value": round(nq / (cl_step_ms / 1e3), 2), "unit": "q-points/s", "ms_per_call": round(cl_step_ms, 3),
              "achieved": round(cl_flops / (cl_step_ms / 1e3) / 1e12, 3), "peak": FP64_PEAK_TFLOPS,
              "unit_roofline": "TFLOP/s", "frac": round(cl_flops / (cl_step_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 4),
              "flops_per_call": cl_flops, "kernel": "closed_build_nu + closed_solve_b64 (64-column block Gauss-Jordan, register sub-panels)"}
    print(f"[bench] closed form: {cl_step_ms:.1f} ms per call", file=sys.stderr, flush=True)
    # --- closed form by FP32 LU (Gauss-Jordan inverse) + FP64 iterative refinement -------
    batched_bse_inputs(ctx, U=U, seed=0)
    ctx.BSE(0, mode=BSE_CLOSED_IR)  # warm-up (allocates its workspace)
    ctx.sync()
    ctx.set_timing(True)
    for _ in range(ncl):
        ctx.BSE(0, mode=BSE_CLOSED_IR)
    ir_ms, _ = ctx.kernel_time(K_BSE_IR)
    ctx.set_timing(False)
    ir_steps, ir_res = ctx.solve_info()
    ir_step_ms = ir_ms / ncl
    refined = {"value": round(nq / (ir_step_ms / 1e3), 2), "unit": "q-points/s", "ms_per_call": round(ir_step_ms, 3),
               "refinement_steps": ir_steps, "residual": ir_res, "residual_bar": 1e-12,
               "vs_direct_fp64": round(cl_step_ms / ir_step_ms, 3),
               "kernel": "closed_build_nu<float2> + closed_solve_b64<float> (FP32 MFMA) + ir_gemm (FP32 / FP64 MFMA) + ir_chi "
                         "+ ir_check",
               "flops_per_call": {"fp32": (12.0 + 8.0 * (1 + ir_steps)) * D ** 3 * 4 * nq,
                                  "fp64": 8.0 * (ir_steps + 1) * D ** 3 * 4 * nq}}
    print(f"[bench] refined closed form: {ir_step_ms:.1f} ms per call, {ir_steps} steps, residual {ir_res:.1e}",
          file=sys.stderr, flush=True)
    # --- convergence check: max |F_r| over the four resident vertices (HBM read) -----------
    from tups_amd.runtime import K_CONV
    nt = nx * nx * nf
    s_new = np.ones(nt, dtype=np.complex128)
    ctx.conv(sigma_old=s_new, sigma=s_new)
    ctx.set_timing(True)
    for _ in range(5):
        ctx.conv(sigma_old=s_new, sigma=s_new)
    cv_ms, cv_n = ctx.kernel_time(K_CONV)
    ctx.set_timing(False)
    cv_bytes = 4.0 * D * D * nq * 16
    conv = {"ms_per_call": round(cv_ms / cv_n, 4), "bytes": cv_bytes,
            "achieved": round(cv_bytes / (cv_ms / cv_n / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(cv_bytes / (cv_ms / cv_n / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernel": "conv_sigma + conv_fmax + conv_final",
            "workload": f"conv() over the four resident F (D={D}, Nq={nq}) and a {nt}-point Sigma pair"}
    print(f"[bench] conv: {cv_ms / cv_n:.3f} ms per call", file=sys.stderr, flush=True)
    ctx.close()

    # split error per channel vs the FP64 form, on 64 q of the same workload
    ec = TupsContext(nx=nx, nf=nf, nl=nl, nq=64, beta=2.0, f_damping=0.5)
    batched_bse_inputs(ec, U=U, seed=5)
    ec.BSE(0)
    ref = [ec.download(GAMMA, r) for r in range(4)]
    batched_bse_inputs(ec, U=U, seed=5)
    ec.set_split(args.moduli, 0)
    ec.BSE(0, mode=BSE_LADDER_SPLIT)
    errs = []
    for r in range(4):
        got = ec.download(GAMMA, r)
        errs.append(float(np.linalg.norm(got - ref[r]) / np.linalg.norm(ref[r])))
    batched_bse_inputs(ec, U=U, seed=5)
    ec.BSE(0, mode=BSE_CLOSED_LU)
    ref_cl = [ec.download(GAMMA, r) for r in range(4)]
    batched_bse_inputs(ec, U=U, seed=5)
    ec.BSE(0, mode=BSE_CLOSED_IR)
    refined["error_vs_fp64_direct"] = [float(np.linalg.norm(ec.download(GAMMA, r) - ref_cl[r])
                                             / np.linalg.norm(ref_cl[r])) for r in range(4)]
    ec.close()
    print(f"[bench] split error per channel: {errs}", file=sys.stderr, flush=True)

    # crossing (calculate_PAE) at Nl = 1 on the 16x16 / Nf = 64 box: the HBM-bound regime
    from tups_amd.profile_workloads import crossing_bytes, crossing_context
    from tups_amd.runtime import K_PAE, K_PAE_GATHER

    cx = crossing_context(16, 64, 1)
    cvar, _ = cx.autotune(K_PAE)
    for _ in range(2):
        cx.calculate_PAE()
    cx.set_timing(True)
    for _ in range(10):
        cx.calculate_PAE()
    pae_ms, pae_n = cx.kernel_time(K_PAE)
    gat_ms, gat_n = cx.kernel_time(K_PAE_GATHER)
    cx.set_timing(False)
    cbytes = crossing_bytes(cx)
    # one calculate_PAE = PhiR re-index pass (unless folded into the variant) + streaming crossing
    per_call_ms = (pae_ms + gat_ms) / pae_n
    gbs = cbytes / (per_call_ms / 1e3) / 1e9
    crossing = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4),
                "traffic": profile_traffic("crossing_pae_nl1_nx16_nf64.md"),
                "kernel": f"pae_stream_nl1 variant {cvar}" + (" (PhiR re-index folded in)" if gat_n == 0
                                                               else " (+ pae_pregather_nl1)"),
                "avg_launch_ms": round(pae_ms / pae_n, 4), "pregather_ms": round(gat_ms / max(1, gat_n), 4),
                "bytes_per_call": cbytes,
                "workload": f"calculate_PAE Nx=16 Nf=64 Nl=1 NR=1, Nq={cx.nq} (all IBZ q), 4 channels"}
    phir1 = phir_leg(cx)
    cx.close()

    # crossing at Nl = 5 (Nx = 8, Nf = 64, NR = 4, all 480 IBZ q): the compute-heavy regime
    from tups_amd.profile_workloads import crossing_flops
    c5 = crossing_context(8, 64, 5)
    c5.calculate_PAE()
    c5.sync()
    c5.set_timing(True)
    for _ in range(2):
        c5.calculate_PAE()
    p5_ms, p5_n = c5.kernel_time(K_PAE)
    y5_ms, _ = c5.kernel_time(K_PAE_GATHER)  # the per-qw gather of PhiR into the GEMM operand
    p5_ms += y5_ms
    c5.set_timing(False)
    fl5 = crossing_flops(c5)
    t5 = p5_ms / p5_n / 1e3
    crossing5 = {"bound": "mfma", "achieved": round(fl5 / t5 / 1e12, 3), "peak": FP64_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": round(fl5 / t5 / 1e12 / FP64_PEAK_TFLOPS, 4),
                 "avg_launch_ms": round(p5_ms / p5_n, 3), "flops_per_call": fl5,
                 "traffic": profile_traffic("crossing_pae_nl5_nx8_nf64.md"),
                 "bytes_per_call": crossing_bytes(c5),
                 "kernel": "pae_ybuild + pae_mm (FP64 MFMA GEMM per bosonic frequency)",
                 "workload": f"calculate_PAE Nx=8 Nf=64 Nl=5 NR={c5.nr}, Nq={c5.nq} (all IBZ q), 4 channels"}
    print(f"[bench] crossing Nl=5: {p5_ms / p5_n:.2f} ms per call", file=sys.stderr, flush=True)
    phir5 = phir_leg(c5)
    c5.close()

    # batching form: 1024 q with Nf_q 64 -> 16 by bosonic frequency and Nl_q in {1, 5, 9}
    from tups_amd.profile_workloads import batch_context, batch_flops
    bt = batch_context()
    bvar, _ = bt.autotune()
    for _ in range(args.warmup):
        bt.BSE()
    bt.sync()
    bt_ms = sum(bt.BSE(timed=True) for _ in range(args.steps)) / args.steps
    bfl = batch_flops(bt)
    dims = bt.dims()
    batching = {"value": round(bt.nq / (bt_ms / 1e3), 2), "unit": "q-points/s", "ms_per_call": round(bt_ms, 3),
                "roofline": {"bound": "mfma", "achieved": round(bfl / (bt_ms / 1e3) / 1e12, 3),
                             "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(bfl / (bt_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                             "traffic": profile_traffic("bse_batch_uneven_nx8_nq1024.md"),
                             "avg_launch_ms": round(bt_ms, 4), "flops_per_launch": bfl},
                "kernel": f"bse_batch_kernel variant {bvar} (one persistent launch, all q and channels)",
                "items": bt.items,
                "workload": f"Nx=8, Nq={bt.nq} (64 momenta x 16 bosonic freq), Nf_q 64->16 by frequency, "
                            f"Nl_q 1/5/9 by momentum, D {min(dims)}..{max(dims)}, FP64"}
    print(f"[bench] batching: {bt_ms:.2f} ms per launch", file=sys.stderr, flush=True)
    bt.close()

    # split form of the same uneven batch: error against its FP64 step from the same inputs, then the rate
    bs = batch_context()
    bs.BSE()
    bs.sync()
    ref = [np.concatenate([x.ravel() for x in bs.download(GAMMA, r)]) for r in range(4)]
    bs.close()
    bs = batch_context()
    bs.BSE_split(args.moduli)
    bs.sync()
    got = [np.concatenate([x.ravel() for x in bs.download(GAMMA, r)]) for r in range(4)]
    berr = [float(np.linalg.norm(got[r] - ref[r]) / np.linalg.norm(ref[r])) for r in range(4)]
    for _ in range(args.warmup):
        bs.BSE_split(args.moduli)
    bs.sync()
    bs_ms = sum(bs.BSE_split(args.moduli, timed=True) for _ in range(args.steps)) / args.steps
    batching_split = {"value": round(bs.nq / (bs_ms / 1e3), 2), "unit": "q-points/s", "ms_per_call": round(bs_ms, 3),
                      "vs_fp64_batch": round(bt_ms / bs_ms, 3), "moduli": args.moduli,
                      "error_vs_fp64": {"max": max(berr), "per_channel": berr},
                      "kernel": "crt_prep_b + crt_rowmax_fold + crt_prep_a + crt_gemm_dma<4> / crt_gemm_big<6>, "
                                "per (Nf_q, Nl_q) class",
                      "workload": "the batching workload above in the split form"}
    print(f"[bench] batching split: {bs_ms:.2f} ms per call, error {max(berr):.2e}", file=sys.stderr, flush=True)
    bs.close()

    # uneven-box iteration (BSE + crossing + G = F - G) replayed as one HIP graph
    from tups_amd.profile_workloads import uneven_context
    un = uneven_context()
    un.batch.autotune()
    un.iteration(graph=True)
    un.sync()
    un.ctx.record(0)
    for _ in range(args.steps):
        un.iteration(graph=True)
    un.ctx.record(1)
    un.sync()
    un_ms = un.ctx.elapsed_ms(0, 1) / args.steps
    udims = un.batch.dims()
    uneven = {"value": round(un.batch.nq / (un_ms / 1e3), 2), "unit": "q-points/s", "ms_per_iteration": round(un_ms, 3),
              "iterations": args.steps,
              "workload": f"Nx=8, Nq={un.batch.nq} (15 IBZ momenta x 32 bosonic freq), Nf_w 64->16 by frequency, "
                          f"Nl_m 1/5/9 by momentum, D {min(udims)}..{max(udims)}; one iteration = batched BSE "
                          f"(one launch) + crossing on the largest box (Nf=64, Nl=9, NR={un.ctx.nr}) + "
                          f"G = F - G, as one replayed HIP graph"}
    print(f"[bench] uneven-box iteration: {un_ms:.2f} ms", file=sys.stderr, flush=True)
    un.close()

    # self-energy (SDE without the use_U2 term) at Nx=8, Nf=64, Nl=5, all 480 IBZ q
    from tups_amd import LAMBDA
    from tups_amd.runtime import K_SDE
    sc = TupsContext(nx=8, nf=64, nl=5, beta=2.0, f_damping=0.5)
    for r in range(4):
        sc.synth_vertex(GAMMA, r, 40 + r, 0.0, 0.1)
    lam = np.zeros((64, 64, sc.nw), dtype=np.complex128)
    lam[:] = U
    sc.upload(LAMBDA, 0, lam)
    lam[:] = -U
    sc.upload(LAMBDA, 1, lam)
    sig = np.zeros(8 * 8 * 64, dtype=np.complex128)
    sc.sde(sig, 0.0, U)
    sc.set_timing(True)
    for _ in range(3):
        sc.sde(sig, 0.0, U)
    sde_ms, sde_n = sc.kernel_time(K_SDE)
    sc.set_timing(False)
    vbytes = 4.0 * sc.nz * sc.nz * sc.nq * 16
    sde = {"ms_per_call": round(sde_ms / sde_n, 3), "vertex_bytes": vbytes,
           "vertex_read_gbs": round(vbytes / (sde_ms / sde_n / 1e3) / 1e9, 1),
           "kernel": "sde_chi_aux + sde_vchi + sde_sigma (FP64)",
           "workload": f"calc_SDE without use_U2, Nx=8 Nf=64 Nl=5, Nq={sc.nq} (all IBZ q), f_range=0"}
    print(f"[bench] sde: {sde_ms / sde_n:.2f} ms per call", file=sys.stderr, flush=True)
    # calc_sus on the same context (F and Phi resident, Sigma = 0 on the device)
    from tups_amd.runtime import K_SUS
    sc.set_sigma(sig, 0.0)
    sc.calc_sus(U)
    sc.set_timing(True)
    for _ in range(3):
        sc.calc_sus(U)
    sus_ms, sus_n = sc.kernel_time(K_SUS)
    sc.set_timing(False)
    nent = sc.nz
    sus_flops = 4.0 * sc.nq * nent * nent * 4 * 8  # 4 channels x Nq x D^2 entries x 4 form factors, complex MAC
    sus = {"ms_per_call": round(sus_ms / sus_n, 3), "kernel": "sus_gg + sus_contract (FP64)",
           "fp64_tflops": round(sus_flops / (sus_ms / sus_n / 1e3) / 1e12, 2),
           "workload": f"calc_sus, Nx=8 Nf=64 Nl=5, Nq={sc.nq}, f_range=0, 4 external form factors"}
    print(f"[bench] sus: {sus_ms / sus_n:.2f} ms per call", file=sys.stderr, flush=True)
    sc.close()

    # calc_Gkw_Grt: mu bisection + G(k, tau) + 2D transform on a coarse-grained grid
    from tups_amd.runtime import K_GTAU
    gc = TupsContext(nx=8, nf=64, nl=1, beta=2.0, f_damping=0.5)
    rng = np.random.default_rng(3)
    sig_g = 0.2 * (rng.standard_normal(8 * 8 * 64) + 1j * rng.standard_normal(8 * 8 * 64))
    gc.set_sigma(sig_g, 0.3, ngrain=4)
    gc.gkw_grt(U, 0.9, 0.3, want_grt=False)
    gc.set_timing(True)
    for _ in range(5):
        mu_g, steps_g, _ = gc.gkw_grt(U, 0.9, 0.3, want_grt=False)
    gt_ms, gt_n = gc.kernel_time(K_GTAU)
    gc.set_timing(False)
    gtau = {"ms_per_call": round(gt_ms / gt_n, 3), "mu_evaluations": steps_g,
            "kernel": "gt_setup + 11 x gt_density (63 candidate mu, tree walk in the last workgroup) + gt_fine + "
                      "gt_dft1 + gt_dft2",
            "workload": "calc_Gkw_Grt, Nx=8 Ngrain=4 (32x32 fine grid), Nf=64, nParticle=0.9, random Sigma"}
    print(f"[bench] gtau: {gt_ms / gt_n:.3f} ms per call, {steps_g} evaluations", file=sys.stderr, flush=True)
    gc.close()

    # the reference's CPU case (4x4, Nf=16, Nl=1) as one full iteration on the device, graph replayed
    from tups_amd import F as FV
    c1 = TupsContext(nx=4, nf=16, nl=1, beta=2.0, f_damping=0.5)
    gam1, fv1, lam1, sig1 = config1_inputs(c1.nq, c1.nz, c1.nf, c1.nw)
    for r in range(4):
        c1.upload(GAMMA, r, gam1[r])
        c1.upload(FV, r, fv1[r])
        c1.upload(LAMBDA, r, lam1[r])
    c1.calc_M()
    c1.set_sigma(sig1, 0.0)
    c1.iteration_full(U, graph=True, use_u2=True, n_particle=1.0)
    c1.sync()
    c1.record(0)
    for _ in range(20):
        c1.iteration_full(U, graph=True, use_u2=True, n_particle=1.0)
    c1.record(1)
    c1.sync()
    c1_ms = c1.elapsed_ms(0, 1) / 20
    mu1, steps1 = c1.get_mu()
    config1 = {"gpu_ms_per_iteration": round(c1_ms, 4), "iterations": 20, "mu": mu1, "mu_evaluations": steps1,
               "workload": CONFIG1_WORKLOAD + ", replayed as one HIP graph"}
    print(f"[bench] config-1 iteration: {c1_ms:.3f} ms", file=sys.stderr, flush=True)
    c1.close()

    # the reference's own parameter file (Nx=16, Nf=64, Nl=1, beta=5, U=2, 100 iterations) run whole
    prms_run = reference_prms_run()

    large = None if args.no_large else large_share(args.moduli)
    dropin = dropin_leg()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(nx, nf, nl, args.cpu_seconds)

    line = {
        "metric": "BSE q-points/s + %MFMA peak: split-FP32 vs FP64 path at rel err<=1e-10",
        "value": round(qps, 2),
        "unit": "q-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_sp / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8 (exact INT32 sums) + f64",
        "data": "synthetic",
        "config": {"workload": f"batched BSE ladder, 4 channels, D={D} (Nl={nl} x Nf={nf}), "
                               f"Nq={nq} (64 momenta x 16 bosonic freq), split form ({args.moduli} CRT moduli)",
                   "nx": nx, "nf": nf, "nl": nl, "nq": nq, "D": D, "path": "int8-crt-mfma",
                   "moduli": args.moduli},
        "split_error": {"max": max(errs), "per_channel": errs, "bar": 1e-10,
                        "vs": "FP64 form of the same step, 64 q of the same workload"},
        "fp64_path": {"value": round(qps64, 2), "unit": "q-points/s", "ms_per_step": round(ms64 / args.steps, 3),
                      "roofline": roof64},
        "split_vs_fp64": round(qps / qps64, 3),
        "fp64_three_mult": {"value": round(qps3, 2), "unit": "q-points/s", "variant": var3,
                            "split_vs": round(qps / qps3, 3),
                            "note": "three-multiplication complex products: meets the complex relative "
                                    "Frobenius bar, but loses relative accuracy on a small imaginary part "
                                    "(7e-9 under cancellation); opt-in, not the parity path"},
        "closed_form": closed,
        "closed_form_refined": refined,
        "uneven_iteration": uneven,
        "dropin_iteration": dropin,
        "roofline": roofline,
        "crossing_roofline": crossing,
        "crossing_nl5": crossing5,
        "phir_nl1": phir1,
        "phir_nl5": phir5,
        "batching": batching,
        "batching_split": batching_split,
        "large_share": large,
        "sde": sde,
        "sus": sus,
        "gtau": gtau,
        "conv": conv,
        "config1_iteration": config1,
        "reference_prms_run": prms_run,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
