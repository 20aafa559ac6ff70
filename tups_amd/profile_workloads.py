"""Named workloads for rocprofv3 runs of the hot-path kernels.

    python -m tups_amd.profile_workloads bse      --steps 3   # batched BSE, D=320, Nq=1024
    python -m tups_amd.profile_workloads split    --steps 3   # same, split (INT8 CRT) form
    python -m tups_amd.profile_workloads crossing --steps 5   # PAE crossing, Nl=1, Nx=16, Nf=64
    python -m tups_amd.profile_workloads phir     --steps 5   # PhiR, Nl=1, Nx=16, Nf=64
    python -m tups_amd.profile_workloads calib    --steps 3   # G = F - G streaming kernel
    python -m tups_amd.profile_workloads batch    --steps 3   # uneven q batch, Nx=8, D 16..576
    python -m tups_amd.profile_workloads uneven   --steps 3   # uneven-box iteration with the crossing
    python -m tups_amd.profile_workloads closed    --steps 1  # direct FP64 elimination, D=320
    python -m tups_amd.profile_workloads closed_ir --steps 1  # FP32 LU + FP64 refinement, D=320

`calib` runs a streaming kernel whose bytes are known exactly (per element:
read 2 x 16 B, write 16 B) so hardware counters (FETCH_SIZE, WRITE_SIZE) can be
scaled to bytes.  Each workload prints its step count and per-launch
algorithmic bytes / flops.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from .runtime import BSE_LADDER_SPLIT, CHI, F, GAMMA, K_PAE, LAMBDA, TupsBatch, TupsContext, TupsUneven
from .synth import batched_bse_inputs


def crossing_context(nx=16, nf=64, nl=1, seed=0):
    """Full-IBZ context with random Phi (device generator), PhiR and M built on the device."""
    c = TupsContext(nx=nx, nf=nf, nl=nl, beta=2.0, f_damping=0.5)
    rng = np.random.default_rng(seed)
    for r in range(4):
        c.synth_vertex(GAMMA, r, seed + r, 0.0, 1.0)
        lam = rng.standard_normal((nf, nf, nf // 2)) + 1j * rng.standard_normal((nf, nf, nf // 2))
        c.upload(LAMBDA, r, lam)
    c.calc_PhiR(0)
    c.calc_M()
    c.sync()
    return c


def crossing_bytes(c: TupsContext) -> float:
    """Minimum HBM bytes of one calculate_PAE launch: Phi read + F write (4 channels),
    PhiR, M and Lambda read once."""
    nz2q = c.nz * c.nz * c.nq
    phir = c.nz * c.nz * c.nw * c.nr
    m = c.nl ** 4 * c.nibz * c.nr
    lam = c.nf * c.nf * c.nw
    return 16.0 * 4 * (2 * nz2q + phir + m + lam)


def crossing_flops(c: TupsContext) -> float:
    """Complex MACs of one launch: every (q, row, col) entry of both target pairs
    takes up to two terms of Nl^2 NR products, each feeding two targets."""
    nf, nw = c.nf, c.nw
    terms = 0
    for qw in range(1, nw + 1):
        n1 = np.add.outer(np.arange(1, nf + 1), np.zeros(nf, int))
        n2 = n1.T
        d = n2 - n1
        ph = ((d >= 0) & (d + 1 <= nw)) | ((d < 0) & (1 - d <= nw))
        w = n1 + n2 + qw - nf - 1
        wt = nf + 3 - qw - n1 - n2
        pp = ((w >= 1) & (w <= nw)) | ((wt >= 2) & (wt <= nw))
        w = qw - n1 - n2 + nf + 1
        wt = n1 + n2 - qw - nf + 1
        php = ((w >= 1) & (w <= nw)) | ((wt >= 2) & (wt <= nw))
        terms += int(ph.sum() + pp.sum() + php.sum() + ph.sum())
    per_term = c.nl ** 2 * c.nr * 2 * 8  # two targets, complex MAC = 8 flops
    return float(terms) * c.nibz * c.nl ** 2 * per_term


def uneven_batch_shape(nx=8, nw=16, nf_max=64, nf_min=16):
    """Per-q (Nf_q, Nl_q) of the uneven batch: q = m nw + w over nx^2 momenta m and nw
    bosonic frequencies w; the fermionic box shrinks linearly with w from nf_max to
    nf_min (kept even) and the form-factor count cycles 1, 5, 9 with m."""
    nf, nl = [], []
    for m in range(nx * nx):
        for w in range(nw):
            f = nf_max - (nf_max - nf_min) * w / max(1, nw - 1)
            nf.append(int(2 * round(f / 2)))
            nl.append((1, 5, 9)[m % 3])
    return nf, nl


def batch_context(nx=8, nw=16, U=2.0, seed=0):
    """TupsBatch of the uneven shape with Gamma = F drawn per q from N(0, (0.3U)^2/D_q)
    plus U on the diagonal, and random bubble blocks of unit scale."""
    nf, nl = uneven_batch_shape(nx, nw)
    b = TupsBatch(nx, nf, nl, beta=2.0, f_damping=0.5)
    rng = np.random.default_rng(seed)
    d = np.array([f * l for f, l in zip(nf, nl)], dtype=np.int64)
    sizes = d * d
    off = np.concatenate(([0], np.cumsum(sizes)))
    scale = np.repeat(0.3 * U / np.sqrt(d), sizes)
    diag = np.concatenate([off[q] + np.arange(d[q]) * (d[q] + 1) for q in range(len(d))])
    for r in range(4):
        v = (rng.standard_normal(b.vertex_elems) + 1j * rng.standard_normal(b.vertex_elems)) * scale
        v[diag] += U
        b.upload_packed(GAMMA, r, v)
        b.upload_packed(F, r, v)
    for p in range(2):
        b.upload_packed(CHI, p, rng.standard_normal(b.chi_elems) + 1j * rng.standard_normal(b.chi_elems))
    return b


def uneven_shape(nx=8, nf_max=64, nf_min=16):
    """Boxes of the uneven-box iteration: Nf_w shrinking linearly (kept even) from
    nf_max at w = 0 to nf_min at the last bosonic frequency, Nl_m cycling 1, 5, 9
    over the IBZ momenta."""
    nw = nf_max // 2
    nxi = (nx + 2) // 2
    nibz = nxi * (nxi + 1) // 2
    nf_w = [int(2 * round((nf_max - (nf_max - nf_min) * w / max(1, nw - 1)) / 2)) for w in range(nw)]
    nl_m = [(1, 5, 9)[m % 3] for m in range(nibz)]
    return nf_w, nl_m


def uneven_context(nx=8, nf_max=64, nf_min=16, U=2.0, seed=0):
    """TupsUneven of that shape: Gamma = F per q from N(0, (0.3U)^2/D_q) plus U on the
    diagonal, bubble blocks of scale 0.1, the parquet-approximation Lambda
    (U, -U, 2U, 0) on the largest box."""
    nf_w, nl_m = uneven_shape(nx, nf_max, nf_min)
    u = TupsUneven(nx, nf_max, nf_w, nl_m, beta=2.0, f_damping=0.5)
    b = u.batch
    rng = np.random.default_rng(seed)
    d = np.array([f * l for f, l in zip(b.nf, b.nl)], dtype=np.int64)
    sizes = d * d
    off = np.concatenate(([0], np.cumsum(sizes)))
    scale = np.repeat(0.3 * U / np.sqrt(d), sizes)
    diag = np.concatenate([off[q] + np.arange(d[q]) * (d[q] + 1) for q in range(len(d))])
    for r in range(4):
        v = (rng.standard_normal(b.vertex_elems) + 1j * rng.standard_normal(b.vertex_elems)) * scale
        v[diag] += U
        b.upload_packed(GAMMA, r, v)
        b.upload_packed(F, r, v)
    for p in range(2):
        b.upload_packed(CHI, p, 0.1 * (rng.standard_normal(b.chi_elems) + 1j * rng.standard_normal(b.chi_elems)))
    nw = nf_max // 2
    for r, val in enumerate((U, -U, 2.0 * U, 0.0)):
        u.ctx.upload(LAMBDA, r, np.full((nf_max, nf_max, nw), val, dtype=np.complex128, order="F"))
    return u


def batch_flops(b: TupsBatch) -> float:
    """Algorithmic flops of one batch launch: 16 D_q^3 per (q, channel)."""
    return float(sum(16.0 * float(d) ** 3 * 4 for d in b.dims()))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["bse", "split", "crossing", "phir", "calib", "batch", "gtau", "config1",
                                         "conv", "sus", "closed", "closed_ir", "uneven"])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--nx", type=int, default=16)
    ap.add_argument("--nf", type=int, default=64)
    ap.add_argument("--nl", type=int, default=1)
    ap.add_argument("--moduli", type=int, default=12)
    ap.add_argument("--qbatch", type=int, default=0, help="split: q points per pass (default: the library's)")
    ap.add_argument("--variant", default="",
                    help="batch: tile variant by name (default: autotune); split: GEMM tile by name (default: the "
                         "library's)")
    args = ap.parse_args(argv)
    if args.workload == "sus":
        # calc_sus on Nx=8, Nf=64, Nl=5 (all 480 IBZ q): random F and Phi, Sigma = 0
        c = TupsContext(nx=8, nf=64, nl=5, beta=2.0, f_damping=0.5)
        rng = np.random.default_rng(5)
        for r in range(4):
            for which in (F, GAMMA):
                c.upload(which, r, 0.1 * (rng.standard_normal((c.nz, c.nz, c.nq))
                                          + 1j * rng.standard_normal((c.nz, c.nz, c.nq))))
        c.set_sigma(np.zeros(8 * 8 * 64, dtype=np.complex128), 0.0)
        for _ in range(args.steps):
            c.calc_sus(2.0)
        print(f"sus: {args.steps} calls, Nq={c.nq}, D={c.nz}")
        c.close()
        sys.stdout.flush()
        return
    if args.workload == "conv":
        # conv over the four resident F of the D=320, Nq=1024 batched workload
        c = TupsContext(nx=8, nf=64, nl=5, nq=1024, beta=2.0, f_damping=0.5)
        batched_bse_inputs(c)
        s = np.ones(8 * 8 * 64, dtype=np.complex128)
        for _ in range(args.steps):
            c.conv(sigma_old=s, sigma=s)
        print(f"conv: {args.steps} calls, F bytes per call {4.0 * c.nz * c.nz * c.nq * 16:.6e}")
        c.close()
        sys.stdout.flush()
        return
    if args.workload == "gtau":
        # calc_Gkw_Grt: Nx=8, Ngrain=4 (32 x 32 fine grid), Nf=64, random Sigma, nParticle 0.9
        c = TupsContext(nx=8, nf=64, nl=1, beta=2.0, f_damping=0.5)
        rng = np.random.default_rng(3)
        sig = 0.2 * (rng.standard_normal(8 * 8 * 64) + 1j * rng.standard_normal(8 * 8 * 64))
        c.set_sigma(sig, 0.3, ngrain=4)
        for _ in range(args.steps):
            mu, steps, _ = c.gkw_grt(2.0, 0.9, 0.3, want_grt=False)
        print(f"gtau: {args.steps} calls, mu {mu:.6f} after {steps} evaluations")
        c.close()
        sys.stdout.flush()
        return
    if args.workload == "config1":
        # one full iteration of Nx=4 Nf=16 Nl=1 (beta=2, U=2, nParticle=1, use_U2), graph replayed
        c = TupsContext(nx=4, nf=16, nl=1, beta=2.0, f_damping=0.5)
        rng = np.random.default_rng(7)
        for r in range(4):
            g = 0.05 * (rng.standard_normal((c.nz, c.nz, c.nq)) + 1j * rng.standard_normal((c.nz, c.nz, c.nq)))
            c.upload(GAMMA, r, g + 0.5 * np.eye(c.nz)[:, :, None])
            c.upload(F, r, g + 0.5 * np.eye(c.nz)[:, :, None])
            c.upload(LAMBDA, r, np.full((c.nf, c.nf, c.nw), (2.0, -2.0, 4.0, 0.0)[r], dtype=np.complex128))
        c.calc_M()
        c.set_sigma(np.zeros(4 * 4 * 16, dtype=np.complex128), 0.0)
        for _ in range(args.steps):
            c.iteration_full(2.0, graph=True, use_u2=True, n_particle=1.0)
        c.sync()
        print(f"config1: {args.steps} graph-replayed iterations, mu {c.get_mu()}")
        c.close()
        sys.stdout.flush()
        return
    if args.workload == "uneven":
        # uneven-box iteration (Nx=8, Nf_w 64 -> 16, Nl_m 1/5/9), graph replayed
        u = uneven_context()
        name, _ = u.batch.autotune()
        for _ in range(args.steps):
            u.iteration(graph=True)
        u.sync()
        print(f"uneven: batch variant {name}, {args.steps} graph-replayed iterations, Nq={u.batch.nq}")
        u.close()
        sys.stdout.flush()
        return
    if args.workload == "batch":
        b = batch_context()
        if args.variant:
            ids = {n: v for v, n in b.variants()}
            b.set_variant(ids[args.variant])
            name = args.variant
        else:
            name, _ = b.autotune()
        print(f"batch: variant {name}")
        for _ in range(args.steps):
            b.BSE()
        b.sync()
        print(f"batch: {args.steps} launches, Nq={b.nq}, items={b.items}, flops/launch {batch_flops(b):.4e}")
        b.close()
        sys.stdout.flush()
        return
    if args.workload in ("bse", "split", "calib", "closed", "closed_ir"):
        c = TupsContext(nx=8, nf=64, nl=5, nq=1024, beta=2.0, f_damping=0.5)
        batched_bse_inputs(c)
        if args.workload == "bse":
            name, _ = c.autotune()
            for _ in range(args.steps):
                c.BSE(0)
            print(f"bse: variant {name}, {args.steps} launches, flops/launch {16.0 * c.nz ** 3 * 4 * c.nq:.4e}")
        elif args.workload == "closed":
            from .runtime import BSE_CLOSED_LU
            for _ in range(args.steps):
                c.BSE(0, mode=BSE_CLOSED_LU)
            print(f"closed: {args.steps} calls, flops/call {12.0 * c.nz ** 3 * 4 * c.nq:.4e}")
        elif args.workload == "closed_ir":
            from .runtime import BSE_CLOSED_IR
            for _ in range(args.steps):
                c.BSE(0, mode=BSE_CLOSED_IR)
            print(f"closed_ir: {args.steps} calls, refinement (steps, residual) {c.solve_info()}")
        elif args.workload == "split":
            c.set_split(args.moduli, args.qbatch)
            if args.variant:
                from .runtime import K_SPLIT_GEMM
                c.set_variant(dict((n, v) for v, n in c.variants(K_SPLIT_GEMM))[args.variant], K_SPLIT_GEMM)
            for _ in range(args.steps):
                c.BSE(0, mode=BSE_LADDER_SPLIT)
            ops = 2.0 * args.moduli * 2.0 * c.nz ** 2 * (2 * c.nz) * 4 * c.nq
            print(f"split: {args.moduli} moduli, {args.steps} steps, INT8 ops/step {ops:.4e}")
        else:
            for _ in range(args.steps):
                c.gamma_update()
            n = c.nz * c.nz * c.nq * 4
            print(f"calib: {args.steps} launches, fetch {32.0 * n:.6e} B, write {16.0 * n:.6e} B per launch")
    else:
        c = crossing_context(args.nx, args.nf, args.nl)
        if args.workload == "crossing" and args.nl == 1:
            name, _ = c.autotune(K_PAE)
            print(f"crossing: variant {name}")
        for _ in range(args.steps):
            if args.workload == "crossing":
                c.calculate_PAE()
            else:
                c.calc_PhiR(0)
        print(f"{args.workload}: {args.steps} launches, Nq={c.nq} Nz={c.nz}, "
              f"crossing bytes/launch {crossing_bytes(c):.6e}")
    c.sync()
    c.close()
    sys.stdout.flush()


if __name__ == "__main__":
    main()
