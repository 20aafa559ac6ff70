"""Oracle: the convergence check of one parquet iteration.

Test infrastructure only: the checker for the GPU path, never shipped or timed.

Restates conv(), SRC/parquet_check.f90:39-150:
* e = |SigmaOld - Sigma| / |Sigma| elementwise; the average |sum(e) / Nt| and the
  maximum, with the first index of the maximum mapped back through
  Index_fermionic (SRC/parquet_util.f90:95-100: idx = ((kx-1) Ny + ky-1) Nf + nu);
* per channel the largest |F_r| and its relative change |1 - max / max_prev|
  against the previous call's maximum;
* converged = (max e < f_sigma) and every F change < f_vert
  (f_sigma = f_vert = 1e-8, SRC/parquet_ini.f90:31-32).

Pinned by (tests/test_oracle_conv.py): Sigma_old = Sigma (1 + eps) gives every
e = |eps|; a single planted discrepancy is found at its (kx, ky, nu); a planted
largest |F| is found whatever its position; F scaled by (1 + d) gives a change
of |d|; the thresholds decide as strict inequalities.
"""
from __future__ import annotations

import numpy as np

F_SIGMA = 1e-8
F_VERT = 1e-8


def conv(sigma_old, sigma, fverts, fmax_prev, nx, ny, nf, f_sigma=F_SIGMA, f_vert=F_VERT):
    """Returns a dict of the fields conv() prints and decides on."""
    e = np.abs(np.asarray(sigma_old) - np.asarray(sigma)) / np.abs(np.asarray(sigma))
    k = int(np.argmax(e))  # first occurrence, as maxloc
    kx, rem = divmod(k, ny * nf)
    ky, nu = divmod(rem, nf)
    fmax = [float(np.max(np.abs(f))) for f in fverts]
    rel = [abs(1.0 - m / p) if p != 0 else float("inf") for m, p in zip(fmax, fmax_prev)]
    out = {
        "sigma_avg": float(abs(np.sum(e) / e.size)),
        "sigma_max": float(np.max(e)),
        "kx": kx + 1, "ky": ky + 1, "nu": nu + 1,
        "f_max": fmax, "f_rel_err": rel,
    }
    out["converged"] = bool(out["sigma_max"] < f_sigma and all(r < f_vert for r in rel))
    return out
