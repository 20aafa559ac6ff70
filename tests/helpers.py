"""Shared helpers for the parity tests (random inputs, error norms)."""
import numpy as np


def crandn(rng, shape, scale=1.0):
    """Complex Gaussian array in Fortran order."""
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return np.asfortranarray(a * scale)


def rel_fro(a, b):
    """||a - b||_F / ||b||_F (0 when both vanish)."""
    nb = np.linalg.norm(np.ravel(b))
    na = np.linalg.norm(np.ravel(a) - np.ravel(b))
    return 0.0 if nb == 0 and na == 0 else na / max(nb, 1e-300)


def random_chi_compact(rng, nl, nf, nq, scale=1.0):
    return crandn(rng, (nl, nl, nf, nq), scale)


def random_vertices(rng, nz, nq, diag=2.0, scale=0.3):
    out = []
    for _ in range(4):
        v = crandn(rng, (nz, nz, nq), scale)
        for q in range(nq):
            v[:, :, q] += diag * np.eye(nz)
        out.append(v)
    return out
