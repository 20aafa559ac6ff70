"""Pins of oracle/conv.py by constructions whose answer is known in advance."""
import numpy as np

from oracle import conv as C


def _rand(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_uniform_relative_change():
    rng = np.random.default_rng(0)
    nx, nf = 4, 8
    s = _rand(rng, nx * nx * nf)
    eps = 3e-7
    out = C.conv(s * (1 + eps), s, [np.ones(4)] * 4, [1.0] * 4, nx, nx, nf)
    assert abs(out["sigma_avg"] - eps) < 1e-15 and abs(out["sigma_max"] - eps) < 1e-15


def test_planted_discrepancy_location():
    rng = np.random.default_rng(1)
    nx, nf = 6, 8
    s = _rand(rng, nx * nx * nf)
    old = s.copy()
    kx, ky, nu = 5, 2, 7
    old[((kx - 1) * nx + ky - 1) * nf + nu - 1] *= 1.5
    out = C.conv(old, s, [np.ones(4)] * 4, [1.0] * 4, nx, nx, nf)
    assert (out["kx"], out["ky"], out["nu"]) == (kx, ky, nu)
    assert abs(out["sigma_max"] - 0.5) < 1e-15
    assert abs(out["sigma_avg"] - 0.5 / s.size) < 1e-15


def test_vertex_maximum_and_change():
    rng = np.random.default_rng(2)
    f = [0.1 * _rand(rng, (16, 16, 5)) for _ in range(4)]
    f[2][3, 11, 4] = 7.0 - 2.0j
    prev = [float(np.max(np.abs(x))) for x in f]
    d = 2.5e-9
    out = C.conv(np.ones(8), np.ones(8), [x * (1 + d) for x in f], prev, 1, 1, 8)
    assert abs(out["f_max"][2] - abs(7.0 - 2.0j) * (1 + d)) < 1e-14
    assert all(abs(r - d) < 1e-15 for r in out["f_rel_err"])
    assert out["converged"]  # 0 < 1e-8 and 2.5e-9 < 1e-8
    out = C.conv(np.ones(8), np.ones(8), [x * (1 + 2e-8) for x in f], prev, 1, 1, 8)
    assert not out["converged"]


def test_first_call_is_not_converged():
    out = C.conv(np.ones(8), np.ones(8), [np.ones(3)] * 4, [0.0] * 4, 1, 1, 8)
    assert out["f_rel_err"] == [float("inf")] * 4 and not out["converged"]
