"""The C-ABI library loads, exports every symbol of include/tups_gpu.h, and its
device-free paths (tables, argument validation) behave — no GPU needed."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle.lattice import Lattice

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "tups_gpu.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*\s*|int\s+)(tups_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from tups_amd import load_library
    from tups_amd.runtime import SIGNATURES

    lib = load_library()
    syms = _declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(SIGNATURES), set(syms) ^ set(SIGNATURES)


@pytest.mark.parametrize("nx,nf,nl", [(4, 8, 1), (4, 8, 5), (4, 8, 9), (8, 16, 5), (16, 8, 9), (6, 4, 9)])
def test_native_tables_match_oracle(nx, nf, nl):
    from tups_amd import build_tables

    ff, ffi, signs, reps, esq, eqs = build_tables(nx, nf, nl)
    lat = Lattice(nx, nf, nl)
    assert np.abs(ff - lat.ff).max() < 1e-15
    assert np.abs(ffi - lat.ff_inv).max() < 1e-15
    assert (signs == lat.signs).all() and (reps == lat.reps).all()
    assert np.abs(esq - lat.expSqR).max() < 1e-15
    assert np.abs(eqs - lat.expqSR).max() < 1e-15


@pytest.mark.parametrize("dims", [(3, 8, 1), (4, 7, 1), (4, 8, 4), (4, 8, 5, 9)])
def test_invalid_dims_rejected_before_device_use(dims):
    from tups_amd.runtime import Dims, _V, load_library

    lib = load_library()
    nx, nf, nl = dims[:3]
    nr = dims[3] if len(dims) > 3 else 0
    d = Dims(nx, nf, nl, nr, 0, 0, 2.0, 0.5)
    h = _V()
    assert lib.tups_create(C.byref(d), C.byref(h)) == -1
    assert lib.tups_last_error()
    assert lib.tups_build_tables(C.byref(d), None, None, None, None, None, None) == -1


def test_fortran_interface_declares_every_symbol():
    text = (ROOT / "tups_amd" / "fortran" / "tups_gpu_iface.f90").read_text().lower()
    for s in _declared_symbols():
        assert f'name="{s}"' in text, s
