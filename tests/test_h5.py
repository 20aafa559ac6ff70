"""The HDF5 vertex dump (libtups_h5.so) read back by the HDF5 library's own h5dump:
the layout write_F produces (SRC/parquet_plot.f90:237-276 through hdf5_write_data,
SRC/hdf5_wrapper.f90:285-287,2410-2445): group F, a compound {"r", "i"} of IEEE
doubles, the Fortran shape (Nf, Nl, Nf, Nl, Nb) as the C dataspace
(Nb, Nl, Nf, Nl, Nf), and the vertex bytes in Fortran order.  Host code only."""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if Path("/opt/conda/bin/h5dump").exists() else None)
LIB = Path(__file__).resolve().parents[1] / "tups_amd" / "libtups_h5.so"

pytestmark = pytest.mark.skipif(H5DUMP is None or not LIB.exists(),
                                reason="needs h5dump and libtups_h5.so (the HDF5 C library)")


def _vertex(rng, nf, nl, nb):
    shape = (nf, nl, nf, nl, nb)
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def test_layout_and_bytes(tmp_path):
    from tups_amd.runtime import h5_write_vertex

    rng = np.random.default_rng(0)
    nf, nl, nb = 4, 2, 3
    f = tmp_path / "F.hdf5"
    arrays = {}
    for i, name in enumerate(("F_d", "F_m", "F_s", "F_t")):
        arrays[name] = _vertex(rng, nf, nl, nb)
        h5_write_vertex(f, "F/" + name, arrays[name], create=(i == 0))
    head = subprocess.run([H5DUMP, "-H", str(f)], capture_output=True, text=True, check=True).stdout
    assert 'GROUP "F"' in head
    for name in arrays:
        assert f'DATASET "{name}"' in head
    assert re.search(r'H5T_COMPOUND\s*\{\s*H5T_IEEE_F64LE\s+"r";\s*H5T_IEEE_F64LE\s+"i";\s*\}', head)
    assert head.count(f"( {nb}, {nl}, {nf}, {nl}, {nf} )") == 2 * len(arrays)
    for name, a in arrays.items():
        txt = subprocess.run([H5DUMP, "-d", f"/F/{name}", "-m", "%.17g", "-w", "0", str(f)], capture_output=True,
                             text=True, check=True).stdout
        body = re.sub(r"\([0-9,]+\):", "", txt[txt.index("DATA {") + 6:])
        vals = np.array([float(x) for x in re.findall(r"[-+]?[0-9][0-9.eE+-]*", body)])
        got = vals[0::2] + 1j * vals[1::2]  # the dataspace in C order: the Fortran memory order
        assert np.array_equal(got, a.ravel(order="F")), name


def test_existing_dataset_is_an_error(tmp_path):
    from tups_amd import TupsError
    from tups_amd.runtime import h5_write_vertex

    a = _vertex(np.random.default_rng(1), 2, 1, 2)
    f = tmp_path / "x.hdf5"
    h5_write_vertex(f, "F/F_d", a, create=True)
    with pytest.raises(TupsError):
        h5_write_vertex(f, "F/F_d", a, create=False)
    with pytest.raises(TupsError):
        h5_write_vertex(tmp_path / "absent" / "y.hdf5", "F/F_d", a, create=True)


def test_library_exports_every_declared_symbol():
    import ctypes

    from tups_amd.runtime import load_library

    load_library()
    lib = ctypes.CDLL(str(LIB))
    header = (Path(__file__).resolve().parents[1] / "include" / "tups_h5.h").read_text()
    names = re.findall(r"\b(tups_h5_\w+)\s*\(", header)
    assert set(names) == {"tups_h5_write_vertex", "tups_h5_write_f", "tups_h5_last_error"}
    for n in names:
        assert hasattr(lib, n), n
