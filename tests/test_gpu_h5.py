"""write_F from the device vertices (tups_h5_write_f): the four resident F_r land in
F/F_d .. F/F_t with the values that were uploaded, as h5dump reads them back."""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.helpers import crandn

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if Path("/opt/conda/bin/h5dump").exists() else None)
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(H5DUMP is None, reason="needs h5dump")]


def test_write_f_round_trip(tmp_path):
    from tups_amd import F, TupsContext

    rng = np.random.default_rng(9)
    with TupsContext(nx=4, nf=8, nl=5, beta=2.0, f_damping=0.5) as c:
        fv = [crandn(rng, (c.nz, c.nz, c.nq), 0.5) for _ in range(4)]
        for r in range(4):
            c.upload(F, r, fv[r])
        path = tmp_path / "F_test_000.hdf5"
        c.write_f(path)
        nf, nl, nq = c.nf, c.nl, c.nq
    head = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True).stdout
    assert head.count(f"( {nq}, {nl}, {nf}, {nl}, {nf} )") == 8
    for r, name in enumerate(("F_d", "F_m", "F_s", "F_t")):
        txt = subprocess.run([H5DUMP, "-d", f"/F/{name}", "-m", "%.17g", "-w", "0", str(path)],
                             capture_output=True, text=True, check=True).stdout
        body = re.sub(r"\([0-9,]+\):", "", txt[txt.index("DATA {") + 6:])
        vals = np.array([float(x) for x in re.findall(r"[-+]?[0-9][0-9.eE+-]*", body)])
        assert np.array_equal(vals[0::2] + 1j * vals[1::2], fv[r].ravel(order="F")), name
