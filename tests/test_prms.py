"""The run-parameter reader (tups_read_prms) against the reference's own parameter
file (tests/golden/prms_reference, the `prms` shipped with the reference) and
hand-written files that exercise the list-directed rules of set_parameters
(SRC/parquet_ini.f90:289-329).  Host code only: no GPU call is made."""
from pathlib import Path

import pytest

from tups_amd import TupsError, read_prms

GOLDEN = Path(__file__).resolve().parent / "golden"

REFERENCE_VALUES = {
    "beta": 5.0, "u": 2.0, "n_particle": 1.0, "nx": 16, "nf": 64, "neighbours": 1, "readinpick": False,
    "pick_file": "inidata/<name_of_list>", "ngrain": 1, "f_range": 0, "f_damping": 0.5, "ite_max": 100,
    "dga": False, "fname_l": "inidata/<name_of_Lambda_file>", "use_sig_dmft": False,
    "fname_sig_dmft": "inidata/<name_of_DMFT_Sigma_file>", "use_u2": True, "save_f": False, "oneshot": False,
    "old": False, "sde_old": False,
}

RECORDS = ["5.0d0", "2.0d0", "1.00d0", "16", "64", "1", ".false.", '"inidata/<name_of_list>"', "1", "0", "0.5d0",
           "100", ".false.", '"inidata/<name_of_Lambda_file>"', ".false.", '"inidata/<name_of_DMFT_Sigma_file>"',
           ".true.", ".false.", ".false.", ".false.", ".false."]


def _write(tmp_path, records, name="prms"):
    p = tmp_path / name
    p.write_text("\n".join(records) + "\n")
    return p


def test_reference_parameter_file():
    got = read_prms(GOLDEN / "prms_reference")
    for k, v in REFERENCE_VALUES.items():
        assert got[k] == v, k
    assert got["dims"] == {"nx": 16, "nf": 64, "nl": 1, "beta": 5.0, "f_damping": 0.5}


def test_config1_parameter_file():
    got = read_prms(GOLDEN / "prms_config1")
    assert (got["beta"], got["u"], got["n_particle"], got["nx"], got["nf"], got["neighbours"]) == (2.0, 2.0, 1.0, 4,
                                                                                                  16, 1)
    assert got["use_u2"] and got["ite_max"] == 1 and got["f_damping"] == 0.5
    assert got["dims"]["nl"] == 1


def test_list_directed_forms(tmp_path):
    rec = list(RECORDS)
    rec[0] = "  5.0D0, 7.0   ! comma ends the value"
    rec[1] = "2e0"
    rec[2] = "3*0.5E-0 / slash ends the record"
    rec[3] = "+8"
    rec[5] = "3"
    rec[6] = "T"
    rec[7] = "'it''s here' trailing words"
    rec[9] = "2,"
    rec[12] = ".TRUE."
    rec[13] = "bare_name,rest"
    rec[16] = "false"
    rec.insert(4, "")  # a blank record is skipped
    got = read_prms(_write(tmp_path, rec))
    assert (got["beta"], got["u"], got["n_particle"], got["nx"], got["nf"]) == (5.0, 2.0, 0.5, 8, 64)
    assert got["neighbours"] == 3 and got["readinpick"] and got["pick_file"] == "it's here"
    assert got["f_range"] == 2 and got["dga"] and got["fname_l"] == "bare_name" and not got["use_u2"]
    assert got["dims"] is None  # a pick list is not supported


def test_neighbours_to_form_factors(tmp_path):
    for nb, nl in ((1, 1), (2, 5), (3, 9)):
        rec = list(RECORDS)
        rec[5] = str(nb)
        assert read_prms(_write(tmp_path, rec, f"p{nb}"))["dims"]["nl"] == nl
    rec = list(RECORDS)
    rec[5] = "4"
    assert read_prms(_write(tmp_path, rec, "p4"))["dims"] is None
    rec = list(RECORDS)
    rec[3], rec[5] = "2", "2"  # on a 2x2 lattice the (1,0) shell lies on the zone edge
    assert read_prms(_write(tmp_path, rec, "p22"))["dims"] is None


def test_long_character_value_is_cut_to_100(tmp_path):
    rec = list(RECORDS)
    rec[13] = '"' + "x" * 150 + '"'
    assert read_prms(_write(tmp_path, rec))["fname_l"] == "x" * 100


@pytest.mark.parametrize("index,value,field", [(3, "16.0", "Nx"), (0, "beta", "beta"), (6, "yes", "readinpick"),
                                               (4, ",", "Nf"), (7, '"open', "pick_file")])
def test_bad_values_name_line_and_field(tmp_path, index, value, field):
    rec = list(RECORDS)
    rec[index] = value
    with pytest.raises(TupsError) as ei:
        read_prms(_write(tmp_path, rec))
    assert f"prms line {index + 1} ({field})" in str(ei.value)


def test_short_or_missing_file(tmp_path):
    with pytest.raises(TupsError) as ei:
        read_prms(_write(tmp_path, RECORDS[:10]))
    assert "f_damping" in str(ei.value) and "ends" in str(ei.value)
    with pytest.raises(TupsError):
        read_prms(tmp_path / "absent")
