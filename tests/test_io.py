"""The reference's file formats on the host (no GPU): the readers against an
independent parse of the reference's own data files, the name rule of
InputRW::getPos, and the config-free CLI argument handling."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import golden_path, read_points

DATA = [f for f in sorted(os.listdir(os.path.dirname(golden_path("x"))))
        if f.endswith((".2d", ".3d", ".tsp"))]


@pytest.mark.parametrize("name", DATA)
def test_reader_matches_independent_parse(pkg, name):
    want = read_points(golden_path(name))
    dim = want.shape[1]
    got = pkg.read_points(golden_path(name), dim)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_2d_data_read_by_a_3d_solution_gets_z_zero(pkg):
    want = read_points(golden_path("1000randomCity.2d"))
    got = pkg.read_points(golden_path("1000randomCity.2d"), 3)
    assert got.shape == (1000, 3)
    assert np.array_equal(got[:, :2], want) and np.all(got[:, 2] == 0)


def test_tsp_read_by_a_3d_solution_gets_z_zero(pkg):
    got = pkg.read_points(golden_path("dj38.tsp"), 3)
    assert got.shape == (38, 3) and np.all(got[:, 2] == 0)
    assert got[0, 0] == np.float32(11003.6111) and got[0, 1] == np.float32(42102.5)


def test_points_list_reader(pkg):
    """.points: 'Width = W Height = H' then W*H points (read here with the
    3 coordinates the writer puts per point)."""
    got = pkg.read_points(golden_path("outypu.points"), 3)
    toks = open(golden_path("outypu.points")).read().split()
    assert toks[:6] == ["Width", "=", "3000", "Height", "=", "1"]
    want = np.array(toks[6:], dtype=np.float32).reshape(3000, 3)
    assert np.array_equal(got, want)


def test_name_rule_and_errors(pkg, tmp_path):
    # getPos cuts at the first '.' or '_': "a_b.2d" is read as TSPLIB "a.tsp"
    shutil.copyfile(golden_path("1000randomCity.2d"), tmp_path / "a_b.2d")
    with pytest.raises(pkg.EmstError) as ei:
        pkg.read_points(str(tmp_path / "a_b.2d"), 2)
    assert ei.value.code == 8
    shutil.copyfile(golden_path("dj38.tsp"), tmp_path / "a.tsp")
    assert pkg.read_points(str(tmp_path / "a_b.2d"), 2).shape == (38, 2)
    # a .3d file needs a 3D solution; a missing file is an I/O error
    with pytest.raises(pkg.EmstError) as ei:
        pkg.read_points(golden_path("1000randomCity.3d"), 2)
    assert ei.value.code == 1
    with pytest.raises(pkg.EmstError) as ei:
        pkg.read_points(str(tmp_path / "missing.2d"), 2)
    assert ei.value.code == 8
    # malformed TSPLIB (no DIMENSION on line 4)
    (tmp_path / "bad.tsp").write_text("NAME : bad\nTYPE : TSP\nCOMMENT : x\nEDGE_WEIGHT_TYPE : EUC_2D\n")
    with pytest.raises(pkg.EmstError) as ei:
        pkg.read_points(str(tmp_path / "bad.tsp"), 2)
    assert ei.value.code == 8


def test_cli_reports_unsupported_modes(pkg, tmp_path):
    """Modes the reference handles outside the EMST path are reported, as its
    CalculateurEMST::run does for unknown modes, before any device work."""
    assert os.path.exists(pkg.CLI_PATH), "emst_cli is not built"
    cfg = tmp_path / "config.cfg"
    cfg.write_text("[coalition_global_param]\nfunctionModeChoice = 2\n")
    r = subprocess.run([pkg.CLI_PATH, golden_path("1000randomCity.2d"), str(tmp_path / "o"),
                        str(tmp_path / "o.stats"), str(cfg)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "UNSUPPORTED FUNCTIONMODE=2 !!" in r.stdout


@pytest.mark.parametrize("name,text", [
    ("huge.2d", "99999999999999\n1 2\n3 4\n"),
    ("neg.3d", "-5\n1 2 3\n"),
    ("short.2d", "3\n1 2\n3 4\n"),
    ("big.points", "Width = 4000000000 Height = 4000000000\n1 2 3\n"),
    ("over.points", "Width = 3037000500 Height = 3037000500\n1 2 3\n"),
])
def test_malformed_headers_are_file_errors(pkg, tmp_path, name, text):
    """A header claiming more points than the file can hold (or a negative or
    overflowing count) is a file error, not an allocation or a crash."""
    path = tmp_path / name
    path.write_text(text)
    dim = 2 if name.endswith(".2d") else 3
    with pytest.raises(pkg.EmstError) as e:
        pkg.read_points(str(path), dim)
    assert e.value.code == 8
