"""The driver's command line and failure paths (driver.cpp:91-196), on the CPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "3ddg_amd", "dg3d_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def run(args, cwd):
    return subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


def test_wrong_argument_count(tmp_path):
    r = run([], tmp_path)
    assert r.returncode == 1
    assert "Not enough command line arguments" in r.stderr


def test_missing_parameter_file(tmp_path):
    r = run([str(tmp_path / "nope.txt"), str(tmp_path / "out")], tmp_path)
    assert r.returncode == 3
    assert "Unable to open parameter file" in r.stderr


def test_unwritable_log(tmp_path):
    r = run([os.path.join(GOLDEN, "parameters1.txt"), "/proc/no_such_dir/out"], tmp_path)
    assert r.returncode == 4
    assert "Unable to open log file" in r.stderr


def test_log_echo_and_unknown_keys(tmp_path):
    pf = tmp_path / "p.txt"
    pf.write_text(open(os.path.join(GOLDEN, "parameters1.txt")).read() + "NOT_A_KEY = 3\nORDER = 9\n")
    r = run([str(pf), str(tmp_path / "out")], tmp_path)
    log = (tmp_path / "out" / "_log.txt").read_text()
    assert log.startswith("Parameters used:\n{\n") and log.rstrip().endswith("}")
    assert "  NUM_BLOCKS = 2 2 1" in log
    assert 'WARNING: IGNORING INVALID PARAMETER LINE: "NOT_A_KEY = 3"' in r.stderr
    assert r.returncode == 2  # ORDER 9 is outside this build's orders


def test_missing_basis_files(tmp_path):
    """Without hardcoded_coeff.txt/matrices.txt in the working directory or next to the driver
    the run stops like the reference's Basis constructor (Basis.cpp:106-127)."""
    import shutil
    d = tmp_path / "pkg"
    d.mkdir()
    shutil.copy(DRIVER, d / "dg3d_driver")
    lib = os.path.join(ROOT, "3ddg_amd", "libdg3d.so")
    shutil.copy(lib, d / "libdg3d.so")
    r = subprocess.run([str(d / "dg3d_driver"), os.path.join(GOLDEN, "parameters1.txt"), str(tmp_path / "o")],
                       cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 3
    assert "hardcoded_coeff.txt" in r.stderr
