"""The wenoss command-line mirror (wenostandalone_amd/wenoss.py) and the
harness's input paths, on the CPU.

* `test` / `test-update` against a stand-in binary that prints the harness's
  checksum lines: a run within rtol passes, one outside fails, a NaN fails,
  test-update writes the golden file in the reference's format and refuses
  an all-zero or non-finite Checksum T or a run without one
  (wenoss.sh:309-414).
* s_read_input (inputs/weno_standalone.f90:111-144) through the drop-in
  module, compiled into a small probe program (no GPU is touched): a missing
  file keeps the defaults and says so, a malformed namelist prints the
  warning, a good one is read."""
import os
import shutil
import stat
import subprocess
import sys

import pytest

from tests.conftest import GOLDEN, ROOT
from wenostandalone_amd import wenoss

FAKE = """#!/bin/bash
# stand-in for the harness: echoes its switches, prints checksum lines
echo "  Kernel mode              : $WENO_KERNEL_MODE"
echo "  Warmup iterations        : $WENO_WARMUP_ITERS"
echo "  Checksum X :   6.5386334223E+08"
echo "  Checksum Y :   6.5400560771E+08"
echo "  Checksum Z :   6.5339390676E+08"
if [ -n "$FAKE_T" ]; then echo "  Checksum T :   $FAKE_T"; fi
exit ${FAKE_RC:-0}
"""


@pytest.fixture
def fake_bin(tmp_path, monkeypatch):
    p = tmp_path / "fake_harness"
    p.write_text(FAKE)
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    monkeypatch.setenv("BIN", str(p))
    monkeypatch.setenv("INPUT_NML", str(tmp_path / "in.nml"))
    monkeypatch.setenv("TEST_GOLDEN_DIR", str(tmp_path / "golden"))
    return p


def _cli(*args, env=None):
    return subprocess.run([sys.executable, "-m", "wenostandalone_amd.wenoss", *args], cwd=ROOT,
                          capture_output=True, text=True, env=env)


def test_extract_and_compare():
    out = "  Checksum X :   1.0E+00\n  Checksum Y :   2.0E+00\n  Checksum Z :   3.0E+00\n  Checksum T :   6.0E+00\n"
    chk = wenoss.extract_checksums(out)
    assert chk == {"X": "1.0E+00", "Y": "2.0E+00", "Z": "3.0E+00", "T": "6.0E+00"}
    lines, ok = wenoss.compare(chk, {"X": 1.0, "Y": 2.0, "Z": 3.0, "T": 6.0}, 1e-10)
    assert ok and all(l.endswith("OK  ") for l in lines)
    lines, ok = wenoss.compare(chk, {"X": 1.0, "Y": 2.0, "Z": 3.0, "T": 6.0 * (1 + 2e-10)}, 1e-10)
    assert not ok and lines[3].endswith("FAIL")
    lines, ok = wenoss.compare({**chk, "T": "NaN"}, {"X": 1.0, "Y": 2.0, "Z": 3.0, "T": 6.0}, 1e-10)
    assert not ok
    lines, ok = wenoss.compare({"X": "1.0"}, {"X": 1.0, "Y": 2.0, "Z": 3.0, "T": 6.0}, 1e-10)
    assert not ok


@pytest.mark.parametrize("t,refused", [("1.9612628567E+09", False), ("0.0000000000E+00", True), ("NaN", True),
                                       ("Infinity", True), (None, True)])
def test_golden_refusal(t, refused):
    chk = {"X": "1", "Y": "1", "Z": "1"}
    if t is not None:
        chk["T"] = t
    assert (wenoss.golden_refusal(chk) is not None) == refused


def test_cli_test_and_update(fake_bin, tmp_path):
    env = dict(os.environ)
    golden = tmp_path / "golden" / "weno5.txt"
    # no golden yet
    r = _cli("test", env=env)
    assert r.returncode == 1 and "golden file not found" in r.stderr
    # update refuses a missing, zero or NaN total, and writes nothing
    for t in (None, "0.0000000000E+00", "NaN"):
        e = dict(env)
        if t is not None:
            e["FAKE_T"] = t
        r = _cli("test-update", env=e)
        assert r.returncode == 1 and ("refusing" in r.stderr or "no 'Checksum T'" in r.stderr), (t, r.stderr)
        assert not golden.exists()
    # a good run becomes the golden, in the reference's "<label> <value>" form
    env["FAKE_T"] = "1.9612628567E+09"
    r = _cli("test-update", env=env)
    assert r.returncode == 0, r.stderr
    assert golden.read_text().split("\n")[3] == "T 1.9612628567E+09"
    assert "Warmup iterations        : 0" in r.stdout     # pinned test settings
    # the same run passes; a different total fails; a failing binary fails
    r = _cli("test", env=env)
    assert r.returncode == 0 and "test passed" in r.stderr, r.stdout + r.stderr
    r = _cli("test", env=dict(env, FAKE_T="1.9612628600E+09"))
    assert r.returncode == 1 and "FAIL" in r.stdout
    r = _cli("test", env=dict(env, FAKE_RC="3"))
    assert r.returncode == 1 and "exited with status 3" in r.stderr
    # the committed golden is the reference's, and the fake prints its X/Y/Z
    r = _cli("test", env=dict(env, TEST_GOLDEN_DIR=GOLDEN))
    assert r.returncode == 0, r.stdout


def test_cli_help_and_unknown():
    r = _cli("help")
    assert r.returncode == 0 and "test-update" in r.stdout
    r = _cli("no-such-command")
    assert r.returncode == 1 and "unknown subcommand" in r.stderr


FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


@pytest.fixture(scope="module")
def nml_probe(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("amdflang is not available")
    lib = os.path.join(ROOT, "wenostandalone_amd", "libweno_mi355.so")
    if not os.path.exists(lib):
        pytest.skip("libweno_mi355.so is not built")
    d = tmp_path_factory.mktemp("nml")
    src = [os.path.join(ROOT, "wenostandalone_amd", "fortran", f) for f in ("m_weno_c.f90", "m_weno_standalone.f90")]
    src.append(os.path.join(ROOT, "tests", "fortran", "nml_probe.f90"))
    exe = str(d / "nml_probe")
    subprocess.run([FLANG, "-O0", "-module-dir", str(d), "-o", exe, *src,
                    "-L" + os.path.dirname(lib), "-lweno_mi355", "-Wl,-rpath," + os.path.dirname(lib)],
                   check=True, capture_output=True, text=True)
    return exe, d


def _probe(exe, path):
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    vals = [l for l in r.stdout.splitlines() if l.startswith("VALUES")][0].split()[1:]
    return r.stdout, vals


def test_missing_namelist_keeps_defaults(nml_probe):
    exe, d = nml_probe
    out, vals = _probe(exe, d / "does_not_exist.nml")
    assert 'not found; using defaults.' in out
    assert vals[:5] == ["199", "199", "199", "8", "6"] and vals[6] == "T"


def test_malformed_namelist_warns(nml_probe):
    exe, d = nml_probe
    bad = d / "bad.nml"
    bad.write_text("&grid_params\n m_cells = 31, n_cells = oops,\n/\n")
    out, vals = _probe(exe, bad)
    assert "WARNING: error reading namelist grid_params; using defaults." in out


def test_good_namelist_is_read(nml_probe):
    exe, d = nml_probe
    good = d / "good.nml"
    good.write_text("&grid_params\n m_cells = 31, n_cells = 15, p_cells = 7, v_size = 3, buff_size = 4,\n"
                    " weno_eps = 1e-12, wenojs = F, teno_CT = 2e-5\n/\n")
    out, vals = _probe(exe, good)
    assert vals[:5] == ["31", "15", "7", "3", "4"] and vals[6] == "F"
    assert float(vals[5]) == 1e-12 and float(vals[7]) == 2e-5
    assert "m_cells (from input)     : 31" in out
