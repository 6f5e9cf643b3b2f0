"""The Fortran drop-in driver (wenostandalone_amd/weno_standalone_mi355) run
the way the reference's `wenoss.sh test` runs its binary (wenoss.sh:330-387):
weno5 mode, full scope, warmup 0, bench 1, default 200^3 namelist, the four
`Checksum X/Y/Z/T` lines compared with tests/golden/weno5.txt at rtol 1e-10.
Every WENO5 route of the module (generic, split, specialised, fused) must
reproduce the golden."""
import os
import re
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "wenostandalone_amd", "weno_standalone_mi355")
NML = """&grid_params
  m_cells = {m}, n_cells = {m}, p_cells = {m}, v_size = 8, buff_size = 6,
  weno_eps = 1e-16, wenojs = T, teno_CT = 1e-5
/
"""


def run_driver(tmp_path, env_extra, m=199):
    nml = tmp_path / "weno_input.nml"
    nml.write_text(NML.format(m=m))
    env = dict(os.environ, WENO_KERNEL_MODE="weno5", WENO5_BENCH_SCOPE="full", WENO5_SPLIT_KERNELS="0",
               WENO5_SPECIALIZED_COMBINED="0", WENO_WARMUP_ITERS="0", WENO_BENCH_ITERS="1")
    env.update(env_extra)
    out = subprocess.run([DRIVER, str(nml)], env=env, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout + out.stderr
    sums = dict(re.findall(r"^  Checksum ([XYZT]) :\s+(\S+)", out.stdout, flags=re.M))
    return {k: float(v) for k, v in sums.items()}, out.stdout


@pytest.mark.parametrize("variant", [
    {},
    {"WENO5_SPLIT_KERNELS": "1"},
    {"WENO5_SPECIALIZED_COMBINED": "1"},
    {"WENO5_FUSED": "1"},
])
def test_driver_golden(tmp_path, variant):
    gold = {k: float(v) for k, v in (l.split() for l in open(os.path.join(GOLDEN, "weno5.txt")))}
    got, log = run_driver(tmp_path, variant)
    for k in "XYZT":
        assert abs(got[k] - gold[k]) <= 1e-10 * abs(gold[k]), (variant, k, got[k], gold[k])
    assert "All kernel families executed with finite non-zero output." in log


def test_driver_all_mode_small(tmp_path):
    got, log = run_driver(tmp_path, {"WENO_KERNEL_MODE": "all", "WENO_BENCH_ITERS": "2"}, m=15)
    assert "All kernel families executed with finite non-zero output." in log
    assert got["T"] > 0


def _wenoss(*args, **env_extra):
    import sys
    env = dict(os.environ, **env_extra)
    return subprocess.run([sys.executable, "-m", "wenostandalone_amd.wenoss", *args], cwd=ROOT, env=env,
                          capture_output=True, text=True, timeout=300)


def test_wenoss_test_against_the_reference_golden():
    """`wenoss test` (wenoss.sh:340-387) with the package's default namelist
    (the reference's 200^3 x 8 configuration) against tests/golden/weno5.txt."""
    r = _wenoss("test")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("OK  ") == 4


def test_wenoss_update_then_test(tmp_path):
    r = _wenoss("test-update", TEST_GOLDEN_DIR=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    new = {k: float(v) for k, v in (l.split() for l in open(tmp_path / "weno5.txt"))}
    gold = {k: float(v) for k, v in (l.split() for l in open(os.path.join(GOLDEN, "weno5.txt")))}
    for k in "XYZT":
        assert abs(new[k] - gold[k]) <= 1e-10 * abs(gold[k])
    r = _wenoss("test", TEST_GOLDEN_DIR=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr


def test_missing_namelist_runs_the_defaults(tmp_path):
    """A missing input file keeps the module defaults (:120-124): the run is
    the reference's default configuration and reproduces its golden."""
    out = subprocess.run([DRIVER, str(tmp_path / "nothing_here.nml")], capture_output=True, text=True, timeout=240,
                         env=dict(os.environ, WENO_KERNEL_MODE="weno5", WENO_WARMUP_ITERS="0", WENO_BENCH_ITERS="1"))
    assert out.returncode == 0, out.stdout + out.stderr
    assert "not found; using defaults." in out.stdout
    gold = {k: float(v) for k, v in (l.split() for l in open(os.path.join(GOLDEN, "weno5.txt")))}
    sums = {k: float(v) for k, v in re.findall(r"^  Checksum ([XYZT]) :\s+(\S+)", out.stdout, flags=re.M)}
    assert abs(sums["T"] - gold["T"]) <= 1e-10 * gold["T"]


def test_no_weight_branch_warns_and_is_not_a_golden(tmp_path):
    """wenojs = F with no other branch leaves alpha unset: the reference's
    outputs are 0/0 = NaN, its driver prints the non-finite warning
    (:1848-1852) and test-update refuses the run (wenoss.sh:389-414)."""
    nml = tmp_path / "nobranch.nml"
    nml.write_text(NML.format(m=15).replace("wenojs = T", "wenojs = F"))
    out = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=240,
                         env=dict(os.environ, WENO_KERNEL_MODE="weno5", WENO_WARMUP_ITERS="0", WENO_BENCH_ITERS="1"))
    assert out.returncode == 0, out.stdout + out.stderr
    assert "WARNING: non-finite values detected in standalone output." in out.stdout
    r = _wenoss("test-update", INPUT_NML=str(nml), TEST_GOLDEN_DIR=str(tmp_path / "g"))
    assert r.returncode == 1 and "refusing to write golden" in r.stderr
    assert not (tmp_path / "g" / "weno5.txt").exists()


def test_driver_metrics_json(tmp_path):
    import json
    nml = tmp_path / "small.nml"
    nml.write_text(NML.format(m=31))
    r = _wenoss("metrics", INPUT_NML=str(nml), WENO_WARMUP_ITERS="1", WENO_BENCH_ITERS="2")
    assert r.returncode == 0, r.stdout + r.stderr
    data = json.loads(r.stdout.strip().splitlines()[-1])
    assert data["status"] == "ok" and data["m_cells"] == 31 and data["bench_iters"] == 2
    assert data["kernel_mode"] == "weno5" and data["wenojs"] is True
    assert abs(data["checksum_t"] - (data["checksum_x"] + data["checksum_y"] + data["checksum_z"])) <= \
        1e-12 * data["checksum_t"]
    assert data["kernel_time_per_iter_s"] > 0
