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
