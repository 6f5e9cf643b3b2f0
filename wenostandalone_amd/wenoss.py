"""Command-line driver of the MI355X harness binary, mirroring the reference's
`wenoss.sh` (build / run / profile / correctness test) for
`weno_standalone_mi355`.

    python -m wenostandalone_amd.wenoss <subcommand>

Subcommands (wenoss.sh:15-28, dispatch :418-435):
  build               compile the library and the harness binary (make)
  run | benchmark     launch the binary with the WENO_* switches
  profile-weno5       rocprofv3 kernel trace + stats of the binary
  list-pmc            the counters rocprofv3 can collect on this GPU
  test | check        run the binary pinned to the test settings and compare
                      its Checksum X/Y/Z/T lines with the golden file
  test-update         rewrite the golden file from a pinned run; refuses an
                      all-zero or non-finite Checksum T
  metrics             run as `run` does and print the binary's JSON metrics
  help                this message

Every knob is an environment variable with the reference's name and default
(wenoss.sh:136-166): INPUT_NML, WENO_KERNEL_MODE, WENO5_BENCH_SCOPE,
WENO5_SPLIT_KERNELS, WENO5_SPECIALIZED_COMBINED, WENO_WARMUP_ITERS,
WENO_BENCH_ITERS, TEST_GOLDEN_DIR, WENO_TEST_RTOL, WENO_TEST_KERNEL_MODE,
WENO_TEST_WARMUP_ITERS, WENO_TEST_BENCH_ITERS, ROCPROFV3, ROCPROFV3_ARGS,
ROCPROFV3_OUTPUT_FORMAT, ROCPROFV3_OUTPUT_DIR, ROCPROFV3_PMC_ARGS, RUNNER,
RESULTS_DIR.  BIN selects the binary (default: the in-tree
weno_standalone_mi355).  There is no compiler argument: one toolchain
builds this path.
"""
from __future__ import annotations

import json
import math
import os
import re
import shlex
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def env(name: str, default: str) -> str:
    v = os.environ.get(name, "")
    return v if v else default


def settings() -> dict:
    results = env("RESULTS_DIR", os.path.join(ROOT, "results"))
    return {
        "BIN": env("BIN", os.path.join(HERE, "weno_standalone_mi355")),
        "INPUT_NML": env("INPUT_NML", os.path.join(HERE, "weno_input.nml")),
        "RUNNER": env("RUNNER", ""),
        "WENO_KERNEL_MODE": env("WENO_KERNEL_MODE", "weno5"),
        "WENO5_BENCH_SCOPE": env("WENO5_BENCH_SCOPE", "full"),
        "WENO5_SPLIT_KERNELS": env("WENO5_SPLIT_KERNELS", "0"),
        "WENO5_SPECIALIZED_COMBINED": env("WENO5_SPECIALIZED_COMBINED", "0"),
        "WENO_WARMUP_ITERS": env("WENO_WARMUP_ITERS", "5"),
        "WENO_BENCH_ITERS": env("WENO_BENCH_ITERS", "10"),
        "TEST_GOLDEN_DIR": env("TEST_GOLDEN_DIR", os.path.join(ROOT, "tests", "golden")),
        "WENO_TEST_RTOL": env("WENO_TEST_RTOL", "1e-10"),
        "WENO_TEST_KERNEL_MODE": env("WENO_TEST_KERNEL_MODE", "weno5"),
        "WENO_TEST_WARMUP_ITERS": env("WENO_TEST_WARMUP_ITERS", "0"),
        "WENO_TEST_BENCH_ITERS": env("WENO_TEST_BENCH_ITERS", "1"),
        "RESULTS_DIR": results,
        "ROCPROFV3": env("ROCPROFV3", "rocprofv3"),
        "ROCPROFV3_ARGS": env("ROCPROFV3_ARGS", "--kernel-trace --stats"),
        "ROCPROFV3_OUTPUT_FORMAT": env("ROCPROFV3_OUTPUT_FORMAT", "csv"),
        "ROCPROFV3_OUTPUT_DIR": env("ROCPROFV3_OUTPUT_DIR", os.path.join(results, "rocprofv3_weno5")),
        "ROCPROFV3_PMC_ARGS": env("ROCPROFV3_PMC_ARGS", ""),
    }


def log(msg: str) -> None:
    print(f">> {msg}", file=sys.stderr, flush=True)


class Die(SystemExit):
    """Fatal error: message on stderr, exit status 1 (wenoss.sh's die)."""

    def __init__(self, msg: str):
        print(f"!! {msg}", file=sys.stderr, flush=True)
        super().__init__(1)


RUN_KEYS = ("WENO_KERNEL_MODE", "WENO5_BENCH_SCOPE", "WENO5_SPLIT_KERNELS", "WENO5_SPECIALIZED_COMBINED",
            "WENO_WARMUP_ITERS", "WENO_BENCH_ITERS")


def run_env(s: dict) -> dict:
    e = dict(os.environ)
    for k in RUN_KEYS:
        e[k] = s[k]
    return e


def pinned_env(s: dict) -> dict:
    """The test settings (wenoss.sh:321-335): deterministic, one iteration."""
    e = dict(os.environ)
    e.update(WENO_KERNEL_MODE=s["WENO_TEST_KERNEL_MODE"], WENO5_BENCH_SCOPE="full", WENO5_SPLIT_KERNELS="0",
             WENO5_SPECIALIZED_COMBINED="0", WENO_WARMUP_ITERS=s["WENO_TEST_WARMUP_ITERS"],
             WENO_BENCH_ITERS=s["WENO_TEST_BENCH_ITERS"])
    return e


def command(s: dict) -> list:
    return shlex.split(s["RUNNER"]) + [s["BIN"], s["INPUT_NML"]]


def check_bin(s: dict) -> None:
    if not os.path.exists(s["BIN"]):
        raise Die(f"binary not found: {s['BIN']}. Build it with: python -m wenostandalone_amd.wenoss build")


# -- checksum parsing and comparison (wenoss.sh:340-387) ----------------------
CHECKSUM_RE = re.compile(r"^  Checksum ([XYZT]) :\s+(\S+)", re.M)


def extract_checksums(stdout: str) -> dict:
    """The `  Checksum X/Y/Z/T : value` lines of a harness run, as written."""
    return {k: v for k, v in CHECKSUM_RE.findall(stdout)}


def read_golden(path: str) -> dict:
    with open(path) as fh:
        return {k: float(v) for k, v in (line.split() for line in fh if line.strip())}


def compare(actual: dict, expected: dict, rtol: float) -> tuple:
    """Lines of the comparison table and whether every checksum is within
    rtol (a NaN or a missing checksum fails)."""
    lines, ok = [], True
    for key in ("X", "Y", "Z", "T"):
        a = float(actual.get(key, "nan"))
        e = float(expected.get(key, "nan"))
        rel = abs(a - e) / max(abs(e), 1e-300)
        good = rel <= rtol and a == a
        lines.append(f"  {key}  expected={e: .10e}  actual={a: .10e}  rel_err={rel:.3e}  {'OK  ' if good else 'FAIL'}")
        ok = ok and good
    return lines, ok


def golden_refusal(checksums: dict):
    """Why a run must not become the golden file, or None (wenoss.sh:389-414):
    no Checksum T line, or T all-zero or non-finite."""
    if "T" not in checksums:
        return "no 'Checksum T' line in binary output -- did the run succeed?"
    try:
        t = float(checksums["T"])
    except ValueError:
        return f"refusing to write golden: Checksum T = {checksums['T']!r}"
    if not math.isfinite(t) or t == 0.0:
        return f"refusing to write golden: Checksum T = {t!r}"
    return None


def golden_file(s: dict) -> str:
    return os.path.join(s["TEST_GOLDEN_DIR"], f"{s['WENO_TEST_KERNEL_MODE']}.txt")


def pinned_run(s: dict) -> str:
    check_bin(s)
    log(f"running test: kernel_mode={s['WENO_TEST_KERNEL_MODE']} warmup={s['WENO_TEST_WARMUP_ITERS']} "
        f"bench={s['WENO_TEST_BENCH_ITERS']}")
    p = subprocess.run(command(s), env=pinned_env(s), capture_output=True, text=True)
    sys.stdout.write(p.stdout)
    sys.stdout.flush()
    if p.returncode != 0:
        sys.stderr.write(p.stderr)
        raise Die(f"{s['BIN']} exited with status {p.returncode}")
    return p.stdout


# -- subcommands ----------------------------------------------------------------
def cmd_build(s: dict) -> int:
    return subprocess.run(["make", "-C", HERE]).returncode


def cmd_run(s: dict) -> int:
    check_bin(s)
    log(f"running {s['BIN']}")
    return subprocess.run(command(s), env=run_env(s)).returncode


def cmd_profile_weno5(s: dict) -> int:
    """rocprofv3 over the binary in weno5 mode (toolchain/profile_weno5_rocprofv3.sh)."""
    check_bin(s)
    os.makedirs(s["ROCPROFV3_OUTPUT_DIR"], exist_ok=True)
    e = run_env(s)
    e["WENO_KERNEL_MODE"] = "weno5"
    cmd = shlex.split(s["RUNNER"]) + [s["ROCPROFV3"], "-d", s["ROCPROFV3_OUTPUT_DIR"], "-o", "weno5_profile"]
    if s["ROCPROFV3_OUTPUT_FORMAT"]:
        cmd += ["-f"] + s["ROCPROFV3_OUTPUT_FORMAT"].split()
    cmd += shlex.split(s["ROCPROFV3_ARGS"]) + shlex.split(s["ROCPROFV3_PMC_ARGS"])
    cmd += ["--", s["BIN"], s["INPUT_NML"]]
    log("rocprofv3 profile -> " + s["ROCPROFV3_OUTPUT_DIR"])
    return subprocess.run(cmd, env=e).returncode


def cmd_list_pmc(s: dict) -> int:
    os.makedirs(s["RESULTS_DIR"], exist_ok=True)
    p = subprocess.run([s["ROCPROFV3"], "--list-avail"], capture_output=True, text=True)
    with open(os.path.join(s["RESULTS_DIR"], "rocprofv3_avail_pmc.txt"), "w") as fh:
        fh.write(p.stdout)
    sys.stdout.write(p.stdout)
    return p.returncode


def cmd_test(s: dict) -> int:
    golden = golden_file(s)
    if not os.path.exists(golden):
        raise Die(f"golden file not found: {golden}. Generate with: python -m wenostandalone_amd.wenoss test-update")
    out = pinned_run(s)
    log(f"comparing against {golden} (rtol={s['WENO_TEST_RTOL']})")
    lines, ok = compare(extract_checksums(out), read_golden(golden), float(s["WENO_TEST_RTOL"]))
    print("\n".join(lines), flush=True)
    if not ok:
        return 1
    log("test passed")
    return 0


def cmd_test_update(s: dict) -> int:
    golden = golden_file(s)
    out = pinned_run(s)
    chk = extract_checksums(out)
    why = golden_refusal(chk)
    if why:
        raise Die(why)
    os.makedirs(s["TEST_GOLDEN_DIR"], exist_ok=True)
    with open(golden, "w") as fh:
        for key in ("X", "Y", "Z", "T"):
            fh.write(f"{key} {chk[key]}\n")
    log(f"wrote golden: {golden}")
    with open(golden) as fh:
        sys.stdout.write(fh.read())
    return 0


def cmd_metrics(s: dict) -> int:
    check_bin(s)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "metrics.json")
        e = run_env(s)
        e["WENO_METRICS_JSON"] = path
        p = subprocess.run(command(s), env=e, capture_output=True, text=True)
        if p.returncode != 0 or not os.path.exists(path):
            sys.stderr.write(p.stdout + p.stderr)
            raise Die(f"{s['BIN']} exited with status {p.returncode} and no metrics")
        with open(path) as fh:
            data = json.load(fh)
    print(json.dumps(data), flush=True)
    return 0 if data.get("status") == "ok" else 1


def cmd_help(s: dict) -> int:
    print(__doc__.strip())
    return 0


COMMANDS = {
    "build": cmd_build, "run": cmd_run, "benchmark": cmd_run, "profile-weno5": cmd_profile_weno5,
    "list-pmc": cmd_list_pmc, "test": cmd_test, "check": cmd_test, "test-update": cmd_test_update,
    "metrics": cmd_metrics, "help": cmd_help, "-h": cmd_help, "--help": cmd_help,
}


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    sub = argv[0] if argv else "help"
    if sub not in COMMANDS:
        raise Die(f"unknown subcommand: {sub}. Try 'python -m wenostandalone_amd.wenoss help'.")
    return COMMANDS[sub](settings())


if __name__ == "__main__":
    sys.exit(main())
