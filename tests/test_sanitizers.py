"""AddressSanitizer + UndefinedBehaviorSanitizer builds on the CPU:

* the oracle (`make -C oracle sanitize`, gcc) running every kernel family,
  the MP limiter, the order-1 copy, the reshape and the adjoint on small
  non-cubic boxes;
* the host side of the C ABI (`make -C wenostandalone_amd sanitize`, the
  device code unchanged) through its argument checks and planners, with no
  device: every entry point must refuse bad input with a status code, never
  with a memory error.

Each runs in a child process with the sanitizer runtime preloaded; any
report ends that process with a non-zero status."""
import os
import subprocess
import sys

import pytest

from tests.conftest import ROOT

ORACLE_ASAN = os.path.join(ROOT, "oracle", "_build", "liboracle_asan.so")
CABI_ASAN = os.path.join(ROOT, "wenostandalone_amd", "_build", "asan", "libweno_mi355_asan.so")


def _gcc_runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _clang_asan():
    import glob
    hits = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def _child(script, preload, extra_env=None):
    env = dict(os.environ, LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=24", PYTHONPATH=ROOT)
    env.update(extra_env or {})
    return subprocess.run([sys.executable, "-c", script], cwd=ROOT, env=env, capture_output=True, text=True,
                          timeout=600)


ORACLE_SCRIPT = r"""
import numpy as np, oracle
oracle.LIB_PATH = %r
from wenostandalone_amd import fields
for (m, n, p, v) in [(9, 12, 7, 2), (6, 6, 6, 1)]:
    for kw in [{}, {"wenojs": False, "teno": True}, {"wenojs": False, "wenoz": True},
               {"wenojs": False, "mapped_weno": True}]:
        og = oracle.grid(m, n, p, v=v, buff=6, **kw)
        f = fields.seeded_field(m, n, p, v, 6, seed=3, recipe=2)
        for d in range(3):
            ns = oracle.dir_extents(og, d)[0]
            v_rs = oracle.reshape(og, d, f)
            cL, cR = oracle.order1_copy(og, d, f)
            for order in (3, 5, 7):
                cs = fields.random_coefficients(ns - 6, seed=order + d)
                eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
                mL, mR = oracle.mp(og, d, v_rs, eL, eR)
                if not kw and order != 3:
                    gL = np.asfortranarray(np.ones_like(eL)); gR = np.asfortranarray(-np.ones_like(eR))
                    oracle.vjp(og, order, d, v_rs, cs, gL, gR)
print("oracle clean")
"""


def test_oracle_under_asan_ubsan():
    rt = _gcc_runtime("libasan.so")
    if rt is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = _child(ORACLE_SCRIPT % ORACLE_ASAN, rt)
    assert r.returncode == 0 and "oracle clean" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]


CABI_SCRIPT = r"""
import ctypes as C, sys
sys.argv = ["x"]
import wenostandalone_amd._lib as L
lib = C.CDLL(%r)
for name, (res, args) in L.PROTOTYPES.items():
    f = getattr(lib, name); f.restype, f.argtypes = res, args
g = L.make_grid(15, v=2)
cf = L.Coeffs()
assert lib.weno_device_count() == 0
assert lib.weno_init(0) == -2
bad = L.make_grid(-1, v=2)
assert lib.weno_reconstruct(C.byref(bad), 5, 0, 0, 0, None, C.byref(cf), None, None, 0) == -1
zero_v = L.make_grid(7, v=0)
assert lib.weno_order1_copy(C.byref(zero_v), 0, 0, None, None, None) == -1
assert lib.weno_reconstruct(C.byref(g), 4, 0, 0, 0, None, C.byref(cf), None, None, 0) == -1
assert lib.weno_reconstruct(C.byref(g), 5, 3, 0, 0, None, C.byref(cf), None, None, 0) == -1
assert lib.weno_reconstruct(C.byref(g), 5, 0, 0, 7, None, C.byref(cf), None, None, 0) == -1
assert lib.weno_reconstruct(C.byref(g), 5, 0, 0, 0, None, C.byref(cf), None, None, 1 << 9) == -1
assert lib.weno_reconstruct(C.byref(g), 5, 0, 0, 0, None, C.byref(cf), None, None, 8 | 16) == -1
assert lib.weno_reconstruct_vjp(C.byref(g), 3, 0, 0, None, C.byref(cf), None, None, None) == -1
fxs = L.Flux(1.0, 0.0, 1.0)
assert lib.weno_reconstruct_flux(C.byref(g), 4, 0, 0, None, C.byref(cf), C.byref(fxs), None, 0) == -1
assert lib.weno_reconstruct_flux(C.byref(g), 5, 0, 0, None, C.byref(cf), None, None, 0) == -1
assert lib.weno_reconstruct_flux(C.byref(L.make_grid(1100, 2, 2, v=1)), 5, 0, 0, None, C.byref(cf), C.byref(fxs),
                                 None, 0) == -1
assert lib.weno_flux_divergence(C.byref(g), 0, None, None, None, None, 0) == -1
assert lib.weno_flux_divergence(C.byref(g), 0, None, None, C.byref(fxs), None, 1 << 9) == -1
blocks = (L.Block * 300)()
assert lib.weno_batched_reconstruct_noalloc(blocks, 300, 0, 2) == -1
assert lib.weno_batched_reconstruct(blocks, 0, 0, 2) == -1
assert lib.weno_field_elems(C.byref(g)) == 34 ** 3 * 2
assert lib.weno_face_elems(C.byref(g), 1) == 28 ** 3 * 2
assert lib.weno_core_count(C.byref(g), 2) == 22
assert lib.weno_set_tile(5, 0, 0, 16, 0, 448, 2) == 0 and lib.weno_set_tile(5, 0, 0, 0, 0, 0, -1) == 0
msg = lib.weno_last_error()
assert msg is not None
print("c-abi clean")
"""


def test_cabi_host_code_under_asan_ubsan():
    import ctypes
    try:
        if ctypes.CDLL(os.path.join(ROOT, "wenostandalone_amd", "libweno_mi355.so")).weno_device_count() > 0:
            pytest.skip("a GPU is visible: these checks cover the no-device paths")
    except OSError:
        pytest.skip("libweno_mi355.so is not built")
    rt = _clang_asan()
    if rt is None:
        pytest.skip("the clang AddressSanitizer runtime is not installed")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "wenostandalone_amd"), "-j8", "sanitize"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    r = _child(CABI_SCRIPT % CABI_ASAN, rt)
    assert r.returncode == 0 and "c-abi clean" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
