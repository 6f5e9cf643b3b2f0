"""Python access to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Wraps oracle/liboracle.so (the C restatement of the reference kernels) and,
where it has been built in this container, oracle/_ref/weno_ref_driver (the
reference's own module compiled from source).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this package, and
only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_DRIVER = os.path.join(REF_DIR, "weno_ref_driver")
REF_PROGRAM = os.path.join(REF_DIR, "weno_standalone")


class OGrid(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("p", C.c_int), ("buff", C.c_int), ("v", C.c_int),
                ("eps", C.c_double), ("wenojs", C.c_int), ("teno", C.c_int), ("teno_ct", C.c_double),
                ("mapped_weno", C.c_int), ("wenoz", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C {HERE}`")
        _lib = C.CDLL(LIB_PATH)
        vp, i, d, G = C.c_void_p, C.c_int, C.c_double, C.POINTER(OGrid)
        protos = {
            "oracle_fill_input_field": [G, vp],
            "oracle_fill_coefficients": [i, vp, vp, vp, vp, vp, d],
            "oracle_reshape": [G, i, vp, vp],
            "oracle_order1_copy": [G, i, vp, vp, vp],
            "oracle_weno3": [G, i, vp, vp, vp, vp, vp, vp, vp, vp],
            "oracle_weno5": [G, i, vp, vp, vp, vp, vp, vp, vp, vp],
            "oracle_weno7": [G, i, vp, vp, vp, vp, vp, vp, vp, vp],
            "oracle_mp": [G, i, vp, vp, vp],
            "oracle_weno_vjp": [G, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            "oracle_set_teno_ratio_out": [vp],
            "oracle_flux_divergence": [G, i, vp, vp, d, d, d, i, vp],
        }
        for name, args in protos.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = None
        _lib.oracle_checksum.argtypes = [vp, C.c_size_t]
        _lib.oracle_checksum.restype = C.c_double
    return _lib


def grid(m, n=None, p=None, v=8, buff=6, eps=1e-16, wenojs=True, teno=False, teno_ct=1e-5,
         mapped_weno=False, wenoz=False) -> OGrid:
    n = m if n is None else n
    p = m if p is None else p
    return OGrid(m, n, p, buff, v, eps, int(bool(wenojs)), int(bool(teno)), teno_ct,
                 int(bool(mapped_weno)), int(bool(wenoz)))


def _p(a: np.ndarray):
    assert a.flags.f_contiguous or a.flags.c_contiguous and a.ndim == 1
    return a.ctypes.data_as(C.c_void_p)


def extents(g: OGrid):
    return g.m + 1 + 2 * g.buff, g.n + 1 + 2 * g.buff, g.p + 1 + 2 * g.buff


def dir_extents(g: OGrid, d: int):
    nx, ny, nz = extents(g)
    return [(nx, ny, nz), (ny, nx, nz), (nz, ny, nx)][d]


def fill_input_field(g: OGrid) -> np.ndarray:
    nx, ny, nz = extents(g)
    out = np.empty((nx + 6, ny + 6, nz + 6, g.v), order="F")
    lib().oracle_fill_input_field(C.byref(g), _p(out))
    return out


def fill_coefficients(nc: int, offset: float):
    polyL = np.empty((nc, 5, 3), order="F")
    polyR = np.empty((nc, 5, 3), order="F")
    dL = np.empty((5, nc), order="F")
    dR = np.empty((5, nc), order="F")
    beta = np.empty((nc, 5, 6), order="F")
    lib().oracle_fill_coefficients(nc, _p(polyL), _p(polyR), _p(dL), _p(dR), _p(beta), offset)
    return polyL, polyR, dL, dR, beta


def reshape(g: OGrid, d: int, v_vf: np.ndarray) -> np.ndarray:
    ns, nk, nl = dir_extents(g, d)
    out = np.empty((ns + 6, nk, nl, g.v), order="F")
    lib().oracle_reshape(C.byref(g), d, _p(np.asfortranarray(v_vf)), _p(out))
    return out


def order1_copy(g: OGrid, d: int, v_vf: np.ndarray):
    ns, nk, nl = dir_extents(g, d)
    vL = np.empty((ns, nk, nl, g.v), order="F")
    vR = np.empty_like(vL, order="F")
    lib().oracle_order1_copy(C.byref(g), d, _p(np.asfortranarray(v_vf)), _p(vL), _p(vR))
    return vL, vR


def weno(g: OGrid, order: int, d: int, v_rs: np.ndarray, coeffs, vL: np.ndarray, vR: np.ndarray):
    """Runs the order-`order` kernel in place on copies of vL, vR."""
    vL = np.array(vL, order="F", copy=True)
    vR = np.array(vR, order="F", copy=True)
    cs = [np.asfortranarray(c, dtype=np.float64) for c in coeffs]
    fn = {3: lib().oracle_weno3, 5: lib().oracle_weno5, 7: lib().oracle_weno7}[order]
    fn(C.byref(g), d, _p(np.asfortranarray(v_rs)), *[_p(c) for c in cs], _p(vL), _p(vR))
    return vL, vR


def teno_ratios(g: OGrid, order: int, d: int, v_rs: np.ndarray, coeffs):
    """(ratios, bases) of the TENO branch of the order-5/7 kernel: the cut-off
    ratios omega_q = alpha_q / sum(alpha) it compares with teno_CT and the
    bases x_q = 1 + tau/beta_q of the powers alpha_q = x_q^6
    (weno_standalone.f90:580-590, :1490-1500), each of shape
    (ns, nk, nl, v, 3 or 4); zero outside the core."""
    assert g.teno and not g.wenojs and order in (5, 7)
    n = 3 if order == 5 else 4
    ns, nk, nl, nv = v_rs.shape[0] - 6, v_rs.shape[1], v_rs.shape[2], v_rs.shape[3]
    flat = np.zeros(2 * n * ns * nk * nl * nv)
    z = np.zeros((ns, nk, nl, nv), order="F")
    lib().oracle_set_teno_ratio_out(_p(flat))
    try:
        weno(g, order, d, v_rs, coeffs, z, z)
    finally:
        lib().oracle_set_teno_ratio_out(None)
    # for output element o: ratios at 2n*o + q, bases at 2n*o + n + q
    both = flat.reshape((2 * n, ns, nk, nl, nv), order="F").transpose(1, 2, 3, 4, 0)
    return both[..., :n], both[..., n:]


def mp(g: OGrid, d: int, v_rs: np.ndarray, vL: np.ndarray, vR: np.ndarray):
    vL = np.array(vL, order="F", copy=True)
    vR = np.array(vR, order="F", copy=True)
    lib().oracle_mp(C.byref(g), d, _p(np.asfortranarray(v_rs)), _p(vL), _p(vR))
    return vL, vR


def vjp(g: OGrid, order: int, d: int, v_rs: np.ndarray, coeffs, gL: np.ndarray, gR: np.ndarray):
    gv = np.empty(v_rs.shape, order="F")
    cs = [np.asfortranarray(c, dtype=np.float64) for c in coeffs]
    lib().oracle_weno_vjp(C.byref(g), d, order, _p(np.asfortranarray(v_rs)), *[_p(c) for c in cs],
                          _p(np.asfortranarray(gL)), _p(np.asfortranarray(gR)), _p(gv))
    return gv


def flux_divergence(g: OGrid, d: int, vL: np.ndarray, vR: np.ndarray, a: float, b: float, inv_dx: float,
                    rhs: np.ndarray | None = None, accumulate: bool = False) -> np.ndarray:
    """rhs of the Rusanov flux divergence over one sweep's faces (cells
    4 .. ns-5); other cells keep rhs's values (zero when rhs is None)."""
    out = np.zeros(vL.shape, order="F") if rhs is None else np.array(rhs, dtype=np.float64, order="F")
    lib().oracle_flux_divergence(C.byref(g), d, _p(np.asfortranarray(vL)), _p(np.asfortranarray(vR)), a, b,
                                 inv_dx, int(accumulate), _p(out))
    return out


def checksum(a: np.ndarray) -> float:
    a = np.asfortranarray(a, dtype=np.float64)
    return float(lib().oracle_checksum(_p(a), a.size))


def sequence(g: OGrid, family: str, d: int, v_vf: np.ndarray, coeffs):
    """vL, vR of direction d after the reference's sequence for `family`:
    order-1 copy, then the reconstruction (weno3/weno5/weno7), then MP for
    mp5/mp7 -- s_prepare_weno5_inputs + kernel (weno_standalone.f90:1696-1730)."""
    v_rs = reshape(g, d, v_vf)
    vL, vR = order1_copy(g, d, v_vf)
    if family == "copy":
        return vL, vR
    order = {"weno3": 3, "weno5": 5, "weno7": 7, "mp5": 5, "mp7": 7}[family]
    vL, vR = weno(g, order, d, v_rs, coeffs, vL, vR)
    if family.startswith("mp"):
        vL, vR = mp(g, d, v_rs, vL, vR)
    return vL, vR


# -- the reference itself (this container only) ------------------------------
def have_ref() -> bool:
    return os.path.exists(REF_DRIVER)


BRANCHES = ("teno", "wenoz", "mapped_weno")


def ref_driver(branch: str | None = None) -> str:
    """The reference oracle driver; `branch` names the build with that
    compile-time weight parameter set (teno, wenoz, mapped_weno)."""
    if branch is None:
        return REF_DRIVER
    if branch not in BRANCHES:
        raise ValueError(f"unknown weight branch {branch!r}")
    return os.path.join(REF_DIR, branch, "weno_ref_driver")


def run_ref(mode: str, m: int, v: int, buff: int, field: np.ndarray | None, eps=1e-16, wenojs=True,
            branch: str | None = None):
    """Runs the reference module's kernel sequence `mode` (cubic grid m=n=p)
    on `field` (None = the harness field) and returns the six output arrays
    (vL_x, vR_x, vL_y, vR_y, vL_z, vR_z) or, for mode='reshape', the three
    reshaped arrays.  `branch` selects a reference build with one weight
    parameter set (see ref_driver); pass wenojs=False to reach it."""
    nx = m + 1 + 2 * buff
    with tempfile.TemporaryDirectory() as td:
        fin = "harness"
        if field is not None:
            fin = os.path.join(td, "field.bin")
            np.asfortranarray(field, dtype=np.float64).ravel(order="F").tofile(fin)
        fout = os.path.join(td, "out.bin")
        # one thread: s_preserve_monotonicity shares its temporaries across the
        # threads of its `parallel do` (weno_standalone.f90:1550-1559), so a
        # multi-threaded host run races
        env = dict(os.environ, OMP_TARGET_OFFLOAD="DISABLED", OMP_NUM_THREADS="1")
        subprocess.run([ref_driver(branch), mode, str(m), str(m), str(m), str(v), str(buff), repr(float(eps)),
                        "T" if wenojs else "F", fin, fout], check=True, env=env,
                       stdout=subprocess.DEVNULL)
        raw = np.fromfile(fout, dtype=np.float64)
    if mode == "reshape":
        shp = (nx + 6, nx, nx, v)
        n = int(np.prod(shp))
        return [raw[i * n:(i + 1) * n].reshape(shp, order="F") for i in range(3)]
    shp = (nx, nx, nx, v)
    n = int(np.prod(shp))
    return [raw[i * n:(i + 1) * n].reshape(shp, order="F") for i in range(6)]
