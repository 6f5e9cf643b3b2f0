"""ctypes binding of libweno_mi355 (include/weno_c.h).

The shared library is built in-tree (``make -C wenostandalone_amd``).  There is
no fallback: if the library is missing or fails to load, importing this module
raises, so nothing can silently run on the CPU instead of the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libweno_mi355.so")

WENO_OK = 0
DIR_X, DIR_Y, DIR_Z = 0, 1, 2
SRC_FIELD, SRC_RESHAPED = 0, 1
F64, F32, BF16 = 0, 1, 2
FUSE_MP, FUSE_COPY, STREAMING, ONLY_L, ONLY_R, ACCUMULATE = 1, 2, 4, 8, 16, 32


class WenoError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"weno error {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    """weno_grid: namelist &grid_params (weno_standalone.f90:57-66) + the
    weight-branch switches teno, mapped_weno, wenoz (compile-time parameters
    of the reference, :72-74)."""

    _fields_ = [
        ("m_cells", C.c_int), ("n_cells", C.c_int), ("p_cells", C.c_int),
        ("v_size", C.c_int), ("buff_size", C.c_int),
        ("weno_eps", C.c_double), ("wenojs", C.c_int), ("teno", C.c_int),
        ("teno_CT", C.c_double), ("mapped_weno", C.c_int), ("wenoz", C.c_int),
    ]


class Coeffs(C.Structure):
    """weno_coeffs: one direction's coefficient arrays (device addresses)."""

    _fields_ = [("polyL", C.c_void_p), ("polyR", C.c_void_p), ("dL", C.c_void_p),
                ("dR", C.c_void_p), ("beta", C.c_void_p)]


class Flux(C.Structure):
    """weno_flux: f(q) = a q + b q^2/2 and 1/dx of the consumer step."""

    _fields_ = [("a", C.c_double), ("b", C.c_double), ("inv_dx", C.c_double)]


class Block(C.Structure):
    """weno_block: one independent sub-domain of a batched launch."""

    _fields_ = [("grid", Grid), ("order", C.c_int), ("v_vf", C.c_void_p), ("coeffs", Coeffs * 3),
                ("vL", C.c_void_p * 3), ("vR", C.c_void_p * 3)]


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
            "(or __graft_entry__.build())")
    return C.CDLL(LIB_PATH)


lib = _load()

_vp, _i, _sz, _d = C.c_void_p, C.c_int, C.c_size_t, C.c_double
_G = C.POINTER(Grid)
_CF = C.POINTER(Coeffs)

# name -> (restype, argtypes); every symbol of include/weno_c.h is listed here
PROTOTYPES = {
    "weno_init": (_i, [_i]),
    "weno_finalize": (_i, []),
    "weno_last_error": (C.c_char_p, []),
    "weno_sync": (_i, []),
    "weno_set_async": (_i, [_i]),
    "weno_set_stream": (_i, [_vp]),
    "weno_get_stream": (_vp, []),
    "weno_device_count": (_i, []),
    "weno_target_enter_data": (_i, [_vp, _sz]),
    "weno_target_update_to": (_i, [_vp, _sz]),
    "weno_target_update_from": (_i, [_vp, _sz]),
    "weno_target_exit_data": (_i, [_vp]),
    "weno_device_address": (_vp, [_vp]),
    "weno_device_alloc": (_vp, [_sz]),
    "weno_device_free": (_i, [_vp]),
    "weno_memcpy_to_device": (_i, [_vp, _vp, _sz]),
    "weno_memcpy_to_host": (_i, [_vp, _vp, _sz]),
    "weno_field_elems": (_sz, [_G]),
    "weno_reshaped_elems": (_sz, [_G, _i]),
    "weno_face_elems": (_sz, [_G, _i]),
    "weno_core_count": (_i, [_G, _i]),
    "weno_reshape": (_i, [_G, _i, _i, _vp, _vp]),
    "weno_order1_copy": (_i, [_G, _i, _i, _vp, _vp, _vp]),
    "weno_reconstruct": (_i, [_G, _i, _i, _i, _i, _vp, _CF, _vp, _vp, _i]),
    "weno_mp_limit": (_i, [_G, _i, _i, _i, _vp, _vp, _vp]),
    "weno_set_tile": (_i, [_i, _i, _i, _i, _i, _i, _i]),
    "weno_reconstruct_vjp": (_i, [_G, _i, _i, _i, _vp, _CF, _vp, _vp, _vp]),
    "weno_stream_copy": (_i, [_vp, _vp, _sz]),
    "weno_reconstruct_flux": (_i, [_G, _i, _i, _i, _vp, _CF, C.POINTER(Flux), _vp, _i]),
    "weno_flux_divergence": (_i, [_G, _i, _vp, _vp, C.POINTER(Flux), _vp, _i]),
    "weno_batched_reconstruct": (_i, [_vp, _i, _i, _i]),
    "weno_batched_reconstruct_noalloc": (_i, [_vp, _i, _i, _i]),
    "weno_batched_graph_create": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "weno_sweep_graph_create": (_i, [_G, _i, _i, _i, _vp, _CF, C.POINTER(_vp), C.POINTER(_vp), _i,
                                     C.POINTER(_vp)]),
    "weno_graph_launch": (_i, [_vp]),
    "weno_graph_destroy": (_i, [_vp]),
    "weno_autotune_sweep": (_i, [_G, _i, _i, _i, _vp, _CF, C.POINTER(_vp), C.POINTER(_vp), _i,
                                 C.POINTER(_i)]),
    "weno_batched_autotune": (_i, [_vp, _i, _i, C.POINTER(_i)]),
}

for _name, (_res, _args) in PROTOTYPES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def last_error() -> str:
    msg = lib.weno_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> int:
    if rc != WENO_OK:
        raise WenoError(rc, last_error())
    return rc


def make_grid(m: int, n: int | None = None, p: int | None = None, v: int = 8, buff: int = 6,
              eps: float = 1e-16, wenojs: bool = True, teno: bool = False,
              teno_ct: float = 1e-5, mapped_weno: bool = False, wenoz: bool = False) -> Grid:
    n = m if n is None else n
    p = m if p is None else p
    return Grid(m, n, p, v, buff, eps, int(bool(wenojs)), int(bool(teno)), teno_ct,
                int(bool(mapped_weno)), int(bool(wenoz)))
