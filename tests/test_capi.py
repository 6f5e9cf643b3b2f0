"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/weno_c.h declares, and rejects bad arguments (or a missing device)
with a status code and a message before launching anything."""
import ctypes as C
import os
import re

import pytest

import wenostandalone_amd as wm
from wenostandalone_amd._lib import PROTOTYPES
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "weno_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(weno_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(wm.lib, n), f"{n} declared in weno_c.h but not exported"
        assert n in PROTOTYPES, f"{n} has no ctypes prototype"


def test_no_device_is_a_clean_error():
    if wm.lib.weno_device_count() > 0:
        pytest.skip("a GPU is visible")
    rc = wm.lib.weno_init(0)
    assert rc == -2
    assert "device" in wm._lib.last_error().lower()


@pytest.mark.parametrize("field,value,msg", [
    ("m_cells", -1, "m/n/p_cells"),
    ("v_size", 0, "v_size"),
    ("buff_size", -2, "buff_size"),
    ("weno_eps", 0.0, "weno_eps"),
])
def test_grid_validation(field, value, msg):
    g = wm.make_grid(7, v=2)
    setattr(g, field, value)
    cf = wm.Coeffs()
    rc = wm.lib.weno_reconstruct(C.byref(g), 5, 0, wm.SRC_FIELD, wm.F64, None, C.byref(cf), None, None, 0)
    assert rc == -1
    assert msg in wm._lib.last_error()


def test_small_extent_rejected():
    g = wm.make_grid(0, v=1, buff=2)        # is extent 5 < 7: empty core
    rc = wm.lib.weno_order1_copy(C.byref(g), 0, wm.F64, None, None, None)
    assert rc == -1 and "core range" in wm._lib.last_error()


@pytest.mark.parametrize("order,d,src,dtype,flags,msg", [
    (4, 0, 0, 0, 0, "order"),
    (5, 3, 0, 0, 0, "dir"),
    (5, 0, 2, 0, 0, "src_kind"),
    (5, 0, 0, 7, 0, "dtype"),
    (5, 0, 0, 0, 64, "flag"),
])
def test_reconstruct_argument_validation(order, d, src, dtype, flags, msg):
    g = wm.make_grid(7, v=2)
    cf = wm.Coeffs()
    rc = wm.lib.weno_reconstruct(C.byref(g), order, d, src, dtype, None, C.byref(cf), None, None, flags)
    assert rc == -1
    assert msg in wm._lib.last_error()


def test_shapes():
    g = wm.make_grid(199, v=8, buff=6)
    assert wm.lib.weno_field_elems(C.byref(g)) == 218 ** 3 * 8
    assert wm.lib.weno_reshaped_elems(C.byref(g), 1) == 218 * 212 * 212 * 8
    assert wm.lib.weno_face_elems(C.byref(g), 2) == 212 ** 3 * 8
    assert wm.lib.weno_core_count(C.byref(g), 0) == 206
    g2 = wm.make_grid(9, 12, 7, v=3, buff=6)     # per-direction extents
    assert wm.lib.weno_face_elems(C.byref(g2), 1) == 25 * 22 * 20 * 3
    assert wm.lib.weno_core_count(C.byref(g2), 2) == 20 - 6


def test_set_tile_validation():
    assert wm.lib.weno_set_tile(5, 0, 0, 65, 0, 0, -1) == -1
    assert wm.lib.weno_set_tile(6, 0, 0, 16, 0, 0, -1) == -1
    assert wm.lib.weno_set_tile(5, 0, 0, 16, 0, 100, -1) == -1     # not a multiple of 64
    assert wm.lib.weno_set_tile(5, 0, 0, 16, 0, 448, 0) == 0
    assert wm.lib.weno_set_tile(5, 0, 0, 0, 0, 0, -1) == 0


def test_noalloc_batched_limit():
    """The allocation-free batched launch refuses more blocks than its static
    scratch holds, before touching the blocks or the GPU."""
    blocks = (wm.Block * 257)()
    rc = wm.lib.weno_batched_reconstruct_noalloc(blocks, 257, 0, wm.FUSE_COPY)
    assert rc == -1 and "at most 256" in wm._lib.last_error()
