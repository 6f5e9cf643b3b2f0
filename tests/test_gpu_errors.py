"""C-ABI error paths that need a device (the CPU-only ones are in
tests/test_capi.py): arrays shorter than the launch would touch are refused
before anything runs, whether they are raw device memory or host arrays
present in the data environment (OpenMP's present check refuses a partly
present array), and the library keeps working afterwards."""
import ctypes as C

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

M, V, B = 9, 2, 6


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield


def _setup(d=0):
    g = wm.make_grid(M, v=V, buff=B)
    f = fields.seeded_field(M, M, M, V, B, seed=5)
    ns, nk, nl = fields.dir_extents(M, M, M, B, d)
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
    devs = [DeviceArray.from_host(c) for c in cs]
    return g, f, (ns, nk, nl, V), cs, devs, wm.Coeffs(*[x.ptr for x in devs])


def test_short_device_buffers_are_refused():
    g, f, shape, cs, devs, coef = _setup()
    vf = DeviceArray.from_host(f)
    full = DeviceArray(shape)
    short = DeviceArray((shape[0] * shape[1] * shape[2] * shape[3] // 2,))
    rc = wm.lib.weno_reconstruct(C.byref(g), 5, 0, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coef), short.vp, full.vp,
                                 wm.FUSE_COPY)
    assert rc == -1 and "vL" in wm._lib.last_error() and "needs" in wm._lib.last_error()
    # a pointer into the middle of an allocation has only the rest of it
    inner = C.c_void_p(full.ptr + (full.nbytes // 2 & ~15))
    rc = wm.lib.weno_order1_copy(C.byref(g), 0, wm.F64, vf.vp, full.vp, inner)
    assert rc == -1 and "vR" in wm._lib.last_error()
    short_cf = DeviceArray((len(cs[0].ravel()) // 2,))
    bad = wm.Coeffs(short_cf.ptr, coef.polyR, coef.dL, coef.dR, coef.beta)
    rc = wm.lib.weno_reconstruct(C.byref(g), 5, 0, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(bad), full.vp, full.vp, 0)
    assert rc == -1 and "polyL" in wm._lib.last_error()
    # the library still works
    L = DeviceArray(shape)
    wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, 0, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coef), L.vp, full.vp,
                                     wm.FUSE_COPY))
    og = oracle.grid(M, v=V, buff=B)
    eL, _ = oracle.sequence(og, "weno5", 0, f, cs)
    assert rel_err(L.download(), eL) <= 1e-13


def test_partly_present_host_array_is_refused():
    g, f, shape, cs, devs, coef = _setup()
    vf = DeviceArray.from_host(f)
    host = np.zeros(shape, order="F")
    half = host.nbytes // 2
    wm.check(wm.lib.weno_target_enter_data(host.ctypes.data_as(C.c_void_p), half))
    try:
        other = DeviceArray(shape)
        rc = wm.lib.weno_reconstruct(C.byref(g), 5, 0, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coef),
                                     host.ctypes.data_as(C.c_void_p), other.vp, wm.FUSE_COPY)
        assert rc == -1 and "present" in wm._lib.last_error()
    finally:
        wm.check(wm.lib.weno_target_exit_data(host.ctypes.data_as(C.c_void_p)))
    # mapped whole, the same call goes through
    wm.check(wm.lib.weno_target_enter_data(host.ctypes.data_as(C.c_void_p), host.nbytes))
    try:
        other = DeviceArray(shape)
        wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, 0, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coef),
                                         host.ctypes.data_as(C.c_void_p), other.vp, wm.FUSE_COPY))
        wm.check(wm.lib.weno_target_update_from(host.ctypes.data_as(C.c_void_p), host.nbytes))
    finally:
        wm.check(wm.lib.weno_target_exit_data(host.ctypes.data_as(C.c_void_p)))
    og = oracle.grid(M, v=V, buff=B)
    eL, _ = oracle.sequence(og, "weno5", 0, f, cs)
    assert rel_err(host, eL) <= 1e-13
