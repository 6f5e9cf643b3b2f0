"""Order-1 copy and reshape through the C-ABI against the oracle (needs the GPU).

Both are permutations of the field (s_initialize_weno_{x,y,z},
inputs/weno_standalone.f90:281-330; s_order1_copy_{x,y,z}, :332-390), so
every case is bitwise:
  * the bench shape (200^3 interior, buff 6, v 2) in x, y and z, FP64;
  * FP32 reshape and copy, BF16 copy (faces widened to FP32, which is exact);
  * odd row lengths (x rows then move one element per lane);
  * outputs that are 8- but not 16-byte aligned (raw device pointers);
  * rows longer than one tile chunk along y and z (448 outputs per chunk).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield


def _bf16(f):
    """Round-to-nearest-even BF16 bits of a float64 field, and their values."""
    f32 = f.astype(np.float32)
    u = f32.view(np.uint32)
    bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return np.asfortranarray(bits), (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _check(m, n, p, v, d, dtype, seed, offset_bytes=0):
    g = wm.make_grid(m, n, p, v=v, buff=6)
    og = oracle.grid(m, n, p, v=v, buff=6)
    f = fields.seeded_field(m, n, p, v, 6, seed=seed)
    if dtype == wm.F64:
        host, ref, odt = f, f, np.float64
    elif dtype == wm.F32:
        host = f.astype(np.float32)
        ref, odt = host.astype(np.float64), np.float32
    else:
        host, ref = _bf16(f)
        odt = np.float32
    vf = DeviceArray.from_host(host)
    ns, nk, nl = fields.dir_extents(m, n, p, 6, d)
    print(f"permute m={m} n={n} p={p} v={v} d={d} dtype={dtype} offset={offset_bytes}", flush=True)
    outs = []
    try:
        if dtype != wm.BF16:
            shape = (ns + 6, nk, nl, v)
            rs = DeviceArray(shape, odt)
            outs.append(rs)
            wm.check(wm.lib.weno_reshape(C.byref(g), d, dtype, vf.vp, rs.vp))
            got = rs.download()
            exp = oracle.reshape(og, d, ref)
            assert np.array_equal(got.astype(np.float64), exp)
        shape = (ns, nk, nl, v)
        size = int(np.prod(shape))
        esz = np.dtype(odt).itemsize
        pad = offset_bytes // esz
        L = DeviceArray((size + pad,), odt)
        R = DeviceArray((size + pad,), odt)
        outs += [L, R]
        L.fill(-1.0)
        R.fill(-2.0)
        pL, pR = C.c_void_p(L.ptr + offset_bytes), C.c_void_p(R.ptr + offset_bytes)
        wm.check(wm.lib.weno_order1_copy(C.byref(g), d, dtype, vf.vp, pL, pR))
        cL, cR = oracle.order1_copy(og, d, ref)
        for ptr, exp in ((pL, cL), (pR, cR)):
            got = np.empty(shape, dtype=odt, order="F")
            wm.check(wm.lib.weno_memcpy_to_host(got.ctypes.data_as(C.c_void_p), ptr, size * esz))
            assert np.array_equal(got.astype(np.float64), exp)
        # the elements before an offset output stay untouched
        if pad:
            assert np.all(L.download()[:pad] == -1.0) and np.all(R.download()[:pad] == -2.0)
    finally:
        for x in outs:
            x.free()
        vf.free()


@pytest.mark.parametrize("d", [0, 1, 2])
def test_bench_shape_fp64(d):
    _check(199, 199, 199, 2, d, wm.F64, seed=11 + d)


@pytest.mark.parametrize("dtype", [wm.F32, wm.BF16])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_reduced_precision_state(dtype, d):
    _check(37, 23, 19, 3, d, dtype, seed=21 + d)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_odd_rows_and_unaligned_outputs(d):
    # 15 cells per axis: rows of 27 faces (33 reshaped), outputs 8 bytes into an allocation;
    # then even x rows (28 faces) whose outputs are 4 but not 8 bytes aligned (FP32)
    _check(14, 14, 14, 2, d, wm.F64, seed=31 + d, offset_bytes=8)
    _check(15, 14, 13, 2, d, wm.F32, seed=34 + d, offset_bytes=4)


@pytest.mark.parametrize("d", [1, 2])
def test_rows_in_two_chunks(d):
    # 471 cells along the sweep: rows of 483 faces in chunks of 242 and 241
    m, n, p = (9, 470, 7) if d == 1 else (9, 7, 470)
    _check(m, n, p, 1, d, wm.F64, seed=41 + d)
