"""HIP path vs the oracle at the configurations' real shapes (needs the GPU).

* Every general-kernel family at the 200^3 bench shape (m=n=p=199, the tile
  plans `bench.py --families` times; v=2 keeps the oracle fast -- the tile
  plan depends on the row length and the pencil counts per variable, not on
  v), x, y and z, seeded field recipe 0 and the harness coefficients:
    WENO3/7-JS, WENO7-JS core-only, WENO7 + fused MP, TENO5/7, WENO-Z3/5/7,
    mapped WENO3/5/7, standalone MP ........ 1e-13 relative (rel_err)
    WENO7 with FP32 / BF16 state ............ 1e-6 relative to the FP64
        oracle on the rounded state
* Rows longer than a workgroup: WENO3/5/7 on a 537-point x axis (543 faces
  per row, more than WENO7's 512-thread limit) and FP32 WENO5/7 on a
  1120-point axis (more than 1024 lanes), with and without the fused copy.
* The batched form at config 3's shape: 16 blocks with sides drawn from
  U{16..128} per axis, v in {5, 8, 10}, WENO5/7 mixed, after
  weno_batched_autotune, fused copy with and without MP, every block against
  its own oracle sequence (1e-13); then the x/y/z graph of the same set,
  replayed with eager work in between, bitwise against the eager launches.
* s_run_all_kernels (inputs/weno_standalone.f90:1651-1682) through the
  module mirror at 200^3 (v=2, the harness field): the core range, MP of
  WENO7, against the oracle's composition of the same sequence (1e-13).  The
  halo cells hold WENO3's outputs there, which the reference computes from
  coefficients read out of bounds (:395-399 vs :413); the oracle and the GPU
  both clamp the coefficient index, and the test holds them to 1e-13 too."""
import ctypes as C

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray
from tests.conftest import rel_err
from tests.test_gpu_batched import Blocks

pytestmark = pytest.mark.gpu

M, V, B = 199, 2, 6
NS = M + 1 + 2 * B


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield


@pytest.fixture(scope="module")
def bench_field():
    f = fields.seeded_field(M, M, M, V, B, seed=0)
    og = oracle.grid(M, v=V, buff=B)
    per_dir = {}
    return f, og, per_dir


def _dir_data(bench_field, d):
    f, og, cache = bench_field
    if d not in cache:
        v_rs = oracle.reshape(og, d, f)
        cL, cR = oracle.order1_copy(og, d, f)
        cs = fields.harness_coefficients(NS - 6, fields.DIR_OFFSETS[d])
        cache[d] = (v_rs, cL, cR, cs)
    return cache[d]


def _coef(cs):
    devs = [DeviceArray.from_host(c, np.float64) for c in cs]
    return devs, wm.Coeffs(*[x.ptr for x in devs])


def _run(g, order, d, src, coef, flags, shape, pre=None, dtype=wm.F64, kind=wm.SRC_FIELD):
    odt = np.float64 if dtype == wm.F64 else np.float32
    vL, vR = DeviceArray(shape, odt), DeviceArray(shape, odt)
    if pre is None:
        vL.fill(7.25)
        vR.fill(-3.5)
    else:
        vL.upload(pre[0])
        vR.upload(pre[1])
    wm.check(wm.lib.weno_reconstruct(C.byref(g), order, d, kind, dtype, src.vp, C.byref(coef), vL.vp, vR.vp,
                                     flags))
    out = vL.download(), vR.download()
    vL.free()
    vR.free()
    return out


# name: (order, grid switches, flags)
FAMILIES = {
    "weno3_js": (3, {}, wm.FUSE_COPY),
    "weno7_js": (7, {}, wm.FUSE_COPY),
    "weno7_js_core_only": (7, {}, 0),
    "weno7_js_fused_mp": (7, {}, wm.FUSE_COPY | wm.FUSE_MP),
    "teno5": (5, {"wenojs": False, "teno": True}, wm.FUSE_COPY),
    "teno7": (7, {"wenojs": False, "teno": True}, wm.FUSE_COPY),
    "wenoz3": (3, {"wenojs": False, "wenoz": True}, wm.FUSE_COPY),
    "wenoz5": (5, {"wenojs": False, "wenoz": True}, wm.FUSE_COPY),
    "wenoz7": (7, {"wenojs": False, "wenoz": True}, wm.FUSE_COPY),
    "mapped3": (3, {"wenojs": False, "mapped_weno": True}, wm.FUSE_COPY),
    "mapped5": (5, {"wenojs": False, "mapped_weno": True}, wm.FUSE_COPY),
    "mapped7": (7, {"wenojs": False, "mapped_weno": True}, wm.FUSE_COPY),
}


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("family", list(FAMILIES))
def test_family_at_bench_shape(bench_field, family, d):
    order, kw, flags = FAMILIES[family]
    f, _, _ = bench_field
    v_rs, cL, cR, cs = _dir_data(bench_field, d)
    og = oracle.grid(M, v=V, buff=B, **kw)
    g = wm.make_grid(M, v=V, buff=B, **kw)
    eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
    if flags & wm.FUSE_MP:
        eL, eR = oracle.mp(og, d, v_rs, eL, eR)
    devs, coef = _coef(cs)
    vf = DeviceArray.from_host(f)
    pre = None if flags & wm.FUSE_COPY else (cL, cR)
    L, R = _run(g, order, d, vf, coef, flags, cL.shape, pre=pre)
    vf.free()
    eLr, eRr = rel_err(L, eL), rel_err(R, eR)
    print(f"{family} dir {d}: rel err L {eLr:.2e} R {eRr:.2e}", flush=True)
    assert eLr <= 1e-13 and eRr <= 1e-13, (family, d)
    if flags & wm.FUSE_COPY and order != 3:
        assert np.array_equal(L[:3], cL[:3]) and np.array_equal(R[-3:], cR[-3:])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_mp_standalone_at_bench_shape(bench_field, d):
    """s_preserve_monotonicity on the reshaped array, as the reference calls
    it (:1544), limiting WENO5 faces in place."""
    v_rs, cL, cR, cs = _dir_data(bench_field, d)
    og = oracle.grid(M, v=V, buff=B)
    g = wm.make_grid(M, v=V, buff=B)
    eL, eR = oracle.weno(og, 5, d, v_rs, cs, cL, cR)
    mL, mR = oracle.mp(og, d, v_rs, eL, eR)
    rs = DeviceArray.from_host(v_rs)
    vL, vR = DeviceArray.from_host(eL), DeviceArray.from_host(eR)
    wm.check(wm.lib.weno_mp_limit(C.byref(g), d, wm.SRC_RESHAPED, wm.F64, rs.vp, vL.vp, vR.vp))
    L, R = vL.download(), vR.download()
    for x in (rs, vL, vR):
        x.free()
    assert rel_err(L, mL) <= 1e-13 and rel_err(R, mR) <= 1e-13, d


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("dtype", [wm.F32, wm.BF16])
def test_weno7_reduced_precision_at_bench_shape(bench_field, dtype, d):
    f, _, _ = bench_field
    f32 = f.astype(np.float32)
    if dtype == wm.F32:
        state, f64r = f32, f32.astype(np.float64)
    else:
        u = f32.view(np.uint32)
        state = np.asfortranarray(((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16))
        f64r = (state.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    og = oracle.grid(M, v=V, buff=B)
    g = wm.make_grid(M, v=V, buff=B)
    cs = fields.harness_coefficients(NS - 6, fields.DIR_OFFSETS[d])
    v_rs = oracle.reshape(og, d, f64r)
    cL, cR = oracle.order1_copy(og, d, f64r)
    eL, eR = oracle.weno(og, 7, d, v_rs, cs, cL, cR)
    del v_rs
    devs, coef = _coef(cs)
    src = DeviceArray.from_host(state)
    L, R = _run(g, 7, d, src, coef, wm.FUSE_COPY, cL.shape, dtype=dtype)
    src.free()
    eLr, eRr = rel_err(L, eL), rel_err(R, eR)
    print(f"weno7 dtype {dtype} dir {d}: rel err L {eLr:.2e} R {eRr:.2e}", flush=True)
    assert eLr <= 1e-6 and eRr <= 1e-6, (dtype, d)


@pytest.mark.parametrize("flags", [0, wm.FUSE_COPY])
@pytest.mark.parametrize("order,m,dtype", [(3, 536, wm.F64), (5, 536, wm.F64), (7, 536, wm.F64),
                                           (5, 1119, wm.F32), (7, 1119, wm.F32)])
def test_rows_longer_than_a_workgroup(order, m, dtype, flags):
    """An x axis of m+1 points with small y/z: one output row per pencil has
    m+1+2*buff faces, more than the kernel's workgroup limit (512 threads for
    WENO7, 1024 otherwise), so the row is reconstructed in several chunks."""
    n = p = 4
    v = 1
    g = wm.make_grid(m, n, p, v=v, buff=B)
    og = oracle.grid(m, n, p, v=v, buff=B)
    f = fields.seeded_field(m, n, p, v, B, seed=70 + order)
    f64r = f if dtype == wm.F64 else f.astype(np.float32).astype(np.float64)
    src = DeviceArray.from_host(f if dtype == wm.F64 else f.astype(np.float32))
    for d in range(3):
        ns = fields.dir_extents(m, n, p, B, d)[0]
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        v_rs = oracle.reshape(og, d, f64r)
        cL, cR = oracle.order1_copy(og, d, f64r)
        eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
        devs, coef = _coef(cs)
        pre = None if flags else ((cL, cR) if dtype == wm.F64 else (cL.astype(np.float32), cR.astype(np.float32)))
        L, R = _run(g, order, d, src, coef, flags, cL.shape, pre=pre, dtype=dtype)
        tol = 1e-13 if dtype == wm.F64 else 1e-6
        assert rel_err(L, eL) <= tol and rel_err(R, eR) <= tol, (order, m, d, flags)


def _config3_specs(seed, n):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(16, 129)) - 1, int(rng.integers(16, 129)) - 1, int(rng.integers(16, 129)) - 1,
             int(rng.choice([5, 8, 10])), int(rng.choice([5, 7]))) for _ in range(n)]


@pytest.fixture(scope="module")
def config3_blocks():
    specs = _config3_specs(2024, 16)
    bl = Blocks(specs, seed=800, coefficients=lambda nc, order, s: fields.harness_coefficients(nc, 0.05))
    yield bl
    for x in bl.keep:
        x.free()
    for fc in bl.faces:
        for L, R in fc:
            L.free()
            R.free()
    for d in range(3):
        for o in (3, 5, 7):
            wm.lib.weno_set_tile(o, wm.F64, d, 0, 0, 0, -1)


@pytest.mark.parametrize("flags", [wm.FUSE_COPY, wm.FUSE_COPY | wm.FUSE_MP])
def test_batched_config3_shape(config3_blocks, flags):
    bl = config3_blocks
    nb = len(bl.specs)
    best = (C.c_int * 3)()
    wm.check(wm.lib.weno_batched_autotune(bl.blocks, nb, flags, best))
    print(f"autotuned threads {best[0]} pencils {best[1]} ({best[2]} us per x+y+z)", flush=True)
    bl.fill(11.5)
    for d in range(3):
        wm.check(wm.lib.weno_batched_reconstruct(bl.blocks, nb, d, flags))
    got = bl.download()
    for b in range(nb):
        for d in range(3):
            eL, eR = bl.expected(b, d, flags)
            assert rel_err(got[b][d][0], eL) <= 1e-13, (b, d, bl.specs[b])
            assert rel_err(got[b][d][1], eR) <= 1e-13, (b, d, bl.specs[b])
    del got
    # the x/y/z graph of the same set, replayed with eager work in between
    h = C.c_void_p()
    eager = bl.download()
    wm.check(wm.lib.weno_batched_graph_create(bl.blocks, nb, flags, C.byref(h)))
    try:
        for rep in range(3):
            bl.fill(float(rep) - 4.0)
            wm.check(wm.lib.weno_graph_launch(h))
            wm.check(wm.lib.weno_sync())
            rg = bl.download()
            for b in range(nb):
                for d in range(3):
                    assert np.array_equal(rg[b][d][0], eager[b][d][0]), (rep, b, d)
                    assert np.array_equal(rg[b][d][1], eager[b][d][1]), (rep, b, d)
            del rg
    finally:
        wm.check(wm.lib.weno_graph_destroy(h))


def test_run_all_kernels_vs_oracle():
    from wenostandalone_amd.standalone import WenoStandalone
    ws = WenoStandalone(m_cells=M, n_cells=M, p_cells=M, v_size=V)
    ws.s_initialize_data()
    faces = [ws.new_faces(d) for d in (0, 0, 1, 1, 2, 2)]
    ws.s_run_all_kernels(*faces)
    got = [a.download() for a in faces]
    for a in faces:
        a.free()
    ws.s_finalize_data()
    og = oracle.grid(M, v=V, buff=B)
    f = fields.harness_field(M, M, M, V, B)
    for d in range(3):
        cs = fields.harness_coefficients(NS - 6, fields.DIR_OFFSETS[d])
        v_rs = oracle.reshape(og, d, f)
        L, R = oracle.order1_copy(og, d, f)
        for order in (3, 5, 7):
            L, R = oracle.weno(og, order, d, v_rs, cs, L, R)
        L, R = oracle.mp(og, d, v_rs, L, R)
        gL, gR = got[2 * d], got[2 * d + 1]
        core = slice(3, NS - 3)
        assert rel_err(gL[core], L[core]) <= 1e-13 and rel_err(gR[core], R[core]) <= 1e-13, d
        assert rel_err(gL, L) <= 1e-13 and rel_err(gR, R) <= 1e-13, d
