"""HIP path vs the oracle and the reference fixtures (needs the GPU).

Tolerances, stated per kernel:
  reshape, order-1 copy, fused halo copy ........ bitwise
  WENO3/5/7-JS FP64, fused MP, standalone MP ..... 1e-13 relative (rel_err)
  TENO FP64 ...................................... 1e-13; the cut-off
      decisions equal the reference's, also with teno_CT placed exactly on
      a ratio the reference forms (test_teno_ct_on_a_reference_ratio)
  FP32 storage ................................... 1e-6 relative to the FP64 oracle
  BF16 state ..................................... 1e-6 relative to the FP64 oracle
      fed the same BF16-rounded inputs (the state's own rounding, 3.9e-3,
      is reported separately by the precision tests)
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

FAMILIES = {"weno3": 3, "weno5": 5, "weno7": 7, "mp5": 5, "mp7": 7}


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield


def upload_coeffs(cs):
    devs = [DeviceArray.from_host(c, np.float64) for c in cs]
    return devs, wm.Coeffs(*[d.ptr for d in devs])


def run_hip(g, order, d, src, src_kind, coef, flags, pre=None, dtype=wm.F64):
    ns, nk, nl = fields.dir_extents(g.m_cells, g.n_cells, g.p_cells, g.buff_size, d)
    odt = np.float64 if dtype == wm.F64 else np.float32
    vL = DeviceArray((ns, nk, nl, g.v_size), odt)
    vR = DeviceArray((ns, nk, nl, g.v_size), odt)
    if pre is None:
        vL.fill(7.25)
        vR.fill(-3.5)
    else:
        vL.upload(pre[0])
        vR.upload(pre[1])
    wm.check(wm.lib.weno_reconstruct(C.byref(g), order, d, src_kind, dtype, src.vp, C.byref(coef),
                                     vL.vp, vR.vp, flags))
    return vL.download(), vR.download()


def test_fixture_sequences_bitwise_shape(small_ref):
    """Reference fixtures (16^3 box, recipe-2 field): the HIP sequence
    copy -> kernel [-> MP] equals the reference's outputs."""
    m, v, b = int(small_ref["m"]), int(small_ref["v"]), int(small_ref["buff"])
    f = small_ref["field"]
    g = wm.make_grid(m, v=v, buff=b)
    vf = DeviceArray.from_host(f)
    for d, name in enumerate("xyz"):
        ns = fields.dir_extents(m, m, m, b, d)[0]
        devs, coef = upload_coeffs(fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d]))
        for fam, order in FAMILIES.items():
            flags = wm.FUSE_COPY | (wm.FUSE_MP if fam.startswith("mp") else 0)
            L, R = run_hip(g, order, d, vf, wm.SRC_FIELD, coef, flags)
            rL, rR = small_ref[f"{fam}_{name}_L"], small_ref[f"{fam}_{name}_R"]
            if fam == "weno3":   # halo cells are undefined in the reference
                L, R, rL, rR = L[3:-3], R[3:-3], rL[3:-3], rR[3:-3]
            assert rel_err(L, rL) <= 1e-13, (fam, name)
            assert rel_err(R, rR) <= 1e-13, (fam, name)


def _grid_and_field(seed, m=9, n=12, p=7, v=3, recipe=0, **kw):
    g = wm.make_grid(m, n, p, v=v, buff=6, **kw)
    og = oracle.grid(m, n, p, v=v, buff=6, **kw)
    f = fields.seeded_field(m, n, p, v, 6, seed=seed, recipe=recipe)
    return g, og, f


@pytest.mark.parametrize("order", [3, 5, 7])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_reconstruct_vs_oracle_noncubic(order, d):
    """Non-cubic box, j-varying coefficients, every source layout and flag."""
    g, og, f = _grid_and_field(seed=100 + order * 3 + d)
    ns = fields.dir_extents(9, 12, 7, 6, d)[0]
    cs = fields.random_coefficients(ns - 6, seed=order * 10 + d)
    devs, coef = upload_coeffs(cs)
    vf = DeviceArray.from_host(f)
    rs = DeviceArray.from_host(oracle.reshape(og, d, f))
    cL, cR = oracle.order1_copy(og, d, f)
    v_rs = oracle.reshape(og, d, f)
    eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
    mL, mR = oracle.mp(og, d, v_rs, eL, eR)
    for src, kind in [(vf, wm.SRC_FIELD), (rs, wm.SRC_RESHAPED)]:
        # plain kernel on top of the order-1 copy
        L, R = run_hip(g, order, d, src, kind, coef, 0, pre=(cL, cR))
        assert rel_err(L, eL) <= 1e-13 and rel_err(R, eR) <= 1e-13
        # fused copy: halo written bitwise, core reconstructed
        L, R = run_hip(g, order, d, src, kind, coef, wm.FUSE_COPY)
        assert rel_err(L, eL) <= 1e-13 and rel_err(R, eR) <= 1e-13
        if order != 3:
            assert np.array_equal(L[:3], cL[:3]) and np.array_equal(R[-3:], cR[-3:])
        # fused MP
        L, R = run_hip(g, order, d, src, kind, coef, wm.FUSE_COPY | wm.FUSE_MP)
        assert rel_err(L, mL) <= 1e-13 and rel_err(R, mR) <= 1e-13


@pytest.mark.parametrize("d", [0, 1, 2])
def test_reshape_copy_mp_standalone(d):
    g, og, f = _grid_and_field(seed=7 + d)
    vf = DeviceArray.from_host(f)
    ns, nk, nl = fields.dir_extents(9, 12, 7, 6, d)
    rs = DeviceArray((ns + 6, nk, nl, 3))
    wm.check(wm.lib.weno_reshape(C.byref(g), d, wm.F64, vf.vp, rs.vp))
    assert np.array_equal(rs.download(), oracle.reshape(og, d, f))
    vL = DeviceArray((ns, nk, nl, 3))
    vR = DeviceArray((ns, nk, nl, 3))
    wm.check(wm.lib.weno_order1_copy(C.byref(g), d, wm.F64, vf.vp, vL.vp, vR.vp))
    cL, cR = oracle.order1_copy(og, d, f)
    assert np.array_equal(vL.download(), cL) and np.array_equal(vR.download(), cR)
    # MP on WENO5 output, standalone kernel reading the reshaped array
    cs = fields.harness_coefficients(ns - 6, 0.05)
    eL, eR = oracle.weno(og, 5, d, oracle.reshape(og, d, f), cs, cL, cR)
    mL, mR = oracle.mp(og, d, oracle.reshape(og, d, f), eL, eR)
    vL.upload(eL)
    vR.upload(eR)
    wm.check(wm.lib.weno_mp_limit(C.byref(g), d, wm.SRC_RESHAPED, wm.F64, rs.vp, vL.vp, vR.vp))
    assert rel_err(vL.download(), mL) <= 1e-13 and rel_err(vR.download(), mR) <= 1e-13


@pytest.mark.parametrize("order", [5, 7])
def test_teno_vs_oracle(order):
    g, og, f = _grid_and_field(seed=21 + order, wenojs=False, teno=True, teno_ct=1e-5)
    for d in range(3):
        ns = fields.dir_extents(9, 12, 7, 6, d)[0]
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs, coef = upload_coeffs(cs)
        vf = DeviceArray.from_host(f)
        cL, cR = oracle.order1_copy(og, d, f)
        eL, eR = oracle.weno(og, order, d, oracle.reshape(og, d, f), cs, cL, cR)
        L, R = run_hip(g, order, d, vf, wm.SRC_FIELD, coef, wm.FUSE_COPY)
        assert rel_err(L, eL) <= 1e-13 and rel_err(R, eR) <= 1e-13


@pytest.mark.parametrize("order", [5, 7])
def test_teno_ct_on_a_reference_ratio(order):
    """teno_CT set exactly to a cut-off ratio alpha_q/sum(alpha) the oracle
    (the reference's arithmetic) forms at one cell: there the reference keeps
    the stencil (ratio < CT is false), and a ratio formed with another
    rounding would fall on either side.  The kernel must send that cell to
    its exact arithmetic and agree with the oracle, at that cell to 1e-12
    element-wise.  For WENO5 the oracle itself is checked to change that
    cell when CT moves one ulp up, so a wrong decision could not pass.
    TENO5's alpha_q = x_q^6 is the C library's pow in the reference and a
    correctly rounded x^6 on the GPU; glibc's pow is not correctly rounded
    for about 0.3 % of these bases (its own last-ulp errors, which no other
    implementation reproduces), so CT is placed only on cells whose three
    powers pow returns correctly rounded (checked with exact rationals)."""
    g0, og0, f = _grid_and_field(seed=41 + order, wenojs=False, teno=True, teno_ct=1e-5)
    d = 0
    ns = fields.dir_extents(9, 12, 7, 6, d)[0]
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
    devs, coef = upload_coeffs(cs)
    vf = DeviceArray.from_host(f)
    v_rs = oracle.reshape(og0, d, f)
    cL, cR = oracle.order1_copy(og0, d, f)
    ratios, bases = oracle.teno_ratios(og0, order, d, v_rs, cs)
    core = ratios[3:-3]
    flat = core.reshape(-1)
    # cells whose ratio is neither ~0 nor ~1 (WENO5) or any non-zero (WENO7)
    cand = np.flatnonzero((flat > 1e-8) & (flat < 0.5)) if order == 5 else np.flatnonzero(flat > 0)
    if order == 5:
        xb = bases[3:-3].reshape(-1, 3)

        def pow_is_cr(cell_stencil):
            return all(math.pow(t, 6.0) == float(Fraction(t) ** 6) for t in xb[cell_stencil // 3])
        cand = np.array([c for c in cand[np.linspace(0, cand.size - 1, 40).astype(int)] if pow_is_cr(c)])
    assert cand.size >= 6
    picks = cand[np.linspace(0, cand.size - 1, 6).astype(int)]
    checked = 0
    for idx in picks:
        rho = float(flat[idx])
        cell = np.unravel_index(idx, core.shape)[:4]
        cell = (cell[0] + 3,) + tuple(cell[1:])
        g, og, _ = _grid_and_field(seed=41 + order, wenojs=False, teno=True, teno_ct=rho)
        eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
        if order == 5:
            _, og_up, _ = _grid_and_field(seed=41 + order, wenojs=False, teno=True,
                                          teno_ct=float(np.nextafter(rho, np.inf)))
            uL, uR = oracle.weno(og_up, order, d, v_rs, cs, cL, cR)
            if uL[cell] == eL[cell] and uR[cell] == eR[cell]:
                continue            # this cell's decision does not move its output
        L, R = run_hip(g, order, d, vf, wm.SRC_FIELD, coef, wm.FUSE_COPY)
        print(f"teno{order} CT on ratio {rho!r} at cell {cell}", flush=True)
        for name, got, ref in (("L", L, eL), ("R", R, eR)):
            if rel_err(got, ref) > 1e-13:
                err = np.abs(np.nan_to_num(got - ref))
                w = np.unravel_index(np.argmax(err), err.shape)
                print(f"  worst {name} at {w}: got {got[w]!r} ref {ref[w]!r} ratios {ratios[w].tolist()!r}",
                      flush=True)
        assert rel_err(L, eL) <= 1e-13 and rel_err(R, eR) <= 1e-13
        for got, ref in ((L[cell], eL[cell]), (R[cell], eR[cell])):
            assert ref != 0.0 and abs(got - ref) <= 1e-12 * abs(ref), (rho, cell, got, ref)
        checked += 1
    assert checked >= 3


def test_no_weight_branch_gives_nan():
    """wenojs=F and teno=F: no branch sets alpha, the reference yields 0/0."""
    g, og, f = _grid_and_field(seed=3, wenojs=False, teno=False)
    cs = fields.harness_coefficients(fields.dir_extents(9, 12, 7, 6, 0)[0] - 6, 0.03)
    devs, coef = upload_coeffs(cs)
    vf = DeviceArray.from_host(f)
    for order in (3, 5, 7):
        L, R = run_hip(g, order, 0, vf, wm.SRC_FIELD, coef, wm.FUSE_COPY)
        core = slice(0, None) if order == 3 else slice(3, -3)
        assert np.isnan(L[core]).all() and np.isnan(R[core]).all()


@pytest.mark.parametrize("order", [5, 7])
def test_fp32_and_bf16_storage(order):
    g, og, f = _grid_and_field(seed=40 + order, recipe=2)
    for d in range(3):
        ns = fields.dir_extents(9, 12, 7, 6, d)[0]
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs, coef = upload_coeffs(cs)
        # FP32 state vs the FP64 oracle on the FP32-rounded field
        f32 = f.astype(np.float32)
        cL, cR = oracle.order1_copy(og, d, f32.astype(np.float64))
        eL, eR = oracle.weno(og, order, d, oracle.reshape(og, d, f32.astype(np.float64)), cs, cL, cR)
        L, R = run_hip(g, order, d, DeviceArray.from_host(f32), wm.SRC_FIELD, coef, wm.FUSE_COPY,
                       dtype=wm.F32)
        assert rel_err(L, eL) <= 1e-6 and rel_err(R, eR) <= 1e-6
        # BF16 state
        bits = (f32.view(np.uint32) + 0x7FFF + ((f32.view(np.uint32) >> 16) & 1)) >> 16
        bf = np.asfortranarray(bits.astype(np.uint16))
        fb = (bf.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        cL, cR = oracle.order1_copy(og, d, fb)
        eL, eR = oracle.weno(og, order, d, oracle.reshape(og, d, fb), cs, cL, cR)
        L, R = run_hip(g, order, d, DeviceArray.from_host(bf), wm.SRC_FIELD, coef, wm.FUSE_COPY,
                       dtype=wm.BF16)
        assert rel_err(L, eL) <= 1e-6 and rel_err(R, eR) <= 1e-6


def test_golden_weno5_full_scope_200():
    """The reference's own test (wenoss.sh:351-387): weno5 mode, full scope,
    default 200^3 namelist, checksums vs tests/golden/weno5.txt at rtol 1e-10;
    run through the Python mirror of the module's call surface."""
    from wenostandalone_amd.standalone import WenoStandalone
    from tests.conftest import GOLDEN
    import os
    gold = dict(l.split() for l in open(os.path.join(GOLDEN, "weno5.txt")))
    ws = WenoStandalone()
    ws.s_initialize_data()
    faces = [ws.new_faces(d) for d in (0, 0, 1, 1, 2, 2)]
    ws.s_run_weno5_kernels(*faces)
    chk = [ws.f_checksum(faces[2 * d]) + ws.f_checksum(faces[2 * d + 1]) for d in range(3)]
    got = {"X": chk[0], "Y": chk[1], "Z": chk[2], "T": sum(chk)}
    for k in "XYZT":
        assert abs(got[k] - float(gold[k])) <= 1e-10 * abs(float(gold[k])), (k, got[k], gold[k])
    for a in faces:
        a.free()
    ws.s_finalize_data()
