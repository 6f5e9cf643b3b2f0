"""The consumer step on the GPU (weno_reconstruct_flux, weno_flux_divergence)
against the oracle's reconstruction followed by its flux divergence
(oracle_flux_divergence, pinned in tests/test_oracle.py).

Tolerances (cells outside j = 4 .. ns-5 must keep their values bitwise):
  * weno_flux_divergence, on the same faces as the oracle: the flux
    arithmetic only, |rhs - ref| <= 1e-13 (|F(j+1/2)| + |F(j-1/2)|) inv_dx;
  * weno_reconstruct_flux: its faces agree with the oracle's to the
    reconstruction's bound, 1e-13 relative to max(|face|, 1e-3 max|face|)
    (conftest.rel_err), and rhs carries that error through the flux:
    |rhs - ref| <= 8e-13 s Q inv_dx, Q = the largest such face magnitude of
    the four faces of the cell's two interfaces, s = |a| + |b| Q (each of
    the four faces moves F by at most 2 s times its error)."""
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


def _flux_scale(vL, vR, a, b, inv_dx):
    def F(qL, qR):
        fl, fr = a * qL + 0.5 * b * qL * qL, a * qR + 0.5 * b * qR * qR
        s = np.maximum(np.abs(a + b * qL), np.abs(a + b * qR))
        return 0.5 * (fl + fr) - 0.5 * s * (qR - qL)
    ns = vL.shape[0]
    Fp = F(vL[4:ns - 4], vR[5:ns - 3])
    Fm = F(vL[3:ns - 5], vR[4:ns - 4])
    return (np.abs(Fp) + np.abs(Fm)) * abs(inv_dx)


def _face_bound(vL, vR, fx):
    ns = vL.shape[0]
    floor = 1e-3 * max(np.abs(vL[3:ns - 3]).max(), np.abs(vR[3:ns - 3]).max())
    Q = np.maximum.reduce([np.abs(vL[3:ns - 5]), np.abs(vL[4:ns - 4]), np.abs(vR[4:ns - 4]), np.abs(vR[5:ns - 3])])
    Q = np.maximum(Q, floor)
    return 8e-13 * (abs(fx.a) + abs(fx.b) * Q) * Q * abs(fx.inv_dx)


def _check(got, ref, vL, vR, fx, pre, same_faces=False):
    ns = got.shape[0]
    err = np.abs(got[4:ns - 4] - ref[4:ns - 4])
    if same_faces:
        tol = 1e-13 * _flux_scale(vL, vR, fx.a, fx.b, fx.inv_dx) + 1e-300
    else:
        tol = _face_bound(vL, vR, fx) + 1e-300
    worst = float(np.max(err / tol))
    assert worst <= 1.0, worst
    assert np.array_equal(got[:4], pre[:4]) and np.array_equal(got[ns - 4:], pre[ns - 4:])


def _coef(cs):
    devs = [DeviceArray.from_host(c) for c in cs]
    return devs, wm.Coeffs(*[x.ptr for x in devs])


CASES = [  # order, grid switches, fused MP
    (5, {}, False), (7, {}, False), (3, {}, False), (5, {}, True), (7, {}, True),
    (5, {"wenojs": False, "teno": True}, False), (7, {"wenojs": False, "wenoz": True}, False),
    (5, {"wenojs": False, "mapped_weno": True}, False),
]


@pytest.mark.parametrize("order,kw,mp", CASES)
def test_fused_flux_vs_oracle(order, kw, mp):
    m, n, p, v = 13, 10, 8, 2
    g = wm.make_grid(m, n, p, v=v, buff=6, **kw)
    og = oracle.grid(m, n, p, v=v, buff=6, **kw)
    f = fields.seeded_field(m, n, p, v, 6, seed=40 + order)
    vf = DeviceArray.from_host(f)
    for d in range(3):
        ns, nk, nl = fields.dir_extents(m, n, p, 6, d)
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs, coef = _coef(cs)
        v_rs = oracle.reshape(og, d, f)
        cL, cR = oracle.order1_copy(og, d, f)
        eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
        if mp:
            eL, eR = oracle.mp(og, d, v_rs, eL, eR)
        rs = DeviceArray.from_host(v_rs)
        for fx in (wm.Flux(1.3, 0.0, 5.0), wm.Flux(-0.4, 1.0, 2.0)):
            pre = np.asfortranarray(np.random.default_rng(d).standard_normal((ns, nk, nl, v)))
            ref = oracle.flux_divergence(og, d, eL, eR, fx.a, fx.b, fx.inv_dx, rhs=pre)
            for src, kind in ((vf, wm.SRC_FIELD), (rs, wm.SRC_RESHAPED)):
                out = DeviceArray.from_host(pre)
                wm.check(wm.lib.weno_reconstruct_flux(C.byref(g), order, d, kind, src.vp, C.byref(coef),
                                                      C.byref(fx), out.vp, wm.FUSE_MP if mp else 0))
                _check(out.download(), ref, eL, eR, fx, pre)
                # accumulate on top of the same call's result
                wm.check(wm.lib.weno_reconstruct_flux(C.byref(g), order, d, kind, src.vp, C.byref(coef),
                                                      C.byref(fx), out.vp,
                                                      wm.ACCUMULATE | (wm.FUSE_MP if mp else 0)))
                ref2 = oracle.flux_divergence(og, d, eL, eR, fx.a, fx.b, fx.inv_dx, rhs=ref, accumulate=True)
                _check(out.download(), ref2, eL, eR, fx, pre)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_flux_divergence_of_faces_vs_oracle(d):
    m, n, p, v = 11, 14, 9, 3
    g = wm.make_grid(m, n, p, v=v, buff=6)
    og = oracle.grid(m, n, p, v=v, buff=6)
    f = fields.seeded_field(m, n, p, v, 6, seed=70 + d)
    ns, nk, nl = fields.dir_extents(m, n, p, 6, d)
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
    cL, cR = oracle.order1_copy(og, d, f)
    eL, eR = oracle.weno(og, 5, d, oracle.reshape(og, d, f), cs, cL, cR)
    fx = wm.Flux(0.2, 0.9, 3.0)
    pre = np.zeros((ns, nk, nl, v), order="F")
    ref = oracle.flux_divergence(og, d, eL, eR, fx.a, fx.b, fx.inv_dx)
    L, R, out = DeviceArray.from_host(eL), DeviceArray.from_host(eR), DeviceArray.from_host(pre)
    wm.check(wm.lib.weno_flux_divergence(C.byref(g), d, L.vp, R.vp, C.byref(fx), out.vp, 0))
    _check(out.download(), ref, eL, eR, fx, pre, same_faces=True)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_fused_flux_at_bench_shape(d):
    """200^3 (v=2), WENO5-JS, Burgers: the fused pass equals
    weno_reconstruct followed by weno_flux_divergence to the tolerance, and
    the oracle."""
    m, v = 199, 2
    g = wm.make_grid(m, v=v, buff=6)
    og = oracle.grid(m, v=v, buff=6)
    f = fields.seeded_field(m, m, m, v, 6, seed=0)
    ns = m + 1 + 12
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
    devs, coef = _coef(cs)
    vf = DeviceArray.from_host(f)
    fx = wm.Flux(0.0, 1.0, float(m + 1))
    shape = (ns, ns, ns, v)
    pre = np.zeros(shape, order="F")
    out = DeviceArray.from_host(pre)
    wm.check(wm.lib.weno_reconstruct_flux(C.byref(g), 5, d, wm.SRC_FIELD, vf.vp, C.byref(coef), C.byref(fx),
                                          out.vp, 0))
    cL, cR = oracle.order1_copy(og, d, f)
    eL, eR = oracle.weno(og, 5, d, oracle.reshape(og, d, f), cs, cL, cR)
    ref = oracle.flux_divergence(og, d, eL, eR, fx.a, fx.b, fx.inv_dx)
    _check(out.download(), ref, eL, eR, fx, pre)
    L, R, two = DeviceArray(shape), DeviceArray(shape), DeviceArray.from_host(pre)
    wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, d, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coef), L.vp, R.vp,
                                     wm.FUSE_COPY))
    wm.check(wm.lib.weno_flux_divergence(C.byref(g), d, L.vp, R.vp, C.byref(fx), two.vp, 0))
    _check(two.download(), ref, eL, eR, fx, pre)


@pytest.mark.parametrize("order,m,mp", [(5, 1017, False), (7, 700, True), (3, 130, False)])
def test_fused_flux_long_rows(order, m, mp):
    """Rows longer than one workgroup's consumer waves cover (WENO5: 15
    waves of 62 cells, WENO7: 7), so each wave walks several segments, and
    the producer's tiles shrink to fewer pencils; x (contiguous rows) and y
    (strided rows) against the oracle."""
    n, p, v = 5, 4, 2
    g = wm.make_grid(m, n, p, v=v, buff=6)
    og = oracle.grid(m, n, p, v=v, buff=6)
    f = fields.seeded_field(m, n, p, v, 6, seed=90 + order)
    vf = DeviceArray.from_host(f)
    fx = wm.Flux(0.3, 0.7, 4.0)
    for d in (0, 1):
        ns, nk, nl = fields.dir_extents(m, n, p, 6, d)
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs, coef = _coef(cs)
        cL, cR = oracle.order1_copy(og, d, f)
        v_rs = oracle.reshape(og, d, f)
        eL, eR = oracle.weno(og, order, d, v_rs, cs, cL, cR)
        if mp:
            eL, eR = oracle.mp(og, d, v_rs, eL, eR)
        pre = np.asfortranarray(np.random.default_rng(d).standard_normal((ns, nk, nl, v)))
        ref = oracle.flux_divergence(og, d, eL, eR, fx.a, fx.b, fx.inv_dx, rhs=pre)
        out = DeviceArray.from_host(pre)
        wm.check(wm.lib.weno_reconstruct_flux(C.byref(g), order, d, wm.SRC_FIELD, vf.vp, C.byref(coef), C.byref(fx),
                                              out.vp, wm.FUSE_MP if mp else 0))
        _check(out.download(), ref, eL, eR, fx, pre)


def test_flux_argument_checks():
    g = wm.make_grid(9, v=1)
    fx = wm.Flux(1.0, 0.0, 1.0)
    cf = wm.Coeffs()
    assert wm.lib.weno_reconstruct_flux(C.byref(g), 4, 0, wm.SRC_FIELD, None, C.byref(cf), C.byref(fx), None, 0) == -1
    assert wm.lib.weno_reconstruct_flux(C.byref(g), 5, 0, wm.SRC_FIELD, None, C.byref(cf), None, None, 0) == -1
    assert wm.lib.weno_reconstruct_flux(C.byref(g), 5, 0, wm.SRC_FIELD, None, C.byref(cf), C.byref(fx), None,
                                        wm.FUSE_COPY) == -1
    assert wm.lib.weno_flux_divergence(C.byref(g), 0, None, None, C.byref(fx), None, wm.FUSE_MP) == -1
    long_g = wm.make_grid(1100, 2, 2, v=1)
    assert wm.lib.weno_reconstruct_flux(C.byref(long_g), 7, 0, wm.SRC_FIELD, None, C.byref(cf), C.byref(fx),
                                        None, 0) == -1
    assert "at most 1024" in wm._lib.last_error()
