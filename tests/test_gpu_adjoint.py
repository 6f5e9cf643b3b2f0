"""Adjoint of WENO5/7-JS on the GPU (weno_reconstruct_vjp), FP64.

Checked two independent ways:
  * against the oracle's FP64 CPU gradient (oracle_weno_vjp, itself pinned by
    a complex-step derivative in tests/test_oracle.py), normwise
    (max|a-b| <= tol * max|b|): each gradient entry sums terms of both signs
    that scale like 1/beta, so entries that nearly cancel carry the rounding
    of the larger terms and an element-wise relative bound is meaningless.
    tol = 1e-11 on smooth data; 1e-10 on the noisy seeded field, whose 1e-3
    noise makes beta ~ 1e-6 and the cancelling terms ~1e4 x the result
    (the oracle divides alpha by sum(alpha) per stencil, the GPU uses one
    shared reciprocal: ulp-level differences in the weights, amplified);
  * against central finite differences (h = 1e-6) of the GPU forward on
    cells of a 16^3 sub-block: 1e-6 relative (FD truncation and roundoff).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray
pytestmark = pytest.mark.gpu


def norm_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))


def field_from_reshaped_grad(gv_rs, d, nx, ny, nz, v):
    """Scatter of the reshape (a permutation of a sub-box of v_vf)."""
    out = np.zeros((nx + 6, ny + 6, nz + 6, v), order="F")
    if d == 0:
        out[:, 3:3 + ny, 3:3 + nz, :] = gv_rs
    elif d == 1:
        out[3:3 + nx, :, 3:3 + nz, :] = gv_rs.transpose(1, 0, 2, 3)
    else:
        out[3:3 + nx, 3:3 + ny, :, :] = gv_rs.transpose(2, 1, 0, 3)
    return out


def cotangents(shape, seed):
    rng = np.random.default_rng(seed)
    gL = np.asfortranarray(rng.standard_normal(shape))
    gR = np.asfortranarray(rng.standard_normal(shape))
    gL[:3] = gL[-3:] = 0.0
    gR[:3] = gR[-3:] = 0.0
    return gL, gR


def smooth_field(m, n, p, v, b):
    nx, ny, nz = fields.extents(m, n, p, b)
    x = np.linspace(0, 1, nx + 6)[:, None, None, None]
    y = np.linspace(0, 1, ny + 6)[None, :, None, None]
    z = np.linspace(0, 1, nz + 6)[None, None, :, None]
    k = np.arange(1, v + 1)[None, None, None, :]
    return np.asfortranarray(1.5 + np.sin(2.0 * x * k) * np.cos(1.3 * y) + 0.3 * np.sin(2.2 * z + k))


@pytest.mark.parametrize("smooth", [True, False])
@pytest.mark.parametrize("order", [5, 7])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_vjp_vs_oracle(order, d, smooth):
    m, n, p, v, b = 9, 12, 7, 2, 6
    g = wm.make_grid(m, n, p, v=v, buff=b)
    og = oracle.grid(m, n, p, v=v, buff=b)
    f = smooth_field(m, n, p, v, b) if smooth else fields.seeded_field(m, n, p, v, b, seed=60 + d)
    tol = 1e-11 if smooth else 1e-10
    nx, ny, nz = fields.extents(m, n, p, b)
    ns, nk, nl = fields.dir_extents(m, n, p, b, d)
    cs = fields.random_coefficients(ns - 6, seed=order + d)
    devs = [DeviceArray.from_host(c) for c in cs]
    coef = wm.Coeffs(*[x.ptr for x in devs])
    gL, gR = cotangents((ns, nk, nl, v), seed=d)
    v_rs = oracle.reshape(og, d, f)
    ref_rs = oracle.vjp(og, order, d, v_rs, cs, gL, gR)
    dgL, dgR = DeviceArray.from_host(gL), DeviceArray.from_host(gR)
    # reshaped input layout
    rs = DeviceArray.from_host(v_rs)
    gv = DeviceArray(v_rs.shape)
    gv.fill(0.0)
    wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, d, wm.SRC_RESHAPED, rs.vp, C.byref(coef),
                                         dgL.vp, dgR.vp, gv.vp))
    assert norm_err(gv.download(), ref_rs) <= tol
    # padded-field layout, accumulated onto a non-zero gradient
    vf = DeviceArray.from_host(f)
    base = np.full(f.shape, 0.5, order="F")
    gvf = DeviceArray.from_host(base)
    wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, d, wm.SRC_FIELD, vf.vp, C.byref(coef),
                                         dgL.vp, dgR.vp, gvf.vp))
    exp = base + field_from_reshaped_grad(ref_rs, d, nx, ny, nz, v)
    assert norm_err(gvf.download() - base, exp - base) <= tol


@pytest.mark.parametrize("order", [5, 7])
def test_vjp_central_differences_16cube(order):
    """A 16^3 block (m=n=p=3, buff 6 -> 16^3 `is` box), x sweep: VJP entries
    against central differences of the GPU forward on 48 sampled cells."""
    m, v, b = 3, 1, 6
    g = wm.make_grid(m, v=v, buff=b)
    f = fields.seeded_field(m, m, m, v, b, seed=77)
    ns, nk, nl = fields.dir_extents(m, m, m, b, 0)
    cs = fields.harness_coefficients(ns - 6, 0.03)
    devs = [DeviceArray.from_host(c) for c in cs]
    coef = wm.Coeffs(*[x.ptr for x in devs])
    gL, gR = cotangents((ns, nk, nl, v), seed=5)
    vf = DeviceArray.from_host(f)
    gv = DeviceArray(f.shape)
    gv.fill(0.0)
    dgL, dgR = DeviceArray.from_host(gL), DeviceArray.from_host(gR)
    wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, 0, wm.SRC_FIELD, vf.vp, C.byref(coef),
                                         dgL.vp, dgR.vp, gv.vp))
    grad = gv.download()
    L = DeviceArray((ns, nk, nl, v))
    R = DeviceArray((ns, nk, nl, v))

    def fwd(field):
        vf.upload(field)
        wm.check(wm.lib.weno_reconstruct(C.byref(g), order, 0, wm.SRC_FIELD, wm.F64, vf.vp, C.byref(coef),
                                         L.vp, R.vp, wm.FUSE_COPY))
        return L.download(), R.download()

    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(48):
        X = int(rng.integers(1, ns + 5))
        Y = int(rng.integers(3, 3 + nk))
        Z = int(rng.integers(3, 3 + nl))
        fp = f.copy(order="F")
        fm = f.copy(order="F")
        fp[X, Y, Z, 0] += h
        fm[X, Y, Z, 0] -= h
        Lp, Rp = fwd(fp)
        Lm, Rm = fwd(fm)
        core = (slice(3, ns - 3), slice(None), slice(None), slice(None))
        fd = (np.sum(gL[core] * (Lp - Lm)[core]) + np.sum(gR[core] * (Rp - Rm)[core])) / (2 * h)
        assert abs(grad[X, Y, Z, 0] - fd) <= 1e-6 * max(1.0, abs(fd)), (X, Y, Z, grad[X, Y, Z, 0], fd)


@pytest.mark.parametrize("order", [5, 7])
@pytest.mark.parametrize("m,n,p,v,dirs", [
    (47, 47, 47, 8, (0, 1, 2)),   # 54-point rows (WENO7: 64-thread workgroups of one lane group)
    (243, 17, 17, 2, (0,)),       # 250-point rows (WENO7: 256-thread workgroups, two per CU)
    (127, 59, 59, 4, (0,)),       # 135-point rows (WENO7: 448-thread workgroups, ~6 tiles each)
])
def test_vjp_many_tiles_per_workgroup(order, m, n, p, v, dirs):
    """Launches where every workgroup runs several tiles in turn, against the
    oracle: WENO5 on the whole-row tile kernel (the producer wave loads tile
    t+1 while the consumer waves work on tile t; each wave owns a segment of
    60 input positions, so the longer rows span several waves), WENO7 on the
    gather kernel with the shapes noted per case."""
    b = 6
    g = wm.make_grid(m, n, p, v=v, buff=b)
    og = oracle.grid(m, n, p, v=v, buff=b)
    f = fields.seeded_field(m, n, p, v, b, seed=90 + order)
    for d in dirs:
        print(f"vjp many tiles: order {order} box {m}x{n}x{p}x{v} dir {d}", flush=True)
        ns, nk, nl = fields.dir_extents(m, n, p, b, d)
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs = [DeviceArray.from_host(c) for c in cs]
        coef = wm.Coeffs(*[x.ptr for x in devs])
        gL, gR = cotangents((ns, nk, nl, v), seed=20 + d)
        v_rs = oracle.reshape(og, d, f)
        ref_rs = oracle.vjp(og, order, d, v_rs, cs, gL, gR)
        dgL, dgR = DeviceArray.from_host(gL), DeviceArray.from_host(gR)
        rs = DeviceArray.from_host(v_rs)
        gv = DeviceArray.from_host(np.full(v_rs.shape, 0.25, order="F"))
        wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, d, wm.SRC_RESHAPED, rs.vp, C.byref(coef),
                                             dgL.vp, dgR.vp, gv.vp))
        assert norm_err(gv.download() - 0.25, ref_rs) <= 1e-10, d


@pytest.mark.parametrize("order,m", [(5, 505), (7, 472)])
def test_vjp_longest_supported_sweep(order, m):
    """The longest sweeps the adjoint supports (WENO5 512, WENO7 479 core
    points) against the oracle; one more point is refused."""
    b, n, p, v, d = 6, 3, 3, 1, 0
    g = wm.make_grid(m, n, p, v=v, buff=b)
    og = oracle.grid(m, n, p, v=v, buff=b)
    f = fields.seeded_field(m, n, p, v, b, seed=95 + order)
    ns, nk, nl = fields.dir_extents(m, n, p, b, d)
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
    devs = [DeviceArray.from_host(c) for c in cs]
    coef = wm.Coeffs(*[x.ptr for x in devs])
    gL, gR = cotangents((ns, nk, nl, v), seed=7)
    v_rs = oracle.reshape(og, d, f)
    ref_rs = oracle.vjp(og, order, d, v_rs, cs, gL, gR)
    dgL, dgR = DeviceArray.from_host(gL), DeviceArray.from_host(gR)
    rs = DeviceArray.from_host(v_rs)
    gv = DeviceArray(v_rs.shape)
    gv.fill(0.0)
    wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, d, wm.SRC_RESHAPED, rs.vp, C.byref(coef),
                                         dgL.vp, dgR.vp, gv.vp))
    assert norm_err(gv.download(), ref_rs) <= 1e-10
    g1 = wm.make_grid(m + 1, n, p, v=v, buff=b)
    rc = wm.lib.weno_reconstruct_vjp(C.byref(g1), order, d, wm.SRC_RESHAPED, rs.vp, C.byref(coef),
                                     dgL.vp, dgR.vp, gv.vp)
    assert rc == -1 and "core points" in wm._lib.last_error()


def _vjp_pair(order, d, f, eps, m, n, p, v, b, seed):
    g = wm.make_grid(m, n, p, v=v, buff=b, eps=eps)
    og = oracle.grid(m, n, p, v=v, buff=b, eps=eps)
    ns, nk, nl = fields.dir_extents(m, n, p, b, d)
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
    devs = [DeviceArray.from_host(c) for c in cs]
    coef = wm.Coeffs(*[x.ptr for x in devs])
    gL, gR = cotangents((ns, nk, nl, v), seed=seed)
    v_rs = oracle.reshape(og, d, f)
    ref = oracle.vjp(og, order, d, v_rs, cs, gL, gR)
    dgL, dgR = DeviceArray.from_host(gL), DeviceArray.from_host(gR)
    rs = DeviceArray.from_host(v_rs)
    gv = DeviceArray(v_rs.shape)
    gv.fill(0.0)
    wm.check(wm.lib.weno_reconstruct_vjp(C.byref(g), order, d, wm.SRC_RESHAPED, rs.vp, C.byref(coef),
                                         dgL.vp, dgR.vp, gv.vp))
    return gv.download(), ref


@pytest.mark.parametrize("order", [5, 7])
@pytest.mark.parametrize("case", ["scale1e22", "scale1e-30", "spikes", "nan"])
def test_vjp_out_of_range_products(order, case):
    """Inputs whose shared weight product leaves the normal range, so lanes
    take the per-stencil IEEE form: a field scaled by 1e22 (overflow) or by
    1e-30 with eps 1e-100 (underflow) in every lane; isolated 1e22 cells in
    an O(1) field (fast and slow lanes in one wavefront: the WENO5 tile kernel
    redoes such a tile's passes); NaN cells (NaN in exactly the gradient
    entries the oracle makes NaN, the rest to the usual bound).  WENO
    reconstruction is homogeneous of degree one in the state, so the
    gradient stays O(1) at every scale: normwise 1e-10 against the oracle."""
    m, n, p, v, b = 37, 11, 9, 2, 6
    rng = np.random.default_rng(400 + order)
    f = fields.seeded_field(m, n, p, v, b, seed=401)
    eps = 1e-16
    if case == "scale1e22":
        f = np.asfortranarray(f * 1e22)
    elif case == "scale1e-30":
        f = np.asfortranarray(f * 1e-30)
        eps = 1e-100
    elif case == "spikes":
        for _ in range(12):
            f[tuple(int(rng.integers(3, s - 3)) for s in f.shape[:3]) + (0,)] = 1e22
    else:
        for _ in range(3):
            f[tuple(int(rng.integers(3, s - 3)) for s in f.shape[:3]) + (1,)] = np.nan
    for d in range(3):
        print(f"vjp out of range: order {order} case {case} dir {d}", flush=True)
        got, ref = _vjp_pair(order, d, f, eps, m, n, p, v, b, seed=410 + d)
        nan = np.isnan(ref)
        assert np.array_equal(np.isnan(got), nan), (d, int(nan.sum()), int(np.isnan(got).sum()))
        if case == "nan":
            assert nan.any(), d
        fin = ~nan
        assert norm_err(got[fin], ref[fin]) <= 1e-10, d
