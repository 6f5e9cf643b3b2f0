"""TENO, WENO-Z and mapped WENO on the producer/consumer sweep kernel
(FP64 state without the fused limiter: sweep5_branch_kernel for WENO5,
sweep7_kernel for WENO7, both planned by plan_sweep5 for rows of at least 96
outputs), against the oracle:

  * extents that leave part-filled tiles in the x, y and z sweeps;
  * rows of two j-chunks (312 cells along each axis in turn);
  * field and reshaped sources; core-only launches keep arbitrary values in
    the halo cells of vL/vR bit for bit, fused-copy launches write the
    order-1 copy there (bitwise);
  * WENO-Z and mapped WENO on j-varying coefficients whose smoothness
    indicators are positive semi-definite (fields.psd_coefficients) and a
    field with a 1e3 jump and exactly constant boxes (recipe 2); TENO on the
    harness coefficients and the seeded field (its cut-off is discontinuous;
    these inputs keep every weight far from teno_CT).

Tolerance on the reconstructed core: 1e-13 relative to max(|face|, largest
|v| of the face's stencil) (rel_err_stencil), as for the same branches on the
general kernel (test_gpu_branches.py)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray
from tests.conftest import rel_err_stencil, stencil_scale

pytestmark = pytest.mark.gpu

BRANCH_KW = {"teno": {"teno": True}, "wenoz": {"wenoz": True}, "mapped_weno": {"mapped_weno": True}}


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield


def _coef(cs):
    devs = [DeviceArray.from_host(c, np.float64) for c in cs]
    return devs, wm.Coeffs(*[d.ptr for d in devs])


def _run(g, order, d, src, kind, coef, flags, shape, pre):
    vL, vR = DeviceArray(shape), DeviceArray(shape)
    vL.upload(pre[0])
    vR.upload(pre[1])
    wm.check(wm.lib.weno_reconstruct(C.byref(g), order, d, kind, wm.F64, src.vp, C.byref(coef), vL.vp, vR.vp,
                                     flags))
    return vL.download(), vR.download()


def _check(branch, order, m, n, p, v, dirs, seed):
    kw = BRANCH_KW[branch]
    g = wm.make_grid(m, n, p, v=v, buff=6, wenojs=False, **kw)
    og = oracle.grid(m, n, p, v=v, buff=6, wenojs=False, **kw)
    recipe = 0 if branch == "teno" else 2
    f = fields.seeded_field(m, n, p, v, 6, seed=seed, recipe=recipe)
    vf = DeviceArray.from_host(f)
    rng = np.random.default_rng(seed)
    for d in dirs:
        ns, nk, nl = fields.dir_extents(m, n, p, 6, d)
        assert ns - 6 >= 96, "rows this short stay on the general kernel"
        if branch == "teno":
            cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        else:
            cs = fields.psd_coefficients(ns - 6, order, seed=seed + 10 * d)
        devs, coef = _coef(cs)
        v_rs = oracle.reshape(og, d, f)
        rs = DeviceArray.from_host(v_rs)
        sc = stencil_scale(v_rs, order // 2)
        shape = (ns, nk, nl, v)
        cL, cR = oracle.order1_copy(og, d, f)
        # core-only launches must leave whatever the halo cells hold
        pL = np.asfortranarray(rng.standard_normal(shape))
        pR = np.asfortranarray(rng.standard_normal(shape))
        eL, eR = oracle.weno(og, order, d, v_rs, cs, pL, pR)
        core = slice(3, ns - 3)
        for src, kind in [(vf, wm.SRC_FIELD), (rs, wm.SRC_RESHAPED)]:
            print(f"branch sweep: {branch} WENO{order} box {m}x{n}x{p}x{v} dir {d} src {kind}", flush=True)
            L, R = _run(g, order, d, src, kind, coef, 0, shape, pre=(pL, pR))
            assert rel_err_stencil(L[core], eL[core], sc[core]) <= 1e-13, (d, kind, "core L")
            assert rel_err_stencil(R[core], eR[core], sc[core]) <= 1e-13, (d, kind, "core R")
            for h in (slice(0, 3), slice(ns - 3, ns)):
                assert np.array_equal(L[h], pL[h]) and np.array_equal(R[h], pR[h]), (d, kind, "halo kept")
            L, R = _run(g, order, d, src, kind, coef, wm.FUSE_COPY, shape, pre=(pL, pR))
            assert rel_err_stencil(L[core], eL[core], sc[core]) <= 1e-13, (d, kind, "copy L")
            assert rel_err_stencil(R[core], eR[core], sc[core]) <= 1e-13, (d, kind, "copy R")
            for h in (slice(0, 3), slice(ns - 3, ns)):
                assert np.array_equal(L[h], cL[h]) and np.array_equal(R[h], cR[h]), (d, kind, "halo copy")


@pytest.mark.parametrize("order", [5, 7])
@pytest.mark.parametrize("branch", list(BRANCH_KW))
def test_branch_sweep_partial_tiles(branch, order):
    # nx, ny, nz = 150, 146, 138 (even, so every sweep takes the kernel): the
    # 16-pencil tiles end part-filled and span two transverse rows
    _check(branch, order, 137, 133, 125, 2, (0, 1, 2), seed=700 + order)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("branch", list(BRANCH_KW))
def test_branch_sweep_two_chunk_rows(branch, axis):
    ext = [15, 17, 13]
    ext[axis] = 299            # 312 cells along the sweep: two tiles along j
    _check(branch, 5, *ext, 1, (axis,), seed=730 + axis)
