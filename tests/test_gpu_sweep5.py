"""The WENO5-JS FP64 producer/consumer kernel (weno_sweep5.hip).

weno_reconstruct routes WENO5-JS launches (both sides, rows of >= 96 outputs,
with or without the fused MP limiter) to this kernel; the other cases stay on
the general sweep kernel.  Checked here, 1e-13 relative (rel_err) throughout,
every case also with the limiter fused (WENO_FUSE_MP, against the oracle's
weno5 followed by its MP):
  * against the oracle on boxes whose three sweeps all take it: field input
    (contiguous x rows loaded in 16-byte pieces, y/z pencils transposed on the
    way into LDS) and reshaped input, core-only and fused halo copy (halo
    bitwise); core-only launches rewrite the halo cells with the values they
    hold, so arbitrary halo contents must come back bit for bit;
  * rows longer than one tile (ns = 312 along each axis in turn: two chunks),
    extents that leave partial tiles (16 pencils consecutive along the
    transverse axis of smallest input stride, so with an odd extent there a
    tile spans two transverse rows);
  * against the general kernel (variant bit 8) on the same inputs;
  * fields scaled by 1e22 and 1e-30 (eps 1e-100): the shared denominator
    (sum_L)(sum_R) over- or underflows, so lanes take the per-stencil IEEE
    path; isolated 1e22 cells mix both paths inside one wavefront;
  * NaN input cells: NaN in exactly the outputs whose stencil touches them;
  * FP32 state and BF16 state (FP32 faces): the kernel widens the state to
    FP64 as it fills LDS and runs the FP64 arithmetic, so its faces are the
    FP64 oracle's faces for the widened (rounded) state, rounded once to FP32:
    within one FP32 ulp (1.2e-7 relative, rel_err); halo cells bitwise.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import wenostandalone_amd as wm
from wenostandalone_amd import fields
from wenostandalone_amd.device import DeviceArray
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    wm.check(wm.lib.weno_init(0))
    yield
    for d in range(3):
        wm.lib.weno_set_tile(5, wm.F64, d, 0, 0, 0, -1)


def _coef(cs):
    devs = [DeviceArray.from_host(c, np.float64) for c in cs]
    return devs, wm.Coeffs(*[x.ptr for x in devs])


def _run(g, d, src, kind, coef, flags, pre=None, dtype=None):
    dtype = wm.F64 if dtype is None else dtype
    ns, nk, nl = fields.dir_extents(g.m_cells, g.n_cells, g.p_cells, g.buff_size, d)
    ft = np.float64 if dtype == wm.F64 else np.float32
    vL = DeviceArray((ns, nk, nl, g.v_size), ft)
    vR = DeviceArray((ns, nk, nl, g.v_size), ft)
    if pre is None:
        vL.fill(7.25)
        vR.fill(-3.5)
    else:
        vL.upload(pre[0])
        vR.upload(pre[1])
    wm.check(wm.lib.weno_reconstruct(C.byref(g), 5, d, kind, dtype, src.vp, C.byref(coef), vL.vp, vR.vp, flags))
    return vL.download(), vR.download()


def _case(m, n, p, v, seed, scale=None, eps=1e-16, nan_cells=()):
    g = wm.make_grid(m, n, p, v=v, buff=6, eps=eps)
    og = oracle.grid(m, n, p, v=v, buff=6, eps=eps)
    f = fields.seeded_field(m, n, p, v, 6, seed=seed)
    if scale is not None:
        f = np.asfortranarray(f * scale)
    for c in nan_cells:
        f[c] = np.nan
    return g, og, f


def _rel(a, b, split=None):
    """rel_err, taken separately over outputs of magnitude above and below
    `split` when a field mixes scales (each part with its own floor)."""
    if split is None:
        return rel_err(a, b)
    big = np.abs(b) > split
    errs = [0.0]
    for m in (big, ~big & ~np.isnan(b)):
        if m.any():
            errs.append(rel_err(a[m], b[m]))
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return np.inf
    return max(errs)


def _check_box(g, og, f, dirs=(0, 1, 2), seed=0, split=None):
    m, n, p = g.m_cells, g.n_cells, g.p_cells
    vf = DeviceArray.from_host(f)
    for d in dirs:
        ns = fields.dir_extents(m, n, p, 6, d)[0]
        cs = fields.random_coefficients(ns - 6, seed=seed + d)
        devs, coef = _coef(cs)
        v_rs = oracle.reshape(og, d, f)
        rs = DeviceArray.from_host(v_rs)
        cL, cR = oracle.order1_copy(og, d, f)
        eL, eR = oracle.weno(og, 5, d, v_rs, cs, cL, cR)
        mL, mR = oracle.mp(og, d, v_rs, eL, eR)
        for src, kind in [(vf, wm.SRC_FIELD), (rs, wm.SRC_RESHAPED)]:
            L, R = _run(g, d, src, kind, coef, 0, pre=(cL, cR))
            assert _rel(L, eL, split) <= 1e-13 and _rel(R, eR, split) <= 1e-13, (d, kind, "core")
            L, R = _run(g, d, src, kind, coef, wm.FUSE_COPY)
            assert _rel(L, eL, split) <= 1e-13 and _rel(R, eR, split) <= 1e-13, (d, kind, "copy")
            assert np.array_equal(L[:3], cL[:3], equal_nan=True)
            assert np.array_equal(R[-3:], cR[-3:], equal_nan=True)
            L, R = _run(g, d, src, kind, coef, wm.FUSE_MP, pre=(cL, cR))
            assert _rel(L, mL, split) <= 1e-13 and _rel(R, mR, split) <= 1e-13, (d, kind, "core+mp")
            L, R = _run(g, d, src, kind, coef, wm.FUSE_COPY | wm.FUSE_MP)
            assert _rel(L, mL, split) <= 1e-13 and _rel(R, mR, split) <= 1e-13, (d, kind, "copy+mp")
            assert np.array_equal(L[:3], cL[:3], equal_nan=True)
            assert np.array_equal(R[-3:], cR[-3:], equal_nan=True)


BOX = (139, 131, 127)      # nx, ny, nz = 152, 144, 140: every sweep takes the kernel


def test_all_three_sweeps_vs_oracle():
    g, og, f = _case(*BOX, 2, seed=501)
    _check_box(g, og, f, seed=11)


def test_core_only_keeps_arbitrary_halo():
    g, og, f = _case(*BOX, 1, seed=505)
    vf = DeviceArray.from_host(f)
    rng = np.random.default_rng(9)
    for d in range(3):
        ns, nk, nl = fields.dir_extents(*BOX, 6, d)
        cs = fields.random_coefficients(ns - 6, seed=70 + d)
        devs, coef = _coef(cs)
        pL = np.asfortranarray(rng.standard_normal((ns, nk, nl, 1)))
        pR = np.asfortranarray(rng.standard_normal((ns, nk, nl, 1)))
        L, R = _run(g, d, vf, wm.SRC_FIELD, coef, 0, pre=(pL, pR))
        halo = np.r_[0:3, ns - 3:ns]
        assert np.array_equal(L[halo], pL[halo]) and np.array_equal(R[halo], pR[halo]), d
        v_rs = oracle.reshape(og, d, f)
        eL, eR = oracle.weno(og, 5, d, v_rs, cs, pL, pR)
        assert rel_err(L, eL) <= 1e-13 and rel_err(R, eR) <= 1e-13, d
        L, R = _run(g, d, vf, wm.SRC_FIELD, coef, wm.FUSE_MP, pre=(pL, pR))
        assert np.array_equal(L[halo], pL[halo]) and np.array_equal(R[halo], pR[halo]), d
        mL, mR = oracle.mp(og, d, v_rs, eL, eR)
        assert rel_err(L, mL) <= 1e-13 and rel_err(R, mR) <= 1e-13, d


def test_odd_extents_and_partial_tiles():
    """nx = 153, ny = 143, nz = 140: partial 16-pencil tiles in every
    sweep, and tiles that span two transverse rows."""
    g, og, f = _case(140, 130, 127, 1, seed=512)
    _check_box(g, og, f, seed=21)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_two_chunk_rows(axis):
    ext = [7, 9, 8]
    ext[axis] = 299           # 312 cells along the sweep: two tiles along j
    g, og, f = _case(*ext, 1, seed=520 + axis)
    _check_box(g, og, f, dirs=(axis,), seed=30 + axis)


def test_matches_general_kernel():
    g, og, f = _case(*BOX, 1, seed=530)
    vf = DeviceArray.from_host(f)
    for d in range(3):
        ns = fields.dir_extents(*BOX, 6, d)[0]
        devs, coef = _coef(fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d]))
        for flags in (wm.FUSE_COPY, wm.FUSE_COPY | wm.FUSE_MP):
            fast = _run(g, d, vf, wm.SRC_FIELD, coef, flags)
            wm.check(wm.lib.weno_set_tile(5, wm.F64, d, 0, 0, 0, 8 | 2))
            try:
                gen = _run(g, d, vf, wm.SRC_FIELD, coef, flags)
            finally:
                wm.check(wm.lib.weno_set_tile(5, wm.F64, d, 0, 0, 0, -1))
            assert rel_err(fast[0], gen[0]) <= 1e-13 and rel_err(fast[1], gen[1]) <= 1e-13, (d, flags)


@pytest.mark.parametrize("scale,eps", [(1e22, 1e-16), (1e-30, 1e-100)])
def test_extreme_scales_take_ieee_path(scale, eps):
    g, og, f = _case(*BOX, 1, seed=540, scale=scale, eps=eps)
    _check_box(g, og, f, seed=41)


def test_spikes_mix_paths_in_one_wavefront():
    """Isolated 1e22 cells: lanes whose window holds one take the IEEE path,
    their neighbours in the same wavefront the shared-reciprocal path."""
    g, og, f = _case(*BOX, 1, seed=550)
    rng = np.random.default_rng(5)
    for _ in range(200):
        f[tuple(int(rng.integers(3, s - 3)) for s in f.shape[:3]) + (0,)] = 1e22
    _check_box(g, og, f, seed=51, split=1e10)


def test_nan_cells_poison_their_stencils_only():
    cells = [(20, 30, 33, 0), (140, 10, 12, 0), (5, 60, 130, 0)]
    g, og, f = _case(*BOX, 1, seed=560, nan_cells=cells)
    _check_box(g, og, f, seed=61)


def _bf16_bits(x):
    """Round to BF16 (nearest even) and return the 16-bit patterns."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _bf16_widen(bits):
    return (np.asarray(bits).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("state", ["f32", "bf16"])
def test_reduced_storage_all_sweeps(state):
    """FP32 / BF16 state on the same boxes: field and reshaped input,
    core-only (arbitrary FP32 halo kept bitwise) and fused copy."""
    g, og, f = _case(*BOX, 2, seed=570)
    if state == "f32":
        st = np.asfortranarray(f.astype(np.float32))
        fw = st.astype(np.float64)
        dt = wm.F32
    else:
        st = np.asfortranarray(_bf16_bits(f))
        fw = _bf16_widen(st)
        dt = wm.BF16
    vf = DeviceArray.from_host(st)
    rng = np.random.default_rng(17)
    m, n, p = BOX
    for d in range(3):
        ns, nk, nl = fields.dir_extents(m, n, p, 6, d)
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs, coef = _coef(cs)
        v_rs = oracle.reshape(og, d, fw)
        rs_host = v_rs.astype(np.float32) if state == "f32" else _bf16_bits(v_rs)
        rs = DeviceArray.from_host(rs_host)
        cL, cR = oracle.order1_copy(og, d, fw)
        pL = np.asfortranarray(rng.standard_normal((ns, nk, nl, 2)).astype(np.float32))
        pR = np.asfortranarray(rng.standard_normal((ns, nk, nl, 2)).astype(np.float32))
        for src, kind in [(vf, wm.SRC_FIELD), (rs, wm.SRC_RESHAPED)]:
            L, R = _run(g, d, src, kind, coef, 0, pre=(pL, pR), dtype=dt)
            halo = np.r_[0:3, ns - 3:ns]
            assert np.array_equal(L[halo], pL[halo]) and np.array_equal(R[halo], pR[halo]), (d, kind)
            eL, eR = oracle.weno(og, 5, d, v_rs, cs, pL.astype(np.float64), pR.astype(np.float64))
            assert rel_err(L, eL) <= 1.2e-7 and rel_err(R, eR) <= 1.2e-7, (d, kind, "core")
            L, R = _run(g, d, src, kind, coef, wm.FUSE_COPY, dtype=dt)
            eL, eR = oracle.weno(og, 5, d, v_rs, cs, cL, cR)
            assert rel_err(L, eL) <= 1.2e-7 and rel_err(R, eR) <= 1.2e-7, (d, kind, "copy")
            assert np.array_equal(L[halo], cL[halo].astype(np.float32))
            assert np.array_equal(R[halo], cR[halo].astype(np.float32))
            L, R = _run(g, d, src, kind, coef, wm.FUSE_COPY | wm.FUSE_MP, dtype=dt)
            mL, mR = oracle.mp(og, d, v_rs, eL, eR)
            assert rel_err(L, mL) <= 1.2e-7 and rel_err(R, mR) <= 1.2e-7, (d, kind, "copy+mp")
            assert np.array_equal(L[halo], cL[halo].astype(np.float32))


def test_fp32_storage_matches_general_kernel_and_nan():
    """FP32 state through both kernels (variant bit 8 keeps the general one,
    whose stencil arithmetic is FP32): agreement to the FP32 form's 1e-6,
    NaN cells poisoning the same outputs."""
    cells = [(20, 30, 33, 0), (140, 10, 12, 0)]
    g, og, f = _case(*BOX, 1, seed=580, nan_cells=cells)
    st = np.asfortranarray(f.astype(np.float32))
    vf = DeviceArray.from_host(st)
    for d in range(3):
        ns = fields.dir_extents(*BOX, 6, d)[0]
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        devs, coef = _coef(cs)
        fast = _run(g, d, vf, wm.SRC_FIELD, coef, wm.FUSE_COPY, dtype=wm.F32)
        wm.check(wm.lib.weno_set_tile(5, wm.F32, d, 0, 0, 0, 8 | 2))
        try:
            gen = _run(g, d, vf, wm.SRC_FIELD, coef, wm.FUSE_COPY, dtype=wm.F32)
        finally:
            wm.check(wm.lib.weno_set_tile(5, wm.F32, d, 0, 0, 0, -1))
        assert rel_err(fast[0], gen[0]) <= 1e-6 and rel_err(fast[1], gen[1]) <= 1e-6, d
        v_rs = oracle.reshape(og, d, st.astype(np.float64))
        cL, cR = oracle.order1_copy(og, d, st.astype(np.float64))
        eL, eR = oracle.weno(og, 5, d, v_rs, cs, cL, cR)
        assert rel_err(fast[0], eL) <= 1.2e-7 and rel_err(fast[1], eR) <= 1.2e-7, d
