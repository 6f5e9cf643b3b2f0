"""Pins the CPU oracle (oracle/weno_oracle.c) before it is trusted as the
checker: against the reference's outputs (fixtures written by the reference's
own kernels, the reference's golden checksums, and -- in this container -- the
live reference build, also built with its TENO / WENO-Z / mapped-WENO
parameters set), and against closed forms (TENO, the WENO7-TENO quirk, MP
invariants, the VJP, which has no reference counterpart)."""
import json
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

import oracle
from wenostandalone_amd import fields
from tests.conftest import GOLDEN

FAMS = ["weno3", "weno5", "weno7", "mp5", "mp7"]


def harness_coefs(g, d):
    ns = oracle.dir_extents(g, d)[0]
    return fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])


def test_fixtures_bitwise(small_ref):
    """Reference outputs (tests/golden/ref_small.npz) reproduced bit for bit;
    weno3 on the core only (its halo is undefined in the reference)."""
    m, v, b = int(small_ref["m"]), int(small_ref["v"]), int(small_ref["buff"])
    g = oracle.grid(m, v=v, buff=b)
    f = small_ref["field"]
    for d, name in enumerate("xyz"):
        cs = harness_coefs(g, d)
        for fam in FAMS:
            L, R = oracle.sequence(g, fam, d, f, cs)
            rL, rR = small_ref[f"{fam}_{name}_L"], small_ref[f"{fam}_{name}_R"]
            if fam == "weno3":
                L, R, rL, rR = L[3:-3], R[3:-3], rL[3:-3], rR[3:-3]
            assert np.array_equal(L, rL), (fam, name)
            assert np.array_equal(R, rR), (fam, name)


def test_checksums_m31():
    """X/Y/Z checksums of the reference at 32^3 x 8 (harness and seeded fields)."""
    ref = json.load(open(os.path.join(GOLDEN, "ref_checksums.json")))
    m, v, b = ref["m"], ref["v"], ref["buff"]
    g = oracle.grid(m, v=v, buff=b)
    fs = {"harness": oracle.fill_input_field(g),
          "seeded0": fields.seeded_field(m, m, m, v, b, seed=ref["seeded_seed"])}
    for key, want in ref["checksums_xyz"].items():
        fname, mode = key.split("_")
        for d in range(3):
            L, R = oracle.sequence(g, mode, d, fs[fname], harness_coefs(g, d))
            got = oracle.checksum(L) + oracle.checksum(R)
            assert abs(got - want[d]) <= 1e-13 * want[d], (key, d, got, want[d])


def test_golden_weno5_200():
    """tests/golden/weno5.txt (the reference's golden, rtol 1e-10,
    wenoss.sh:365-385) and the x-only values of the reference's roofline run
    (results/amd_mp_roof_ex/log.txt:37-39: X = 6.5386334223E+08, order-1
    copy Y = Z = 6.5364085459E+08)."""
    gold = {k: float(v) for k, v in (l.split() for l in open(os.path.join(GOLDEN, "weno5.txt")))}
    g = oracle.grid(199, v=8, buff=6)
    f = oracle.fill_input_field(g)
    np.testing.assert_array_equal(f[:, :, :, :2], fields.harness_field(199, 199, 199, 2, 6))
    chk = []
    for d in range(3):
        L, R = oracle.sequence(g, "weno5", d, f, oracle.fill_coefficients(206, fields.DIR_OFFSETS[d]))
        chk.append(oracle.checksum(L) + oracle.checksum(R))
        if d == 1:
            cL, cR = oracle.order1_copy(g, d, f)
            copy_chk = oracle.checksum(cL) + oracle.checksum(cR)
        del L, R
    for k, val in zip("XYZT", chk + [sum(chk)]):
        assert abs(val - gold[k]) <= 1e-10 * gold[k], (k, val, gold[k])
    assert f"{chk[0]:.10E}" == "6.5386334223E+08"
    assert f"{copy_chk:.10E}" == "6.5364085459E+08"


@pytest.mark.skipif(not oracle.have_ref(), reason="reference build absent (only in the build container)")
def test_live_reference_random_fields():
    """Fresh seeded fields through the live reference build, all families."""
    for seed, recipe in [(101, 0), (202, 2)]:
        m, v, b = 4, 2, 6
        g = oracle.grid(m, v=v, buff=b)
        f = fields.seeded_field(m, m, m, v, b, seed=seed, recipe=recipe)
        rs = oracle.run_ref("reshape", m, v, b, f)
        for d in range(3):
            assert np.array_equal(rs[d], oracle.reshape(g, d, f))
        for fam in FAMS:
            ref = oracle.run_ref(fam, m, v, b, f)
            for d in range(3):
                L, R = oracle.sequence(g, fam, d, f, harness_coefs(g, d))
                s = slice(3, -3) if fam == "weno3" else slice(None)
                assert np.array_equal(L[s], ref[2 * d][s]) and np.array_equal(R[s], ref[2 * d + 1][s])


BRANCH_KW = {"teno": {"teno": True}, "wenoz": {"wenoz": True}, "mapped_weno": {"mapped_weno": True}}


def test_branch_fixtures_bitwise():
    """TENO, WENO-Z and mapped WENO reproduced bit for bit against the
    reference built with the matching compile-time parameter set
    (tests/golden/ref_branches.npz, core range)."""
    ref = np.load(os.path.join(GOLDEN, "ref_branches.npz"))
    m, v, b = int(ref["m"]), int(ref["v"]), int(ref["buff"])
    f = ref["field"]
    n = 0
    for branch, kw in BRANCH_KW.items():
        g = oracle.grid(m, v=v, buff=b, wenojs=False, **kw)
        for fam in ["weno3", "weno5", "weno7"]:
            if f"{branch}_{fam}_x_L" not in ref.files:
                continue
            for d, name in enumerate("xyz"):
                L, R = oracle.sequence(g, fam, d, f, harness_coefs(g, d))
                assert np.array_equal(L[3:-3], ref[f"{branch}_{fam}_{name}_L"]), (branch, fam, name)
                assert np.array_equal(R[3:-3], ref[f"{branch}_{fam}_{name}_R"]), (branch, fam, name)
                n += 1
    assert n == 3 * 8


def test_weight_branch_precedence():
    """wenojs wins over every other branch (:552), mapped_weno over wenoz and
    teno, wenoz over teno (:557, :572, :578); WENO3 has no TENO branch, so
    teno alone leaves it NaN."""
    m, v, b = 3, 1, 6
    f = fields.seeded_field(m, m, m, v, b, seed=4, recipe=2)
    def run(order, **kw):
        g = oracle.grid(m, v=v, buff=b, **kw)
        return oracle.sequence(g, {3: "weno3", 5: "weno5", 7: "weno7"}[order], 0, f, harness_coefs(g, 0))
    for order in (3, 5, 7):
        js = run(order)
        allon = run(order, wenojs=True, mapped_weno=True, wenoz=True, teno=True)
        assert all(np.array_equal(a, c) for a, c in zip(js, allon))
        mp = run(order, wenojs=False, mapped_weno=True)
        assert all(np.array_equal(a, c) for a, c in zip(mp, run(order, wenojs=False, mapped_weno=True, wenoz=True,
                                                              teno=True)))
        z = run(order, wenojs=False, wenoz=True)
        assert all(np.array_equal(a, c) for a, c in zip(z, run(order, wenojs=False, wenoz=True, teno=True)))
        assert not np.array_equal(js[0][3:-3], z[0][3:-3]) and not np.array_equal(js[0][3:-3], mp[0][3:-3])
    L, R = run(3, wenojs=False, teno=True)
    assert np.isnan(L[3:-3]).all() and np.isnan(R[3:-3]).all()


@pytest.mark.skipif(not all(os.path.exists(oracle.ref_driver(b)) for b in oracle.BRANCHES),
                    reason="reference branch builds absent (only in the build container)")
def test_live_reference_branches():
    """Fresh seeded fields through the live reference branch builds."""
    for seed, recipe in [(303, 0), (404, 2)]:
        m, v, b = 4, 2, 6
        f = fields.seeded_field(m, m, m, v, b, seed=seed, recipe=recipe)
        for branch, kw in BRANCH_KW.items():
            g = oracle.grid(m, v=v, buff=b, wenojs=False, **kw)
            for fam in ["weno3", "weno5", "weno7", "mp5", "mp7"]:
                ref = oracle.run_ref(fam, m, v, b, f, wenojs=False, branch=branch)
                for d in range(3):
                    L, R = oracle.sequence(g, fam, d, f, harness_coefs(g, d))
                    s = slice(3, -3) if fam == "weno3" else slice(None)
                    assert np.array_equal(L[s], ref[2 * d][s], equal_nan=True), (branch, fam, d)
                    assert np.array_equal(R[s], ref[2 * d + 1][s], equal_nan=True), (branch, fam, d)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def _pencil_field(values_along_x, v=1, m=3, b=6):
    """v_vf whose every x-pencil carries `values_along_x` (len nx+6)."""
    nx = m + 1 + 2 * b
    a = np.empty((nx + 6, nx + 6, nx + 6, v), order="F")
    a[...] = np.asarray(values_along_x, dtype=np.float64)[:, None, None, None]
    return a


def _run_x(order, vals, teno=False, wenojs=True, m=3):
    g = oracle.grid(m, v=1, buff=6, wenojs=wenojs, teno=teno)
    f = _pencil_field(vals, m=m)
    cs = harness_coefs(g, 0)
    v_rs = oracle.reshape(g, 0, f)
    cL, cR = oracle.order1_copy(g, 0, f)
    L, R = oracle.weno(g, order, 0, v_rs, cs, cL, cR)
    return L[:, 0, 0, 0], R[:, 0, 0, 0], cs


def test_constant_field_reproduces_constant():
    c = 1.2345678901234567
    for order in (3, 5, 7):
        L, R, _ = _run_x(order, np.full(22, c))
        core = slice(0, None) if order == 3 else slice(3, -3)
        assert np.max(np.abs(L[core] - c)) <= 4 * np.spacing(c)
        assert np.max(np.abs(R[core] - c)) <= 4 * np.spacing(c)


def _exact_js(order, slope, v0, cs, side):
    """Exact rational value of the JS reconstruction of a linear field
    (all first differences equal `slope`), from the published WENO-JS
    definition: w_r = (d_r/b_r^2)/sum(d/b^2), value = sum w_r p_r."""
    polyL, polyR, dL, dR, beta = cs
    poly = polyL if side == "L" else polyR
    dd = dL if side == "L" else dR
    s = Fr(slope)
    eps = Fr(1e-16)
    nst = {3: 2, 5: 3, 7: 4}[order]
    nterm = {3: 1, 5: 2, 7: 3}[order]
    nb = {3: 1, 5: 3, 7: 6}[order]
    num = Fr(0)
    den = Fr(0)
    for q in range(nst):
        row = q if order != 7 else 3 - q     # WENO7: candidate q uses coefficient row 3-q
        p = Fr(v0) + sum(Fr(poly[0, row, t]) * s for t in range(nterm))
        bq = sum(Fr(beta[0, row, t]) for t in range(nb)) * s * s + eps
        a = Fr(dd[q, 0]) / (bq * bq)
        num += a * p
        den += a
    return num / den


@pytest.mark.parametrize("order", [3, 5, 7])
def test_linear_field_exact(order):
    slope, v0 = 0.013671875, 0.75
    vals = v0 + slope * (np.arange(22) - 9)  # exact binary values
    L, R, cs = _run_x(order, vals)
    for j in range(3, 13):
        vj = vals[j + 3]
        for got, side in ((L[j], "L"), (R[j], "R")):
            exact = _exact_js(order, slope, vj, cs, side)
            assert abs(Fr(got) - exact) <= Fr(1e-15) * abs(exact), (order, j, side)


def test_teno5_smooth_keeps_all_stencils():
    """Linear data: every TENO weight is ~1/3 >= CT, so delta=1 for all and the
    result is sum(d_r p_r)/sum(d_r) (the optimal linear weights)."""
    slope = 0.01
    vals = 1.0 + slope * np.arange(22)
    L, R, cs = _run_x(5, vals, teno=True, wenojs=False)
    polyL, polyR, dL, dR, _ = cs
    for j in range(3, 13):
        vj = vals[j + 3]
        for got, poly, dd in ((L[j], polyL, dL), (R[j], polyR, dR)):
            p = [vj + poly[0, r, 0] * slope + poly[0, r, 1] * slope for r in range(3)]
            exp = sum(dd[r, 0] * p[r] for r in range(3)) / sum(dd[r, 0] for r in range(3))
            assert abs(got - exp) <= 1e-14 * abs(exp)


def test_teno5_drops_stencils_across_a_kink():
    """Values 1,1,1,2,3 around j: only the left stencil (r=2) is smooth;
    the others get omega ~ 1e-96 < CT and are cut, so vL = vR = p2 = 1."""
    vals = np.ones(22)
    vals[12:] = 1.0 + np.arange(10)   # v(J)=1 up to J=12, then rising
    L, R, _ = _run_x(5, vals, teno=True, wenojs=False)
    j = 12 - 3 - 1                    # stencil j-2..j+2 = J 10..14 -> 1,1,1,2,3
    assert L[j] == 1.0 and R[j] == 1.0


def test_weno7_teno_quirk():
    """beta(4) stays weno_eps and alpha(4) is never cut (weno_standalone.f90:
    1490-1503).  Linear data: tau = beta(3) >> eps makes alpha(4) dominate, all
    four deltas are 0 and the output is exactly 0.  Constant data: tau = 0,
    every alpha = 1, all deltas 1, output c*sum(d)/(sum(d)+1)."""
    vals = 1.0 + 0.01 * np.arange(22)
    L, R, _ = _run_x(7, vals, teno=True, wenojs=False)
    assert np.all(L[3:-3] == 0.0) and np.all(R[3:-3] == 0.0)
    c = 2.5
    L, R, cs = _run_x(7, np.full(22, c), teno=True, wenojs=False)
    dL, dR = cs[2], cs[3]
    for got, dd in ((L, dL), (R, dR)):
        sd = sum(dd[q, 0] for q in range(4))
        exp = c * sd / (sd + 1.0)
        assert np.max(np.abs(got[3:-3] - exp)) <= 1e-15 * exp


def test_no_weight_branch_nan():
    for order in (3, 5, 7):
        L, R, _ = _run_x(order, 1.0 + 0.01 * np.arange(22), wenojs=False, teno=False)
        core = slice(0, None) if order == 3 else slice(3, -3)
        assert np.isnan(L[core]).all() and np.isnan(R[core]).all()


def test_mp_invariants():
    """MP leaves a face value inside its bounds unchanged (the cell value
    v_j always is), and is idempotent."""
    m, v, b = 5, 2, 6
    g = oracle.grid(m, v=v, buff=b)
    f = fields.seeded_field(m, m, m, v, b, seed=9, recipe=2)
    for d in range(3):
        v_rs = oracle.reshape(g, d, f)
        cL, cR = oracle.order1_copy(g, d, f)
        L, R = oracle.mp(g, d, v_rs, cL, cR)
        assert np.array_equal(L, cL) and np.array_equal(R, cR)
        eL, eR = oracle.weno(g, 5, d, v_rs, harness_coefs(g, d), cL, cR)
        # push faces far outside so every cell is clipped
        L1, R1 = oracle.mp(g, d, v_rs, eL * 3.0 + 1.0, eR * 3.0 - 1.0)
        L2, R2 = oracle.mp(g, d, v_rs, L1, R1)
        # the clip x + (hi - x) lands within a few ulps of x of the bound
        assert np.allclose(L2, L1, rtol=1e-15, atol=0) and np.allclose(R2, R1, rtol=1e-15, atol=0)
        assert not np.array_equal(L1[3:-3], (eL * 3.0 + 1.0)[3:-3])


# ---------------------------------------------------------------------------
# VJP of the oracle
# ---------------------------------------------------------------------------
def _weno_js_pencil_complex(v, order, cs, j, side):
    """Pure-Python WENO-JS on one pencil in complex arithmetic (for the
    complex-step derivative); v is indexed by the reshaped J."""
    polyL, polyR, dL, dR, beta = cs
    poly = polyL if side == "L" else polyR
    dd = dL if side == "L" else dR
    J = j + 3
    jc = j - 3
    D = {k: v[J + k + 1] - v[J + k] for k in range(-4, 4)}
    out_n, out_d = 0, 0
    R = 3 if order == 5 else 4
    for q in range(R):
        if order == 5:
            row, offs = q, [1 - q, -q]
            bt = [(0, 1 - q, 1 - q), (1, 1 - q, -q), (2, -q, -q)]
        else:
            row, offs = 3 - q, [q - 1, q - 2, q - 3]
            A, B, Cc = q - 1, q - 2, q - 3
            bt = [(0, A, A), (1, A, B), (2, A, Cc), (3, B, B), (4, B, Cc), (5, Cc, Cc)]
        p = v[J] + sum(poly[jc, row, t] * D[o] for t, o in enumerate(offs))
        bq = sum(beta[jc, row, t] * D[a] * D[c] for t, a, c in bt) + 1e-16
        a = dd[q, jc] / (bq * bq)
        out_n += a * p
        out_d += a
    return out_n / out_d


@pytest.mark.parametrize("order", [5, 7])
def test_vjp_complex_step_and_fd(order):
    m, v, b = 2, 1, 6
    g = oracle.grid(m, v=v, buff=b)
    f = fields.seeded_field(m, m, m, v, b, seed=31)
    d = 0
    v_rs = oracle.reshape(g, d, f)
    ns = oracle.dir_extents(g, d)[0]
    cs = fields.random_coefficients(ns - 6, seed=3)
    rng = np.random.default_rng(0)
    gL = np.asfortranarray(rng.standard_normal((ns,) + v_rs.shape[1:]))
    gR = np.asfortranarray(rng.standard_normal((ns,) + v_rs.shape[1:]))
    gL[:3] = gL[-3:] = 0.0
    gR[:3] = gR[-3:] = 0.0
    gv = oracle.vjp(g, order, d, v_rs, cs, gL, gR)
    # complex step on one pencil (k=l=0): d/dv_m sum_j gL_j vL_j + gR_j vR_j
    pen = v_rs[:, 0, 0, 0].astype(np.complex128)
    h = 1e-30
    for mcell in range(v_rs.shape[0]):
        pv = pen.copy()
        pv[mcell] += 1j * h
        tot = 0
        for j in range(3, ns - 3):
            tot += gL[j, 0, 0, 0] * _weno_js_pencil_complex(pv, order, cs, j, "L")
            tot += gR[j, 0, 0, 0] * _weno_js_pencil_complex(pv, order, cs, j, "R")
        cstep = tot.imag / h
        assert abs(gv[mcell, 0, 0, 0] - cstep) <= 1e-10 * max(1.0, abs(cstep)), (mcell, gv[mcell, 0, 0, 0], cstep)
    # central finite differences of the oracle forward on a few cells
    for mcell in (4, 7, 10):
        def J_of(x):
            vv = v_rs.copy()
            vv[mcell, 1, 1, 0] = x
            L0, R0 = np.zeros((ns,) + v_rs.shape[1:], order="F"), np.zeros((ns,) + v_rs.shape[1:], order="F")
            L, R = oracle.weno(g, order, d, vv, cs, L0, R0)
            return float(np.sum(gL * L) + np.sum(gR * R))
        x0 = v_rs[mcell, 1, 1, 0]
        hh = 1e-6
        fd = (J_of(x0 + hh) - J_of(x0 - hh)) / (2 * hh)
        assert abs(gv[mcell, 1, 1, 0] - fd) <= 1e-5 * max(1.0, abs(fd))


@pytest.mark.parametrize("order", [5, 7])
def test_teno_ratio_hook(order):
    """oracle.teno_ratios reports the weights the cut-off compares: for TENO5
    they are alpha_q / sum of the same alphas, so they sum to 1 to rounding;
    for TENO7 the uncut alpha_4 takes the rest of the sum (below 1).  The
    reported bases reproduce the ratios: alpha_q = pow(x_q, 6) (TENO5, the C
    library's pow, as the oracle) or (x_q^3)^2 as products (TENO7)."""
    og = oracle.grid(9, 12, 7, v=2, buff=6, wenojs=False, teno=True)
    f = fields.seeded_field(9, 12, 7, 2, 6, seed=5)
    ns = fields.dir_extents(9, 12, 7, 6, 1)[0]
    cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[1])
    r, x = oracle.teno_ratios(og, order, 1, oracle.reshape(og, 1, f), cs)
    core = r[3:-3]
    assert np.all(x[3:-3] >= 1.0) and np.all(x[:3] == 0)   # x = 1 + tau/beta, beta > 0
    assert np.all(r[:3] == 0) and np.all(r[-3:] == 0)
    assert np.all(core >= 0)
    total = core.sum(axis=-1)
    if order == 5:
        assert np.max(np.abs(total - 1.0)) <= 4e-16
        import math
        a = np.vectorize(lambda t: math.pow(t, 6.0))(x[3:-3])   # the C library's pow
        s = a[..., 0] + a[..., 1] + a[..., 2]
        assert np.array_equal(a / s[..., None], core)
    else:
        assert np.all(total <= 1.0)


# ---- consumer step (no reference counterpart): oracle_flux_divergence ------
def _faces(m=9, n=7, p=6, v=2, seed=5):
    import oracle as O
    from wenostandalone_amd import fields
    og = O.grid(m, n, p, v=v, buff=6)
    f = fields.seeded_field(m, n, p, v, 6, seed=seed)
    out = []
    for d in range(3):
        ns = O.dir_extents(og, d)[0]
        cs = fields.harness_coefficients(ns - 6, fields.DIR_OFFSETS[d])
        cL, cR = O.order1_copy(og, d, f)
        eL, eR = O.weno(og, 5, d, O.reshape(og, d, f), cs, cL, cR)
        out.append((eL, eR))
    return og, out


@pytest.mark.parametrize("a", [1.7, -0.6])
def test_flux_linear_advection_is_upwind(a):
    """b = 0: the Rusanov flux is the upwind flux, F = a qL for a > 0 and
    F = a qR for a < 0, so rhs(j) = -a (vL(j) - vL(j-1)) / dx or
    -a (vR(j+1) - vR(j)) / dx -- which pins which face of which cell meets
    at an interface."""
    import oracle as O
    og, faces = _faces()
    for d, (vL, vR) in enumerate(faces):
        rhs = O.flux_divergence(og, d, vL, vR, a, 0.0, 4.0)
        ns = vL.shape[0]
        if a > 0:
            exp = -a * (vL[4:ns - 4] - vL[3:ns - 5]) * 4.0
        else:
            exp = -a * (vR[5:ns - 3] - vR[4:ns - 4]) * 4.0
        scale = 8.0 * abs(a) * np.maximum(np.abs(vL[3:ns - 4]).max(), np.abs(vR[3:ns - 3]).max())
        assert np.max(np.abs(rhs[4:ns - 4] - exp)) <= 1e-15 * scale, d
        assert not rhs[:4].any() and not rhs[ns - 4:].any()


def test_flux_burgers_constant_state_is_steady():
    import oracle as O
    og, faces = _faces()
    for d, (vL, _) in enumerate(faces):
        c = np.full(vL.shape, 1.375, order="F")
        rhs = O.flux_divergence(og, d, c, c, 0.0, 1.0, 3.0)
        assert not rhs.any()


def test_flux_divergence_telescopes():
    """Summed over a pencil, the divergence leaves the two boundary
    interfaces: sum_j rhs(j) dx = -(F(ns-5 + 1/2) - F(4 - 1/2)), with the
    interfaces j+1/2 = (vL(j), vR(j+1)).  A flux formed from any other pair
    of faces would not cancel between neighbouring cells."""
    import oracle as O
    og, faces = _faces()
    a, b, idx = 0.3, 0.8, 2.5

    def F(qL, qR):
        fl, fr = a * qL + 0.5 * b * qL * qL, a * qR + 0.5 * b * qR * qR
        s = np.maximum(np.abs(a + b * qL), np.abs(a + b * qR))
        return 0.5 * (fl + fr) - 0.5 * s * (qR - qL)
    for d, (vL, vR) in enumerate(faces):
        ns = vL.shape[0]
        rhs = O.flux_divergence(og, d, vL, vR, a, b, idx)
        total = rhs[4:ns - 4].sum(axis=0) / idx
        exp = -(F(vL[ns - 5], vR[ns - 4]) - F(vL[3], vR[4]))
        assert np.max(np.abs(total - exp)) <= 1e-12 * max(1.0, np.abs(exp).max()), d


def test_flux_accumulate_and_untouched_cells():
    import oracle as O
    og, faces = _faces()
    vL, vR = faces[1]
    base = np.full(vL.shape, -2.0, order="F")
    once = O.flux_divergence(og, 1, vL, vR, 0.5, 1.0, 1.0, rhs=base)
    twice = O.flux_divergence(og, 1, vL, vR, 0.5, 1.0, 1.0, rhs=once, accumulate=True)
    ns = vL.shape[0]
    assert np.all(once[:4] == -2.0) and np.all(once[ns - 4:] == -2.0)
    assert np.allclose(twice[4:ns - 4], 2.0 * once[4:ns - 4], rtol=1e-14, atol=0)
