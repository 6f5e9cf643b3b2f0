/*
 * CPU oracle for the WENO reconstruction path -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement of the reference harness (inputs/weno_standalone.f90) in C,
 * kept in the reference's evaluation order (left-to-right sums, x**2 as x*x,
 * IEEE divisions, no FMA contraction: build with -ffp-contract=off) so that
 * its outputs can be compared bit for bit with the reference's own kernels
 * built from source (oracle/_ref).  Each function cites the lines it follows.
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   fill_input_field / fill_coefficients / reshape / order1_copy / weno5 /
 *   weno7 / weno3 (core) / mp / checksum: bitwise against oracle/_ref on seeded
 *   inputs (committed fixtures) + tests/golden/weno5.txt checksums + closed
 *   forms (constant and linear fields, MP in-bounds idempotence).
 *   TENO (weno5/weno7), WENO-Z and mapped WENO (weno3/weno5/weno7): the
 *   reference selects them with compile-time parameters (teno, wenoz,
 *   mapped_weno, :72-74), all .false. as shipped.  oracle/Makefile builds the
 *   reference once per branch with that one parameter set to .true.
 *   (oracle/_ref/<name>/); the committed fixtures tests/golden/ref_branches.npz
 *   hold its outputs, compared bitwise.  TENO is also pinned by hand-derived
 *   closed forms on crafted stencils (tests/test_oracle.py::test_teno_*).
 *   weno3 halo cells (j outside the core range): the reference reads its
 *   coefficient arrays out of bounds there (:395-399 vs :413), so nothing
 *   pins them; this oracle clamps the coefficient index to the core range
 *   (documented choice, "parity unpinned" for those cells).
 *   weno_vjp: no reference counterpart; pinned by central finite differences
 *   and a complex-step derivative of this file's forward (tests).
 */
#include "weno_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PAD 3 /* weno_polyn_max, weno_standalone.f90:69 */

static inline int ext_x(const oracle_grid* g) { return g->m + 1 + 2 * g->buff; }
static inline int ext_y(const oracle_grid* g) { return g->n + 1 + 2 * g->buff; }
static inline int ext_z(const oracle_grid* g) { return g->p + 1 + 2 * g->buff; }

void oracle_dir_extents(const oracle_grid* g, int dir, int* ns, int* nk, int* nl)
{
    int nx = ext_x(g), ny = ext_y(g), nz = ext_z(g);
    if (dir == 0) { *ns = nx; *nk = ny; *nl = nz; }
    else if (dir == 1) { *ns = ny; *nk = nx; *nl = nz; }
    else { *ns = nz; *nk = ny; *nl = nx; }
}

/* v_vf index: X,Y,Z are 0-based positions inside the padded array */
static inline size_t vf_idx(const oracle_grid* g, int X, int Y, int Z, int i)
{
    size_t NX = (size_t)ext_x(g) + 6, NY = (size_t)ext_y(g) + 6, NZ = (size_t)ext_z(g) + 6;
    return (size_t)X + NX * ((size_t)Y + NY * ((size_t)Z + NZ * (size_t)i));
}

/* weno_standalone.f90:221-244 */
void oracle_fill_input_field(const oracle_grid* g, double* v_vf)
{
    int nx = ext_x(g), ny = ext_y(g), nz = ext_z(g);
    int is1_beg = -g->buff, is2_beg = -g->buff, is3_beg = -g->buff;
    int is1_end = g->m + g->buff, is2_end = g->n + g->buff, is3_end = g->p + g->buff;
    double dx = (double)(is1_end - is1_beg + 3 > 1 ? is1_end - is1_beg + 3 : 1);
    double dy = (double)(is2_end - is2_beg + 3 > 1 ? is2_end - is2_beg + 3 : 1);
    double dz = (double)(is3_end - is3_beg + 3 > 1 ? is3_end - is3_beg + 3 : 1);
    for (int l = 1; l <= g->v; ++l) {
#pragma omp parallel for schedule(static)
        for (int Z = 0; Z < nz + 6; ++Z) {
            int k = Z + is3_beg - PAD;
            double z = (double)(k - is3_beg + 1) / dz;
            for (int Y = 0; Y < ny + 6; ++Y) {
                int j = Y + is2_beg - PAD;
                double y = (double)(j - is2_beg + 1) / dy;
                for (int X = 0; X < nx + 6; ++X) {
                    int i = X + is1_beg - PAD;
                    double x = (double)(i - is1_beg + 1) / dx;
                    double val = sin(2.1 * x * (double)l) + cos(3.7 * y) + 0.25 * sin(4.3 * z)
                                 + 0.02 * (double)(i + 2 * j - k + l);
                    v_vf[vf_idx(g, X, Y, Z, l - 1)] = val;
                }
            }
        }
    }
}

/* weno_standalone.f90:246-279 */
void oracle_fill_coefficients(int nc, double* polyL, double* polyR, double* dL,
                              double* dR, double* beta, double offset)
{
    static const double baseL[5] = {0.30, 0.27, 0.20, 0.14, 0.09};
    static const double baseR[5] = {0.09, 0.14, 0.20, 0.27, 0.30};
    for (int jc = 0; jc < nc; ++jc) {
        for (int r = 0; r < 5; ++r) {
            dL[r + 5 * jc] = baseL[r] + offset * 0.10;
            dR[r + 5 * jc] = baseR[r] + offset * 0.10;
            for (int s = 0; s < 3; ++s) {
                polyL[jc + nc * (r + 5 * s)] = offset + 0.015 * (double)((r + 1) * (s + 1));
                polyR[jc + nc * (r + 5 * s)] = -offset + 0.012 * (double)(2 * r + s + 1);
            }
            for (int s = 0; s < 6; ++s) beta[jc + nc * (r + 5 * s)] = 0.0;
            beta[jc + nc * (r + 5 * 0)] = 0.8 + offset + 0.05 * (double)(r + 1);
            beta[jc + nc * (r + 5 * 2)] = 0.9 + offset + 0.04 * (double)(r + 1);
            beta[jc + nc * (r + 5 * 3)] = 1.0 + offset + 0.03 * (double)(r + 1);
            beta[jc + nc * (r + 5 * 5)] = 1.1 + offset + 0.02 * (double)(r + 1);
        }
    }
}

/* map (sweep index J in [0,ns+6), transverse k,l) of direction dir to v_vf */
static inline size_t dir_src(const oracle_grid* g, int dir, int J, int k, int l, int i)
{
    if (dir == 0) return vf_idx(g, J, k + PAD, l + PAD, i);     /* v_vf(k,l,q,j)  :290 */
    if (dir == 1) return vf_idx(g, k + PAD, J, l + PAD, i);     /* v_vf(l,k,q,j)  :307 */
    return vf_idx(g, l + PAD, k + PAD, J, i);                   /* v_vf(q,l,k,j)  :324 */
}

/* weno_standalone.f90:281-330 */
void oracle_reshape(const oracle_grid* g, int dir, const double* v_vf, double* v_rs)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    size_t NS = (size_t)ns + 6;
    for (int i = 0; i < g->v; ++i)
#pragma omp parallel for schedule(static)
        for (int l = 0; l < nl; ++l)
            for (int k = 0; k < nk; ++k)
                for (int J = 0; J < ns + 6; ++J)
                    v_rs[J + NS * (k + (size_t)nk * (l + (size_t)nl * i))] =
                        v_vf[dir_src(g, dir, J, k, l, i)];
}

/* weno_standalone.f90:332-390 */
void oracle_order1_copy(const oracle_grid* g, int dir, const double* v_vf,
                        double* vL, double* vR)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    for (int i = 0; i < g->v; ++i)
#pragma omp parallel for schedule(static)
        for (int l = 0; l < nl; ++l)
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < ns; ++j) {
                    size_t o = j + (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                    double val = v_vf[dir_src(g, dir, j + PAD, k, l, i)];
                    vL[o] = val;
                    vR[o] = val;
                }
}

/* The mapping of mapped WENO (weno_standalone.f90:436-437, :561-569,
 * :1481-1482): alpha_r <- (d_r (1 + d_r - 3 w_r) + w_r**2) *
 * (w_r / (d_r**2 + w_r (1 - 2 d_r))), with w = alpha / sum(alpha) taken over
 * the stencils in array order.  Real exponent 2 is a product (x*x). */
static inline double map_one(double d, double w)
{
    return (d * (1.0 + d - 3.0 * w) + w * w) * (w / (d * d + w * (1.0 - 2.0 * d)));
}
static void map2(const double* d, double* a0, double* a1)
{
    double s = *a0 + *a1;
    double w0 = *a0 / s, w1 = *a1 / s;
    *a0 = map_one(d[0], w0);
    *a1 = map_one(d[1], w1);
}
static void map3(const double* d, double* a0, double* a1, double* a2)
{
    double s = *a0 + *a1 + *a2;
    double w0 = *a0 / s, w1 = *a1 / s, w2 = *a2 / s;
    *a0 = map_one(d[0], w0);
    *a1 = map_one(d[1], w1);
    *a2 = map_one(d[2], w2);
}

#define RS(J) v_rs[(size_t)(J) + base_in]

/* weno_standalone.f90:392-475 (WENO3; j runs over the whole `is` range :413).
 * Coefficients for j outside the core range are read out of bounds by the
 * reference; here the index is clamped to the core range. */
void oracle_weno3(const oracle_grid* g, int dir, const double* v_rs,
                  const double* polyL, const double* polyR, const double* dL,
                  const double* dR, const double* beta, double* vL, double* vR)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    int nc = ns - 6;
    const double eps = g->eps;
    for (int l = 0; l < nl; ++l)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < ns; ++j)
                for (int i = 0; i < g->v; ++i) {
                    size_t base_in = (size_t)(ns + 6) * (k + (size_t)nk * (l + (size_t)nl * i));
                    size_t o = j + (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                    int J = j + PAD;
                    int jc = j - PAD;
                    if (jc < 0) jc = 0;
                    if (jc > nc - 1) jc = nc - 1;
                    double alpha0 = 0.0, alpha1 = 0.0, omega0, omega1;
                    double dvd0 = RS(J + 1) - RS(J);
                    double dvdm1 = RS(J) - RS(J - 1);
                    double poly0 = RS(J) + polyL[jc + nc * 0] * dvd0;
                    double poly1 = RS(J) + polyL[jc + nc * 1] * dvdm1;
                    double beta0 = beta[jc + nc * 0] * dvd0 * dvd0 + eps;
                    double beta1 = beta[jc + nc * 1] * dvdm1 * dvdm1 + eps;
                    double tau = 0.0;
                    if (g->wenojs) {
                        alpha0 = dL[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dL[1 + 5 * jc] / (beta1 * beta1);
                    } else if (g->mapped_weno) {
                        /* :433-437 */
                        alpha0 = dL[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dL[1 + 5 * jc] / (beta1 * beta1);
                        map2(dL + 5 * jc, &alpha0, &alpha1);
                    } else if (g->wenoz) {
                        /* :439-441 */
                        tau = fabs(beta1 - beta0);
                        alpha0 = dL[0 + 5 * jc] * (1.0 + tau / beta0);
                        alpha1 = dL[1 + 5 * jc] * (1.0 + tau / beta1);
                    }
                    double sum_alpha = alpha0 + alpha1;
                    omega0 = alpha0 / sum_alpha;
                    omega1 = alpha1 / sum_alpha;
                    vL[o] = omega0 * poly0 + omega1 * poly1;

                    poly0 = RS(J) + polyR[jc + nc * 0] * dvd0;
                    poly1 = RS(J) + polyR[jc + nc * 1] * dvdm1;
                    if (g->wenojs) {
                        alpha0 = dR[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dR[1 + 5 * jc] / (beta1 * beta1);
                    } else if (g->mapped_weno) {
                        /* :456-460 */
                        alpha0 = dR[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dR[1 + 5 * jc] / (beta1 * beta1);
                        map2(dR + 5 * jc, &alpha0, &alpha1);
                    } else if (g->wenoz) {
                        /* :462-463, tau of the left side */
                        alpha0 = dR[0 + 5 * jc] * (1.0 + tau / beta0);
                        alpha1 = dR[1 + 5 * jc] * (1.0 + tau / beta1);
                    }
                    sum_alpha = alpha0 + alpha1;
                    omega0 = alpha0 / sum_alpha;
                    omega1 = alpha1 / sum_alpha;
                    vR[o] = omega0 * poly0 + omega1 * poly1;
                }
}

/* Test hook: when set, the TENO branches of oracle_weno5 / oracle_weno7 store
 * each cell's cut-off ratios omega_q and the bases x_q = 1 + tau/beta_q of
 * the powers (N = 3 or 4 stencils; for output element o, omega_q at
 * 2N*o + q and x_q at 2N*o + N + q), so that a test can place teno_CT
 * exactly on a ratio the reference forms.  NULL (the default) stores
 * nothing. */
static double* teno_ratio_out = NULL;
void oracle_set_teno_ratio_out(double* p) { teno_ratio_out = p; }

/* weno_standalone.f90:477-669 (s_weno5_kernel; identical maths in :671-1397) */
void oracle_weno5(const oracle_grid* g, int dir, const double* v_rs,
                  const double* polyL, const double* polyR, const double* dL,
                  const double* dR, const double* beta, double* vL, double* vR)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    int nc = ns - 6;
    const double eps = g->eps, ct = g->teno_ct;
    for (int l = 0; l < nl; ++l)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nk; ++k)
            for (int j = PAD; j <= ns - 1 - PAD; ++j)
                for (int i = 0; i < g->v; ++i) {
                    size_t base_in = (size_t)(ns + 6) * (k + (size_t)nk * (l + (size_t)nl * i));
                    size_t o = j + (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                    int J = j + PAD, jc = j - PAD;
#define PL(r, s) polyL[jc + nc * ((r) + 5 * (s))]
#define PR(r, s) polyR[jc + nc * ((r) + 5 * (s))]
#define BC(r, s) beta[jc + nc * ((r) + 5 * (s))]
                    double alpha0 = 0.0, alpha1 = 0.0, alpha2 = 0.0;
                    double delta0 = 0.0, delta1 = 0.0, delta2 = 0.0;
                    double vj = RS(J), vjp2 = RS(J + 2), vjp1 = RS(J + 1);
                    double vjm1 = RS(J - 1), vjm2 = RS(J - 2);
                    double dvd_p1 = vjp2 - vjp1;
                    double dvd_0 = vjp1 - vj;
                    double dvd_m1 = vj - vjm1;
                    double dvd_m2 = vjm1 - vjm2;
                    double poly0 = vj + PL(0, 0) * dvd_p1 + PL(0, 1) * dvd_0;
                    double poly1 = vj + PL(1, 0) * dvd_0 + PL(1, 1) * dvd_m1;
                    double poly2 = vj + PL(2, 0) * dvd_m1 + PL(2, 1) * dvd_m2;
                    double beta0 = BC(0, 0) * dvd_p1 * dvd_p1 + BC(0, 1) * dvd_p1 * dvd_0
                                   + BC(0, 2) * dvd_0 * dvd_0 + eps;
                    double beta1 = BC(1, 0) * dvd_0 * dvd_0 + BC(1, 1) * dvd_0 * dvd_m1
                                   + BC(1, 2) * dvd_m1 * dvd_m1 + eps;
                    double beta2 = BC(2, 0) * dvd_m1 * dvd_m1 + BC(2, 1) * dvd_m1 * dvd_m2
                                   + BC(2, 2) * dvd_m2 * dvd_m2 + eps;
                    double sum_alpha, omega0, omega1, omega2, tau = 0.0;
                    if (g->wenojs) {
                        alpha0 = dL[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dL[1 + 5 * jc] / (beta1 * beta1);
                        alpha2 = dL[2 + 5 * jc] / (beta2 * beta2);
                    } else if (g->mapped_weno) {
                        /* :557-570 */
                        alpha0 = dL[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dL[1 + 5 * jc] / (beta1 * beta1);
                        alpha2 = dL[2 + 5 * jc] / (beta2 * beta2);
                        map3(dL + 5 * jc, &alpha0, &alpha1, &alpha2);
                    } else if (g->wenoz) {
                        /* :572-576 */
                        tau = fabs(beta2 - beta0);
                        alpha0 = dL[0 + 5 * jc] * (1.0 + (tau / beta0));
                        alpha1 = dL[1 + 5 * jc] * (1.0 + (tau / beta1));
                        alpha2 = dL[2 + 5 * jc] * (1.0 + (tau / beta2));
                    } else if (g->teno) {
                        /* :578-606 */
                        tau = fabs(beta2 - beta0);
                        alpha0 = pow(1.0 + tau / beta0, 6.0);
                        alpha1 = pow(1.0 + tau / beta1, 6.0);
                        alpha2 = pow(1.0 + tau / beta2, 6.0);
                        if (teno_ratio_out) {
                            teno_ratio_out[6 * o + 3] = 1.0 + tau / beta0;
                            teno_ratio_out[6 * o + 4] = 1.0 + tau / beta1;
                            teno_ratio_out[6 * o + 5] = 1.0 + tau / beta2;
                        }
                        sum_alpha = alpha0 + alpha1 + alpha2;
                        omega0 = alpha0 / sum_alpha;
                        omega1 = alpha1 / sum_alpha;
                        omega2 = alpha2 / sum_alpha;
                        delta0 = (omega0 < ct) ? 0.0 : 1.0;
                        delta1 = (omega1 < ct) ? 0.0 : 1.0;
                        delta2 = (omega2 < ct) ? 0.0 : 1.0;
                        if (teno_ratio_out) {
                            teno_ratio_out[6 * o + 0] = omega0;
                            teno_ratio_out[6 * o + 1] = omega1;
                            teno_ratio_out[6 * o + 2] = omega2;
                        }
                        alpha0 = delta0 * dL[0 + 5 * jc];
                        alpha1 = delta1 * dL[1 + 5 * jc];
                        alpha2 = delta2 * dL[2 + 5 * jc];
                    }
                    sum_alpha = alpha0 + alpha1 + alpha2;
                    omega0 = alpha0 / sum_alpha;
                    omega1 = alpha1 / sum_alpha;
                    omega2 = alpha2 / sum_alpha;
                    vL[o] = omega0 * poly0 + omega1 * poly1 + omega2 * poly2;

                    poly0 = vj + PR(0, 0) * dvd_p1 + PR(0, 1) * dvd_0;
                    poly1 = vj + PR(1, 0) * dvd_0 + PR(1, 1) * dvd_m1;
                    poly2 = vj + PR(2, 0) * dvd_m1 + PR(2, 1) * dvd_m2;
                    if (g->wenojs) {
                        alpha0 = dR[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dR[1 + 5 * jc] / (beta1 * beta1);
                        alpha2 = dR[2 + 5 * jc] / (beta2 * beta2);
                    } else if (g->mapped_weno) {
                        /* :632-645 */
                        alpha0 = dR[0 + 5 * jc] / (beta0 * beta0);
                        alpha1 = dR[1 + 5 * jc] / (beta1 * beta1);
                        alpha2 = dR[2 + 5 * jc] / (beta2 * beta2);
                        map3(dR + 5 * jc, &alpha0, &alpha1, &alpha2);
                    } else if (g->wenoz) {
                        /* :647-650, tau of the left side */
                        alpha0 = dR[0 + 5 * jc] * (1.0 + (tau / beta0));
                        alpha1 = dR[1 + 5 * jc] * (1.0 + (tau / beta1));
                        alpha2 = dR[2 + 5 * jc] * (1.0 + (tau / beta2));
                    } else if (g->teno) {
                        alpha0 = delta0 * dR[0 + 5 * jc];
                        alpha1 = delta1 * dR[1 + 5 * jc];
                        alpha2 = delta2 * dR[2 + 5 * jc];
                    }
                    sum_alpha = alpha0 + alpha1 + alpha2;
                    omega0 = alpha0 / sum_alpha;
                    omega1 = alpha1 / sum_alpha;
                    omega2 = alpha2 / sum_alpha;
                    vR[o] = omega0 * poly0 + omega1 * poly1 + omega2 * poly2;
                }
}

/* weno_standalone.f90:1419-1542 (WENO7, with the reversed stencil pairing
 * and the TENO quirk that beta(4) stays weno_eps and alpha(4) is never
 * cleared by delta) */
void oracle_weno7(const oracle_grid* g, int dir, const double* v_rs,
                  const double* polyL, const double* polyR, const double* dL,
                  const double* dR, const double* beta, double* vL, double* vR)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    int nc = ns - 6;
    const double eps = g->eps, ct = g->teno_ct;
    for (int l = 0; l < nl; ++l)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nk; ++k)
            for (int j = PAD; j <= ns - 1 - PAD; ++j)
                for (int i = 0; i < g->v; ++i) {
                    size_t base_in = (size_t)(ns + 6) * (k + (size_t)nk * (l + (size_t)nl * i));
                    size_t o = j + (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                    int J = j + PAD, jc = j - PAD;
                    double alpha[5] = {0, 0, 0, 0, 0}, omega[5], delta[5] = {0, 0, 0, 0, 0};
                    double beta_[5] = {eps, eps, eps, eps, eps}, poly[5] = {0, 0, 0, 0, 0};
                    double d2 = RS(J + 3) - RS(J + 2);
                    double d1 = RS(J + 2) - RS(J + 1);
                    double d0 = RS(J + 1) - RS(J);
                    double dm1 = RS(J) - RS(J - 1);
                    double dm2 = RS(J - 1) - RS(J - 2);
                    double dm3 = RS(J - 2) - RS(J - 3);
                    double vj = RS(J);
                    poly[3] = vj + PL(0, 0) * d2 + PL(0, 1) * d1 + PL(0, 2) * d0;
                    poly[2] = vj + PL(1, 0) * d1 + PL(1, 1) * d0 + PL(1, 2) * dm1;
                    poly[1] = vj + PL(2, 0) * d0 + PL(2, 1) * dm1 + PL(2, 2) * dm2;
                    poly[0] = vj + PL(3, 0) * dm1 + PL(3, 1) * dm2 + PL(3, 2) * dm3;
                    beta_[3] = BC(0, 0) * d2 * d2 + BC(0, 1) * d2 * d1 + BC(0, 2) * d2 * d0
                               + BC(0, 3) * d1 * d1 + BC(0, 4) * d1 * d0 + BC(0, 5) * d0 * d0 + eps;
                    beta_[2] = BC(1, 0) * d1 * d1 + BC(1, 1) * d1 * d0 + BC(1, 2) * d1 * dm1
                               + BC(1, 3) * d0 * d0 + BC(1, 4) * d0 * dm1 + BC(1, 5) * dm1 * dm1 + eps;
                    beta_[1] = BC(2, 0) * d0 * d0 + BC(2, 1) * d0 * dm1 + BC(2, 2) * d0 * dm2
                               + BC(2, 3) * dm1 * dm1 + BC(2, 4) * dm1 * dm2 + BC(2, 5) * dm2 * dm2 + eps;
                    beta_[0] = BC(3, 0) * dm1 * dm1 + BC(3, 1) * dm1 * dm2 + BC(3, 2) * dm1 * dm3
                               + BC(3, 3) * dm2 * dm2 + BC(3, 4) * dm2 * dm3 + BC(3, 5) * dm3 * dm3 + eps;
                    double tau = 0.0;
                    if (g->wenojs) {
                        for (int q = 0; q < 4; ++q) alpha[q] = dL[q + 5 * jc] / (beta_[q] * beta_[q]);
                    } else if (g->mapped_weno) {
                        /* :1478-1482: the sum runs over all five entries (alpha(4) = 0) */
                        for (int q = 0; q < 4; ++q) alpha[q] = dL[q + 5 * jc] / (beta_[q] * beta_[q]);
                        double s = alpha[0] + alpha[1] + alpha[2] + alpha[3] + alpha[4];
                        for (int q = 0; q < 5; ++q) omega[q] = alpha[q] / s;
                        for (int q = 0; q < 4; ++q) alpha[q] = map_one(dL[q + 5 * jc], omega[q]);
                    } else if (g->wenoz) {
                        /* :1484-1488, wenoz_q = 2 */
                        tau = fabs(beta_[3] - beta_[0]);
                        for (int q = 0; q < 4; ++q) {
                            double t = tau / beta_[q];
                            alpha[q] = dL[q + 5 * jc] * (1.0 + t * t);
                        }
                    } else if (g->teno) {
                        /* :1490-1503 */
                        tau = fabs(beta_[4] - beta_[3]);
                        for (int q = 0; q < 5; ++q) alpha[q] = 1.0 + tau / beta_[q];
                        if (teno_ratio_out)
                            for (int q = 0; q < 4; ++q) teno_ratio_out[8 * o + 4 + q] = alpha[q];
                        /* (alpha**3._wp)**2._wp: the reference build evaluates
                         * the real power 3 as alpha*alpha*alpha (bitwise
                         * against oracle/_ref/teno) and the power 6 of
                         * WENO5 (:580) as pow */
                        for (int q = 0; q < 5; ++q) {
                            double t = alpha[q] * alpha[q] * alpha[q];
                            alpha[q] = t * t;
                        }
                        double s = alpha[0] + alpha[1] + alpha[2] + alpha[3] + alpha[4];
                        for (int q = 0; q < 5; ++q) omega[q] = alpha[q] / s;
                        for (int q = 0; q < 4; ++q) {
                            delta[q] = (omega[q] < ct) ? 0.0 : 1.0;
                            if (teno_ratio_out) teno_ratio_out[8 * o + q] = omega[q];
                            alpha[q] = delta[q] * dL[q + 5 * jc];
                        }
                    }
                    double s = alpha[0] + alpha[1] + alpha[2] + alpha[3] + alpha[4];
                    for (int q = 0; q < 5; ++q) omega[q] = alpha[q] / s;
                    vL[o] = omega[0] * poly[0] + omega[1] * poly[1] + omega[2] * poly[2] + omega[3] * poly[3];

                    poly[3] = vj + PR(0, 0) * d2 + PR(0, 1) * d1 + PR(0, 2) * d0;
                    poly[2] = vj + PR(1, 0) * d1 + PR(1, 1) * d0 + PR(1, 2) * dm1;
                    poly[1] = vj + PR(2, 0) * d0 + PR(2, 1) * dm1 + PR(2, 2) * dm2;
                    poly[0] = vj + PR(3, 0) * dm1 + PR(3, 1) * dm2 + PR(3, 2) * dm3;
                    if (g->wenojs) {
                        for (int q = 0; q < 4; ++q) alpha[q] = dR[q + 5 * jc] / (beta_[q] * beta_[q]);
                    } else if (g->mapped_weno) {
                        /* :1517-1521 */
                        for (int q = 0; q < 4; ++q) alpha[q] = dR[q + 5 * jc] / (beta_[q] * beta_[q]);
                        double s2 = alpha[0] + alpha[1] + alpha[2] + alpha[3] + alpha[4];
                        for (int q = 0; q < 5; ++q) omega[q] = alpha[q] / s2;
                        for (int q = 0; q < 4; ++q) alpha[q] = map_one(dR[q + 5 * jc], omega[q]);
                    } else if (g->wenoz) {
                        /* :1523-1526, tau of the left side */
                        for (int q = 0; q < 4; ++q) {
                            double t = tau / beta_[q];
                            alpha[q] = dR[q + 5 * jc] * (1.0 + t * t);
                        }
                    } else if (g->teno) {
                        for (int q = 0; q < 4; ++q) alpha[q] = delta[q] * dR[q + 5 * jc];
                    }
                    s = alpha[0] + alpha[1] + alpha[2] + alpha[3] + alpha[4];
                    for (int q = 0; q < 5; ++q) omega[q] = alpha[q] / s;
                    vR[o] = omega[0] * poly[0] + omega[1] * poly[1] + omega[2] * poly[2] + omega[3] * poly[3];
                }
}
#undef PL
#undef PR
#undef BC

/* Fortran SIGN(a,b) for real arguments: |a| with the sign bit of b */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }
static inline double min3(double a, double b, double c) { return fmin(fmin(a, b), c); }
static inline double max3(double a, double b, double c) { return fmax(fmax(a, b), c); }
static inline double min4(double a, double b, double c, double d) { return fmin(fmin(fmin(a, b), c), d); }

/* the minmod-4 expression of :1568-1571 with a=d(-1), b=d(0) (and its reuses) */
static inline double mp_minmod4(double a, double b)
{
    return (fsign(1.0, 4.0 * a - b) + fsign(1.0, 4.0 * b - a))
           * fabs((fsign(1.0, 4.0 * a - b) + fsign(1.0, a)) * (fsign(1.0, 4.0 * a - b) + fsign(1.0, b)))
           * min4(fabs(4.0 * a - b), fabs(a), fabs(4.0 * b - a), fabs(b)) / 8.0;
}

/* weno_standalone.f90:1544-1615 */
void oracle_mp(const oracle_grid* g, int dir, const double* v_rs, double* vL, double* vR)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    const double alpha_mp = 2.0;
    const double beta_mp = 4.0 / 3.0;
    for (int l = 0; l < nl; ++l)
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nk; ++k)
            for (int j = PAD; j <= ns - 1 - PAD; ++j)
                for (int i = 0; i < g->v; ++i) {
                    size_t base_in = (size_t)(ns + 6) * (k + (size_t)nk * (l + (size_t)nl * i));
                    size_t o = j + (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                    int J = j + PAD;
                    double dm1 = RS(J) + RS(J - 2) - 2.0 * RS(J - 1);
                    double d0 = RS(J + 1) + RS(J - 1) - 2.0 * RS(J);
                    double dp1 = RS(J + 2) + RS(J) - 2.0 * RS(J + 1);

                    double d_MD = mp_minmod4(dm1, d0);
                    double d_LC = mp_minmod4(d0, dp1);
                    double vL_UL = RS(J) - alpha_mp * (RS(J + 1) - RS(J));
                    double vL_MD = (RS(J) + RS(J - 1) - d_MD) * 5.e-1;
                    double vL_LC = RS(J) - (RS(J + 1) - RS(J)) * 5.e-1 + beta_mp * d_LC;
                    double vL_min = fmax(min3(RS(J), RS(J - 1), vL_MD), min3(RS(J), vL_UL, vL_LC));
                    double vL_max = fmin(max3(RS(J), RS(J - 1), vL_MD), max3(RS(J), vL_UL, vL_LC));
                    double x = vL[o];
                    vL[o] = x + (fsign(5.e-1, vL_min - x) + fsign(5.e-1, vL_max - x))
                                    * fmin(fabs(vL_min - x), fabs(vL_max - x));

                    double vR_UL = RS(J) + alpha_mp * (RS(J) - RS(J - 1));
                    d_MD = mp_minmod4(d0, dp1);
                    d_LC = mp_minmod4(dm1, d0);
                    double vR_MD = (RS(J) + RS(J + 1) - d_MD) * 5.e-1;
                    double vR_LC = RS(J) + (RS(J) - RS(J - 1)) * 5.e-1 + beta_mp * d_LC;
                    double vR_min = fmax(min3(RS(J), RS(J + 1), vR_MD), min3(RS(J), vR_UL, vR_LC));
                    double vR_max = fmin(max3(RS(J), RS(J + 1), vR_MD), max3(RS(J), vR_UL, vR_LC));
                    x = vR[o];
                    vR[o] = x + (fsign(5.e-1, vR_min - x) + fsign(5.e-1, vR_max - x))
                                    * fmin(fabs(vR_min - x), fabs(vR_max - x));
                }
}

/* weno_standalone.f90:1617-1620 (column-major sequential sum) */
double oracle_checksum(const double* a, size_t n)
{
    double s = 0.0;
    for (size_t t = 0; t < n; ++t) s += fabs(a[t]);
    return s;
}

/* ---------------------------------------------------------------------------
 * VJP of WENO5/7-JS.  For each core output j (both sides share beta):
 *   vL = sum_r w_r p_r,  w_r = a_r / S,  a_r = d_r / b_r^2
 *   dvL/dp_r = w_r,   dvL/db_r = -2 w_r (p_r - vL) / b_r
 * and p_r, b_r are linear / quadratic in the divided differences
 * D[k] = v[J+k+1] - v[J+k].  Contributions are scattered into gv.
 * ------------------------------------------------------------------------- */
void oracle_weno_vjp(const oracle_grid* g, int dir, int order, const double* v_rs,
                     const double* polyL, const double* polyR, const double* dL,
                     const double* dR, const double* beta, const double* gL,
                     const double* gR, double* gv)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    int nc = ns - 6;
    const double eps = g->eps;
    size_t total = (size_t)(ns + 6) * nk * nl * g->v;
    memset(gv, 0, total * sizeof(double));
    int R = (order == 5) ? 3 : 4;   /* stencils */
    int W = R - 1;                  /* D offsets from -R .. R-2 */
    for (int i = 0; i < g->v; ++i)
#pragma omp parallel for schedule(static)
        for (int l = 0; l < nl; ++l)
            for (int k = 0; k < nk; ++k) {
                size_t base_in = (size_t)(ns + 6) * (k + (size_t)nk * (l + (size_t)nl * i));
                for (int j = PAD; j <= ns - 1 - PAD; ++j) {
                    size_t o = j + (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                    int J = j + PAD, jc = j - PAD;
                    double D[8], gD[8];   /* D[k+W+1] = v[J+k+1]-v[J+k], k=-R..R-2 */
                    for (int kk = -R; kk <= R - 2; ++kk) {
                        D[kk + R] = RS(J + kk + 1) - RS(J + kk);
                        gD[kk + R] = 0.0;
                    }
                    (void)W;
                    double vj = RS(J);
                    /* stencil q: coefficient row, D offsets used by poly and beta */
                    double pL[4], pR[4], b[4], aL[4], aR[4];
                    int row[4], off[4][3];
                    int nt = (order == 5) ? 2 : 3;
                    for (int q = 0; q < R; ++q) {
                        if (order == 5) { row[q] = q; off[q][0] = 1 - q; off[q][1] = -q; }
                        else { row[q] = 3 - q; off[q][0] = q - 1; off[q][1] = q - 2; off[q][2] = q - 3; }
                        int rr = row[q];
                        pL[q] = vj; pR[q] = vj;
                        for (int s = 0; s < nt; ++s) {
                            pL[q] += polyL[jc + nc * (rr + 5 * s)] * D[off[q][s] + R];
                            pR[q] += polyR[jc + nc * (rr + 5 * s)] * D[off[q][s] + R];
                        }
                        double bb = eps;
                        if (order == 5) {
                            double A = D[off[q][0] + R], B = D[off[q][1] + R];
                            bb += beta[jc + nc * (rr + 0)] * A * A + beta[jc + nc * (rr + 5)] * A * B
                                  + beta[jc + nc * (rr + 10)] * B * B;
                        } else {
                            double A = D[off[q][0] + R], B = D[off[q][1] + R], C = D[off[q][2] + R];
                            bb += beta[jc + nc * (rr + 0)] * A * A + beta[jc + nc * (rr + 5)] * A * B
                                  + beta[jc + nc * (rr + 10)] * A * C + beta[jc + nc * (rr + 15)] * B * B
                                  + beta[jc + nc * (rr + 20)] * B * C + beta[jc + nc * (rr + 25)] * C * C;
                        }
                        b[q] = bb;
                        aL[q] = dL[q + 5 * jc] / (bb * bb);
                        aR[q] = dR[q + 5 * jc] / (bb * bb);
                    }
                    double SL = 0, SR = 0, vLo = 0, vRo = 0;
                    for (int q = 0; q < R; ++q) { SL += aL[q]; SR += aR[q]; }
                    for (int q = 0; q < R; ++q) { vLo += aL[q] / SL * pL[q]; vRo += aR[q] / SR * pR[q]; }
                    double cL = gL[o], cR = gR[o];
                    double gvj = 0.0;
                    for (int q = 0; q < R; ++q) {
                        int rr = row[q];
                        double wL = aL[q] / SL, wR = aR[q] / SR;
                        double gpL = cL * wL, gpR = cR * wR;
                        double gb = -2.0 * (cL * wL * (pL[q] - vLo) + cR * wR * (pR[q] - vRo)) / b[q];
                        gvj += gpL + gpR;
                        for (int s = 0; s < nt; ++s)
                            gD[off[q][s] + R] += gpL * polyL[jc + nc * (rr + 5 * s)]
                                                 + gpR * polyR[jc + nc * (rr + 5 * s)];
                        if (order == 5) {
                            int ia = off[q][0] + R, ib = off[q][1] + R;
                            double A = D[ia], B = D[ib];
                            double b0 = beta[jc + nc * (rr + 0)], b1 = beta[jc + nc * (rr + 5)],
                                   b2 = beta[jc + nc * (rr + 10)];
                            gD[ia] += gb * (2.0 * b0 * A + b1 * B);
                            gD[ib] += gb * (b1 * A + 2.0 * b2 * B);
                        } else {
                            int ia = off[q][0] + R, ib = off[q][1] + R, ic = off[q][2] + R;
                            double A = D[ia], B = D[ib], C = D[ic];
                            double b0 = beta[jc + nc * (rr + 0)], b1 = beta[jc + nc * (rr + 5)],
                                   b2 = beta[jc + nc * (rr + 10)], b3 = beta[jc + nc * (rr + 15)],
                                   b4 = beta[jc + nc * (rr + 20)], b5 = beta[jc + nc * (rr + 25)];
                            gD[ia] += gb * (2.0 * b0 * A + b1 * B + b2 * C);
                            gD[ib] += gb * (b1 * A + 2.0 * b3 * B + b4 * C);
                            gD[ic] += gb * (b2 * A + b4 * B + 2.0 * b5 * C);
                        }
                    }
                    gv[base_in + J] += gvj;
                    for (int kk = -R; kk <= R - 2; ++kk) {
                        gv[base_in + J + kk + 1] += gD[kk + R];
                        gv[base_in + J + kk] -= gD[kk + R];
                    }
                }
            }
}
#undef RS

/* Rusanov interface flux of f(q) = a q + b q^2 / 2 (see weno_oracle.h) */
static double rusanov_flux(double qL, double qR, double a, double b)
{
    double fL = a * qL + 0.5 * b * qL * qL;
    double fR = a * qR + 0.5 * b * qR * qR;
    double sL = fabs(a + b * qL), sR = fabs(a + b * qR);
    double s = sL > sR ? sL : sR;
    return 0.5 * (fL + fR) - 0.5 * s * (qR - qL);
}

void oracle_flux_divergence(const oracle_grid* g, int dir, const double* vL, const double* vR, double a,
                            double b, double inv_dx, int accumulate, double* rhs)
{
    int ns, nk, nl;
    oracle_dir_extents(g, dir, &ns, &nk, &nl);
    for (int i = 0; i < g->v; ++i)
        for (int l = 0; l < nl; ++l)
#pragma omp parallel for schedule(static)
            for (int k = 0; k < nk; ++k) {
                size_t row = (size_t)ns * (k + (size_t)nk * (l + (size_t)nl * i));
                for (int j = PAD + 1; j <= ns - 2 - PAD; ++j) {
                    double Fp = rusanov_flux(vL[row + j], vR[row + j + 1], a, b);
                    double Fm = rusanov_flux(vL[row + j - 1], vR[row + j], a, b);
                    double r = -(Fp - Fm) * inv_dx;
                    rhs[row + j] = accumulate ? rhs[row + j] + r : r;
                }
            }
}
