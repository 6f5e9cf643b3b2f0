/*
 * CPU oracle for the WENO reconstruction path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference harness's kernels
 * (inputs/weno_standalone.f90 of the reference), written operation by
 * operation in the reference's evaluation order so that, compiled without
 * FMA contraction, it reproduces the reference build bit for bit.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product library never links it.
 *
 * Layout conventions (0-based, column-major like the Fortran arrays):
 *   nx = m+1+2*buff, ny = n+1+2*buff, nz = p+1+2*buff   (the `is` boxes)
 *   v_vf          (nx+6, ny+6, nz+6, V)                  weno_standalone.f90:163-165
 *   reshaped dir  (ns+6, nk, nl, V)   x:(nx,ny,nz) y:(ny,nx,nz) z:(nz,ny,nx)
 *   vL/vR dir     (ns,   nk, nl, V)                      :343,:363,:383
 *   polyL/polyR   (nc, 5, 3)   nc = ns-6                 :171-176
 *   dL/dR         (5, nc)                                :178-183
 *   beta_coef     (nc, 5, 6)                             :185-187
 */
#ifndef WENO_ORACLE_H
#define WENO_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_grid {
    int m, n, p;          /* m_cells, n_cells, p_cells */
    int buff;             /* buff_size */
    int v;                /* v_size */
    double eps;           /* weno_eps */
    int wenojs;           /* wenojs  */
    int teno;             /* teno (a compile-time parameter in the reference) */
    double teno_ct;       /* teno_CT */
    int mapped_weno;      /* mapped_weno (a compile-time parameter, :72) */
    int wenoz;            /* wenoz (a compile-time parameter, :73; wenoz_q = 2, :75) */
} oracle_grid;

/* extents: sweep length ns, transverse nk, nl for direction dir (0=x,1=y,2=z) */
void oracle_dir_extents(const oracle_grid* g, int dir, int* ns, int* nk, int* nl);

/* A3: weno_standalone.f90:221-244 / :246-279 */
void oracle_fill_input_field(const oracle_grid* g, double* v_vf);
void oracle_fill_coefficients(int nc, double* polyL, double* polyR, double* dL,
                              double* dR, double* beta, double offset);

/* A4: :281-330 */
void oracle_reshape(const oracle_grid* g, int dir, const double* v_vf, double* v_rs);
/* A5: :332-390 */
void oracle_order1_copy(const oracle_grid* g, int dir, const double* v_vf,
                        double* vL, double* vR);
/* A7/A8/A9: the 8-argument family, input is the reshaped array */
void oracle_weno3(const oracle_grid* g, int dir, const double* v_rs,
                  const double* polyL, const double* polyR, const double* dL,
                  const double* dR, const double* beta, double* vL, double* vR);
void oracle_weno5(const oracle_grid* g, int dir, const double* v_rs,
                  const double* polyL, const double* polyR, const double* dL,
                  const double* dR, const double* beta, double* vL, double* vR);
void oracle_weno7(const oracle_grid* g, int dir, const double* v_rs,
                  const double* polyL, const double* polyR, const double* dL,
                  const double* dR, const double* beta, double* vL, double* vR);
/* A10: :1544-1615 */
void oracle_mp(const oracle_grid* g, int dir, const double* v_rs, double* vL, double* vR);
/* A6: :1617-1620 */
/* Consumer of one sweep's faces (no reference counterpart: the harness stops
 * at vL/vR; MFC's Riemann solvers pair qL(j) with qR(j+1) at interface
 * j+1/2).  Per variable, scalar flux f(q) = a q + b q^2 / 2, Rusanov (local
 * Lax-Friedrichs) interface flux
 *   F(qL, qR) = (f(qL) + f(qR)) / 2 - max(|f'(qL)|, |f'(qR)|) (qR - qL) / 2,
 * and rhs(j) = -(F(vL(j), vR(j+1)) - F(vL(j-1), vR(j))) * inv_dx for the
 * cells j = 4 .. ns-5 whose two interfaces lie in the reconstructed core;
 * accumulate != 0 adds to rhs.  Face layout (ns, nk, nl, v) for vL, vR, rhs. */
void oracle_flux_divergence(const oracle_grid* g, int dir, const double* vL, const double* vR, double a,
                            double b, double inv_dx, int accumulate, double* rhs);

double oracle_checksum(const double* a, size_t n);

/* test hook: per-cell TENO cut-off ratios of the next oracle_weno5/7 calls */
void oracle_set_teno_ratio_out(double* p);

/* Vector-Jacobian product of WENO5/7-JS (no reference counterpart; the
 * derivative of oracle_weno5 / oracle_weno7 with wenojs=T).
 * gv has the reshaped-input shape; it is overwritten. */
void oracle_weno_vjp(const oracle_grid* g, int dir, int order, const double* v_rs,
                     const double* polyL, const double* polyR, const double* dL,
                     const double* dR, const double* beta, const double* gL,
                     const double* gR, double* gv);

#ifdef __cplusplus
}
#endif
#endif
