// The WENO5-JS and WENO7-JS face pairs of one cell in the shared-reciprocal
// form used by the sweep kernels (the reference's JS branches of s_weno5,
// inputs/weno_standalone.f90:477-1417, and s_weno7, :1419-1542).  WENO5:
//     D_i = v(i+1) - v(i),  S_i = D_i^2,  X_i = D_i D_(i-1)
//     beta_r = B_r0 S + B_r1 X + B_r2 S' + eps              (3 FMA each)
//     u_r    = prod_{s != r} beta_s^2                     (6 multiplies)
//     g_r    = (d_r a_r) D + (d_r b_r) D'                 (p_r = v_j + a_r D + b_r D')
//     out    = v_j + (sum_r u_r g_r) / (sum_r d_r u_r)
// which is  sum_r w_r p_r,  w_r = (d_r/beta_r^2)/sum_s(...),  regrouped so
// that both sides share one reciprocal of (sum_L)(sum_R).  WENO7 is the same
// with four stencils, three differences per candidate and six terms per beta
// (the 15 products D_i^2, D_i D_(i-1), D_i D_(i-2) shared by all four).
// cell5/cell7 report when the shared product leaves the normal range; the
// caller then redoes the cell with cell5_slow_inl/cell7_slow_inl (per-stencil IEEE
// divisions, which also carry NaN and infinity through as the reference
// does).
#pragma once

#include "weno_device.h"

namespace wm {
namespace {

struct Coef5 {
    double eL[3], fL[3], eR[3], fR[3];   // d_r * poly(r, 0..1)
    double b[9];                         // beta_coef(j, r, s), r-major
    double dL[3], dR[3];
};

__device__ __forceinline__ void load_coef5(const Coefs& c, int jc, Coef5& k)
{
    const int nc = c.nc;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double dl = c.dL[r + 5 * jc], dr = c.dR[r + 5 * jc];
        k.dL[r] = dl;
        k.dR[r] = dr;
        k.eL[r] = dl * c.polyL[jc + nc * r];
        k.fL[r] = dl * c.polyL[jc + nc * (r + 5)];
        k.eR[r] = dr * c.polyR[jc + nc * r];
        k.fR[r] = dr * c.polyR[jc + nc * (r + 5)];
#pragma unroll
        for (int s = 0; s < 3; ++s) k.b[3 * r + s] = c.beta[jc + nc * (r + 5 * s)];
    }
}

// positive normal (class bits 8 = +normal)
__device__ __forceinline__ bool pos_normal(double x) { return __builtin_amdgcn_class(x, 1 << 8); }

// Fast path; returns false when the shared denominator is out of range.
// NEWTON steps refine the hardware reciprocal (about 2^-23 relative): two
// for FP64 faces, one (2^-44) for FP32 faces.
template <int NEWTON>
__device__ __forceinline__ bool cell5(double v0, double v1, double v2, double v3, double v4, const Coef5& k,
                                      double eps, double& oL, double& oR)
{
    const double D3 = v4 - v3, D2 = v3 - v2, D1 = v2 - v1, D0 = v1 - v0;
    const double S3 = D3 * D3, X3 = D3 * D2, S2 = D2 * D2, X2 = D2 * D1, S1 = D1 * D1, X1 = D1 * D0, S0 = D0 * D0;
    const double b0 = fma(k.b[0], S3, fma(k.b[1], X3, fma(k.b[2], S2, eps)));
    const double b1 = fma(k.b[3], S2, fma(k.b[4], X2, fma(k.b[5], S1, eps)));
    const double b2 = fma(k.b[6], S1, fma(k.b[7], X1, fma(k.b[8], S0, eps)));
    const double m12 = b1 * b2, m02 = b0 * b2, m01 = b0 * b1;
    const double u0 = m12 * m12, u1 = m02 * m02, u2 = m01 * m01;
    const double gL0 = fma(k.eL[0], D3, k.fL[0] * D2), gL1 = fma(k.eL[1], D2, k.fL[1] * D1),
                 gL2 = fma(k.eL[2], D1, k.fL[2] * D0);
    const double gR0 = fma(k.eR[0], D3, k.fR[0] * D2), gR1 = fma(k.eR[1], D2, k.fR[1] * D1),
                 gR2 = fma(k.eR[2], D1, k.fR[2] * D0);
    const double NL = fma(u0, gL0, fma(u1, gL1, u2 * gL2));
    const double NR = fma(u0, gR0, fma(u1, gR1, u2 * gR2));
    const double SL = fma(k.dL[0], u0, fma(k.dL[1], u1, k.dL[2] * u2));
    const double SR = fma(k.dR[0], u0, fma(k.dR[1], u1, k.dR[2] * u2));
    const double T = SL * SR;
    double r = __builtin_amdgcn_rcp(T);
#pragma unroll
    for (int it = 0; it < NEWTON; ++it) {
        const double e = fma(-T, r, 1.0);
        r = fma(r, e, r);
    }
    oL = fma(NL, SR * r, v2);
    oR = fma(NR, SL * r, v2);
    // both tests without a branch
    return (static_cast<int>(pos_normal(T)) & static_cast<int>(pos_normal(r))) != 0;
}

// Per-stencil IEEE form (the reference's alpha_r = d_r / beta_r^2).  Always
// inlined: an out-of-line version takes the coefficient struct's address,
// and the compiler then keeps the caller's coefficient registers in scratch
// (written to memory at every launch).
__device__ __forceinline__ void cell5_slow_inl(double v0, double v1, double v2, double v3, double v4, const Coef5& k,
                                        double eps, double& oL, double& oR)
{
    const double D3 = v4 - v3, D2 = v3 - v2, D1 = v2 - v1, D0 = v1 - v0;
    const double b0 = fma(k.b[0], D3 * D3, fma(k.b[1], D3 * D2, fma(k.b[2], D2 * D2, eps)));
    const double b1 = fma(k.b[3], D2 * D2, fma(k.b[4], D2 * D1, fma(k.b[5], D1 * D1, eps)));
    const double b2 = fma(k.b[6], D1 * D1, fma(k.b[7], D1 * D0, fma(k.b[8], D0 * D0, eps)));
    const double y0 = 1.0 / (b0 * b0), y1 = 1.0 / (b1 * b1), y2 = 1.0 / (b2 * b2);
    const double NL = fma(k.eL[0], D3, k.fL[0] * D2) * y0 + fma(k.eL[1], D2, k.fL[1] * D1) * y1 +
                      fma(k.eL[2], D1, k.fL[2] * D0) * y2;
    const double NR = fma(k.eR[0], D3, k.fR[0] * D2) * y0 + fma(k.eR[1], D2, k.fR[1] * D1) * y1 +
                      fma(k.eR[2], D1, k.fR[2] * D0) * y2;
    const double SL = k.dL[0] * y0 + k.dL[1] * y1 + k.dL[2] * y2;
    const double SR = k.dR[0] * y0 + k.dR[1] * y1 + k.dR[2] * y2;
    oL = v2 + NL / SL;
    oR = v2 + NR / SR;
}

// WENO7: stencil q (0..3) of the reference's pairing (:1457-1473) uses the
// poly row 3-q, the differences D_(q+2), D_(q+1), D_q of the 7-point window
// (D_i = v(i+1) - v(i)), beta(q) and d(q).
struct Coef7 {
    double eL[4][3], eR[4][3];   // d_q * poly(3-q, 0..2)
    double b[4][6];              // beta_coef(j, 3-q, 0..5)
    double dL[4], dR[4];
};

__device__ __forceinline__ void load_coef7(const Coefs& c, int jc, Coef7& k)
{
    const int nc = c.nc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 3 - q;
        const double dl = c.dL[q + 5 * jc], dr = c.dR[q + 5 * jc];
        k.dL[q] = dl;
        k.dR[q] = dr;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            k.eL[q][s] = dl * c.polyL[jc + nc * (r + 5 * s)];
            k.eR[q][s] = dr * c.polyR[jc + nc * (r + 5 * s)];
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) k.b[q][s] = c.beta[jc + nc * (r + 5 * s)];
    }
}

// beta(q) from the shared products: the reference's six terms
// (:1462-1472) for stencil row 3-q, whose window starts at D_q.
__device__ __forceinline__ void betas7(const double (&D)[6], const Coef7& k, double eps, double (&b)[4])
{
    double sq[6], n1[5], n2[4];
#pragma unroll
    for (int i = 0; i < 6; ++i) sq[i] = D[i] * D[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) n1[i] = D[i + 1] * D[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) n2[i] = D[i + 2] * D[i];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        b[q] = fma(k.b[q][0], sq[q + 2],
                   fma(k.b[q][1], n1[q + 1],
                       fma(k.b[q][2], n2[q], fma(k.b[q][3], sq[q + 1], fma(k.b[q][4], n1[q], fma(k.b[q][5], sq[q], eps))))));
}

template <int NEWTON>
__device__ __forceinline__ bool cell7(const double (&v)[7], const Coef7& k, double eps, double& oL, double& oR)
{
    double D[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) D[i] = v[i + 1] - v[i];
    double b[4];
    betas7(D, k, eps, b);
    const double q0 = b[0] * b[0], q1 = b[1] * b[1], q2 = b[2] * b[2], q3 = b[3] * b[3];
    const double q01 = q0 * q1, q23 = q2 * q3;
    const double u[4] = {q1 * q23, q0 * q23, q01 * q3, q01 * q2};   // prod_{s != q} beta_s^2
    double NL = 0.0, NR = 0.0, SL = 0.0, SR = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double gL = fma(k.eL[q][0], D[q + 2], fma(k.eL[q][1], D[q + 1], k.eL[q][2] * D[q]));
        const double gR = fma(k.eR[q][0], D[q + 2], fma(k.eR[q][1], D[q + 1], k.eR[q][2] * D[q]));
        NL = fma(u[q], gL, NL);
        NR = fma(u[q], gR, NR);
        SL = fma(k.dL[q], u[q], SL);
        SR = fma(k.dR[q], u[q], SR);
    }
    const double T = SL * SR;
    double r = __builtin_amdgcn_rcp(T);
#pragma unroll
    for (int it = 0; it < NEWTON; ++it) {
        const double e = fma(-T, r, 1.0);
        r = fma(r, e, r);
    }
    oL = fma(NL, SR * r, v[3]);
    oR = fma(NR, SL * r, v[3]);
    return (static_cast<int>(pos_normal(T)) & static_cast<int>(pos_normal(r))) != 0;
}

__device__ __forceinline__ void cell7_slow_inl(const double (&v)[7], const Coef7& k, double eps, double& oL,
                                               double& oR)
{
    double D[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) D[i] = v[i + 1] - v[i];
    double b[4];
    betas7(D, k, eps, b);
    double NL = 0.0, NR = 0.0, SL = 0.0, SR = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double y = 1.0 / (b[q] * b[q]);
        NL += fma(k.eL[q][0], D[q + 2], fma(k.eL[q][1], D[q + 1], k.eL[q][2] * D[q])) * y;
        NR += fma(k.eR[q][0], D[q + 2], fma(k.eR[q][1], D[q + 1], k.eR[q][2] * D[q])) * y;
        SL += k.dL[q] * y;
        SR += k.dR[q] * y;
    }
    oL = v[3] + NL / SL;
    oR = v[3] + NR / SR;
}

}  // namespace
}  // namespace wm
