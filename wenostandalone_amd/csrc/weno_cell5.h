// The WENO5-JS face pair of one cell in the shared-reciprocal form used by
// the sweep kernels (the reference's s_weno5 JS branch,
// inputs/weno_standalone.f90:477-1417):
//     D_i = v(i+1) - v(i),  S_i = D_i^2,  X_i = D_i D_(i-1)
//     beta_r = B_r0 S + B_r1 X + B_r2 S' + eps              (3 FMA each)
//     u_r    = prod_{s != r} beta_s^2                     (6 multiplies)
//     g_r    = (d_r a_r) D + (d_r b_r) D'                 (p_r = v_j + a_r D + b_r D')
//     out    = v_j + (sum_r u_r g_r) / (sum_r d_r u_r)
// which is  sum_r w_r p_r,  w_r = (d_r/beta_r^2)/sum_s(...),  regrouped so
// that both sides share one reciprocal of (sum_L)(sum_R).  cell5 reports
// when that product leaves the normal range; the caller then redoes the cell
// with cell5_slow (per-stencil IEEE divisions, which also carry NaN and
// infinity through as the reference does).
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

// Per-stencil IEEE form (the reference's alpha_r = d_r / beta_r^2).  The
// inline form is for kernels that take it per cell; cell5_slow keeps it out
// of line for kernels that redo whole tiles.
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

__device__ __noinline__ void cell5_slow(double v0, double v1, double v2, double v3, double v4, const Coef5& k,
                                        double eps, double& oL, double& oR)
{
    cell5_slow_inl(v0, v1, v2, v3, v4, k, eps, oL, oR);
}

}  // namespace
}  // namespace wm
