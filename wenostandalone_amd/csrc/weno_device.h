// Device-side building blocks shared by the sweep and adjoint kernels:
// storage types, coefficient access, the WENO3/5/7 (JS, TENO) cell updates
// and the monotonicity-preserving limiter.
#pragma once
#include "weno_internal.h"
#include <cfloat>
#include <cmath>

namespace wm {


struct bf16_t {
    uint16_t bits;
};

template <typename TIN>
struct CT {
    using type = float;
};
template <>
struct CT<double> {
    using type = double;
};

__device__ __forceinline__ double ld_in(const double* p) { return *p; }
__device__ __forceinline__ float ld_in(const float* p) { return *p; }
__device__ __forceinline__ float ld_in(const bf16_t* p)
{
    return __uint_as_float(static_cast<uint32_t>(p->bits) << 16);
}

template <typename T, typename V>
__device__ __forceinline__ void st_out(T* p, V v, bool nt)
{
    T x = static_cast<T>(v);
    if (nt)
        __builtin_nontemporal_store(x, p);
    else
        *p = x;
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
// contiguous range of tiles so that neighbouring tiles (which share boundary
// cache lines) meet in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int per = nblocks >> 3, rem = nblocks & 7;
    const int xcd = bid & 7, idx = bid >> 3;
    return xcd * per + min(xcd, rem) + idx;
}

// n/d for positive, finite, normal d: hardware reciprocal, two Newton steps
// and one residual correction (within one ulp of the IEEE quotient).
__device__ __forceinline__ double fast_div(double n, double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    double q = n * r;
    double res = fma(-d, q, n);
    return fma(res, r, q);
}

// 1/d to about one ulp: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp2(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

// ---------------------------------------------------------------------------
// per-j coefficients (reference layouts: poly (nc,5,3), d (5,nc), beta (nc,5,6))
// ---------------------------------------------------------------------------
// The coefficient arrays are device global memory.  Read through
// address-space-1 pointers they are global loads; through a generic pointer
// they would be flat loads, which also count against the LDS wait counter, so
// every later LDS wait would wait for them too.
typedef __attribute__((address_space(1))) const double GlbCoef;
__device__ __forceinline__ const GlbCoef* gcoef(const double* p) { return (const GlbCoef*)p; }

template <int ORDER, typename C>
struct Cf;
template <typename C>
struct Cf<3, C> {
    C pl[2], pr[2];
};
template <typename C>
struct Cf<5, C> {
    C pl[3][2], pr[3][2];
};
template <typename C>
struct Cf<7, C> {
    C pl[4][3], pr[4][3];
};

template <typename C>
__device__ __forceinline__ void load_cf(const Coefs& c, int jc, Cf<3, C>& f)
{
    const int nc = c.nc;
    const GlbCoef *pl = gcoef(c.polyL), *pr = gcoef(c.polyR);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        f.pl[r] = static_cast<C>(pl[jc + nc * r]);
        f.pr[r] = static_cast<C>(pr[jc + nc * r]);
    }
}
template <typename C>
__device__ __forceinline__ void load_cf(const Coefs& c, int jc, Cf<5, C>& f)
{
    const int nc = c.nc;
    const GlbCoef *pl = gcoef(c.polyL), *pr = gcoef(c.polyR);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            f.pl[r][s] = static_cast<C>(pl[jc + nc * (r + 5 * s)]);
            f.pr[r][s] = static_cast<C>(pr[jc + nc * (r + 5 * s)]);
        }
    }
}
template <typename C>
__device__ __forceinline__ void load_cf(const Coefs& c, int jc, Cf<7, C>& f)
{
    const int nc = c.nc;
    const GlbCoef *pl = gcoef(c.polyL), *pr = gcoef(c.polyR);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            f.pl[r][s] = static_cast<C>(pl[jc + nc * (r + 5 * s)]);
            f.pr[r][s] = static_cast<C>(pr[jc + nc * (r + 5 * s)]);
        }
    }
}

// Smoothness-indicator coefficients beta_coef(j, r, s) and the linear weights
// d_cbL/d_cbR(r, j) of one lane, kept in an LDS table [lane][entry] shared by
// the block's lane groups (they hold the same j): entries r*NS+s for beta,
// then NR entries of dL and NR of dR, rows padded to an odd number of doubles
// (BetaTraits::LD) so that the lanes of a wave hit distinct banks.  A lane's
// entries sit at constant offsets from its row, so they are read with
// immediate-offset ds_read2_b64 and no address arithmetic.  Reading them at
// use keeps them out of the register budget.
// LDS-resident doubles: loads through such a pointer are ds_read (a generic
// pointer would make them flat loads, which also wait on the wave's
// outstanding global stores).
typedef __attribute__((address_space(3))) const double LdsF64;
__device__ __forceinline__ const LdsF64* lds_ptr(const double* p) { return (const LdsF64*)p; }

template <int NS>
struct BetaRef {
    const LdsF64* p;   // the lane's row
    int nb;            // number of beta entries (NS*NR)
    int nr;            // stencils
    __device__ __forceinline__ double operator()(int r, int s) const { return p[r * NS + s]; }
    __device__ __forceinline__ double dl(int r) const { return p[nb + r]; }
    __device__ __forceinline__ double dr(int r) const { return p[nb + nr + r]; }
};
template <int ORDER>
struct BetaTraits {
    static constexpr int NS = ORDER == 3 ? 1 : (ORDER == 5 ? 3 : 6);   // terms per stencil
    static constexpr int NR = ORDER == 3 ? 2 : (ORDER == 5 ? 3 : 4);   // stencils
    static constexpr int N = NS * NR;
    static constexpr int ENTRIES = N + 2 * NR;
    static constexpr int LD = (ENTRIES % 2) ? ENTRIES : ENTRIES + 1;   // table row stride
};

// The same table held in registers (loaded once per lane); fits the
// 128-VGPR budget for WENO3/5.
template <int ORDER>
struct BetaRegs {
    static constexpr int NS = BetaTraits<ORDER>::NS, NR = BetaTraits<ORDER>::NR, N = BetaTraits<ORDER>::N;
    double b[N], l[NR], r[NR];
    __device__ __forceinline__ void load(const Coefs& c, int jc)
    {
#pragma unroll
        for (int t = 0; t < N; ++t) {
            const int rr = t / NS, ss = t - (t / NS) * NS;
            b[t] = gcoef(c.beta)[jc + c.nc * (rr + 5 * ss)];
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            l[q] = gcoef(c.dL)[q + 5 * jc];
            r[q] = gcoef(c.dR)[q + 5 * jc];
        }
    }
    __device__ __forceinline__ double operator()(int rr, int ss) const { return b[rr * NS + ss]; }
    __device__ __forceinline__ double dl(int q) const { return l[q]; }
    __device__ __forceinline__ double dr(int q) const { return r[q]; }
};

// P_r = prod_{s != r} q_s
template <int N>
__device__ __forceinline__ void excl_products(const double (&q)[N], double (&P)[N])
{
    if constexpr (N == 2) {
        P[0] = q[1];
        P[1] = q[0];
    } else if constexpr (N == 3) {
        P[0] = q[1] * q[2];
        P[1] = q[0] * q[2];
        P[2] = q[0] * q[1];
    } else {
        const double q01 = q[0] * q[1], q23 = q[2] * q[3];
        P[0] = q[1] * q23;
        P[1] = q[0] * q23;
        P[2] = q01 * q[3];
        P[3] = q01 * q[2];
    }
}

// JS: sum_r w_r p_r with w_r = (d_r/q_r) / sum_s (d_s/q_s)
template <int N, typename C>
__device__ __forceinline__ double js_combine(const double (&q)[N], const double (&P)[N],
                                             const double (&d)[N], const C (&p)[N])
{
    double S = 0.0, num = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double a = d[r] * P[r];
        S += a;
        num = fma(a, static_cast<double>(p[r]), num);
    }
    if (!(S >= DBL_MIN && S <= DBL_MAX)) {
        S = 0.0;
        num = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const double a = d[r] / q[r];
            S += a;
            num = fma(a, static_cast<double>(p[r]), num);
        }
        return num / S;
    }
    return fast_div(num, S);
}

// Both sides of a JS face pair with one reciprocal (the weno3/5/7 default;
// SHARED_RCP = false normalises each side on its own, which measured 3 %
// faster in the batched kernel): a_r = d_r P_r per side,
// out = num_L S_R / (S_L S_R) and num_R S_L / (S_L S_R), the reciprocal of
// S_L S_R refined to about one ulp.  A lane whose product is not a positive
// normal number (or whose reciprocal is not) normalises each side on its own
// (js_combine, which also carries NaN/Inf through as the reference does).
template <int N, typename C>
__device__ __forceinline__ void js_combine2(const double (&q)[N], const double (&P)[N], const double (&dl)[N],
                                            const double (&dr)[N], const C (&pL)[N], const C (&pR)[N],
                                            double& oL, double& oR)
{
    double SL = 0.0, SR = 0.0, nL = 0.0, nR = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double aL = dl[r] * P[r], aR = dr[r] * P[r];
        SL += aL;
        SR += aR;
        nL = fma(aL, static_cast<double>(pL[r]), nL);
        nR = fma(aR, static_cast<double>(pR[r]), nR);
    }
    const double T = SL * SR;
    const double rT = rcp2(T);
    if (__builtin_amdgcn_class(T, 1 << 8) && __builtin_amdgcn_class(rT, 1 << 8)) {
        oL = nL * (SR * rT);
        oR = nR * (SL * rT);
    } else {
        oL = js_combine<N, C>(q, P, dl, pL);
        oR = js_combine<N, C>(q, P, dr, pR);
    }
}

template <int N, typename C>
__device__ __forceinline__ double masked_combine(const double (&a)[N], const C (&p)[N], double extra)
{
    double S = 0.0, num = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        S += a[r];
        num = fma(a[r], static_cast<double>(p[r]), num);
    }
    const double D = S + extra;
    return (D >= DBL_MIN && D <= DBL_MAX) ? fast_div(num, D) : num / D;
}


// TENO cut-off decisions a_r / s < ct (weno_standalone.f90:584-590,
// :1494-1500), screened.  The weights a_r = (1 + tau/beta_r)^6 enter the
// result only through these decisions, so they are first formed with
// reciprocal-based quotients and plain products: with every
// x_r = 1 + tau/beta_r >= 1/2 (no cancellation) and a finite sum, each ratio
// is within about 100 ulps (2.2e-14 relative) of the one the reference forms
// with IEEE divisions and pow.  The decisions are taken from these ratios
// when every ratio is more than 1e-12 ct away from ct; otherwise (or on any
// NaN) the function returns false and the caller decides with the exact
// arithmetic.  `extra` is added to the sum as is (TENO7's uncut a_4).
template <int N>
__device__ __forceinline__ bool teno_screen(const double (&b)[N], double tau, double extra, double ct,
                                            double (&del)[N])
{
    double a[N], s = extra;
    bool ok = true;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double x = fma(tau, rcp2(b[r]), 1.0);
        ok = ok && (x >= 0.5);
        const double t = x * x * x;
        a[r] = t * t;
        s += a[r];
    }
    ok = ok && (s <= DBL_MAX);
    const double rs = rcp2(s), m = 1e-12 * fabs(ct);
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double ratio = a[r] * rs;
        ok = ok && (fabs(ratio - ct) > m);
        del[r] = ratio < ct ? 0.0 : 1.0;
    }
    return ok;
}

// x**6 for TENO5's (1 + tau/beta)**6._wp (weno_standalone.f90:580-582), which
// the reference evaluates with the C library's pow.  Here x^2 = p2 + e2 and
// x^3 = p3 + e3 are carried as unevaluated double-double sums (Dekker's
// products through FMA), x^6 = (x^3)^2 likewise, and the sum is rounded once:
// relative error about 2^-100 before that rounding, so the result is x^6
// correctly rounded (pow's result to within its own < 1 ulp) at about a tenth
// of pow's instruction count.  When x^6 overflows the plain product (inf) is
// returned, as pow does.  Contraction is off: fused into the final sum, the
// product p6 = p3*p3 would count its own rounding error twice (p3*p3 + e6
// instead of p6 + e6).
__device__ __forceinline__ double pow6(double x)
{
    _Pragma("clang fp contract(off)")
    const double p2 = x * x, e2 = fma(x, x, -p2);
    const double p3 = p2 * x, e3 = fma(p2, x, -p3) + e2 * x;
    const double p6 = p3 * p3, e6 = fma(p3, p3, -p6) + 2.0 * p3 * e3;
    return fabs(p6) <= DBL_MAX ? p6 + e6 : p6;
}

__device__ __forceinline__ bool normal_mag(double x) { return fabs(x) >= DBL_MIN && fabs(x) <= DBL_MAX; }

// Mapped WENO (weno_standalone.f90:433-437, :557-570, :1478-1482): the JS
// weights w_r = a_r / sum(a), a_r = d_r / beta_r^2, mapped through
//   g_r(w) = (d_r (1 + d_r - 3 w) + w^2) * (w / (d_r^2 + w (1 - 2 d_r)))
// (Henrick, Aslam & Powers 2005), then normalised again.  The quotients are
// formed with reciprocals refined to about one ulp (rcp2; the reciprocals of
// beta_r^2 are the same for both sides), which moves the face by a few ulps
// against the reference's IEEE divisions; a lane whose denominators leave the
// normal range uses IEEE divisions throughout.
template <int N, typename C>
__device__ __forceinline__ double mapped_combine(const double (&q)[N], const double (&d)[N], const C (&p)[N])
{
    double a[N], s = 0.0;
    bool ok = true;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        ok = ok && normal_mag(q[r]);
        a[r] = d[r] * rcp2(q[r]);
        s += a[r];
    }
    ok = ok && normal_mag(s);
    if (ok) {
        const double rs = rcp2(s);
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const double w = a[r] * rs;
            const double den = d[r] * d[r] + w * (1.0 - 2.0 * d[r]);
            ok = ok && normal_mag(den);
            a[r] = (d[r] * (1.0 + d[r] - 3.0 * w) + w * w) * (w * rcp2(den));
        }
    }
    if (!ok) {
        s = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            a[r] = d[r] / q[r];
            s += a[r];
        }
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const double w = a[r] / s;
            a[r] = (d[r] * (1.0 + d[r] - 3.0 * w) + w * w) * (w / (d[r] * d[r] + w * (1.0 - 2.0 * d[r])));
        }
    }
    return masked_combine<N, C>(a, p, 0.0);
}

// WENO-Z (weno_standalone.f90:439-441, :572-576, :1484-1488):
// a_r = d_r (1 + (tau / beta_r)^Q), Q = 1 for WENO3/5 and wenoz_q = 2 for
// WENO7 as the reference writes them.  IEEE divisions: the quotients are
// shared by both sides, and a reciprocal-based form with its normal-range
// fallback measured slower here (more registers than the kernel holds).
template <int N, int Q, typename C>
__device__ __forceinline__ double wenoz_combine(const double (&b)[N], double tau, const double (&d)[N],
                                                const C (&p)[N])
{
    double a[N];
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double t = tau / b[r];
        a[r] = d[r] * (1.0 + (Q == 2 ? t * t : t));
    }
    return masked_combine<N, C>(a, p, 0.0);
}

// ---- stencil terms: candidate polynomials and smoothness indicators --------
// Differences and candidates (DIFF/CAND) and smoothness indicators (BETA) of
// each order, written in the reference's evaluation order (left-to-right
// sums and products, :421-428, :529-550, :1450-1473).
#define WM_DIFF3 const C d0 = v[2] - v[1], dm1 = v[1] - v[0];
#define WM_CAND3                                                                            \
    pL[0] = v[1] + f.pl[0] * d0;                                                            \
    pL[1] = v[1] + f.pl[1] * dm1;                                                           \
    pR[0] = v[1] + f.pr[0] * d0;                                                            \
    pR[1] = v[1] + f.pr[1] * dm1;
#define WM_BETA3                                                                            \
    const double D0 = d0, Dm1 = dm1;                                                        \
    b0 = B(0, 0) * D0 * D0 + eps;                                                           \
    b1 = B(1, 0) * Dm1 * Dm1 + eps;
#define WM_DIFF5                                                                            \
    const C vj = v[2];                                                                      \
    const C dp1 = v[4] - v[3], d0 = v[3] - v[2], dm1 = v[2] - v[1], dm2 = v[1] - v[0];
#define WM_CAND5                                                                            \
    pL[0] = vj + f.pl[0][0] * dp1 + f.pl[0][1] * d0;                                        \
    pL[1] = vj + f.pl[1][0] * d0 + f.pl[1][1] * dm1;                                        \
    pL[2] = vj + f.pl[2][0] * dm1 + f.pl[2][1] * dm2;                                       \
    pR[0] = vj + f.pr[0][0] * dp1 + f.pr[0][1] * d0;                                        \
    pR[1] = vj + f.pr[1][0] * d0 + f.pr[1][1] * dm1;                                        \
    pR[2] = vj + f.pr[2][0] * dm1 + f.pr[2][1] * dm2;
#define WM_BETA5                                                                            \
    const double Dp1 = dp1, D0 = d0, Dm1 = dm1, Dm2 = dm2;                                  \
    b0 = B(0, 0) * Dp1 * Dp1 + B(0, 1) * Dp1 * D0 + B(0, 2) * D0 * D0 + eps;                \
    b1 = B(1, 0) * D0 * D0 + B(1, 1) * D0 * Dm1 + B(1, 2) * Dm1 * Dm1 + eps;                \
    b2 = B(2, 0) * Dm1 * Dm1 + B(2, 1) * Dm1 * Dm2 + B(2, 2) * Dm2 * Dm2 + eps;
// WENO7: poly(q)/beta(q) use coefficient row 3-q (:1457-1473)
#define WM_DIFF7                                                                            \
    const C vj = v[3];                                                                      \
    const C d2 = v[6] - v[5], d1 = v[5] - v[4], d0 = v[4] - v[3];                           \
    const C dm1 = v[3] - v[2], dm2 = v[2] - v[1], dm3 = v[1] - v[0];
#define WM_CAND7                                                                            \
    pL[0] = vj + f.pl[3][0] * dm1 + f.pl[3][1] * dm2 + f.pl[3][2] * dm3;                    \
    pL[1] = vj + f.pl[2][0] * d0 + f.pl[2][1] * dm1 + f.pl[2][2] * dm2;                     \
    pL[2] = vj + f.pl[1][0] * d1 + f.pl[1][1] * d0 + f.pl[1][2] * dm1;                      \
    pL[3] = vj + f.pl[0][0] * d2 + f.pl[0][1] * d1 + f.pl[0][2] * d0;                       \
    pR[0] = vj + f.pr[3][0] * dm1 + f.pr[3][1] * dm2 + f.pr[3][2] * dm3;                    \
    pR[1] = vj + f.pr[2][0] * d0 + f.pr[2][1] * dm1 + f.pr[2][2] * dm2;                     \
    pR[2] = vj + f.pr[1][0] * d1 + f.pr[1][1] * d0 + f.pr[1][2] * dm1;                      \
    pR[3] = vj + f.pr[0][0] * d2 + f.pr[0][1] * d1 + f.pr[0][2] * d0;
#define WM_BETA7                                                                            \
    const double D2 = d2, D1 = d1, D0 = d0, Dm1 = dm1, Dm2 = dm2, Dm3 = dm3;                \
    b[3] = B(0, 0) * D2 * D2 + B(0, 1) * D2 * D1 + B(0, 2) * D2 * D0 + B(0, 3) * D1 * D1 +  \
           B(0, 4) * D1 * D0 + B(0, 5) * D0 * D0 + eps;                                     \
    b[2] = B(1, 0) * D1 * D1 + B(1, 1) * D1 * D0 + B(1, 2) * D1 * Dm1 + B(1, 3) * D0 * D0 + \
           B(1, 4) * D0 * Dm1 + B(1, 5) * Dm1 * Dm1 + eps;                                  \
    b[1] = B(2, 0) * D0 * D0 + B(2, 1) * D0 * Dm1 + B(2, 2) * D0 * Dm2 + B(2, 3) * Dm1 * Dm1 + \
           B(2, 4) * Dm1 * Dm2 + B(2, 5) * Dm2 * Dm2 + eps;                                 \
    b[0] = B(3, 0) * Dm1 * Dm1 + B(3, 1) * Dm1 * Dm2 + B(3, 2) * Dm1 * Dm3 + B(3, 3) * Dm2 * Dm2 + \
           B(3, 4) * Dm2 * Dm3 + B(3, 5) * Dm3 * Dm3 + eps;

// _fused: everything contracted (the JS weights, well conditioned in beta).
// _exact: the smoothness indicators without contraction (`fp contract(off)`:
// HIP's default contraction mode honours the pragma), for the mapped, WENO-Z
// and TENO branches, whose weights take differences of betas
// (tau = |beta_a - beta_b|) or cut at a threshold, so one-ulp differences in
// beta would be amplified; evaluated exactly as the reference does, beta and
// tau are bit-identical to it.  The candidates enter only the final weighted
// sum, so they are contracted in both (three FMAs each).
#define WM_TERMS_FNS(O, ARGS)                                                               \
    template <typename C, typename BT>                                                      \
    __device__ __forceinline__ void terms##O##_fused(const C* v, const Cf<O, C>& f, const BT& B, double eps, ARGS) \
    {                                                                                       \
        WM_DIFF##O WM_CAND##O WM_BETA##O                                                    \
    }                                                                                       \
    template <typename C, typename BT>                                                      \
    __device__ __forceinline__ void terms##O##_exact(const C* v, const Cf<O, C>& f, const BT& B, double eps, ARGS) \
    {                                                                                       \
        WM_DIFF##O WM_CAND##O                                                               \
        {                                                                                   \
            _Pragma("clang fp contract(off)")                                               \
            WM_BETA##O                                                                      \
        }                                                                                   \
    }
#define WM_COMMA ,
WM_TERMS_FNS(3, C (&pL)[2] WM_COMMA C (&pR)[2] WM_COMMA double& b0 WM_COMMA double& b1)
WM_TERMS_FNS(5, C (&pL)[3] WM_COMMA C (&pR)[3] WM_COMMA double& b0 WM_COMMA double& b1 WM_COMMA double& b2)
WM_TERMS_FNS(7, C (&pL)[4] WM_COMMA C (&pR)[4] WM_COMMA double (&b)[4])
#undef WM_TERMS_FNS
#undef WM_COMMA
#undef WM_DIFF3
#undef WM_CAND3
#undef WM_BETA3
#undef WM_DIFF5
#undef WM_CAND5
#undef WM_BETA5
#undef WM_DIFF7
#undef WM_CAND7
#undef WM_BETA7

template <bool EXACT, typename C, typename BT>
__device__ __forceinline__ void terms3(const C* v, const Cf<3, C>& f, const BT& B, double eps, C (&pL)[2],
                                       C (&pR)[2], double& b0, double& b1)
{
    if constexpr (EXACT) terms3_exact(v, f, B, eps, pL, pR, b0, b1);
    else terms3_fused(v, f, B, eps, pL, pR, b0, b1);
}
template <bool EXACT, typename C, typename BT>
__device__ __forceinline__ void terms5(const C* v, const Cf<5, C>& f, const BT& B, double eps, C (&pL)[3],
                                       C (&pR)[3], double& b0, double& b1, double& b2)
{
    if constexpr (EXACT) terms5_exact(v, f, B, eps, pL, pR, b0, b1, b2);
    else terms5_fused(v, f, B, eps, pL, pR, b0, b1, b2);
}
// WENO7 terms of the JS form (contracted): the 15 distinct products of the
// differences D_-3..D_2 (squares, neighbours at distance 1 and 2) are formed
// once and shared by the four smoothness indicators, 15 products + 24 FMA
// instead of 24 products + 24 FMA.  Row r of the coefficients uses the
// differences (A, B, C) = (D[h], D[h-1], D[h-2]), h = 5 - r, into b[3 - r],
// the reference's pairing (:1457-1473).
template <typename C, typename BT>
__device__ __forceinline__ void terms7_shared(const C* v, const Cf<7, C>& f, const BT& B, double eps,
                                              C (&pL)[4], C (&pR)[4], double (&b)[4])
{
    const C vj = v[3];
    const C d2 = v[6] - v[5], d1 = v[5] - v[4], d0 = v[4] - v[3];
    const C dm1 = v[3] - v[2], dm2 = v[2] - v[1], dm3 = v[1] - v[0];
    pL[0] = vj + f.pl[3][0] * dm1 + f.pl[3][1] * dm2 + f.pl[3][2] * dm3;
    pL[1] = vj + f.pl[2][0] * d0 + f.pl[2][1] * dm1 + f.pl[2][2] * dm2;
    pL[2] = vj + f.pl[1][0] * d1 + f.pl[1][1] * d0 + f.pl[1][2] * dm1;
    pL[3] = vj + f.pl[0][0] * d2 + f.pl[0][1] * d1 + f.pl[0][2] * d0;
    pR[0] = vj + f.pr[3][0] * dm1 + f.pr[3][1] * dm2 + f.pr[3][2] * dm3;
    pR[1] = vj + f.pr[2][0] * d0 + f.pr[2][1] * dm1 + f.pr[2][2] * dm2;
    pR[2] = vj + f.pr[1][0] * d1 + f.pr[1][1] * d0 + f.pr[1][2] * dm1;
    pR[3] = vj + f.pr[0][0] * d2 + f.pr[0][1] * d1 + f.pr[0][2] * d0;
    const double D[6] = {dm3, dm2, dm1, d0, d1, d2};   // D[i] = D_(i-3)
    double sq[6], n1[5], n2[4];
#pragma unroll
    for (int i = 0; i < 6; ++i) sq[i] = D[i] * D[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) n1[i] = D[i + 1] * D[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) n2[i] = D[i + 2] * D[i];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int h = 5 - r;
        b[3 - r] = fma(B(r, 0), sq[h],
                       fma(B(r, 1), n1[h - 1],
                           fma(B(r, 2), n2[h - 2],
                               fma(B(r, 3), sq[h - 1], fma(B(r, 4), n1[h - 2], fma(B(r, 5), sq[h - 2], eps))))));
    }
}

template <bool EXACT, typename C, typename BT>
__device__ __forceinline__ void terms7(const C* v, const Cf<7, C>& f, const BT& B, double eps, C (&pL)[4],
                                       C (&pR)[4], double (&b)[4])
{
    if constexpr (EXACT) terms7_exact(v, f, B, eps, pL, pR, b);
    else terms7_shared(v, f, B, eps, pL, pR, b);
}

// ---- WENO3: v[0..2] = j-1, j, j+1 (weno_standalone.f90:392-475) ----------
template <int MODE, typename C, bool SHARED_RCP = true, typename BT>
__device__ __forceinline__ void weno3(const C* v, const Cf<3, C>& f, const BT& B, double eps,
                                      double, double& oL, double& oR)
{
    // no TENO branch for WENO3 in the reference: teno alone leaves NaN
    if constexpr (MODE == MODE_NONE || MODE == MODE_TENO) {
        oL = oR = __builtin_nan("");
        return;
    } else {
        C pL[2], pR[2];
        double b0, b1;
        terms3<MODE != MODE_JS>(v, f, B, eps, pL, pR, b0, b1);
        const double q[2] = {b0 * b0, b1 * b1};
        const double dl[2] = {B.dl(0), B.dl(1)}, dr[2] = {B.dr(0), B.dr(1)};
        if constexpr (MODE == MODE_JS) {
            double P[2];
            excl_products<2>(q, P);
            if constexpr (SHARED_RCP) {
                js_combine2<2, C>(q, P, dl, dr, pL, pR, oL, oR);
            } else {
                oL = js_combine<2, C>(q, P, dl, pL);
                oR = js_combine<2, C>(q, P, dr, pR);
            }
        } else if constexpr (MODE == MODE_MAPPED) {
            oL = mapped_combine<2, C>(q, dl, pL);
            oR = mapped_combine<2, C>(q, dr, pR);
        } else {
            const double bb[2] = {b0, b1};
            const double tau = fabs(b1 - b0);
            oL = wenoz_combine<2, 1, C>(bb, tau, dl, pL);
            oR = wenoz_combine<2, 1, C>(bb, tau, dr, pR);
        }
    }
}

// ---- WENO5: v[0..4] = j-2..j+2 (weno_standalone.f90:477-669) -------------
template <int MODE, typename C, bool SHARED_RCP = true, typename BT>
__device__ __forceinline__ void weno5(const C* v, const Cf<5, C>& f, const BT& B, double eps,
                                      double ct, double& oL, double& oR)
{
    if constexpr (MODE == MODE_NONE) {
        oL = oR = __builtin_nan("");
        return;
    } else {
        C pL[3], pR[3];
        double b0, b1, b2;
        terms5<MODE != MODE_JS>(v, f, B, eps, pL, pR, b0, b1, b2);
        if constexpr (MODE == MODE_JS) {
            const double q[3] = {b0 * b0, b1 * b1, b2 * b2};
            double P[3];
            excl_products<3>(q, P);
            const double dl[3] = {B.dl(0), B.dl(1), B.dl(2)}, dr[3] = {B.dr(0), B.dr(1), B.dr(2)};
            if constexpr (SHARED_RCP) {
                js_combine2<3, C>(q, P, dl, dr, pL, pR, oL, oR);
            } else {
                oL = js_combine<3, C>(q, P, dl, pL);
                oR = js_combine<3, C>(q, P, dr, pR);
            }
        } else if constexpr (MODE == MODE_MAPPED) {
            const double q[3] = {b0 * b0, b1 * b1, b2 * b2};
            const double dl[3] = {B.dl(0), B.dl(1), B.dl(2)}, dr[3] = {B.dr(0), B.dr(1), B.dr(2)};
            oL = mapped_combine<3, C>(q, dl, pL);
            oR = mapped_combine<3, C>(q, dr, pR);
        } else if constexpr (MODE == MODE_WENOZ) {
            const double bb[3] = {b0, b1, b2};
            const double tau = fabs(b2 - b0);
            const double dl[3] = {B.dl(0), B.dl(1), B.dl(2)}, dr[3] = {B.dr(0), B.dr(1), B.dr(2)};
            oL = wenoz_combine<3, 1, C>(bb, tau, dl, pL);
            oR = wenoz_combine<3, 1, C>(bb, tau, dr, pR);
        } else {
            // :578-606: decisions screened (teno_screen), else IEEE
            // divisions and pow6 as the reference, no product fused into a
            // sum
            const double tau = fabs(b2 - b0);
            const double bb[3] = {b0, b1, b2};
            double del[3];
            if (!teno_screen<3>(bb, tau, 0.0, ct, del)) {
                _Pragma("clang fp contract(off)")
                const double a0 = pow6(1.0 + tau / b0);
                const double a1 = pow6(1.0 + tau / b1);
                const double a2 = pow6(1.0 + tau / b2);
                const double s = a0 + a1 + a2;
                del[0] = (a0 / s < ct) ? 0.0 : 1.0;
                del[1] = (a1 / s < ct) ? 0.0 : 1.0;
                del[2] = (a2 / s < ct) ? 0.0 : 1.0;
            }
            const double aL[3] = {del[0] * B.dl(0), del[1] * B.dl(1), del[2] * B.dl(2)};
            const double aR[3] = {del[0] * B.dr(0), del[1] * B.dr(1), del[2] * B.dr(2)};
            oL = masked_combine<3, C>(aL, pL, 0.0);
            oR = masked_combine<3, C>(aR, pR, 0.0);
        }
    }
}

// ---- WENO7: v[0..6] = j-3..j+3 (weno_standalone.f90:1419-1542) -----------
// poly(q)/beta(q) use coefficient row 3-q, alpha(q) uses d(q): the
// reference's pairing is kept as is.
template <int MODE, typename C, bool SHARED_RCP = true, typename BT>
__device__ __forceinline__ void weno7(const C* v, const Cf<7, C>& f, const BT& B, double eps,
                                      double ct, double& oL, double& oR)
{
    if constexpr (MODE == MODE_NONE) {
        oL = oR = __builtin_nan("");
        return;
    } else {
        C pL[4], pR[4];
        double b[4];
        terms7<MODE != MODE_JS>(v, f, B, eps, pL, pR, b);
        if constexpr (MODE == MODE_JS) {
            const double q[4] = {b[0] * b[0], b[1] * b[1], b[2] * b[2], b[3] * b[3]};
            double P[4];
            excl_products<4>(q, P);
            const double dl[4] = {B.dl(0), B.dl(1), B.dl(2), B.dl(3)};
            const double dr[4] = {B.dr(0), B.dr(1), B.dr(2), B.dr(3)};
            if constexpr (SHARED_RCP) {
                js_combine2<4, C>(q, P, dl, dr, pL, pR, oL, oR);
            } else {
                oL = js_combine<4, C>(q, P, dl, pL);
                oR = js_combine<4, C>(q, P, dr, pR);
            }
        } else if constexpr (MODE == MODE_MAPPED) {
            // alpha(4) = 0 joins the sums of :1480 and changes nothing
            const double q[4] = {b[0] * b[0], b[1] * b[1], b[2] * b[2], b[3] * b[3]};
            const double dl[4] = {B.dl(0), B.dl(1), B.dl(2), B.dl(3)};
            const double dr[4] = {B.dr(0), B.dr(1), B.dr(2), B.dr(3)};
            oL = mapped_combine<4, C>(q, dl, pL);
            oR = mapped_combine<4, C>(q, dr, pR);
        } else if constexpr (MODE == MODE_WENOZ) {
            const double tau = fabs(b[3] - b[0]);
            const double dl[4] = {B.dl(0), B.dl(1), B.dl(2), B.dl(3)};
            const double dr[4] = {B.dr(0), B.dr(1), B.dr(2), B.dr(3)};
            oL = wenoz_combine<4, 2, C>(b, tau, dl, pL);
            oR = wenoz_combine<4, 2, C>(b, tau, dr, pR);
        } else {
            // :1490-1503: beta(4) keeps weno_eps, alpha(4) is never cut.  The
            // real power 3 is the product x*x*x, as the reference build
            // evaluates it (the cut-off is discontinuous in these values).
            // The decisions are screened (teno_screen); a_4, which also
            // enters the result, is formed as the reference forms it.
            const double tau = fabs(eps - b[3]);
            double a4;
            {
                _Pragma("clang fp contract(off)")
                const double x = 1.0 + tau / eps;
                const double t = x * x * x;
                a4 = t * t;
            }
            double del[4];
            if (!teno_screen<4>(b, tau, a4, ct, del)) {
                // as the reference rounds it: no product fused into the sum
                _Pragma("clang fp contract(off)")
                double a[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double x = 1.0 + tau / b[q];
                    const double t = x * x * x;
                    a[q] = t * t;
                }
                const double s = a[0] + a[1] + a[2] + a[3] + a4;
#pragma unroll
                for (int q = 0; q < 4; ++q) del[q] = (a[q] / s < ct) ? 0.0 : 1.0;
            }
            const double aL[4] = {del[0] * B.dl(0), del[1] * B.dl(1), del[2] * B.dl(2), del[3] * B.dl(3)};
            const double aR[4] = {del[0] * B.dr(0), del[1] * B.dr(1), del[2] * B.dr(2), del[3] * B.dr(3)};
            oL = masked_combine<4, C>(aL, pL, a4);
            oR = masked_combine<4, C>(aR, pR, a4);
        }
    }
}

// ---- monotonicity-preserving limiter (weno_standalone.f90:1544-1615) ------
// v[0..4] = j-2..j+2.  The two minmod4 values of :1568-1576 are computed once
// and used in swapped roles for the two sides (:1590-1598).
template <typename C>
__device__ __forceinline__ C mp_minmod4(C a, C b)
{
    const C one = C(1);
    const C s4ab = copysign(one, C(4) * a - b);
    return (s4ab + copysign(one, C(4) * b - a)) *
           fabs((s4ab + copysign(one, a)) * (s4ab + copysign(one, b))) *
           fmin(fmin(fmin(fabs(C(4) * a - b), fabs(a)), fabs(C(4) * b - a)), fabs(b)) / C(8);
}

template <typename C>
__device__ __forceinline__ C mp_clip(C x, C lo, C hi)
{
    const C h = C(0.5);
    return x + (copysign(h, lo - x) + copysign(h, hi - x)) * fmin(fabs(lo - x), fabs(hi - x));
}

template <typename C>
__device__ __forceinline__ void mp_limit(const C* v, C& vL, C& vR)
{
    const C vm2 = v[0], vm1 = v[1], v0 = v[2], vp1 = v[3], vp2 = v[4];
    const C dm1 = v0 + vm2 - C(2) * vm1;
    const C d0 = vp1 + vm1 - C(2) * v0;
    const C dp1 = vp2 + v0 - C(2) * vp1;
    const C m_lo = mp_minmod4(dm1, d0);   // d_MD of the L side, d_LC of the R side
    const C m_hi = mp_minmod4(d0, dp1);   // d_LC of the L side, d_MD of the R side
    const C beta_mp = C(4.0 / 3.0);
    {
        const C ul = v0 - C(2) * (vp1 - v0);
        const C md = (v0 + vm1 - m_lo) * C(0.5);
        const C lc = v0 - (vp1 - v0) * C(0.5) + beta_mp * m_hi;
        const C lo = fmax(fmin(fmin(v0, vm1), md), fmin(fmin(v0, ul), lc));
        const C hi = fmin(fmax(fmax(v0, vm1), md), fmax(fmax(v0, ul), lc));
        vL = mp_clip(vL, lo, hi);
    }
    {
        const C ul = v0 + C(2) * (v0 - vm1);
        const C md = (v0 + vp1 - m_hi) * C(0.5);
        const C lc = v0 + (v0 - vm1) * C(0.5) + beta_mp * m_lo;
        const C lo = fmax(fmin(fmin(v0, vp1), md), fmin(fmin(v0, ul), lc));
        const C hi = fmin(fmax(fmax(v0, vp1), md), fmax(fmax(v0, ul), lc));
        vR = mp_clip(vR, lo, hi);
    }
}


}  // namespace wm
