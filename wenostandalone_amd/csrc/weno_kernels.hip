// HIP kernels for gfx950: the sweep kernel behind reshape, order-1 copy,
// WENO3/5/7 (JS and TENO), the fused or standalone monotonicity-preserving
// limiter, in FP64, FP32 and BF16-state storage.
//
// Design (one kernel template, one workgroup = one tile):
//   * A tile is NP pencils x one j-chunk of one variable.  The pencils of a
//     tile are consecutive in the input's fastest transverse coordinate, so
//     the load phase is coalesced for every sweep direction:
//       - sweep along the contiguous axis (x, or a reshaped array): lanes run
//         along the pencil;
//       - sweep along a strided axis (y or z of the padded field): lanes run
//         across the NP pencils, NP*8 bytes contiguous per row.
//     The tile lands in LDS as [pencil][J] (row stride padded to an odd
//     number of elements), i.e. the y/z transposes happen on the way into LDS
//     and no reshaped copy of the field is ever written to HBM.
//   * Compute phase: one lane per output j.  The lane loads the per-j stencil
//     coefficients once into registers and reuses them for all NP pencils,
//     reading its window from LDS.  Outputs of a pencil are contiguous along
//     j in every direction (vL_x(x,y,z), vL_y(y,x,z), vL_z(z,y,x)), so stores
//     are coalesced.  Halo cells (j outside the core range) can be filled with
//     the order-1 copy in the same pass.
//   * Weights: the JS weights are normalised with one division per side:
//       w_r = a_r / sum(a),  a_r = d_r / b_r^2
//       => sum_r w_r p_r = (sum_r d_r P_r p_r) / (sum_r d_r P_r),
//          P_r = prod_{s != r} b_s^2
//     which replaces the reference's 2*(N+N) IEEE divisions per cell by two.
//     If the products leave the normal range the lane falls back to the
//     per-stencil reciprocals.  TENO keeps the reference's divisions and
//     powers because its cut-off decision is discontinuous.
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

// ---------------------------------------------------------------------------
// per-j coefficients (reference layouts: poly (nc,5,3), d (5,nc), beta (nc,5,6))
// ---------------------------------------------------------------------------
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
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        f.pl[r] = static_cast<C>(c.polyL[jc + nc * r]);
        f.pr[r] = static_cast<C>(c.polyR[jc + nc * r]);
    }
}
template <typename C>
__device__ __forceinline__ void load_cf(const Coefs& c, int jc, Cf<5, C>& f)
{
    const int nc = c.nc;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            f.pl[r][s] = static_cast<C>(c.polyL[jc + nc * (r + 5 * s)]);
            f.pr[r][s] = static_cast<C>(c.polyR[jc + nc * (r + 5 * s)]);
        }
    }
}
template <typename C>
__device__ __forceinline__ void load_cf(const Coefs& c, int jc, Cf<7, C>& f)
{
    const int nc = c.nc;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            f.pl[r][s] = static_cast<C>(c.polyL[jc + nc * (r + 5 * s)]);
            f.pr[r][s] = static_cast<C>(c.polyR[jc + nc * (r + 5 * s)]);
        }
    }
}

// Smoothness-indicator coefficients beta_coef(j, r, s) and the linear weights
// d_cbL/d_cbR(r, j) of one lane, kept in an LDS table [entry][lane] shared by
// the block's lane groups (they hold the same j): entries r*NS+s for beta,
// then NR entries of dL and NR of dR.  Reading them at use keeps them out of
// the register budget.
template <int NS>
struct BetaRef {
    const double* p;
    int stride;
    int nb;   // number of beta entries (NS*NR)
    int nr;   // stencils
    __device__ __forceinline__ double operator()(int r, int s) const { return p[(r * NS + s) * stride]; }
    __device__ __forceinline__ double dl(int r) const { return p[(nb + r) * stride]; }
    __device__ __forceinline__ double dr(int r) const { return p[(nb + nr + r) * stride]; }
};
template <int ORDER>
struct BetaTraits {
    static constexpr int NS = ORDER == 3 ? 1 : (ORDER == 5 ? 3 : 6);   // terms per stencil
    static constexpr int NR = ORDER == 3 ? 2 : (ORDER == 5 ? 3 : 4);   // stencils
    static constexpr int N = NS * NR;
    static constexpr int ENTRIES = N + 2 * NR;
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

template <int N, typename C>
__device__ __forceinline__ double masked_combine(const double (&a)[N], const C (&p)[N], double extra)
{
    double S = 0.0, num = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
        S += a[r];
        num = fma(a[r], static_cast<double>(p[r]), num);
    }
    return num / (S + extra);
}

// ---- WENO3: v[0..2] = j-1, j, j+1 (weno_standalone.f90:392-475) ----------
template <int MODE, typename C>
__device__ __forceinline__ void weno3(const C* v, const Cf<3, C>& f, const BetaRef<1>& B, double eps,
                                      double, double& oL, double& oR)
{
    if constexpr (MODE != MODE_JS) {
        oL = oR = __builtin_nan("");
        return;
    } else {
        const C d0 = v[2] - v[1], dm1 = v[1] - v[0];
        const C pL[2] = {v[1] + f.pl[0] * d0, v[1] + f.pl[1] * dm1};
        const C pR[2] = {v[1] + f.pr[0] * d0, v[1] + f.pr[1] * dm1};
        const double D0 = d0, Dm1 = dm1;
        const double b0 = B(0, 0) * D0 * D0 + eps;
        const double b1 = B(1, 0) * Dm1 * Dm1 + eps;
        const double q[2] = {b0 * b0, b1 * b1};
        double P[2];
        excl_products<2>(q, P);
        const double dl[2] = {B.dl(0), B.dl(1)}, dr[2] = {B.dr(0), B.dr(1)};
        oL = js_combine<2, C>(q, P, dl, pL);
        oR = js_combine<2, C>(q, P, dr, pR);
    }
}

// ---- WENO5: v[0..4] = j-2..j+2 (weno_standalone.f90:477-669) -------------
template <int MODE, typename C>
__device__ __forceinline__ void weno5(const C* v, const Cf<5, C>& f, const BetaRef<3>& B, double eps,
                                      double ct, double& oL, double& oR)
{
    if constexpr (MODE == MODE_NONE) {
        oL = oR = __builtin_nan("");
        return;
    } else {
        const C vj = v[2];
        const C dp1 = v[4] - v[3], d0 = v[3] - v[2], dm1 = v[2] - v[1], dm2 = v[1] - v[0];
        const C pL[3] = {vj + f.pl[0][0] * dp1 + f.pl[0][1] * d0,
                         vj + f.pl[1][0] * d0 + f.pl[1][1] * dm1,
                         vj + f.pl[2][0] * dm1 + f.pl[2][1] * dm2};
        const C pR[3] = {vj + f.pr[0][0] * dp1 + f.pr[0][1] * d0,
                         vj + f.pr[1][0] * d0 + f.pr[1][1] * dm1,
                         vj + f.pr[2][0] * dm1 + f.pr[2][1] * dm2};
        const double Dp1 = dp1, D0 = d0, Dm1 = dm1, Dm2 = dm2;
        const double b0 = B(0, 0) * Dp1 * Dp1 + B(0, 1) * Dp1 * D0 + B(0, 2) * D0 * D0 + eps;
        const double b1 = B(1, 0) * D0 * D0 + B(1, 1) * D0 * Dm1 + B(1, 2) * Dm1 * Dm1 + eps;
        const double b2 = B(2, 0) * Dm1 * Dm1 + B(2, 1) * Dm1 * Dm2 + B(2, 2) * Dm2 * Dm2 + eps;
        if constexpr (MODE == MODE_JS) {
            const double q[3] = {b0 * b0, b1 * b1, b2 * b2};
            double P[3];
            excl_products<3>(q, P);
            const double dl[3] = {B.dl(0), B.dl(1), B.dl(2)}, dr[3] = {B.dr(0), B.dr(1), B.dr(2)};
            oL = js_combine<3, C>(q, P, dl, pL);
            oR = js_combine<3, C>(q, P, dr, pR);
        } else {
            // :578-606, kept with IEEE divisions and pow
            const double tau = fabs(b2 - b0);
            const double a0 = pow(1.0 + tau / b0, 6.0);
            const double a1 = pow(1.0 + tau / b1, 6.0);
            const double a2 = pow(1.0 + tau / b2, 6.0);
            const double s = a0 + a1 + a2;
            const double del0 = (a0 / s < ct) ? 0.0 : 1.0;
            const double del1 = (a1 / s < ct) ? 0.0 : 1.0;
            const double del2 = (a2 / s < ct) ? 0.0 : 1.0;
            const double aL[3] = {del0 * B.dl(0), del1 * B.dl(1), del2 * B.dl(2)};
            const double aR[3] = {del0 * B.dr(0), del1 * B.dr(1), del2 * B.dr(2)};
            oL = masked_combine<3, C>(aL, pL, 0.0);
            oR = masked_combine<3, C>(aR, pR, 0.0);
        }
    }
}

// ---- WENO7: v[0..6] = j-3..j+3 (weno_standalone.f90:1419-1542) -----------
// poly(q)/beta(q) use coefficient row 3-q, alpha(q) uses d(q): the
// reference's pairing is kept as is.
template <int MODE, typename C>
__device__ __forceinline__ void weno7(const C* v, const Cf<7, C>& f, const BetaRef<6>& B, double eps,
                                      double ct, double& oL, double& oR)
{
    if constexpr (MODE == MODE_NONE) {
        oL = oR = __builtin_nan("");
        return;
    } else {
        const C vj = v[3];
        const C d2 = v[6] - v[5], d1 = v[5] - v[4], d0 = v[4] - v[3];
        const C dm1 = v[3] - v[2], dm2 = v[2] - v[1], dm3 = v[1] - v[0];
        const C pL[4] = {vj + f.pl[3][0] * dm1 + f.pl[3][1] * dm2 + f.pl[3][2] * dm3,
                         vj + f.pl[2][0] * d0 + f.pl[2][1] * dm1 + f.pl[2][2] * dm2,
                         vj + f.pl[1][0] * d1 + f.pl[1][1] * d0 + f.pl[1][2] * dm1,
                         vj + f.pl[0][0] * d2 + f.pl[0][1] * d1 + f.pl[0][2] * d0};
        const C pR[4] = {vj + f.pr[3][0] * dm1 + f.pr[3][1] * dm2 + f.pr[3][2] * dm3,
                         vj + f.pr[2][0] * d0 + f.pr[2][1] * dm1 + f.pr[2][2] * dm2,
                         vj + f.pr[1][0] * d1 + f.pr[1][1] * d0 + f.pr[1][2] * dm1,
                         vj + f.pr[0][0] * d2 + f.pr[0][1] * d1 + f.pr[0][2] * d0};
        const double D2 = d2, D1 = d1, D0 = d0, Dm1 = dm1, Dm2 = dm2, Dm3 = dm3;
        double b[4];
        {
            auto c = [&](int t) { return B(0, t); };
            b[3] = c(0) * D2 * D2 + c(1) * D2 * D1 + c(2) * D2 * D0 + c(3) * D1 * D1 + c(4) * D1 * D0 +
                   c(5) * D0 * D0 + eps;
        }
        {
            auto c = [&](int t) { return B(1, t); };
            b[2] = c(0) * D1 * D1 + c(1) * D1 * D0 + c(2) * D1 * Dm1 + c(3) * D0 * D0 + c(4) * D0 * Dm1 +
                   c(5) * Dm1 * Dm1 + eps;
        }
        {
            auto c = [&](int t) { return B(2, t); };
            b[1] = c(0) * D0 * D0 + c(1) * D0 * Dm1 + c(2) * D0 * Dm2 + c(3) * Dm1 * Dm1 +
                   c(4) * Dm1 * Dm2 + c(5) * Dm2 * Dm2 + eps;
        }
        {
            auto c = [&](int t) { return B(3, t); };
            b[0] = c(0) * Dm1 * Dm1 + c(1) * Dm1 * Dm2 + c(2) * Dm1 * Dm3 + c(3) * Dm2 * Dm2 +
                   c(4) * Dm2 * Dm3 + c(5) * Dm3 * Dm3 + eps;
        }
        if constexpr (MODE == MODE_JS) {
            const double q[4] = {b[0] * b[0], b[1] * b[1], b[2] * b[2], b[3] * b[3]};
            double P[4];
            excl_products<4>(q, P);
            const double dl[4] = {B.dl(0), B.dl(1), B.dl(2), B.dl(3)};
            const double dr[4] = {B.dr(0), B.dr(1), B.dr(2), B.dr(3)};
            oL = js_combine<4, C>(q, P, dl, pL);
            oR = js_combine<4, C>(q, P, dr, pR);
        } else {
            // :1490-1503: beta(4) keeps weno_eps, alpha(4) is never cut
            const double tau = fabs(eps - b[3]);
            double a[5];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double t = pow(1.0 + tau / b[q], 3.0);
                a[q] = t * t;
            }
            {
                const double t = pow(1.0 + tau / eps, 3.0);
                a[4] = t * t;
            }
            const double s = a[0] + a[1] + a[2] + a[3] + a[4];
            double del[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) del[q] = (a[q] / s < ct) ? 0.0 : 1.0;
            const double aL[4] = {del[0] * B.dl(0), del[1] * B.dl(1), del[2] * B.dl(2), del[3] * B.dl(3)};
            const double aR[4] = {del[0] * B.dr(0), del[1] * B.dr(1), del[2] * B.dr(2), del[3] * B.dr(3)};
            oL = masked_combine<4, C>(aL, pL, a[4]);
            oR = masked_combine<4, C>(aR, pR, a[4]);
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

constexpr int kMaxThreads = 1024;
constexpr int kMaxNP = 64;
constexpr int kMaxSlots = 6;

template <int ORDER>
struct OrderTraits {
    static constexpr int R = ORDER / 2;
};

// WENO7 keeps 56 coefficients per lane in registers: give its waves the
// 256-VGPR budget of two waves per SIMD (512-thread blocks); everything else
// runs four waves per SIMD (1024-thread blocks, 128 VGPRs).
template <int OP, int ORDER>
struct BlockTraits {
    static constexpr int kThreads = (OP == OP_WENO && ORDER == 7) ? 512 : kMaxThreads;
};

// ---------------------------------------------------------------------------
// the sweep kernel: persistent, double-buffered
//
//   block = G lane groups; group g owns pencils [g*npg, (g+1)*npg) of a tile
//   and lane jj of the group owns output j = jlo + jj for the whole launch, so
//   the per-j coefficients are loaded once per thread.
//   Each block walks a contiguous-per-XCD range of tiles.  While tile k is
//   computed from LDS buffer k&1, tile k+1 is already in flight from HBM into
//   registers and is written to the other buffer afterwards: one barrier per
//   tile.  Pencil address tables are built two tiles ahead (3 slots).
// ---------------------------------------------------------------------------
template <int OP, int ORDER, int MODE, bool MP, typename TIN, typename TOUT>
__global__ void __launch_bounds__((BlockTraits<OP, ORDER>::kThreads)) sweep_kernel(const TileArgs a, const int nt)
{
    using C = typename CT<TIN>::type;
    constexpr int RW = (OP == OP_WENO) ? OrderTraits<ORDER>::R : 0;
    constexpr int R = (OP == OP_MP || (OP == OP_WENO && MP)) ? (RW > 2 ? RW : 2) : RW;
    constexpr int W = 2 * R + 1;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* const lds0 = reinterpret_cast<C*>(smem_raw);
    __shared__ long long s_in[3][kMaxNP];
    __shared__ long long s_out[3][kMaxNP];

    const int tid = threadIdx.x;
    const int T = blockDim.x;
    const int np = a.np;
    const int ldj = a.ldj;
    C* const lds1 = lds0 + np * ldj;
    constexpr int OB = (ORDER == 3 || ORDER == 5 || ORDER == 7) ? ORDER : 5;
    constexpr int NS = BetaTraits<OB>::NS;
    constexpr int NB = (OP == OP_WENO) ? BetaTraits<OB>::N : 0;
    constexpr int NR = BetaTraits<OB>::NR;
    double* const beta_lds =
        reinterpret_cast<double*>(smem_raw + ((2 * static_cast<size_t>(np) * ldj * sizeof(C) + 15) & ~size_t(15)));

    // ---- which chunk and which tiles ----
    const int bpc = a.bpc;
    const int ch = blockIdx.x / bpc;
    const int bin = blockIdx.x - ch * bpc;
    const int ntile = a.tiles_per_var * a.nv;
    const int xcd = bin & 7, rank = bin >> 3, nbx = bpc >> 3;
    const int per = ntile >> 3, rem = ntile & 7;
    const int t_first = xcd * per + min(xcd, rem) + rank;
    const int t_end = (xcd + 1) * per + min(xcd + 1, rem);
    const int t_stride = nbx;
    if (t_first >= t_end) return;

    // ---- output range of the chunk and the input range it needs ----
    const int jo0 = ch * a.chunk;
    const int jo1 = min(a.out_len, jo0 + a.chunk);
    const int c0 = max(jo0, a.j_beg), c1 = min(jo1 - 1, a.j_end);
    const bool has_core = (OP == OP_WENO || OP == OP_MP) && c0 <= c1;
    int lo, hi, jlo, jhi;
    if (a.copy_halo) {
        lo = jo0 + a.J0f;
        hi = jo1 - 1 + a.J0f;
        if (has_core) {
            lo = min(lo, c0 + a.J0f - R);
            hi = max(hi, c1 + a.J0f + R);
        }
        jlo = jo0;
        jhi = jo1 - 1;
    } else {
        if (!has_core) return;
        lo = c0 + a.J0f - R;
        hi = c1 + a.J0f + R;
        jlo = c0;
        jhi = c1;
    }
    const int nJ = hi - lo + 1;

    // ---- fixed per-thread roles ----
    const int gq = tid / a.jn;
    const int j = jlo + (tid - gq * a.jn);
    const bool active = gq < a.groups && j <= jhi;
    const bool core = active && has_core && j >= c0 && j <= c1;
    const bool do_mp = MP && core && j >= a.mp_beg && j <= a.mp_end;
    const int col = j + a.J0f - lo;
    const int q0 = gq * a.npg;

    // load slots of this thread (same for every tile): (q << 16) | J, or -1
    int slot[kMaxSlots];
    const int nel = np * nJ;
#pragma unroll
    for (int s = 0; s < kMaxSlots; ++s) {
        const int e = tid + s * T;
        slot[s] = -1;
        if (s < a.slots && e < nel) {
            int q, J;
            if (a.s_j == 1) {
                q = e / nJ;
                J = e - q * nJ;
            } else {
                J = e / np;
                q = e - J * np;
            }
            slot[s] = (q << 16) | J;
        }
    }

    Cf<OB, C> f;
    const int jj = tid - gq * a.jn;
    if constexpr (OP == OP_WENO) {
        if (core) {
            int jc = j - 3;
            jc = jc < 0 ? 0 : (jc > a.c.nc - 1 ? a.c.nc - 1 : jc);
            load_cf(a.c, jc, f);
            if (gq == 0) {
#pragma unroll
                for (int t = 0; t < NB; ++t) {
                    const int r = t / NS, s = t - (t / NS) * NS;
                    beta_lds[t * a.jn + jj] = a.c.beta[jc + a.c.nc * (r + 5 * s)];
                }
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    beta_lds[(NB + r) * a.jn + jj] = a.c.dL[r + 5 * jc];
                    beta_lds[(NB + NR + r) * a.jn + jj] = a.c.dR[r + 5 * jc];
                }
            }
        }
    }

    const TIN* const src = static_cast<const TIN*>(a.src);
    TOUT* const outL = static_cast<TOUT*>(a.outL);
    TOUT* const outR = static_cast<TOUT*>(a.outR);
    const int npen = a.na * a.nb;
    const bool ntb = nt != 0;
    const bool wl = (a.sides & 1) != 0, wr = (a.sides & 2) != 0;

    auto tables = [&](int t, int slot) {
        if (tid < np) {
            const int iv = t / a.tiles_per_var;
            const int P = (t - iv * a.tiles_per_var) * np + tid;
            if (P < npen) {
                const int pa = P % a.na, pb = P / a.na;
                s_in[slot][tid] = a.src_off + pa * a.s_a + pb * a.s_b + iv * a.s_i + lo * a.s_j;
                const int k = a.a_is_k ? pa : pb;
                const int l = a.a_is_k ? pb : pa;
                s_out[slot][tid] = k * a.o_k + l * a.o_l + iv * a.o_i;
            } else {
                s_in[slot][tid] = -1;
                s_out[slot][tid] = -1;
            }
        }
    };
    C regs[kMaxSlots];
    auto issue = [&](int slot_id) {
#pragma unroll
        for (int s = 0; s < kMaxSlots; ++s) {
            regs[s] = C(0);
            // opaque copy: keeps per-slot 64-bit addresses from being hoisted
            // out of the tile loop into registers
            int sl = slot[s];
            asm volatile("" : "+v"(sl));
            if (sl >= 0) {
                const long long b = s_in[slot_id][sl >> 16];
                if (b >= 0) regs[s] = static_cast<C>(ld_in(src + b + static_cast<long long>(sl & 0xffff) * a.s_j));
            }
        }
    };
    auto commit = [&](C* buf) {
#pragma unroll
        for (int s = 0; s < kMaxSlots; ++s)
            if (slot[s] >= 0) buf[(slot[s] >> 16) * ldj + (slot[s] & 0xffff)] = regs[s];
    };

    // ---- prologue ----
    tables(t_first, 0);
    if (t_first + t_stride < t_end) tables(t_first + t_stride, 1);
    __syncthreads();
    issue(0);
    commit(lds0);
    __syncthreads();

    int k = 0;
    for (int t = t_first; t < t_end; t += t_stride, ++k) {
        C* const cur = (k & 1) ? lds1 : lds0;
        C* const nxt = (k & 1) ? lds0 : lds1;
        const int tslot = k % 3;
        const bool more = t + t_stride < t_end;
        if (more) issue((k + 1) % 3);
        if (t + 2 * t_stride < t_end) tables(t + 2 * t_stride, (k + 2) % 3);

        if (active) {
#pragma unroll 1
            for (int qq = 0; qq < a.npg; ++qq) {
                const int q = q0 + qq;
                const long long ob = s_out[tslot][q];
                if (ob < 0) continue;
                const C* row = cur + q * ldj + col;
                if (!core) {
                    const C val = row[0];
                    if (wl) st_out(outL + ob + j, val, ntb);
                    if constexpr (OP != OP_RESHAPE)
                        if (wr) st_out(outR + ob + j, val, ntb);
                    continue;
                }
                if constexpr (OP == OP_WENO) {
                    // re-read the LDS coefficient table per pencil: an opaque
                    // pointer keeps the loads from being hoisted into registers
                    const double* bp = beta_lds + jj;
                    asm volatile("" : "+v"(bp));
                    const BetaRef<NS> B{bp, a.jn, NB, NR};
                    C v[W];
#pragma unroll
                    for (int o = 0; o < W; ++o) v[o] = row[o - R];
                    double oL, oR;
                    const C* vc = v + (R - RW);
                    if constexpr (ORDER == 3)
                        weno3<MODE, C>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                    else if constexpr (ORDER == 5)
                        weno5<MODE, C>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                    else
                        weno7<MODE, C>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                    if constexpr (MP) {
                        C xl = static_cast<C>(oL), xr = static_cast<C>(oR);
                        if (do_mp) mp_limit<C>(v + (R - 2), xl, xr);
                        if (wl) st_out(outL + ob + j, xl, ntb);
                        if (wr) st_out(outR + ob + j, xr, ntb);
                    } else {
                        if (wl) st_out(outL + ob + j, oL, ntb);
                        if (wr) st_out(outR + ob + j, oR, ntb);
                    }
                } else if constexpr (OP == OP_MP) {
                    C v[W];
#pragma unroll
                    for (int o = 0; o < W; ++o) v[o] = row[o - R];
                    C xl = static_cast<C>(outL[ob + j]);
                    C xr = static_cast<C>(outR[ob + j]);
                    mp_limit<C>(v, xl, xr);
                    st_out(outL + ob + j, xl, false);
                    st_out(outR + ob + j, xr, false);
                }
            }
        }
        if (more) commit(nxt);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
static int cu_count()
{
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
    }
    return n;
}

template <int OP, int ORDER, int MODE, bool MP, typename TIN, typename TOUT>
static hipError_t go(TileArgs a, const Launch& l, hipStream_t st)
{
    using C = typename CT<TIN>::type;
    auto kern = sweep_kernel<OP, ORDER, MODE, MP, TIN, TOUT>;
    constexpr int NB = (OP == OP_WENO) ? BetaTraits<ORDER>::ENTRIES : 0;
    const size_t lds = ((2 * static_cast<size_t>(a.np) * a.ldj * sizeof(C) + 15) & ~size_t(15)) +
                       static_cast<size_t>(NB) * a.jn * sizeof(double);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, a.threads, lds) != hipSuccess || occ < 1) occ = 1;
    const int ntile = a.tiles_per_var * a.nv;
    int bpc = cu_count() * occ;
    if (l.blocks_per_cu > 0) bpc = cu_count() * l.blocks_per_cu;
    const int need = (ntile + 1) / 2;    // at least two tiles per block
    if (bpc > need) bpc = need;
    bpc = (bpc + 7) & ~7;
    a.bpc = bpc;
    hipLaunchKernelGGL(kern, dim3(bpc * a.chunks), dim3(a.threads), lds, st, a, l.nontemporal);
    return hipGetLastError();
}

template <typename TIN, typename TOUT>
static hipError_t go_weno(const TileArgs& a, const Launch& l, hipStream_t st)
{
#define WM_CASE(O, M)                                                             \
    if (l.order == O && l.mode == M)                                              \
        return l.mp ? go<OP_WENO, O, M, true, TIN, TOUT>(a, l, st)                \
                    : go<OP_WENO, O, M, false, TIN, TOUT>(a, l, st);
    WM_CASE(3, MODE_NONE)
    WM_CASE(3, MODE_JS)
    WM_CASE(5, MODE_NONE)
    WM_CASE(5, MODE_JS)
    WM_CASE(5, MODE_TENO)
    WM_CASE(7, MODE_NONE)
    WM_CASE(7, MODE_JS)
    WM_CASE(7, MODE_TENO)
#undef WM_CASE
    return hipErrorInvalidValue;
}

int max_threads(int op, int order);

hipError_t launch_tile(const TileArgs& a, const Launch& l, hipStream_t st)
{
    if (a.tiles_per_var <= 0 || a.nv <= 0) return hipSuccess;
    if (a.np > kMaxNP || a.slots > kMaxSlots || a.threads > max_threads(l.op, l.order) || a.threads < 64 ||
        a.ldj > 0xffff)
        return hipErrorInvalidConfiguration;
    switch (l.op) {
    case OP_RESHAPE:
        if (l.dtype == DT_F64) return go<OP_RESHAPE, 5, MODE_JS, false, double, double>(a, l, st);
        if (l.dtype == DT_F32) return go<OP_RESHAPE, 5, MODE_JS, false, float, float>(a, l, st);
        return hipErrorInvalidValue;
    case OP_COPY:
        if (l.dtype == DT_F64) return go<OP_COPY, 5, MODE_JS, false, double, double>(a, l, st);
        if (l.dtype == DT_F32) return go<OP_COPY, 5, MODE_JS, false, float, float>(a, l, st);
        return go<OP_COPY, 5, MODE_JS, false, bf16_t, float>(a, l, st);
    case OP_MP:
        if (l.dtype == DT_F64) return go<OP_MP, 5, MODE_JS, true, double, double>(a, l, st);
        if (l.dtype == DT_F32) return go<OP_MP, 5, MODE_JS, true, float, float>(a, l, st);
        return go<OP_MP, 5, MODE_JS, true, bf16_t, float>(a, l, st);
    case OP_WENO:
        if (l.dtype == DT_F64) return go_weno<double, double>(a, l, st);
        if (l.dtype == DT_F32) return go_weno<float, float>(a, l, st);
        return go_weno<bf16_t, float>(a, l, st);
    }
    return hipErrorInvalidValue;
}

int kernel_max_slots() { return kMaxSlots; }
int max_threads(int op, int order) { return (op == OP_WENO && order == 7) ? 512 : kMaxThreads; }
int kernel_max_np() { return kMaxNP; }

}  // namespace wm
