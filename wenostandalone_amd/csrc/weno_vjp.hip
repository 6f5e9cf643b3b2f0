// Adjoint (vector-Jacobian product) of the WENO5/7-JS reconstruction.
//
// For cotangents gL, gR on the core outputs of one sweep, accumulates
//   gv += (dvL/dv)^T gL + (dvR/dv)^T gR
// into an array with the input's layout.  Nothing from the forward pass is
// stored: beta, the weights and the candidate polynomials are recomputed per
// output.  Per output j (both sides share beta):
//   vL = sum_r w_r p_r,  w_r = a_r / sum(a),  a_r = d_r / b_r^2
//   dvL/dp_r = w_r,  dvL/db_r = -2 w_r (p_r - vL) / b_r,
// p_r is linear and b_r quadratic in the differences D_k = v(J+k+1) - v(J+k).
// Gather formulation, no atomics: a tile of pencils sits in LDS as in the
// forward sweep; each lane writes its output's 2R+1 local contributions to
// LDS, then each input position sums the contributions of the outputs whose
// stencils cover it; the tile's gradient leaves through the same coalesced
// slot mapping the forward uses for loads (read-modify-write of gv).
#include "weno_device.h"
#include "weno_cells.h"
#include "weno_tiles.h"

#include <algorithm>

namespace wm {

constexpr int kVjpThreads = 512;
constexpr int kVjpSlots = 8;
constexpr int kVjpMaxNP = 32;

// contributions c[0..2R] of one output to v(J-R .. J+R)
template <int ORDER>
__device__ __forceinline__ void weno_vjp_cell(const double* v, const Cf<ORDER, double>& f,
                                              const BetaRef<BetaTraits<ORDER>::NS>& B, double eps, double gl,
                                              double gr, double* c)
{
    constexpr int R = ORDER / 2;
    constexpr int NR = R + 1;     // candidate stencils
    constexpr int NT = R;         // polynomial terms per stencil
    constexpr int ND = 2 * R;     // differences D_k, k = -R..R-1, stored at k+R
    double D[ND], gD[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        D[i] = v[i + 1] - v[i];
        gD[i] = 0.0;
    }
    const double vj = v[R];
    double pL[NR], pR[NR], b[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        // WENO5: candidate q uses row q and D_{1-q}, D_{-q};
        // WENO7: candidate q uses row 3-q and D_{q-1}, D_{q-2}, D_{q-3}
        const int row = (ORDER == 5) ? q : 3 - q;
        double sL = vj, sR = vj;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int k = (ORDER == 5) ? (1 - q - t) : (q - 1 - t);
            sL = fma(f.pl[row][t], D[k + R], sL);
            sR = fma(f.pr[row][t], D[k + R], sR);
        }
        pL[q] = sL;
        pR[q] = sR;
        if constexpr (ORDER == 5) {
            const double A = D[1 - q + R], Bv = D[-q + R];
            b[q] = B(row, 0) * A * A + B(row, 1) * A * Bv + B(row, 2) * Bv * Bv + eps;
        } else {
            const double A = D[q - 1 + R], Bv = D[q - 2 + R], Cv = D[q - 3 + R];
            b[q] = B(row, 0) * A * A + B(row, 1) * A * Bv + B(row, 2) * A * Cv + B(row, 3) * Bv * Bv +
                   B(row, 4) * Bv * Cv + B(row, 5) * Cv * Cv + eps;
        }
    }
    // Normalised weights and 1/b_r.  Shared form: w_r = d_r P_r / S with
    // P_r = prod_{s != r} b_s^2; WENO5 normalises both sides by one
    // reciprocal of S_L S_R and takes 1/b_r = (prod_{s != r} b_s) / prod_s b_s
    // (one more reciprocal); WENO7 (register-bound) divides per side and per
    // stencil.  Lanes whose products leave the normal range use the
    // per-stencil form 1/b_r^2 with IEEE-accurate divisions.
    double qq[NR], P[NR], wL[NR], wR[NR], ib[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) qq[q] = b[q] * b[q];
    excl_products<NR>(qq, P);
    double SL = 0.0, SR = 0.0;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        wL[q] = B.dl(q) * P[q];
        wR[q] = B.dr(q) * P[q];
        SL += wL[q];
        SR += wR[q];
    }
    double iL, iR;
    if constexpr (ORDER == 7) {
        if (!(SL >= DBL_MIN && SL <= DBL_MAX && SR >= DBL_MIN && SR <= DBL_MAX)) {
            SL = SR = 0.0;
#pragma unroll
            for (int q = 0; q < NR; ++q) {
                wL[q] = B.dl(q) / qq[q];
                wR[q] = B.dr(q) / qq[q];
                SL += wL[q];
                SR += wR[q];
            }
        }
        iL = fast_div(1.0, SL);
        iR = fast_div(1.0, SR);
    } else {
    double mb[NR];
    excl_products<NR>(b, mb);
    const double T = SL * SR, Q = mb[0] * b[0];
    double rT = __builtin_amdgcn_rcp(T), rQ = __builtin_amdgcn_rcp(Q);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        rT = fma(rT, fma(-T, rT, 1.0), rT);
        rQ = fma(rQ, fma(-Q, rQ, 1.0), rQ);
    }
    const bool shared = __builtin_amdgcn_class(T, 1 << 8) & __builtin_amdgcn_class(rT, 1 << 8) &
                        __builtin_amdgcn_class(Q, 1 << 8) & __builtin_amdgcn_class(rQ, 1 << 8);
    if (shared) {
        iL = SR * rT;
        iR = SL * rT;
#pragma unroll
        for (int q = 0; q < NR; ++q) ib[q] = mb[q] * rQ;
    } else {
        SL = SR = 0.0;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            wL[q] = B.dl(q) / qq[q];
            wR[q] = B.dr(q) / qq[q];
            SL += wL[q];
            SR += wR[q];
            ib[q] = 1.0 / b[q];
        }
        iL = 1.0 / SL;
        iR = 1.0 / SR;
    }
    }
    double vLo = 0.0, vRo = 0.0;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        wL[q] *= iL;
        wR[q] *= iR;
        vLo = fma(wL[q], pL[q], vLo);
        vRo = fma(wR[q], pR[q], vRo);
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int row = (ORDER == 5) ? q : 3 - q;
        const double gpL = gl * wL[q], gpR = gr * wR[q];
        double gb;
        if constexpr (ORDER == 7)
            gb = -2.0 * fast_div(fma(gpL, pL[q] - vLo, gpR * (pR[q] - vRo)), b[q]);
        else
            gb = -2.0 * fma(gpL, pL[q] - vLo, gpR * (pR[q] - vRo)) * ib[q];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int k = (ORDER == 5) ? (1 - q - t) : (q - 1 - t);
            gD[k + R] = fma(gpL, f.pl[row][t], fma(gpR, f.pr[row][t], gD[k + R]));
        }
        if constexpr (ORDER == 5) {
            const int ia = 1 - q + R, ib = -q + R;
            const double A = D[ia], Bv = D[ib];
            gD[ia] = fma(gb, 2.0 * B(row, 0) * A + B(row, 1) * Bv, gD[ia]);
            gD[ib] = fma(gb, B(row, 1) * A + 2.0 * B(row, 2) * Bv, gD[ib]);
        } else {
            const int ia = q - 1 + R, ib = q - 2 + R, ic = q - 3 + R;
            const double A = D[ia], Bv = D[ib], Cv = D[ic];
            gD[ia] = fma(gb, 2.0 * B(row, 0) * A + B(row, 1) * Bv + B(row, 2) * Cv, gD[ia]);
            gD[ib] = fma(gb, B(row, 1) * A + 2.0 * B(row, 3) * Bv + B(row, 4) * Cv, gD[ib]);
            gD[ic] = fma(gb, B(row, 2) * A + B(row, 4) * Bv + 2.0 * B(row, 5) * Cv, gD[ic]);
        }
    }
#pragma unroll
    for (int o = 0; o <= 2 * R; ++o) c[o] = 0.0;
    // sum_r w_r = 1 on each side: the centre value's direct derivative
    c[R] = gl + gr;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        c[i + 1] += gD[i];
        c[i] -= gD[i];
    }
}

// a.src: v; a.outL/a.outR: gL/gR (read); grad: gv (input layout, accumulated)
//
// WENO5 pipelines the tiles of a workgroup: the loads of tile t's gradient
// and of tile t+1's input are issued before tile t's arithmetic and land in
// registers (the input in a second LDS tile buffer after the sums), so one
// memory round trip per tile overlaps the arithmetic; tile addresses rotate
// through a ring of three.  WENO7 (about 255 VGPRs) has no registers for the
// loads in flight and runs the tile stages in sequence.
template <int ORDER>
__global__ void __launch_bounds__(kVjpThreads) vjp_kernel(const TileArgs a, double* __restrict__ grad)
{
    constexpr int R = ORDER / 2;
    constexpr int W = 2 * R + 1;
    constexpr int NS = BetaTraits<ORDER>::NS;
    constexpr int NB = BetaTraits<ORDER>::N;
    constexpr int LDB = BetaTraits<ORDER>::LD;
    constexpr int NR = BetaTraits<ORDER>::NR;
    constexpr bool ALL = ORDER == 5;   // pipelined tiles, cotangents fetched ahead (registers)

    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int np = a.np, ldj = a.ldj, G = a.groups, jn = a.jn;
    double* const tile0 = reinterpret_cast<double*>(smem_raw);    // [np][ldj] inputs
    double* const tile1 = tile0 + (ALL ? np * ldj : 0);            // second input buffer (WENO5)
    double* const gtile = tile1 + np * ldj;                        // [np][ldj] gradient
    // contributions: per lane group a slot of jn + 4R rows of W doubles, output jj's W values in row jj + 2R, and 2R
    // zero rows at either end, so that the sum for an input position reads
    // its W entries without bounds checks
    const int cbs = (jn + 4 * R) * W;
    double* const cbuf = gtile + np * ldj;
    double* const beta_lds = cbuf + G * cbs;                       // coefficient table
    constexpr int NRING = ALL ? 3 : 1;
    __shared__ long long s_in[NRING][kVjpMaxNP];
    __shared__ long long s_out[NRING][kVjpMaxNP];

    const int tid = threadIdx.x, T = blockDim.x;
    const int ntile = a.tiles_per_var * a.nv;
    const int bpc = a.bpc;
    const int xcd = blockIdx.x & 7, rank = blockIdx.x >> 3, nbx = bpc >> 3;
    const int per = ntile >> 3, rem = ntile & 7;
    const int t_first = xcd * per + min(xcd, rem) + rank;
    const int t_end = (xcd + 1) * per + min(xcd + 1, rem);
    if (t_first >= t_end) return;

    const int c0 = a.j_beg, c1 = a.j_end;            // one chunk: the whole core
    const int lo = c0 + a.J0f - R, hi = c1 + a.J0f + R;
    const int nJ = hi - lo + 1;
    const int gq = tid / jn;
    const int jj = tid - gq * jn;
    const int j = c0 + jj;
    const bool grp = gq < G;
    const bool core = grp && j <= c1;

    Cf<ORDER, double> f;
    if (core) {
        const int jc = j - 3;
        load_cf(a.c, jc, f);
        if (gq == 0) {
            for (int t = 0; t < NB; ++t) {
                const int r = t / NS, s = t - (t / NS) * NS;
                beta_lds[jj * LDB + t] = a.c.beta[jc + a.c.nc * (r + 5 * s)];
            }
            for (int r = 0; r < NR; ++r) {
                beta_lds[jj * LDB + NB + r] = a.c.dL[r + 5 * jc];
                beta_lds[jj * LDB + NB + NR + r] = a.c.dR[r + 5 * jc];
            }
        }
    }
    {
        const int pad = 2 * R, per = 2 * pad * W, nslot = G;
        for (int e = tid; e < nslot * per; e += T) {
            const int sl = e / per, rem = e - sl * per, r = rem / W, o = rem - r * W;
            cbuf[sl * cbs + (r < pad ? r : jn + r) * W + o] = 0.0;
        }
    }
    const int npen = a.na * a.nb;
    const double* const src = static_cast<const double*>(a.src);
    const double* const gL = static_cast<const double*>(a.outL);
    const double* const gR = static_cast<const double*>(a.outR);

    // this lane's tile element s (the same for every tile): e = tid + s*T is
    // (pencil q, row J) with e = q*nJ + J for contiguous rows, e = J*np + q
    // otherwise (consecutive lanes read consecutive addresses); packed as
    // (q << 16) | J, -1 for none.  WENO5 holds the table in registers for
    // the whole launch; WENO7 recomputes it where used (held, it would push
    // the kernel past 256 VGPRs).
    auto slot = [&](int s) -> int {
        const int e = tid + s * T;
        if (s >= a.slots || e >= np * nJ) return -1;
        int q, J;
        if (a.s_j == 1) {
            q = e / nJ;
            J = e - q * nJ;
        } else {
            J = e / np;
            q = e - J * np;
        }
        return (q << 16) | J;
    };
    // addresses of tile t's pencils into ring entry k (threads < np)
    auto plan = [&](int t, int k) {
        const int iv = t / a.tiles_per_var;
        const int P = (t - iv * a.tiles_per_var) * np + tid;
        if (P < npen) {
            const int pa = P % a.na, pb = P / a.na;
            s_in[k][tid] = a.src_off + pa * a.s_a + pb * a.s_b + iv * a.s_i + lo * a.s_j;
            const int kk = a.a_is_k ? pa : pb;
            const int l = a.a_is_k ? pb : pa;
            s_out[k][tid] = kk * a.o_k + l * a.o_l + iv * a.o_i;
        } else {
            s_in[k][tid] = -1;
            s_out[k][tid] = -1;
        }
    };
    // the whole slot table of lane `me` with one division: consecutive
    // slots are T elements apart.  `me` is the thread index passed through
    // an empty asm statement, so that the table is rebuilt where it is used
    // rather than hoisted out of the tile loop into registers held for the
    // whole launch.
    auto slot_table = [&](int me, int* sl) {
        const bool rows = a.s_j == 1;           // e = q*nJ + J, else e = J*np + q
        const int div = rows ? nJ : np, lim = rows ? np : nJ;
        const int step_hi = T / div, step_lo = T - step_hi * div;
        int hi = me / div, lo = me - hi * div;
#pragma unroll
        for (int s = 0; s < kVjpSlots; ++s) {
            sl[s] = (s < a.slots && hi < lim) ? (rows ? (hi << 16) | lo : (lo << 16) | hi) : -1;
            hi += step_hi;
            lo += step_lo;
            if (lo >= div) {
                lo -= div;
                ++hi;
            }
        }
    };
    // WENO5 helpers over a slot table: every slot's load is issued before
    // the first use, so a lane waits for one memory latency per tile array
    auto load_slots = [&](const int* sl, const double* arr, const long long* sin, double* r) {
#pragma unroll
        for (int s = 0; s < kVjpSlots; ++s) {
            r[s] = 0.0;
            if (sl[s] >= 0) {
                const long long b = sin[sl[s] >> 16];
                if (b >= 0) r[s] = arr[b + static_cast<long long>(sl[s] & 0xffff) * a.s_j];
            }
        }
    };
    auto put_slots = [&](const int* sl, double* dst, const double* r) {
#pragma unroll
        for (int s = 0; s < kVjpSlots; ++s)
            if (sl[s] >= 0) dst[(sl[s] >> 16) * ldj + (sl[s] & 0xffff)] = r[s];
    };
    // gv += gtile over the tile (each input element belongs to one pencil,
    // so no other lane or workgroup touches it)
    auto store_grad = [&](const int* sl, const long long* sin, const double* g) {
#pragma unroll
        for (int s = 0; s < kVjpSlots; ++s) {
            if (sl[s] >= 0) {
                const int q = sl[s] >> 16, J = sl[s] & 0xffff;
                const long long b = sin[q];
                if (b >= 0) grad[b + static_cast<long long>(J) * a.s_j] = g[s] + gtile[q * ldj + J];
            }
        }
    };
    // one pencil per lane group at a time: the contributions of its outputs,
    // one barrier, the sums per input position into gtile, one barrier.  WENO5
    // fetches each pencil's cotangents one pencil ahead (gl_n, gr_n hold the
    // first pencil's on entry); WENO7 (register-bound) loads them per cell.
    auto reconstruct = [&](const double* tile, const long long* sout, double gl_n, double gr_n) {
        for (int q0 = 0; q0 < a.npg; ++q0) {
            const int q1 = q0 + 1;
#pragma unroll 1
            for (int qq = q0; qq < q1; ++qq) {
                const int q = gq * a.npg + qq;
                double* const cb = cbuf + gq * cbs + (jj + 2 * R) * W;
                if (core) {
                    const long long ob = sout[q];
                    double gl = gl_n, gr = gr_n;
                    if constexpr (ALL) {
                        gl_n = gr_n = 0.0;
                        if (qq + 1 < a.npg) {
                            const long long obn = sout[q + 1];
                            if (obn >= 0) {
                                gl_n = gL[obn + j];
                                gr_n = gR[obn + j];
                            }
                        }
                    } else if (ob >= 0) {
                        gl = gL[ob + j];
                        gr = gR[ob + j];
                    }
                    double c[W];
                    if (ob >= 0) {
                        const LdsF64* bp = lds_ptr(beta_lds + jj * LDB);
                        asm volatile("" : "+v"(bp));
                        const BetaRef<NS> Bref{bp, NB, NR};
                        weno_vjp_cell<ORDER>(tile + q * ldj + jj, f, Bref, a.eps, gl, gr, c);
                    } else {
#pragma unroll
                        for (int o = 0; o < W; ++o) c[o] = 0.0;
                    }
#pragma unroll
                    for (int o = 0; o < W; ++o) cb[o] = c[o];
                }
            }
            __syncthreads();
            if (grp) {
                for (int qq = q0; qq < q1; ++qq) {
                    const int q = gq * a.npg + qq;
                    const double* const cb = cbuf + gq * cbs;
                    for (int m = jj; m < nJ; m += jn) {
                        // position m receives c[o] of output m - o, held in
                        // row m - o + 2R: entry (W-1-o)(W-1) from p (zero
                        // rows outside the outputs)
                        const double* const pm = cb + m * W + (W - 1);
                        double sum = 0.0;
#pragma unroll
                        for (int o = 0; o < W; ++o) sum += pm[(W - 1 - o) * (W - 1)];
                        gtile[q * ldj + m] = sum;
                    }
                }
            }
            __syncthreads();
        }
    };
    auto first_cot = [&](const long long* sout, double& gl, double& gr) {
        gl = gr = 0.0;
        if (ALL && core) {
            const long long ob = sout[gq * a.npg];
            if (ob >= 0) {
                gl = gL[ob + j];
                gr = gR[ob + j];
            }
        }
    };

    if constexpr (ALL) {
        if (tid < np) {
            plan(t_first, 0);
            if (t_first + nbx < t_end) plan(t_first + nbx, 1);
        }
        __syncthreads();
        {
            double r[kVjpSlots];
            int sl[kVjpSlots];
            slot_table(tid, sl);
            load_slots(sl, src, s_in[0], r);
            put_slots(sl, tile0, r);
        }
        __syncthreads();
        int k = 0, b = 0;
        for (int t = t_first; t < t_end; t += nbx) {
            const int k1 = k == 2 ? 0 : k + 1, k2 = k1 == 2 ? 0 : k1 + 1;
            const bool more = t + nbx < t_end;
            double* const cur = b ? tile1 : tile0;
            double* const nxt = b ? tile0 : tile1;
            double g[kVjpSlots], r[kVjpSlots];
            {
                int me = tid, sl[kVjpSlots];
                asm volatile("" : "+v"(me));
                slot_table(me, sl);
                load_slots(sl, grad, s_in[k], g);
                if (more) load_slots(sl, src, s_in[k1], r);
            }
            double gl0, gr0;
            first_cot(s_out[k], gl0, gr0);
            reconstruct(cur, s_out[k], gl0, gr0);
            // tile t's gradient out, tile t+1's input into the other buffer,
            // tile t+2's addresses into the ring entry tile t-1 used
            {
                int me = tid, sl[kVjpSlots];
                asm volatile("" : "+v"(me));
                slot_table(me, sl);
                store_grad(sl, s_in[k], g);
                if (more) put_slots(sl, nxt, r);
            }
            if (tid < np && t + 2 * nbx < t_end) plan(t + 2 * nbx, k2);
            __syncthreads();
            k = k1;
            b ^= 1;
        }
    } else {
        for (int t = t_first; t < t_end; t += nbx) {
            if (tid < np) plan(t, 0);
            __syncthreads();
            {
                double r[kVjpSlots];
                int sl[kVjpSlots];
#pragma unroll
                for (int s = 0; s < kVjpSlots; ++s) {
                    r[s] = 0.0;
                    sl[s] = slot(s);
                    if (sl[s] >= 0) {
                        const long long b = s_in[0][sl[s] >> 16];
                        if (b >= 0) r[s] = src[b + static_cast<long long>(sl[s] & 0xffff) * a.s_j];
                    }
                }
#pragma unroll
                for (int s = 0; s < kVjpSlots; ++s)
                    if (sl[s] >= 0) tile0[(sl[s] >> 16) * ldj + (sl[s] & 0xffff)] = r[s];
            }
            __syncthreads();
            reconstruct(tile0, s_out[0], 0.0, 0.0);
            {
                double g[kVjpSlots];
                int sl[kVjpSlots];
#pragma unroll
                for (int s = 0; s < kVjpSlots; ++s) {
                    g[s] = 0.0;
                    sl[s] = slot(s);
                    if (sl[s] >= 0) {
                        const long long b = s_in[0][sl[s] >> 16];
                        if (b >= 0) g[s] = grad[b + static_cast<long long>(sl[s] & 0xffff) * a.s_j];
                    }
                }
#pragma unroll
                for (int s = 0; s < kVjpSlots; ++s) {
                    if (sl[s] >= 0) {
                        const int q = sl[s] >> 16, J = sl[s] & 0xffff;
                        const long long b = s_in[0][q];
                        if (b >= 0) grad[b + static_cast<long long>(J) * a.s_j] = g[s] + gtile[q * ldj + J];
                    }
                }
            }
            __syncthreads();
        }
    }
}

hipError_t launch_vjp(TileArgs a, int order, double* grad, hipStream_t st)
{
    // tile elements are packed as (pencil << 16) | row
    if (a.ldj > 0xffff || a.np > kVjpMaxNP || a.slots > kVjpSlots) return hipErrorInvalidConfiguration;
    const size_t lds = ((order == 5 ? 3 : 2) * static_cast<size_t>(a.np) * a.ldj +
                        static_cast<size_t>(a.groups) * order * (a.jn + 2 * (order - 1)) +
                        static_cast<size_t>(order == 5 ? BetaTraits<5>::LD : BetaTraits<7>::LD) * a.jn) *
                       sizeof(double);
    const void* kern = order == 5 ? reinterpret_cast<const void*>(vjp_kernel<5>)
                                  : reinterpret_cast<const void*>(vjp_kernel<7>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, a.threads, lds) != hipSuccess || occ < 1) occ = 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const int ntile = a.tiles_per_var * a.nv;
    int bpc = std::min(cus * occ, ntile);
    bpc = (bpc + 7) & ~7;
    a.bpc = bpc;
    if (order == 5)
        hipLaunchKernelGGL(vjp_kernel<5>, dim3(bpc), dim3(a.threads), lds, st, a, grad);
    else
        hipLaunchKernelGGL(vjp_kernel<7>, dim3(bpc), dim3(a.threads), lds, st, a, grad);
    return hipGetLastError();
}

// ---- WENO5-JS on whole-row tiles ------------------------------------------
//
// The adjoint of cell5's shared-reciprocal form.  Forward, with
// D_i = v(i+1) - v(i) over the window v0..v4: smoothness b_r (b_0 on D3, D2;
// b_1 on D2, D1; b_2 on D1, D0), m_r = prod_{s != r} b_s, u_r = m_r^2,
// g^X_r = e^X_r D + f^X_r D' (d^X_r times candidate r's deviation from v2),
// oX = v2 + N_X / S_X with N_X = sum_r u_r g^X_r, S_X = sum_r d^X_r u_r
// (X = L, R), both sides normalised by one reciprocal of S_L S_R.
// Backward, for cotangents gX of oX, with A_X = gX / S_X and
// q_X = N_X / S_X = oX - v2:
//   d/dg^X_r = A_X u_r,
//   d/du_r   = A_L (g^L_r - q_L d^L_r) + A_R (g^R_r - q_R d^R_r) =: mu_r,
//   d/db_0   = 2 (m_1 mu_1 b_2 + m_2 mu_2 b_1), and cyclically,
// and the differences collect both paths (the candidate rows and the
// smoothness quadratic forms).  About 125 FP64 operations per face pair
// and one reciprocal.  Returns false when the shared product T = S_L S_R or
// its reciprocal leaves the normal range; the lane then takes
// vjp_cell5_slow.  With T normal nothing in the backward overflows for
// finite cotangents: S_L and S_R weight the same u_r by the linear weights,
// so each is within a small factor of sqrt(T), above 1e-154.
__device__ __forceinline__ bool vjp_cell5(double v0, double v1, double v2, double v3, double v4, const Coef5& k,
                                          const double (&blo2)[3], double eps, double gl, double gr, double (&c)[5])
{
    const double D3 = v4 - v3, D2 = v3 - v2, D1 = v2 - v1, D0 = v1 - v0;
    // b_r = B_hi hi^2 + B_x hi lo + B_lo lo^2 + eps over (hi, lo) = (D3, D2),
    // (D2, D1), (D1, D0); with p_r = B_hi hi + B_x lo, b_r = hi p_r + B_lo lo^2
    // + eps, db_r/dhi = p_r + B_hi hi, db_r/dlo = B_x hi + 2 B_lo lo
    const double p0 = fma(k.b[0], D3, k.b[1] * D2), p1 = fma(k.b[3], D2, k.b[4] * D1),
                 p2 = fma(k.b[6], D1, k.b[7] * D0);
    const double b0 = fma(D3, p0, fma(k.b[2], D2 * D2, eps));
    const double b1 = fma(D2, p1, fma(k.b[5], D1 * D1, eps));
    const double b2 = fma(D1, p2, fma(k.b[8], D0 * D0, eps));
    const double m0 = b1 * b2, m1 = b0 * b2, m2 = b0 * b1;
    const double u0 = m0 * m0, u1 = m1 * m1, u2 = m2 * m2;
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
    for (int it = 0; it < 2; ++it) r = fma(r, fma(-T, r, 1.0), r);
    const double iL = SR * r, iR = SL * r;   // 1 / S_L, 1 / S_R
    const double qL = NL * iL, qR = NR * iR;
    const double AL = gl * iL, AR = gr * iR;
    const double AL2 = AL + AL, AR2 = AR + AR;
    // 2 mu_r, then m_r * 2 mu_r
    const double n0 = m0 * fma(AL2, fma(-qL, k.dL[0], gL0), AR2 * fma(-qR, k.dR[0], gR0));
    const double n1 = m1 * fma(AL2, fma(-qL, k.dL[1], gL1), AR2 * fma(-qR, k.dR[1], gR1));
    const double n2 = m2 * fma(AL2, fma(-qL, k.dL[2], gL2), AR2 * fma(-qR, k.dR[2], gR2));
    // d/db_r
    const double E0 = fma(n1, b2, n2 * b1), E1 = fma(n0, b2, n2 * b0), E2 = fma(n0, b1, n1 * b0);
    // d/dg^X_r
    const double aL0 = AL * u0, aL1 = AL * u1, aL2 = AL * u2;
    const double aR0 = AR * u0, aR1 = AR * u1, aR2 = AR * u2;
    double gD3 = fma(aL0, k.eL[0], aR0 * k.eR[0]);
    gD3 = fma(E0, fma(k.b[0], D3, p0), gD3);
    double gD2 = fma(aL0, k.fL[0], fma(aR0, k.fR[0], fma(aL1, k.eL[1], aR1 * k.eR[1])));
    gD2 = fma(E0, fma(blo2[0], D2, k.b[1] * D3), gD2);
    gD2 = fma(E1, fma(k.b[3], D2, p1), gD2);
    double gD1 = fma(aL1, k.fL[1], fma(aR1, k.fR[1], fma(aL2, k.eL[2], aR2 * k.eR[2])));
    gD1 = fma(E1, fma(blo2[1], D1, k.b[4] * D2), gD1);
    gD1 = fma(E2, fma(k.b[6], D1, p2), gD1);
    double gD0 = fma(aL2, k.fL[2], aR2 * k.fR[2]);
    gD0 = fma(E2, fma(blo2[2], D0, k.b[7] * D1), gD0);
    c[0] = -gD0;
    c[1] = gD0 - gD1;
    c[2] = (gD1 - gD2) + (gl + gr);
    c[3] = gD2 - gD3;
    c[4] = gD3;
    return (static_cast<int>(pos_normal(T)) & static_cast<int>(pos_normal(r))) != 0;
}

// The same contributions in the per-stencil form with IEEE divisions
// (alpha_r = d_r / b_r^2, w_r = alpha_r / sum alpha, candidate deviation
// g_r / d_r): d oX / d g^X_r = 1 / (b_r^2 sum alpha), d oX / d b_r =
// -2 w_r (g_r / d_r - q_X) / b_r.  For lanes whose shared product leaves the
// normal range; carries NaN and Inf as the forward does.  Always inlined (an
// out-of-line call would keep the caller's coefficient registers in scratch).
__device__ __forceinline__ void vjp_cell5_slow(double v0, double v1, double v2, double v3, double v4, const Coef5& k,
                                               double eps, double gl, double gr, double (&c)[5])
{
    const double D3 = v4 - v3, D2 = v3 - v2, D1 = v2 - v1, D0 = v1 - v0;
    const double b0 = fma(k.b[0], D3 * D3, fma(k.b[1], D3 * D2, fma(k.b[2], D2 * D2, eps)));
    const double b1 = fma(k.b[3], D2 * D2, fma(k.b[4], D2 * D1, fma(k.b[5], D1 * D1, eps)));
    const double b2 = fma(k.b[6], D1 * D1, fma(k.b[7], D1 * D0, fma(k.b[8], D0 * D0, eps)));
    const double y0 = 1.0 / (b0 * b0), y1 = 1.0 / (b1 * b1), y2 = 1.0 / (b2 * b2);
    double E0 = 0.0, E1 = 0.0, E2 = 0.0;
    // one side: coefficients by value, so that nothing takes the address of
    // the caller's coefficient registers
    auto side = [&](double gx, double d0, double d1, double d2, double g0, double g1, double g2, double& a0,
                    double& a1, double& a2) {
        const double al0 = d0 * y0, al1 = d1 * y1, al2 = d2 * y2;
        const double S = al0 + al1 + al2;
        const double e0 = g0 / d0, e1 = g1 / d1, e2 = g2 / d2;   // candidate deviations from v2
        const double qx = (al0 * e0 + al1 * e1 + al2 * e2) / S;
        a0 = gx * y0 / S;
        a1 = gx * y1 / S;
        a2 = gx * y2 / S;
        E0 -= 2.0 * gx * (al0 / S) * (e0 - qx) / b0;
        E1 -= 2.0 * gx * (al1 / S) * (e1 - qx) / b1;
        E2 -= 2.0 * gx * (al2 / S) * (e2 - qx) / b2;
    };
    double ga[2][3];
    side(gl, k.dL[0], k.dL[1], k.dL[2], fma(k.eL[0], D3, k.fL[0] * D2), fma(k.eL[1], D2, k.fL[1] * D1),
         fma(k.eL[2], D1, k.fL[2] * D0), ga[0][0], ga[0][1], ga[0][2]);
    side(gr, k.dR[0], k.dR[1], k.dR[2], fma(k.eR[0], D3, k.fR[0] * D2), fma(k.eR[1], D2, k.fR[1] * D1),
         fma(k.eR[2], D1, k.fR[2] * D0), ga[1][0], ga[1][1], ga[1][2]);
    double gD3 = ga[0][0] * k.eL[0] + ga[1][0] * k.eR[0];
    double gD2 = ga[0][0] * k.fL[0] + ga[1][0] * k.fR[0] + ga[0][1] * k.eL[1] + ga[1][1] * k.eR[1];
    double gD1 = ga[0][1] * k.fL[1] + ga[1][1] * k.fR[1] + ga[0][2] * k.eL[2] + ga[1][2] * k.eR[2];
    double gD0 = ga[0][2] * k.fL[2] + ga[1][2] * k.fR[2];
    gD3 += E0 * (2.0 * k.b[0] * D3 + k.b[1] * D2);
    gD2 += E0 * (k.b[1] * D3 + 2.0 * k.b[2] * D2) + E1 * (2.0 * k.b[3] * D2 + k.b[4] * D1);
    gD1 += E1 * (k.b[4] * D2 + 2.0 * k.b[5] * D1) + E2 * (2.0 * k.b[6] * D1 + k.b[7] * D0);
    gD0 += E2 * (k.b[7] * D1 + 2.0 * k.b[8] * D0);
    c[0] = -gD0;
    c[1] = gD0 - gD1;
    c[2] = (gD1 - gD2) + (gl + gr);
    c[3] = gD2 - gD3;
    c[4] = gD3;
}

namespace {

// Input position p (lane p of the wave) collects c[2] of its own face, c[3],
// c[4] of the faces p-1, p-2 and c[1], c[0] of the faces p+1, p+2: four
// lane shifts (the wave's two end lanes on either side get partial sums).
// Call it with every lane of the wave active: a lane shift reads 0 from a
// lane that is switched off.
__device__ __forceinline__ double lane_sum(const double (&c)[5])
{
    const double A = c[3] + from_prev_lane(c[4]);
    const double B = c[1] + from_next_lane(c[0]);
    return c[2] + (from_prev_lane(A) + from_next_lane(B));
}

constexpr int kTileSlots = 28;      // producer registers per array (doubles per lane)
constexpr int kTileMaxNP = 16;      // pencils per tile
constexpr int kTileConsumers = 7;   // consumer waves (at most)
constexpr int kSegment = 60;        // inputs owned per consumer wave
constexpr int kMinConsumers = 7;    // consumer waves (those past the last segment share the gradient write-back)
static_assert(64 * kTileSlots <= 2048, "span element indices must stay exact in FP32");

struct Vjp5Args {
    TileArgs a;
    double* grad;
    int lo;     // input J of a tile row's first element (the core's first face - 2)
    int nJ;     // inputs per row: the core and two on either side
    int rows;   // rows laid end to end per pass
};

// One workgroup = up to 7 consumer waves + 1 producer wave, persistent over
// tiles of np pencils (of one plane) x whole rows of inputs J = lo ..
// lo+nJ-1, i.e. every input the core's faces read.  LDS: two input tiles
// and two gradient tiles, [pencil][J] with an odd row stride.  One barrier
// per tile.  The consumer's 27 coefficients and the adjoint cell take about
// 150 VGPRs (at 128 the coefficients spill and are reloaded every pass):
// 2 waves per SIMD, one workgroup per CU.
//   producer : loads tile k+1's input and commits it to LDS (it never
//              stores to global memory).
//   consumers: the inputs of R rows, end to end, in segments of 60 per
//              wave: wave w, lane i takes input position 60 w + i - 2 and
//              the face there (the two inputs at either end of a row have
//              no face).  Each lane forms its face's five contributions;
//              lanes 2..61 own their input and sum the contributions of
//              the faces two either side by four lane shifts, then write
//              the sum to the gradient tile.  No atomics, the same order of
//              additions in every run.  Lanes 0, 1, 62 and 63 repeat the
//              neighbouring segments' faces.  The cotangents of the next
//              pass (or of the next tile's first pass) are loaded one pass
//              ahead.  In tile k's interval the consumer waves also add
//              tile k-1's gradient tile to gv, in the producer's coalesced
//              slot mapping split over the waves: the gv loads are issued
//              before tile k's passes, the stores after them.  Every
//              element of a gradient tile is written in its tile's interval
//              (each input has one owner lane), so the tiles are never
//              cleared.
__global__ void __attribute__((amdgpu_flat_work_group_size(64, 64 * (kTileConsumers + 1)), amdgpu_waves_per_eu(2)))
vjp5_tile_kernel(const Vjp5Args va)
{
    const TileArgs& a = va.a;
    extern __shared__ __align__(16) double lds[];
    const int np = a.np, nJ = va.nJ, ld = a.ldj, lo = va.lo;
    double* const vb0 = lds;
    double* const vb1 = lds + np * ld;
    double* const gb0 = lds + 2 * np * ld;
    double* const gb1 = lds + 3 * np * ld;
    const int tid = threadIdx.x;
    const int nwave = blockDim.x >> 6;
    const bool producer = __builtin_amdgcn_readfirstlane(tid >> 6) == nwave - 1;

    // tiles t_first, t_first + tstep, ... < t_end; strided rows share input
    // lines with the neighbouring tiles: one contiguous range per XCD
    const int ntile = a.tiles_per_var * a.nv;
    int t_first = blockIdx.x, t_end = ntile, tstep = gridDim.x;
    if (a.s_j != 1 && (gridDim.x & 7) == 0) {
        const int xcd = blockIdx.x & 7, rank = blockIdx.x >> 3;
        const int per = ntile >> 3, rem = ntile & 7;
        t_first = xcd * per + min(xcd, rem) + rank;
        t_end = (xcd + 1) * per + min(xcd + 1, rem);
        tstep = gridDim.x >> 3;
    }
    if (t_first >= t_end) return;

    if (producer) {
        double r[kTileSlots];
        tile_commit(a, nJ, ld, tile_issue(a, lo, nJ, t_first, r), r, vb0);
        __syncthreads();
        int k = 0;
        for (int t = t_first; t < t_end; t += tstep, ++k) {
            const int tn = t + tstep;
            if (tn < t_end) tile_commit(a, nJ, ld, tile_issue(a, lo, nJ, tn, r), r, (k & 1) ? vb0 : vb1);
            __syncthreads();
        }
        return;
    }

    const int i = tid & 63, cw = tid >> 6, nc = nwave - 1;
    const int R = va.rows, npos = R * nJ;
    const int pos = kSegment * cw + i - 2;
    const int pc = min(max(pos, 0), npos - 1);
    const int rr = pc / nJ, ii = pc - rr * nJ;
    const bool inrange = pos >= 0 && pos < npos;
    const bool own = inrange && i >= 2 && i < kSegment + 2;
    const bool face = inrange && ii >= 2 && ii < nJ - 2;
    const int j = lo - a.J0f + ii;   // output index of the face at this input
    Coef5 k5;
    {
        int jc = j - 3;
        jc = jc < 0 ? 0 : (jc > a.c.nc - 1 ? a.c.nc - 1 : jc);
        load_coef5(a.c, jc, k5);
    }
    const double blo2[3] = {k5.b[2] + k5.b[2], k5.b[5] + k5.b[5], k5.b[8] + k5.b[8]};   // 2 B_lo
    const int col = min(max(ii, 2), nJ - 3);
    const long long o_a = a.a_is_k ? a.o_k : a.o_l, o_b = a.a_is_k ? a.o_l : a.o_k;
    const double eps = a.eps;
    // cotangents through buffer loads at 32-bit byte offsets (the face
    // arrays are below 4 GiB, checked by the launcher); offsets past the end
    // read 0, so lanes without a face in a pass load nothing and compute
    // zero contributions
    const int fbytes = static_cast<int>(static_cast<unsigned long long>(a.o_i) * a.nv * sizeof(double));
    const __amdgpu_buffer_rsrc_t rgL = __builtin_amdgcn_make_buffer_rsrc(a.outL, 0, fbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rgR = __builtin_amdgcn_make_buffer_rsrc(a.outR, 0, fbytes, 0x00020000);
    const unsigned pass_step = static_cast<unsigned>(R * o_a * sizeof(double));
    // byte offset of this lane's face in row rr of tile t (first pass), and the tile's pencil count
    auto cot_off = [&](int t, int& cnt) {
        int iv, pa0, pb;
        plane_tile(a, t, iv, pa0, pb, cnt);
        return static_cast<unsigned>((pa0 * o_a + pb * o_b + iv * a.o_i + j + rr * o_a) * sizeof(double));
    };
    auto cot = [&](__amdgpu_buffer_rsrc_t rs, unsigned off, bool valid) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, valid ? off : 0xffffffffu, 0, 0));
    };
    // this lane's gradient slots (wave cw takes slots cw, cw + nc, ... of the
    // producer's mapping), the same in every tile: pencil q in the high 16
    // bits and LDS index q * ld + J in the low 16 (no slot: pencil 0x7fff,
    // which no tile has, and index 0), and the element offset from the
    // tile's first element
    constexpr int PER = (kTileSlots + kMinConsumers - 1) / kMinConsumers;   // gradient slots per consumer lane
    int sl_lds[PER];
    unsigned sl_off[PER];
    {
        const int G = 64 / np, sq = i % np, J0 = i / np;
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int s = cw + nc * m;
            int q, J;
            unsigned off;
            if (a.s_j == 1) {
                const int e = i + 64 * s, sa = static_cast<int>(a.s_a);
                q = e / sa;
                J = e - q * sa;
                off = static_cast<unsigned>(e);
            } else {
                q = J0 < G ? sq : np;   // lanes past np G idle
                J = J0 + G * s;
                off = static_cast<unsigned>(sq + J * a.s_j);
            }
            const bool valid = s < kTileSlots && J < nJ && q < np;
            sl_lds[m] = valid ? (q << 16) | (q * ld + J) : 0x7fff << 16;
            sl_off[m] = off;
        }
    }
    double gv[PER];
    auto grad_base = [&](int t, int& cnt) {
        int iv, pa0, pb;
        plane_tile(a, t, iv, pa0, pb, cnt);
        return va.grad + (a.src_off + pa0 * a.s_a + pb * a.s_b + iv * a.s_i + static_cast<long long>(lo) * a.s_j);
    };
    using GD = __attribute__((address_space(1))) double;
    int cnt;
    unsigned off = cot_off(t_first, cnt);
    double gl = cot(rgL, off, face && rr < cnt), gr = cot(rgR, off, face && rr < cnt);
    __syncthreads();
    int k = 0, tp = -1;
    for (int t = t_first; t < t_end; t += tstep, ++k) {
        const double* const cur = (k & 1) ? vb1 : vb0;
        double* const gt = (k & 1) ? gb1 : gb0;
        double* const gprev = (k & 1) ? gb0 : gb1;
        // gv over tile k-1's rows, in flight during tile k's passes
        int cntp = 0;
        GD* const bp = reinterpret_cast<GD*>(reinterpret_cast<unsigned long long>(tp >= 0 ? grad_base(tp, cntp) : va.grad));
        if (tp >= 0) {
#pragma unroll
            for (int m = 0; m < PER; ++m)
                if ((sl_lds[m] >> 16) < cntp) gv[m] = bp[sl_off[m]];
        }
        const int tn = t + tstep;
        int cntn = 0;
        const unsigned offn = tn < t_end ? cot_off(tn, cntn) : 0u;
        const unsigned off0 = off;
        bool allok = true;
        for (int q0 = 0; q0 < cnt; q0 += R) {
            const int q = q0 + rr;
            // cotangents one pass ahead: this tile's next pass, or the next tile's first
            double gln, grn;
            if (q0 + R < cnt) {
                gln = cot(rgL, off + pass_step, face && q + R < cnt);
                grn = cot(rgR, off + pass_step, face && q + R < cnt);
            } else {
                gln = cot(rgL, offn, face && rr < cntn);
                grn = cot(rgR, offn, face && rr < cntn);
            }
            const double* const rw = cur + min(q, cnt - 1) * ld + col;
            const double v0 = rw[-2], v1 = rw[-1], v2 = rw[0], v3 = rw[1], v4 = rw[2];
            double c[5];
            // lanes without a face have zero cotangents, so a cell in range
            // gives them exactly zero contributions
            allok &= vjp_cell5(v0, v1, v2, v3, v4, k5, blo2, eps, gl, gr, c);
            // the lane shifts run with every lane of the wave active
            const double sum = lane_sum(c);
            if (own && q < cnt) gt[q * ld + ii] = sum;
            gl = gln;
            gr = grn;
            off += pass_step;
        }
        // Rare: a lane's shared product left the normal range.  The wave
        // redoes the tile's passes (it owns the positions it wrote): the
        // slow form for faces, zeros for the lanes without one.
        if (__builtin_amdgcn_ballot_w64(!allok) != 0) {
            unsigned offr = off0;
            for (int q0 = 0; q0 < cnt; q0 += R) {
                const int q = q0 + rr;
                const bool live = face && q < cnt;
                const double glr = cot(rgL, offr, live), grr = cot(rgR, offr, live);
                const double* const rw = cur + min(q, cnt - 1) * ld + col;
                double x0 = rw[-2], x1 = rw[-1], x2 = rw[0], x3 = rw[1], x4 = rw[2];
                double c[5];
                const bool ok = vjp_cell5(x0, x1, x2, x3, x4, k5, blo2, eps, glr, grr, c);
                asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4));
                if (!ok) {
                    if (live) {
                        vjp_cell5_slow(x0, x1, x2, x3, x4, k5, eps, glr, grr, c);
                    } else {
#pragma unroll
                        for (int o = 0; o < 5; ++o) c[o] = 0.0;
                    }
                }
                const double sum = lane_sum(c);
                if (own && q < cnt) gt[q * ld + ii] = sum;
                offr += pass_step;
            }
        }
        if (tp >= 0) {
            // every (pencil < cnt, J < nJ) of a gradient tile has one owner
            // lane that writes it in its tile's interval: nothing to clear
            double gs[PER];
#pragma unroll
            for (int m = 0; m < PER; ++m) gs[m] = gprev[sl_lds[m] & 0xffff];
#pragma unroll
            for (int m = 0; m < PER; ++m)
                if ((sl_lds[m] >> 16) < cntp) bp[sl_off[m]] = gv[m] + gs[m];
        }
        tp = t;
        off = offn;
        cnt = cntn;
        __syncthreads();
    }
    // the last tile's gradient, after its barrier
    {
        int cntp = 0;
        GD* const bp = reinterpret_cast<GD*>(reinterpret_cast<unsigned long long>(grad_base(tp, cntp)));
        double* const gprev = (k & 1) ? gb0 : gb1;
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            if ((sl_lds[m] >> 16) < cntp) bp[sl_off[m]] = bp[sl_off[m]] + gprev[sl_lds[m] & 0xffff];
        }
    }
}

}  // namespace

// Rows of at most kTileConsumers * kSegment inputs (core + 4) whose tiles
// fit the producer's registers; returns hipErrorNotSupported otherwise (the
// caller then takes vjp_kernel).
hipError_t launch_vjp5_tiles(TileArgs a, double* grad, hipStream_t st)
{
    Vjp5Args va;
    va.lo = a.j_beg - 2 + a.J0f;
    va.nJ = (a.j_end - a.j_beg + 1) + 4;
    va.grad = grad;
    if (va.nJ > kTileConsumers * kSegment || (a.s_j != 1 && a.s_a != 1)) return hipErrorNotSupported;
    // 32-bit byte offsets into the face arrays, 32-bit element offsets of a
    // tile's gradient slots
    if (static_cast<unsigned long long>(a.o_i) * a.nv * sizeof(double) > 0xffffffffull ||
        static_cast<unsigned long long>(va.nJ + 1) * static_cast<unsigned long long>(a.s_j) + 64 > 0xffffffffull)
        return hipErrorNotSupported;
    // pencils per tile and rows per pass: the tile's pencils take
    // ceil(np / R) passes of R rows over max(nseg, kMinConsumers) waves;
    // take the shape with the most face lanes per lane-pass (then the
    // larger tile) among those whose tile fits the producer's registers
    const int nf = va.nJ - 4;
    double best = 0.0;
    a.np = 0;
    for (int np = 1; np <= kTileMaxNP; ++np) {
        const bool rows = a.s_j == 1;
        const long long need = rows ? ((np - 1) * a.s_a + va.nJ + 63) / 64
                                    : (va.nJ + 64 / np - 1) / (64 / np);
        if (need > kTileSlots) continue;
        const int R = std::max(1, std::min(np, kTileConsumers * kSegment / va.nJ));
        const int waves = std::max((R * va.nJ + kSegment - 1) / kSegment, kMinConsumers);
        const double use = double(np) * nf / (double((np + R - 1) / R) * 64.0 * waves);
        if (use > best + 1e-9 || (use > best - 1e-9 && np > a.np)) {
            best = use;
            a.np = np;
            va.rows = R;
        }
    }
    if (a.np < 1) return hipErrorNotSupported;
    const int nseg = (va.rows * va.nJ + kSegment - 1) / kSegment;
    // waves past the last segment own no input and share the gradient write-back
    a.threads = 64 * (std::max(nseg, kMinConsumers) + 1);
    a.ldj = va.nJ | 1;
    a.chunks = (a.na + a.np - 1) / a.np;   // tiles per plane
    a.tiles_per_var = a.chunks * a.nb;
    va.a = a;
    const size_t lds = 4 * static_cast<size_t>(a.np) * a.ldj * sizeof(double);
    const void* kern = reinterpret_cast<const void*>(vjp5_tile_kernel);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, a.threads, lds) != hipSuccess || occ < 1) occ = 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const long long ntile = static_cast<long long>(a.tiles_per_var) * a.nv;
    long long grid = std::min<long long>(ntile, 1LL * cus * occ);
    if (a.s_j != 1 && grid >= 8) grid &= ~7LL;   // whole XCD ranges
    hipLaunchKernelGGL(vjp5_tile_kernel, dim3(static_cast<unsigned>(std::max<long long>(1, grid))), dim3(a.threads),
                       lds, st, va);
    return hipGetLastError();
}

int vjp_max_threads() { return kVjpThreads; }
int vjp_max_slots() { return kVjpSlots; }
int vjp_max_np() { return kVjpMaxNP; }

}  // namespace wm
