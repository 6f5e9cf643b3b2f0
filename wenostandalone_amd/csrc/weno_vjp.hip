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

int vjp_max_threads() { return kVjpThreads; }
int vjp_max_slots() { return kVjpSlots; }
int vjp_max_np() { return kVjpMaxNP; }

}  // namespace wm
