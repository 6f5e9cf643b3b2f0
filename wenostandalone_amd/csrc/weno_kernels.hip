// HIP kernels for gfx950: the general sweep kernel behind WENO3/5/7 (every
// weight branch), the fused or standalone monotonicity-preserving limiter,
// in FP64, FP32 and BF16-state storage (the reshape and the order-1 copy run
// on weno_permute.hip).  WENO5-JS launches that
// write both sides over the full reconstructed range (rows of at least 96
// outputs, aligned arrays) run on the producer/consumer kernel of
// weno_sweep5.hip instead (launch_tile hands them over).
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
#include "weno_device.h"

namespace wm {

constexpr int kMaxThreads = 1024;
constexpr int kMaxNP = 64;
constexpr int kMaxSlots = 6;

template <int ORDER>
struct OrderTraits {
    static constexpr int R = ORDER / 2;
};

// WENO7 FP64 needs more than 128 VGPRs: its blocks are 512 threads and its
// waves may use up to 256 VGPRs; FP32/BF16 WENO7 is held to 128 VGPRs (two
// 512-thread blocks per CU); everything else runs four waves per SIMD
// (1024-thread blocks, 128 VGPRs).
template <int OP, int ORDER, typename TIN = double>
struct BlockTraits {
    static constexpr int kThreads = (OP == OP_WENO && ORDER == 7) ? 512 : kMaxThreads;
    // second __launch_bounds__ argument: minimum waves per SIMD (4 -> 128 VGPRs)
    static constexpr int kMinWaves = (OP == OP_WENO && ORDER == 7 && sizeof(TIN) < sizeof(double)) ? 4 : 1;
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
template <int OP, int ORDER, int MODE, bool MP, typename TIN, typename TOUT, bool CREG>
__global__ void __launch_bounds__((BlockTraits<OP, ORDER, TIN>::kThreads), (BlockTraits<OP, ORDER, TIN>::kMinWaves))
    sweep_kernel(const TileArgs a, const int nt)
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
    constexpr int LDB = BetaTraits<OB>::LD;
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
    BetaRegs<OB> breg;
    const int jj = tid - gq * a.jn;
    if constexpr (OP == OP_WENO) {
        if (core) {
            int jc = j - 3;
            jc = jc < 0 ? 0 : (jc > a.c.nc - 1 ? a.c.nc - 1 : jc);
            load_cf(a.c, jc, f);
            if constexpr (CREG) breg.load(a.c, jc);
            if (!CREG && gq == 0) {
#pragma unroll
                for (int t = 0; t < NB; ++t) {
                    const int r = t / NS, s = t - (t / NS) * NS;
                    beta_lds[jj * LDB + t] = a.c.beta[jc + a.c.nc * (r + 5 * s)];
                }
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    beta_lds[jj * LDB + NB + r] = a.c.dL[r + 5 * jc];
                    beta_lds[jj * LDB + NB + NR + r] = a.c.dR[r + 5 * jc];
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
        // prefetch: the next tile's loads are in flight during this compute
        if (more && a.prefetch) issue((k + 1) % 3);
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
                    const LdsF64* bp = lds_ptr(beta_lds + jj * LDB);
                    if constexpr (!CREG) asm volatile("" : "+v"(bp));
                    const BetaRef<NS> Bl{bp, NB, NR};
                    C v[W];
#pragma unroll
                    for (int o = 0; o < W; ++o) v[o] = row[o - R];
                    double oL, oR;
                    const C* vc = v + (R - RW);
                    auto cell = [&](const auto& B) {
                        if constexpr (ORDER == 3)
                            weno3<MODE, C>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                        else if constexpr (ORDER == 5)
                            weno5<MODE, C>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                        else
                            weno7<MODE, C>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                    };
                    if constexpr (CREG)
                        cell(breg);
                    else
                        cell(Bl);
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
        if (more && a.prefetch) commit(nxt);
        __syncthreads();
        if (more && !a.prefetch) {
            issue((k + 1) % 3);
            commit(nxt);
            __syncthreads();
        }
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

template <int OP, int ORDER, int MODE, bool MP, typename TIN, typename TOUT, bool CREG = false>
static hipError_t go(TileArgs a, const Launch& l, hipStream_t st)
{
    using C = typename CT<TIN>::type;
    auto kern = sweep_kernel<OP, ORDER, MODE, MP, TIN, TOUT, CREG>;
    constexpr int NB = (OP == OP_WENO && !CREG) ? BetaTraits<ORDER>::LD : 0;
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
    if (l.one_tile) bpc = ntile;         // one tile per workgroup: no persistence
    bpc = (bpc + 7) & ~7;
    a.bpc = bpc;
    hipLaunchKernelGGL(kern, dim3(bpc * a.chunks), dim3(a.threads), lds, st, a, l.nontemporal);
    return hipGetLastError();
}

template <typename TIN, typename TOUT>
static hipError_t go_weno(const TileArgs& a, const Launch& l, hipStream_t st)
{
#define WM_CASE(O, M)                                                             \
    if (l.order == O && l.mode == M) {                                            \
        if ((O == 5 || (O == 7 && sizeof(TIN) == 8)) && l.creg)                  \
            return l.mp ? go<OP_WENO, O, M, true, TIN, TOUT, true>(a, l, st)      \
                        : go<OP_WENO, O, M, false, TIN, TOUT, true>(a, l, st);    \
        return l.mp ? go<OP_WENO, O, M, true, TIN, TOUT>(a, l, st)                \
                    : go<OP_WENO, O, M, false, TIN, TOUT>(a, l, st);              \
    }
    WM_CASE(3, MODE_NONE)
    WM_CASE(3, MODE_JS)
    WM_CASE(3, MODE_MAPPED)
    WM_CASE(3, MODE_WENOZ)
    WM_CASE(5, MODE_NONE)
    WM_CASE(5, MODE_JS)
    WM_CASE(5, MODE_TENO)
    WM_CASE(5, MODE_MAPPED)
    WM_CASE(5, MODE_WENOZ)
    WM_CASE(7, MODE_NONE)
    WM_CASE(7, MODE_JS)
    WM_CASE(7, MODE_TENO)
    WM_CASE(7, MODE_MAPPED)
    WM_CASE(7, MODE_WENOZ)
#undef WM_CASE
    return hipErrorInvalidValue;
}

int max_threads(int op, int order);

hipError_t launch_tile(const TileArgs& a, const Launch& l, hipStream_t st)
{
    if (a.tiles_per_var <= 0 || a.nv <= 0) return hipSuccess;
    if (l.fast5 > 0) return launch_sweep5(a, l.order, l.mode, l.fast5 - 1, l.dtype, l.mp, st);
    if (a.np > kMaxNP || a.slots > kMaxSlots || a.threads > max_threads(l.op, l.order) || a.threads < 64 ||
        a.ldj > 0xffff)
        return hipErrorInvalidConfiguration;
    // the reshape and the order-1 copy run on launch_permute
    switch (l.op) {
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
