// Batched reconstruction: many independent blocks (sub-domains) of different
// sizes, variable counts and WENO orders in one persistent launch per sweep
// direction (no reference counterpart: the harness reconstructs one box).
//
// Work items are (block, tile) pairs, ordered by block cost; workgroups pull
// items from an atomic counter until none are left, so the launch drains on
// its own and large blocks do not leave a tail.  Each item loads its tile of
// pencils into LDS exactly like the single-box sweep kernel (direction-aware,
// coalesced), then its lanes reconstruct with the block's own order, weights
// and coefficients.  The counters are reset by a one-thread kernel (no
// memset), so a whole x/y/z sweep can be captured in one hipGraph.
#include "weno_device.h"
#include "weno_cells.h"
#include "weno_capi_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>
#include <type_traits>

namespace wm {

struct BatchDesc {
    TileArgs a;
    int order;
    int mode;
    int item0;
    int pad;
};

constexpr int kBatchThreads = 512;
constexpr int kBatchMaxNP = 256;
constexpr int kBatchLdsDescs = 1024;

// reconstruct the tile in LDS with order ORDER (after the tile barrier)
typedef __attribute__((address_space(3))) double LdsMutF64;
typedef __attribute__((address_space(3))) const long long LdsI64;
// the descriptors come from memory, so their pointers are generic: casting
// them to global keeps loads and stores off the flat path (flat accesses
// count against both wait counters, so LDS waits would include them)
typedef __attribute__((address_space(1))) double GlbF64;
typedef __attribute__((address_space(1))) const double GlbCF64;

// Geometry of a block's items along the sweep: input rows [lo, lo+nJ) land in
// LDS, lanes cover outputs [jlo, jhi], the core (reconstructed) range is
// [c0, c1].
struct ItemGeom {
    int lo, nJ, jlo, jhi, c0, c1;
    bool has_core;
};

__device__ __forceinline__ ItemGeom item_geom(const TileArgs& a, int R)
{
    ItemGeom g;
    g.c0 = a.j_beg;
    g.c1 = a.j_end;
    g.has_core = g.c0 <= g.c1;
    int hi;
    if (a.copy_halo) {
        g.lo = min(a.J0f, g.c0 + a.J0f - R);
        hi = max(a.out_len - 1 + a.J0f, g.c1 + a.J0f + R);
        g.jlo = 0;
        g.jhi = a.out_len - 1;
    } else {
        g.lo = g.c0 + a.J0f - R;
        hi = g.c1 + a.J0f + R;
        g.jlo = g.c0;
        g.jhi = g.c1;
    }
    g.nJ = hi - g.lo + 1;
    return g;
}

// Item t of a block: pencil addresses into s_in/s_out (threads < np), a
// barrier, then the tile into LDS.  Element e of the tile is (pencil q, row
// J): e = q*nJ + J when rows are contiguous (s_j == 1), e = J*np + q
// otherwise, so that consecutive lanes read consecutive addresses.  Each lane
// takes elements tid, tid+T, ...; (q, J) advance by the divmod of T instead of
// dividing per element, and kFill loads are issued before the first LDS
// store, so a lane waits for one memory latency per kFill elements instead of
// one per element.  The caller's next barrier publishes the tile.
__device__ __forceinline__ void fill_tile(const TileArgs& a, const ItemGeom& g, int t, double* tile,
                                          long long* s_in, long long* s_out)
{
    const int tid = threadIdx.x, T = blockDim.x;
    const int np = a.np, ldj = a.ldj, nJ = g.nJ;
    if (tid < np) {
        const int iv = t / a.tiles_per_var;
        const int P = (t - iv * a.tiles_per_var) * np + tid;
        if (P < a.na * a.nb) {
            const int pa = P % a.na, pb = P / a.na;
            s_in[tid] = a.src_off + pa * a.s_a + pb * a.s_b + iv * a.s_i + g.lo * a.s_j;
            const int k = a.a_is_k ? pa : pb;
            const int l = a.a_is_k ? pb : pa;
            s_out[tid] = k * a.o_k + l * a.o_l + iv * a.o_i;
        } else {
            s_in[tid] = -1;
            s_out[tid] = -1;
        }
    }
    __syncthreads();
    GlbCF64* const src = (GlbCF64*)a.src;
    constexpr int kFill = 8;
    const int n = np * nJ;
    const bool rows = a.s_j == 1;
    const int inner = rows ? nJ : np;      // extent of the fast index
    const int d_hi = T / inner, d_lo = T - d_hi * inner;
    int hi_i = tid / inner, lo_i = tid - hi_i * inner;
    for (int e0 = tid; e0 < n; e0 += kFill * T) {
        double r[kFill];
        int dst[kFill];
#pragma unroll
        for (int u = 0; u < kFill; ++u) {
            const int q = rows ? hi_i : lo_i, J = rows ? lo_i : hi_i;
            dst[u] = -1;
            r[u] = 0.0;
            if (e0 + u * T < n) {
                const long long b = s_in[q];
                dst[u] = q * ldj + J;
                if (b >= 0) r[u] = src[b + J * a.s_j];
            }
            lo_i += d_lo;
            hi_i += d_hi;
            if (lo_i >= inner) {
                lo_i -= inner;
                ++hi_i;
            }
        }
#pragma unroll
        for (int u = 0; u < kFill; ++u)
            if (dst[u] >= 0) tile[dst[u]] = r[u];
    }
}

// A run of consecutive items of one block (descriptor D), starting at `item`.
// The block's coefficients are loaded once for the run.  JS weights (WENO5
// and WENO7) take the sweep kernels' shared-reciprocal cells (cell5/cell7)
// with every coefficient in registers (64-block batch 7.58 -> 6.49 ms per
// x+y+z; the workgroup alternates load and compute phases, so the cells'
// fewer instructions shorten it directly); the other branches keep the
// candidate coefficients in registers and beta and d in an LDS table
// [lane][entry] after the tile.  Per item: pencil addresses, tile, one
// barrier, the claim of the next item (thread 0; its round trip overlaps the
// reconstruction), reconstruction and stores, one barrier after which every
// thread reads the claimed item from s_item.  Returns when that item belongs
// to another block (or there are no more items); s_item holds it.
template <int ORDER, int MODE, bool MP>
__device__ void batch_block(const BatchDesc& D, int item, double* tile, long long* s_in, long long* s_out,
                            int* s_item, int* counter, int& next)
{
    const TileArgs& a = D.a;
    constexpr int RW = ORDER / 2;
    constexpr int R = MP ? (RW > 2 ? RW : 2) : RW;
    constexpr int W = 2 * R + 1;
    constexpr int NS = BetaTraits<ORDER>::NS, NB = BetaTraits<ORDER>::N, NR = BetaTraits<ORDER>::NR;
    constexpr int LDB = BetaTraits<ORDER>::LD;
    const ItemGeom g = item_geom(a, R);
    const int item_end = D.item0 + a.tiles_per_var * a.nv;
    const int tid = threadIdx.x;
    const int gq = tid / a.jn;
    const int jj = tid - gq * a.jn;
    const int j = g.jlo + jj;
    const bool active = gq < a.groups && j <= g.jhi;
    const bool core = active && g.has_core && j >= g.c0 && j <= g.c1;
    const bool do_mp = MP && core && j >= a.mp_beg && j <= a.mp_end;
    LdsMutF64* const table = (LdsMutF64*)(tile + a.np * a.ldj);
    constexpr bool CELL = MODE == MODE_JS && ORDER != 3;
    std::conditional_t<ORDER == 7, Coef7, Coef5> kc;
    Cf<ORDER, double> f;
    if (core) {
        int jc = j - 3;
        jc = jc < 0 ? 0 : (jc > a.c.nc - 1 ? a.c.nc - 1 : jc);
        if constexpr (CELL) {
            if constexpr (ORDER == 5) load_coef5(a.c, jc, kc);
            else load_coef7(a.c, jc, kc);
        } else {
            load_cf(a.c, jc, f);
        }
        if (!CELL && gq == 0) {   // the groups share the table (same j per lane index)
            const GlbCoef *beta = gcoef(a.c.beta), *dL = gcoef(a.c.dL), *dR = gcoef(a.c.dR);
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int r = t / NS, s = t - (t / NS) * NS;
                table[jj * LDB + t] = beta[jc + a.c.nc * (r + 5 * s)];
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                table[jj * LDB + NB + r] = dL[r + 5 * jc];
                table[jj * LDB + NB + NR + r] = dR[r + 5 * jc];
            }
        }
    }
    GlbF64* const outL = (GlbF64*)a.outL;
    GlbF64* const outR = (GlbF64*)a.outR;
    const int col = j + a.J0f - g.lo;
    const BetaRef<NS> B{(const LdsF64*)table + jj * LDB, NB, NR};
    const LdsF64* const ltile = lds_ptr(tile);
    const LdsI64* const lout = (const LdsI64*)s_out;
    for (;;) {
        fill_tile(a, g, item - D.item0, tile, s_in, s_out);
        __syncthreads();
        if (tid == 0) next = atomicAdd(counter, 1);
        if (active) {
            for (int qq = 0; qq < a.npg; ++qq) {
                const int q = gq * a.npg + qq;
                const long long ob = lout[q];
                if (ob < 0) continue;
                const LdsF64* row = ltile + q * a.ldj + col;
                if (!core) {
                    outL[ob + j] = row[0];
                    outR[ob + j] = row[0];
                    continue;
                }
                double v[W];
#pragma unroll
                for (int o = 0; o < W; ++o) v[o] = row[o - R];
                double oL, oR;
                const double* vc = v + (R - RW);
                if constexpr (CELL && ORDER == 5) {
                    if (!cell5<2>(vc[0], vc[1], vc[2], vc[3], vc[4], kc, a.eps, oL, oR))
                        cell5_slow_inl(vc[0], vc[1], vc[2], vc[3], vc[4], kc, a.eps, oL, oR);
                } else if constexpr (CELL && ORDER == 7) {
                    const double w[7] = {vc[0], vc[1], vc[2], vc[3], vc[4], vc[5], vc[6]};
                    if (!cell7<2>(w, kc, a.eps, oL, oR)) cell7_slow_inl(w, kc, a.eps, oL, oR);
                } else if constexpr (ORDER == 3)
                    weno3<MODE, double, false>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                else if constexpr (ORDER == 5)
                    weno5<MODE, double, false>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                else
                    weno7<MODE, double, false>(vc, f, B, a.eps, a.teno_ct, oL, oR);
                if constexpr (MP) {
                    if (do_mp) mp_limit<double>(v + (R - 2), oL, oR);
                }
                outL[ob + j] = oL;
                outR[ob + j] = oR;
            }
        }
        if (tid == 0) *s_item = next;
        __syncthreads();
        item = *s_item;
        if (item >= item_end) return;
    }
}

template <bool MP>
__global__ void __launch_bounds__(kBatchThreads)
    batched_kernel(const BatchDesc* __restrict__ descs, int ndesc, int total, int* __restrict__ counter)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* const tile = reinterpret_cast<double*>(smem_raw);
    __shared__ long long s_in[kBatchMaxNP], s_out[kBatchMaxNP];
    __shared__ int s_item, s_desc;
    // first item of each descriptor, for the item -> descriptor search
    __shared__ int s_item0[kBatchLdsDescs];
    const int tid = threadIdx.x, T = blockDim.x;
    const bool lds_search = ndesc <= kBatchLdsDescs;
    if (lds_search)
        for (int i = tid; i < ndesc; i += T) s_item0[i] = descs[i].item0;
    // thread 0 holds the claimed item in `next` (batch_block claims the
    // following one while it reconstructs)
    int next = 0;
    if (tid == 0) {
        next = atomicAdd(counter, 1);
        s_item = next;
    }
    for (;;) {
        // the block of the claimed item (s_item, written by thread 0 before
        // the barrier that every thread passed since its last read)
        if (tid == 0) {
            int lo = 0, hi = ndesc - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((lds_search ? s_item0[mid] : descs[mid].item0) <= next) lo = mid;
                else hi = mid - 1;
            }
            s_desc = lo;
        }
        __syncthreads();
        const int item = s_item;
        if (item >= total) break;
        const BatchDesc& D = descs[s_desc];
        const int order = D.order, mode = D.mode;
#define WM_BATCH(O, M) batch_block<O, M, MP>(D, item, tile, s_in, s_out, &s_item, counter, next)
        if (order == 5) {
            if (mode == MODE_JS) WM_BATCH(5, MODE_JS);
            else if (mode == MODE_TENO) WM_BATCH(5, MODE_TENO);
            else if (mode == MODE_MAPPED) WM_BATCH(5, MODE_MAPPED);
            else if (mode == MODE_WENOZ) WM_BATCH(5, MODE_WENOZ);
            else WM_BATCH(5, MODE_NONE);
        } else if (order == 7) {
            if (mode == MODE_JS) WM_BATCH(7, MODE_JS);
            else if (mode == MODE_TENO) WM_BATCH(7, MODE_TENO);
            else if (mode == MODE_MAPPED) WM_BATCH(7, MODE_MAPPED);
            else if (mode == MODE_WENOZ) WM_BATCH(7, MODE_WENOZ);
            else WM_BATCH(7, MODE_NONE);
        } else {
            if (mode == MODE_JS) WM_BATCH(3, MODE_JS);
            else if (mode == MODE_MAPPED) WM_BATCH(3, MODE_MAPPED);
            else if (mode == MODE_WENOZ) WM_BATCH(3, MODE_WENOZ);
            else WM_BATCH(3, MODE_NONE);
        }
#undef WM_BATCH
    }
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

hipError_t launch_batched(const BatchDesc* descs_dev, int ndesc, int total, int* counter, size_t lds, int threads,
                          int mp, hipStream_t st)
{
    const void* kern = mp ? reinterpret_cast<const void*>(batched_kernel<true>)
                          : reinterpret_cast<const void*>(batched_kernel<false>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, lds) != hipSuccess || occ < 1) occ = 1;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const int grid = std::max(1, std::min(cus * occ, total));
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, st, counter);
    if (mp)
        hipLaunchKernelGGL(batched_kernel<true>, dim3(grid), dim3(threads), lds, st, descs_dev, ndesc, total, counter);
    else
        hipLaunchKernelGGL(batched_kernel<false>, dim3(grid), dim3(threads), lds, st, descs_dev, ndesc, total, counter);
    return hipGetLastError();
}

int batch_max_threads() { return kBatchThreads; }
int batch_max_np() { return kBatchMaxNP; }

}  // namespace wm
