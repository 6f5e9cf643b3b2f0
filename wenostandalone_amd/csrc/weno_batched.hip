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
#include "weno_capi_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace wm {

struct BatchDesc {
    TileArgs a;
    int order;
    int mode;
    int item0;
    int pad;
};

constexpr int kBatchThreads = 512;
constexpr int kBatchMaxNP = 64;
constexpr int kBatchLdsDescs = 1024;

// reconstruct the tile in LDS with order ORDER (after the tile barrier)
typedef __attribute__((address_space(3))) double LdsMutF64;
typedef __attribute__((address_space(3))) const long long LdsI64;
// the descriptors come from memory, so their pointers are generic: casting
// them to global keeps loads and stores off the flat path (flat accesses
// count against both wait counters, so LDS waits would include them)
typedef __attribute__((address_space(1))) double GlbF64;
typedef __attribute__((address_space(1))) const double GlbCF64;

// tile, table and s_out live in LDS; the address-space-qualified pointers
// keep their accesses ds_read/ds_write when this function is not inlined
template <int ORDER, int MODE, bool MP>
__device__ void batch_compute(const TileArgs& a, const LdsF64* tile, LdsMutF64* table, const LdsI64* s_out,
                              int lo, int c0, int c1, int jlo, int jhi, bool has_core, int* counter, int& next)
{
    constexpr int RW = ORDER / 2;
    constexpr int R = MP ? (RW > 2 ? RW : 2) : RW;
    constexpr int W = 2 * R + 1;
    constexpr int NS = BetaTraits<ORDER>::NS, NB = BetaTraits<ORDER>::N, NR = BetaTraits<ORDER>::NR;
    constexpr int LDB = BetaTraits<ORDER>::LD;
    const int tid = threadIdx.x;
    const int gq = tid / a.jn;
    const int jj = tid - gq * a.jn;
    const int j = jlo + jj;
    const bool active = gq < a.groups && j <= jhi;
    const bool core = active && has_core && j >= c0 && j <= c1;
    const bool do_mp = MP && core && j >= a.mp_beg && j <= a.mp_end;
    Cf<ORDER, double> f;
    if (core) {
        int jc = j - 3;
        jc = jc < 0 ? 0 : (jc > a.c.nc - 1 ? a.c.nc - 1 : jc);
        load_cf(a.c, jc, f);
        if (gq == 0) {
            for (int t = 0; t < NB; ++t) {
                const int r = t / NS, s = t - (t / NS) * NS;
                table[jj * LDB + t] = a.c.beta[jc + a.c.nc * (r + 5 * s)];
            }
            for (int r = 0; r < NR; ++r) {
                table[jj * LDB + NB + r] = a.c.dL[r + 5 * jc];
                table[jj * LDB + NB + NR + r] = a.c.dR[r + 5 * jc];
            }
        }
    }
    __syncthreads();
    // claim the next item now: its round trip overlaps the reconstruction
    // (thread 0's coefficient loads have completed, its stores come later)
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    if (!active) return;
    GlbF64* const outL = (GlbF64*)a.outL;
    GlbF64* const outR = (GlbF64*)a.outR;
    const int col = j + a.J0f - lo;
    const BetaRef<NS> B{table + jj * LDB, NB, NR};
    for (int qq = 0; qq < a.npg; ++qq) {
        const int q = gq * a.npg + qq;
        const long long ob = s_out[q];
        if (ob < 0) continue;
        const LdsF64* row = tile + q * a.ldj + col;
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
        if constexpr (ORDER == 3)
            weno3<MODE, double>(vc, f, B, a.eps, a.teno_ct, oL, oR);
        else if constexpr (ORDER == 5)
            weno5<MODE, double>(vc, f, B, a.eps, a.teno_ct, oL, oR);
        else
            weno7<MODE, double>(vc, f, B, a.eps, a.teno_ct, oL, oR);
        if constexpr (MP) {
            if (do_mp) mp_limit<double>(v + (R - 2), oL, oR);
        }
        outL[ob + j] = oL;
        outR[ob + j] = oR;
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
    // the next item is claimed while the current one is computed, so the
    // atomic's round trip is hidden (thread 0 holds it; batch_compute claims)
    int next = 0;
    if (tid == 0) next = atomicAdd(counter, 1);
    __syncthreads();

    for (;;) {
        if (tid == 0) {
            const int it = next;
            int lo = 0, hi = ndesc - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((lds_search ? s_item0[mid] : descs[mid].item0) <= it) lo = mid;
                else hi = mid - 1;
            }
            s_item = it;
            s_desc = lo;
        }
        __syncthreads();
        const int item = s_item;
        if (item >= total) break;
        const BatchDesc& D = descs[s_desc];
        const TileArgs& a = D.a;
        const int order = D.order, mode = D.mode;
        const int t = item - D.item0;
        const int np = a.np, ldj = a.ldj;
        const int RW = order / 2;
        const int R = MP ? (RW > 2 ? RW : 2) : RW;
        const int c0 = a.j_beg, c1 = a.j_end;
        const bool has_core = c0 <= c1;
        int lo, hi, jlo, jhi;
        if (a.copy_halo) {
            lo = min(a.J0f, c0 + a.J0f - R);
            hi = max(a.out_len - 1 + a.J0f, c1 + a.J0f + R);
            jlo = 0;
            jhi = a.out_len - 1;
        } else {
            lo = c0 + a.J0f - R;
            hi = c1 + a.J0f + R;
            jlo = c0;
            jhi = c1;
        }
        const int nJ = hi - lo + 1;
        if (tid < np) {
            const int iv = t / a.tiles_per_var;
            const int P = (t - iv * a.tiles_per_var) * np + tid;
            if (P < a.na * a.nb) {
                const int pa = P % a.na, pb = P / a.na;
                s_in[tid] = a.src_off + pa * a.s_a + pb * a.s_b + iv * a.s_i + lo * a.s_j;
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
        for (int e = tid; e < np * nJ; e += T) {
            int q, J;
            if (a.s_j == 1) {
                q = e / nJ;
                J = e - q * nJ;
            } else {
                J = e / np;
                q = e - J * np;
            }
            const long long b = s_in[q];
            tile[q * ldj + J] = b >= 0 ? src[b + J * a.s_j] : 0.0;
        }
        __syncthreads();
        double* const table = tile + np * ldj;
#define WM_BATCH(O, M)                                                                                      \
    batch_compute<O, M, MP>(a, lds_ptr(tile), (LdsMutF64*)table, (const LdsI64*)s_out, lo, c0, c1, jlo, jhi, \
                            has_core, counter, next)
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
        __syncthreads();
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
