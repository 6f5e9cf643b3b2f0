// Order-1 copy (s_order1_copy_{x,y,z}, inputs/weno_standalone.f90:332-390)
// and reshape (s_initialize_weno_{x,y,z}, :281-330): the field permuted into
// rows along the sweep direction.  No arithmetic: one read and one (reshape)
// or two (copy) writes per element, bound by HBM.
//
// * x sweeps (rows contiguous in v_vf): a flat streaming copy, each lane
//   moving two adjacent elements of one output row (16-byte stores for FP64
//   faces; one 16-byte load where the input pair is aligned, as for the
//   reshape, or the aligned pair one element back plus the next lane's
//   element, as for the copy's rows); the row and pencil of a pair come
//   from multiply-high divisions by the row length and extents.  Odd row
//   lengths and outputs that are not aligned to a pair take one element per
//   lane, one wave per row at a time.
// * y and z sweeps: tiles of 16 pencils that are consecutive in x (the
//   field's unit-stride coordinate) by a chunk of the row along the sweep,
//   transposed through LDS.  Loads: lane -> (pencil t & 15, row t >> 4), so a
//   wave instruction reads 4 rows of 16 pencils (128 contiguous bytes per
//   row for FP64); eight loads per lane are in flight before the first LDS
//   store.  Stores: 64 pairs of consecutive outputs of one row per wave
//   instruction (1 KB for FP64; 64 single outputs where rows are odd).
//   Neighbouring tiles share the input lines at their edges (a 16-pencil
//   segment starts 24 bytes into a line), so tiles go to the XCDs in groups
//   of 32 consecutive tiles and a shared line is fetched into one L2.
#include "weno_device.h"

#include <algorithm>
#include <type_traits>

namespace wm {

namespace {

constexpr int kPermThreads = 256;
constexpr int kPermNP = 16;       // pencils per tile (y/z)
constexpr int kPermMaxChunk = 448; // outputs per tile along the sweep (LDS <= 57.5 KB for FP64)
constexpr int kPermGroup = 32;    // consecutive tiles per XCD
constexpr int kPermFill = 8;      // loads in flight per lane

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

template <typename T>
struct Pair;
template <>
struct Pair<double> {
    using type = f64x2;
};
template <>
struct Pair<float> {
    using type = f32x2;
};

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// (k, l, i) of output row r (rows are k-fastest) -> pencil (a, b) of the input
__device__ __forceinline__ void row_pencil(const TileArgs& a, long long r, int& pa, int& pb, int& iv,
                                           long long& ob)
{
    const int nk = a.a_is_k ? a.na : a.nb;
    const int nl = a.a_is_k ? a.nb : a.na;
    const long long kl = static_cast<long long>(nk) * nl;
    iv = static_cast<int>(r / kl);
    const int rest = static_cast<int>(r - iv * kl);
    const int l = rest / nk, k = rest - l * nk;
    pa = a.a_is_k ? k : l;
    pb = a.a_is_k ? l : k;
    ob = k * a.o_k + l * a.o_l + iv * a.o_i;
}

// Divisions by a launch constant d (1 <= d < 2^32) as a multiply-high:
// n / d = (umulhi(n, m) + n) >> s for every 32-bit n, with s = ceil(log2 d)
// and m = floor(2^32 (2^s - d) / d) + 1 (Granlund and Montgomery, "Division
// by invariant integers using multiplication", 1994, section 4).
struct FastDiv {
    unsigned m;
    int s;
};
FastDiv fast_div_of(unsigned d)
{
    int s = 0;
    while ((1ull << s) < d) ++s;
    const unsigned long long m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
    return FastDiv{static_cast<unsigned>(m), s};
}
__device__ __forceinline__ unsigned fdiv(unsigned n, FastDiv f)
{
    return static_cast<unsigned>((static_cast<unsigned long long>(__umulhi(n, f.m)) + n) >> f.s);
}

// Lane i receives lane i+1's value (DPP wave_shl:1; lane 63 receives 0).
template <typename T>
__device__ __forceinline__ T from_next_lane(T x)
{
    static_assert(sizeof(T) == 8 || sizeof(T) == 4, "32- or 64-bit values");
    if constexpr (sizeof(T) == 8) {
        const long long b = __builtin_bit_cast(long long, x);
        const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), 0x130, 0xf, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x130, 0xf, 0xf, true);
        return __builtin_bit_cast(T, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
    }
}

// How a lane reads its input pair: two element loads (storage and face types
// differ), one aligned pair load, or (pairs starting one element past an
// aligned pair, as the order-1 copy's rows do) the aligned pair one element
// back, whose second element is this pair's first; the pair's second element
// is the next lane's first, passed by a lane shift, and loaded directly by
// the lanes whose next lane holds another row (or no pair).
enum PairLoad { PL_ELEMENTS = 0, PL_ALIGNED = 1, PL_SHIFTED = 2 };

// x sweeps, even row lengths and pair-aligned outputs: pair c of the flat
// output -> row r = c / (out_len/2) -> (k, l, i).
template <typename TS, typename TO, int PL>
__global__ void __launch_bounds__(kPermThreads)
    permute_flat_kernel(const TileArgs a, unsigned npairs, FastDiv dhalf, FastDiv dk, FastDiv dl)
{
    const unsigned half = a.out_len / 2;
    const int nk = a.a_is_k ? a.na : a.nb;
    const int nl = a.a_is_k ? a.nb : a.na;
    const TS* const src = static_cast<const TS*>(a.src);
    TO* const o0 = static_cast<TO*>(a.outL);
    TO* const o1 = static_cast<TO*>(a.outR);
    const bool two = (a.sides & 2) && a.outR != a.outL;
    using P = typename Pair<TO>::type;
    const int lane = threadIdx.x & 63;
    for (unsigned c = blockIdx.x * kPermThreads + threadIdx.x; c < npairs; c += gridDim.x * kPermThreads) {
        const unsigned r = fdiv(c, dhalf);
        const int j = 2 * static_cast<int>(c - r * half);
        const unsigned t = fdiv(r, dk);
        const int k = static_cast<int>(r - t * nk);
        const unsigned iv = fdiv(t, dl);
        const int l = static_cast<int>(t - iv * nl);
        const int pa = a.a_is_k ? k : l, pb = a.a_is_k ? l : k;
        const TS* const in =
            src + a.src_off + pa * a.s_a + pb * a.s_b + static_cast<long long>(iv) * a.s_i + a.J0f + j;
        P x;
        if constexpr (PL == PL_ALIGNED && std::is_same_v<TS, TO>) {
            x = *reinterpret_cast<const P*>(in);
        } else if constexpr (PL == PL_SHIFTED && std::is_same_v<TS, TO>) {
            const P b = *reinterpret_cast<const P*>(in - 1);
            x.x = b.y;
            x.y = from_next_lane(b.x);
            if (lane == 63 || j + 2 == a.out_len || c + 1 >= npairs) x.y = in[1];
        } else {
            x.x = static_cast<TO>(ld_in(in));
            x.y = static_cast<TO>(ld_in(in + 1));
        }
        const long long ob = static_cast<long long>(r) * a.out_len + j;
        __builtin_nontemporal_store(x, reinterpret_cast<P*>(o0 + ob));
        if (two) __builtin_nontemporal_store(x, reinterpret_cast<P*>(o1 + ob));
    }
}

// x sweeps otherwise: one wave per output row at a time, one element per lane.
template <typename TS, typename TO>
__global__ void __launch_bounds__(kPermThreads) permute_rows_kernel(const TileArgs a, long long nrows)
{
    const int lane = threadIdx.x & 63;
    const long long w0 = static_cast<long long>(blockIdx.x) * (kPermThreads / 64) + uniform(threadIdx.x >> 6);
    const long long nw = static_cast<long long>(gridDim.x) * (kPermThreads / 64);
    const TS* const src = static_cast<const TS*>(a.src);
    TO* const o0 = static_cast<TO*>(a.outL);
    TO* const o1 = static_cast<TO*>(a.outR);
    const bool two = (a.sides & 2) && a.outR != a.outL;
    const int n = a.out_len;
    for (long long r = w0; r < nrows; r += nw) {
        int pa, pb, iv;
        long long ob;
        row_pencil(a, r, pa, pb, iv, ob);
        const TS* const in = src + a.src_off + pa * a.s_a + pb * a.s_b + iv * a.s_i + a.J0f;
        for (int j = lane; j < n; j += 64) {
            const TO x = static_cast<TO>(ld_in(in + j));
            __builtin_nontemporal_store(x, o0 + ob + j);
            if (two) __builtin_nontemporal_store(x, o1 + ob + j);
        }
    }
}

// y and z sweeps: 16-pencil tiles transposed through LDS.  Tile t = (variable,
// pencil group, chunk), chunk fastest; LDS rows of ld elements (odd, or 2 mod
// 4 with PAIRS, so that the 16 pencils of a load instruction fall in distinct
// banks).  PAIRS: each lane stores two adjacent outputs (even rows and chunks,
// pair-aligned outputs).
template <typename TS, typename TO, bool PAIRS>
__global__ void __launch_bounds__(kPermThreads)
    permute_tiles_kernel(const TileArgs a, int chunk, int chunks, int groups, int ntile, int ld)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TO* const tile = reinterpret_cast<TO*>(smem_raw);
    __shared__ long long s_ob[kPermNP];
    const int b = blockIdx.x, xcd = b & 7, idx = b >> 3;
    const int t = kPermGroup * (xcd + 8 * (idx / kPermGroup)) + idx % kPermGroup;
    if (t >= ntile) return;
    const int c = t % chunks, tg = t / chunks;
    const int g = tg % groups, iv = tg / groups;
    const int j0 = c * chunk, len = min(chunk, a.out_len - j0);
    const int tid = threadIdx.x;
    const int q = tid & (kPermNP - 1), jr = tid >> 4;
    // pencils past the end repeat the last one (same values, same addresses)
    const int P = min(g * kPermNP + q, a.na * a.nb - 1);
    const int pbi = P / a.na, pa = P - pbi * a.na;
    if (tid < kPermNP) {
        const int k = a.a_is_k ? pa : pbi, l = a.a_is_k ? pbi : pa;
        s_ob[tid] = k * a.o_k + l * a.o_l + iv * a.o_i + j0;
    }
    const TS* const in =
        static_cast<const TS*>(a.src) + a.src_off + pa * a.s_a + pbi * a.s_b + iv * a.s_i + (j0 + a.J0f) * a.s_j;
    constexpr int kRowsPerPass = kPermThreads / kPermNP;   // 16
    for (int jb = jr; jb < len; jb += kPermFill * kRowsPerPass) {
        TO r[kPermFill];
#pragma unroll
        for (int u = 0; u < kPermFill; ++u) {
            const int jj = jb + u * kRowsPerPass;
            r[u] = jj < len ? static_cast<TO>(ld_in(in + jj * a.s_j)) : TO(0);
        }
#pragma unroll
        for (int u = 0; u < kPermFill; ++u) {
            const int jj = jb + u * kRowsPerPass;
            if (jj < len) tile[q * ld + jj] = r[u];
        }
    }
    __syncthreads();
    TO* const o0 = static_cast<TO*>(a.outL);
    TO* const o1 = static_cast<TO*>(a.outR);
    const bool two = (a.sides & 2) && a.outR != a.outL;
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int qq = 0; qq < kPermNP / (kPermThreads / 64); ++qq) {
        const int qs = w + (kPermThreads / 64) * qq;
        const long long ob = s_ob[qs];
        if constexpr (PAIRS) {
            using P = typename Pair<TO>::type;
            for (int j = 2 * lane; j < len; j += 128) {
                const P x = *reinterpret_cast<const P*>(tile + qs * ld + j);
                __builtin_nontemporal_store(x, reinterpret_cast<P*>(o0 + ob + j));
                if (two) __builtin_nontemporal_store(x, reinterpret_cast<P*>(o1 + ob + j));
            }
        } else {
            for (int j = lane; j < len; j += 64) {
                const TO x = tile[qs * ld + j];
                __builtin_nontemporal_store(x, o0 + ob + j);
                if (two) __builtin_nontemporal_store(x, o1 + ob + j);
            }
        }
    }
}

template <typename TS, typename TO>
hipError_t launch_permute_t(const TileArgs& a, hipStream_t st)
{
    const long long nrows = static_cast<long long>(a.na) * a.nb * a.nv;
    if (nrows <= 0 || a.out_len <= 0) return hipSuccess;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    if (a.s_j == 1) {
        const int nk = a.a_is_k ? a.na : a.nb, nl = a.a_is_k ? a.nb : a.na;
        const long long npairs = nrows * (a.out_len / 2);
        const bool pairs = a.out_len % 2 == 0 && npairs <= 0x7fffffffLL &&
                           reinterpret_cast<uintptr_t>(a.outL) % (2 * sizeof(TO)) == 0 &&
                           reinterpret_cast<uintptr_t>(a.outR) % (2 * sizeof(TO)) == 0;
        if (pairs) {
            // pair parity is the same on every row when the strides are even
            const bool even = a.s_a % 2 == 0 && a.s_b % 2 == 0 && a.s_i % 2 == 0 &&
                              reinterpret_cast<uintptr_t>(a.src) % (2 * sizeof(TS)) == 0;
            const int pl = !std::is_same_v<TS, TO> || !even ? PL_ELEMENTS
                           : (a.src_off + a.J0f) % 2 == 0   ? PL_ALIGNED
                           : a.src_off + a.J0f > 0          ? PL_SHIFTED
                                                            : PL_ELEMENTS;
            const unsigned grid = static_cast<unsigned>(
                std::max(1LL, std::min((npairs + kPermThreads - 1) / kPermThreads, static_cast<long long>(cus) * 32)));
            const FastDiv dh = fast_div_of(a.out_len / 2), dk = fast_div_of(nk), dl = fast_div_of(nl);
            const unsigned np2 = static_cast<unsigned>(npairs);
            if (pl == PL_ALIGNED)
                hipLaunchKernelGGL((permute_flat_kernel<TS, TO, PL_ALIGNED>), dim3(grid), dim3(kPermThreads), 0, st,
                                   a, np2, dh, dk, dl);
            else if (pl == PL_SHIFTED)
                hipLaunchKernelGGL((permute_flat_kernel<TS, TO, PL_SHIFTED>), dim3(grid), dim3(kPermThreads), 0, st,
                                   a, np2, dh, dk, dl);
            else
                hipLaunchKernelGGL((permute_flat_kernel<TS, TO, PL_ELEMENTS>), dim3(grid), dim3(kPermThreads), 0, st,
                                   a, np2, dh, dk, dl);
            return hipGetLastError();
        }
        const long long waves = kPermThreads / 64;
        const long long grid = std::max(1LL, std::min((nrows + waves - 1) / waves, static_cast<long long>(cus) * 32));
        hipLaunchKernelGGL((permute_rows_kernel<TS, TO>), dim3(static_cast<unsigned>(grid)), dim3(kPermThreads), 0,
                           st, a, nrows);
        return hipGetLastError();
    }
    // the transposed tiles read the pencils' unit-stride coordinate
    if (a.s_a != 1) return hipErrorInvalidValue;
    const int chunks = (a.out_len + kPermMaxChunk - 1) / kPermMaxChunk;
    // even chunks of even rows (at most kPermMaxChunk)
    const int chunk = a.out_len % 2 == 0 ? (((a.out_len + chunks - 1) / chunks + 1) & ~1)
                                         : (a.out_len + chunks - 1) / chunks;
    const bool pairs = a.out_len % 2 == 0 && reinterpret_cast<uintptr_t>(a.outL) % (2 * sizeof(TO)) == 0 &&
                       reinterpret_cast<uintptr_t>(a.outR) % (2 * sizeof(TO)) == 0;
    const int ld = !pairs ? chunk + 1 : (chunk % 4 == 2 ? chunk : chunk + 2);
    const long long groups = (static_cast<long long>(a.na) * a.nb + kPermNP - 1) / kPermNP;
    const long long ntile = groups * a.nv * chunks;
    const long long span = 8LL * kPermGroup;
    const long long grid = (ntile + span - 1) / span * span;
    if (ntile > 0x7fffffffLL || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    const size_t lds = static_cast<size_t>(kPermNP) * ld * sizeof(TO);
    if (pairs)
        hipLaunchKernelGGL((permute_tiles_kernel<TS, TO, true>), dim3(static_cast<unsigned>(grid)), dim3(kPermThreads),
                           lds, st, a, chunk, chunks, static_cast<int>(groups), static_cast<int>(ntile), ld);
    else
        hipLaunchKernelGGL((permute_tiles_kernel<TS, TO, false>), dim3(static_cast<unsigned>(grid)),
                           dim3(kPermThreads), lds, st, a, chunk, chunks, static_cast<int>(groups),
                           static_cast<int>(ntile), ld);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_permute(const TileArgs& a, int dtype, hipStream_t st)
{
    if (dtype == DT_F64) return launch_permute_t<double, double>(a, st);
    if (dtype == DT_F32) return launch_permute_t<float, float>(a, st);
    if (dtype == DT_BF16) return launch_permute_t<bf16_t, float>(a, st);
    return hipErrorInvalidValue;
}

}  // namespace wm
