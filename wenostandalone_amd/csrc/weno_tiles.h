// Whole-row tiles for the persistent producer/consumer kernels that walk
// pencils one plane at a time (the consumer step and the adjoint): tile t of
// a launch is np pencils pa0 .. pa0+cnt-1 of plane pb of variable iv, and
// the rows J = lo .. lo+nJ-1 of those pencils go to LDS as [pencil][J] with
// row stride ld.
//
// The producer wave loads a tile into kSlots registers per lane, then
// commits it to LDS:
//   contiguous rows (s_j = 1: x sweeps of the field, reshaped arrays): the
//     tile's rows are s_a elements apart, so the producer loads the span from
//     the first row's J = lo to the last row's J = lo+nJ-1 (gaps included),
//     64 consecutive elements per instruction; element e of the span is LDS
//     row e / s_a, column e % s_a;
//   strided rows (y, z sweeps of the field, s_a = 1): the tile's pencils are
//     consecutive in memory; with G = 64 / np rows per instruction, lane
//     l < np G holds pencil l % np of rows J = l / np + G s (the other lanes
//     idle when np does not divide 64).
#pragma once

#include "weno_internal.h"

namespace wm {
namespace {

template <typename T>
using GlobalIn = __attribute__((address_space(1))) const T;

// Tile t of the launch: variable iv, plane pb, pencils pa0 .. pa0 + cnt - 1
// (tiles never cross a plane: a.chunks tiles per plane).
__device__ __forceinline__ void plane_tile(const TileArgs& a, int t, int& iv, int& pa0, int& pb, int& cnt)
{
    iv = t / a.tiles_per_var;
    const int tv = t - iv * a.tiles_per_var;
    pb = tv / a.chunks;
    pa0 = (tv - pb * a.chunks) * a.np;
    cnt = min(a.np, a.na - pa0);
}

// Loads of tile t into r (producer lanes); returns the tile's pencil count.
// The per-slot offsets do not change from tile to tile; the opaque copy of
// the lane index keeps the compiler from holding all of them in registers
// across the tile loop.
template <int SLOTS>
__device__ __forceinline__ int tile_issue(const TileArgs& a, int lo, int nJ, int t, double (&r)[SLOTS])
{
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    const int np = a.np, G = 64 / np;
    const int sq = l % np, J0 = l / np;
    int iv, pa0, pb, cnt;
    plane_tile(a, t, iv, pa0, pb, cnt);
    const unsigned long long bm =
        reinterpret_cast<unsigned long long>(a.src) +
        8ull * (a.src_off + pa0 * a.s_a + pb * a.s_b + iv * a.s_i + static_cast<long long>(lo) * a.s_j);
    if (a.s_j == 1) {
        const int span = (cnt - 1) * static_cast<int>(a.s_a) + nJ;
#pragma unroll
        for (int s = 0; s < SLOTS; ++s)
            if (l + 64 * s < span)
                r[s] = *reinterpret_cast<GlobalIn<double>*>(bm + 8u * static_cast<unsigned>(l + 64 * s));
    } else {
        const unsigned off = static_cast<unsigned>(8 * (sq + J0 * a.s_j));
        const unsigned long long stepb = 8ull * G * a.s_j;
#pragma unroll
        for (int s = 0; s < SLOTS; ++s)
            if (sq < cnt && J0 < G && J0 + G * s < nJ) r[s] = *reinterpret_cast<GlobalIn<double>*>(bm + stepb * s + off);
    }
    return cnt;
}

// The loaded tile into LDS rows b[q * ld + J].
template <int SLOTS>
__device__ __forceinline__ void tile_commit(const TileArgs& a, int nJ, int ld, int cnt, const double (&r)[SLOTS],
                                            double* b)
{
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    const int np = a.np, G = 64 / np;
    const int sq = l % np, J0 = l / np;
    if (a.s_j == 1) {
        const int sa = static_cast<int>(a.s_a);
        const float inv_sa = 1.0f / static_cast<float>(sa);
        const int span = (cnt - 1) * sa + nJ;
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            // element e of the span -> row q = e / s_a, exact in FP32 for
            // e < 2^11 (the quotient is never within 0.5/s_a of an integer it
            // does not equal)
            const int e = l + 64 * s;
            const int q = static_cast<int>((static_cast<float>(e) + 0.5f) * inv_sa), J = e - q * sa;
            if (e < span && J < nJ) b[q * ld + J] = r[s];
        }
    } else {
#pragma unroll
        for (int s = 0; s < SLOTS; ++s)
            if (sq < cnt && J0 < G && J0 + G * s < nJ) b[sq * ld + J0 + G * s] = r[s];
    }
}

// Read-add-write of tile t of the array at `base` (the input's layout),
// shared by `parts` waves: wave `part` takes the slots part, part + parts,
// ... of the producer's slot mapping (PER of them, covering SLOTS), so the
// accesses stay as coalesced as the producer's loads.  STORE = false loads
// the elements into r; STORE = true adds b[q * ld + J] to r, stores the
// element back and clears b[q * ld + J] for the next tile.  Only the tile's
// own elements are stored (not the gaps between contiguous rows), so tiles
// of one launch never write the same element.
template <int SLOTS, int PER, bool STORE>
__device__ __forceinline__ void tile_part(const TileArgs& a, void* base, int lo, int nJ, int ld, int t, int part,
                                          int parts, double (&r)[PER], double* b)
{
    int l = threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    const int np = a.np, G = 64 / np;
    const int sq = l % np, J0 = l / np;
    int iv, pa0, pb, cnt;
    plane_tile(a, t, iv, pa0, pb, cnt);
    const unsigned long long bm =
        reinterpret_cast<unsigned long long>(base) +
        8ull * (a.src_off + pa0 * a.s_a + pb * a.s_b + iv * a.s_i + static_cast<long long>(lo) * a.s_j);
    using GD = __attribute__((address_space(1))) double;
    if (a.s_j == 1) {
        const int sa = static_cast<int>(a.s_a);
        const float inv_sa = 1.0f / static_cast<float>(sa);
        const int span = (cnt - 1) * sa + nJ;
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int s = part + parts * m;
            const int e = l + 64 * s;
            // element e of the span -> row q = e / s_a (exact in FP32, e < 2^11)
            const int q = static_cast<int>((static_cast<float>(e) + 0.5f) * inv_sa), J = e - q * sa;
            if (s < SLOTS && e < span && J < nJ) {
                GD* const p = reinterpret_cast<GD*>(bm + 8u * static_cast<unsigned>(e));
                if constexpr (STORE) {
                    double* const g = b + q * ld + J;
                    *p = r[m] + *g;
                    *g = 0.0;
                } else {
                    r[m] = *p;
                }
            }
        }
    } else {
        const unsigned off = static_cast<unsigned>(8 * (sq + J0 * a.s_j));
        const unsigned long long stepb = 8ull * G * a.s_j;
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int s = part + parts * m;
            if (s < SLOTS && sq < cnt && J0 < G && J0 + G * s < nJ) {
                GD* const p = reinterpret_cast<GD*>(bm + stepb * s + off);
                if constexpr (STORE) {
                    double* const g = b + sq * ld + J0 + G * s;
                    *p = r[m] + *g;
                    *g = 0.0;
                } else {
                    r[m] = *p;
                }
            }
        }
    }
}

// Lane i receives lane i+1's value (DPP wave_shl:1) / lane i-1's value
// (DPP wave_shr:1); the lanes at the wave's ends receive 0.
__device__ __forceinline__ double from_next_lane(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x130, 0xf, 0xf, true);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ __forceinline__ double from_prev_lane(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), 0x138, 0xf, 0xf, true);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

}  // namespace

// Host side: the largest power-of-two pencil count (at most max_np) whose
// tile of rows of nJ inputs fits `slots` registers per producer lane; 0 if
// none does.
inline int tile_pencils(int nJ, long long s_a, bool rows, int slots, int max_np)
{
    for (int np = max_np; np >= 1; np >>= 1) {
        const long long need = rows ? ((np - 1) * s_a + nJ + 63) / 64 : (nJ + 64 / np - 1) / (64 / np);
        if (need <= slots) return np;
    }
    return 0;
}

}  // namespace wm
