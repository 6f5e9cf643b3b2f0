// The WENO5-JS sweep kernel for gfx950 (the reference's hot path:
// s_weno5_*_kernel, inputs/weno_standalone.f90:477-1417, in its JS form),
// for FP64 state and faces, FP32 state and faces, and BF16 state with FP32
// faces.  The tile is converted to FP64 on its way into LDS, so the
// arithmetic below is FP64 for every storage type: on gfx950 an FP32 FMA
// (packed FP32 is not used) takes the same issue slot as an FP64 FMA, and
// FP32 polynomial terms would need the weights normalised before the sum
// (six more multiplies and six conversions per face pair).
//
// One workgroup = 7 consumer waves + 1 producer wave, persistent over a
// sequence of tiles.  A tile is 16 pencils x one j-chunk of at most 224
// outputs; tiles flow through two LDS buffers.  Tiles go round-robin over
// the whole grid, so that all workgroups stream through one narrow window of
// the input and output arrays; in y and z sweeps each round hands groups of
// 32 neighbouring tiles, which share input lines, to the workgroups of one
// XCD.  Per 200^3 x 8 sweep: x 3 % faster than one tile range per XCD; y
// 408 us with one range per XCD, 396 us round-robin, 393 us in groups;
// z 461 us with tiles of 8 x-values x 2 y-rows (adjacent output rows) and
// one range per XCD, 448 us with 16 x-values per tile in groups.
//
//   producer wave : vector loads of tile k+1 from HBM into registers, then
//                   into LDS as [pencil][J] (the y/z transposes happen here);
//                   it never stores to global memory, so its waits on the
//                   vector memory counter cover its own loads only.
//   consumer waves: two lane groups, one lane per output j of the chunk
//                   row; group g takes the tile's pencils 2i+g, so the
//                   wavefront that holds the end of group 0 and the start of
//                   group 1 writes the end of one output row and the start
//                   of the next, which are adjacent in memory for x and y.
//                   The per-j coefficients live in registers for the whole
//                   launch; each lane reads its 5-point window from LDS and
//                   writes vL and vR straight to HBM (non-temporal: the
//                   outputs are not re-read, the inputs' shared lines are),
//                   never waiting on its stores.
//   one barrier per tile hands buffer k+1 to the consumers and buffer k back
//   to the producer.
//
// With MP the monotonicity-preserving limiter (s_preserve_monotonicity,
// weno_standalone.f90:1544-1615) is applied in the same pass: its 5-point
// window is the WENO5 window the lane already holds, so the limited faces are
// written once and the limiter costs arithmetic only (FP64, before the face
// is rounded to its storage type).
//
// Output rows are always written whole.  Outputs outside the reconstructed
// range (j < 3, j > ns-4) get the order-1 copy with WENO_FUSE_COPY; without
// it they are rewritten with the values they already hold (the producer loads
// them with the tile), which leaves them unchanged: MI355X writes partly
// filled 128-byte lines back at about 60% of the rate of full ones.
//
// Per face pair the JS weights take the shared-reciprocal form of cell5
// (weno_cells.h): both sides share one reciprocal, and lanes whose product
// leaves the normal range redo the cell with per-stencil IEEE divisions.
#include "weno_device.h"
#include "weno_cells.h"

#include <algorithm>
#include <type_traits>

namespace wm {

namespace {

constexpr int kNP = 16;         // pencils per tile
constexpr int kNPG = 8;         // pencils per lane group (two groups)
constexpr int kConsumers = 448; // 7 waves
constexpr int kThreads5 = kConsumers + 64;
constexpr int kMaxOut = 224;    // outputs per chunk row
constexpr int kRPI = 64 / kNP;  // strided rows: rows per wave instruction (4)
constexpr int kMaxLoads = 54;   // strided rows: producer registers, 4 x 54 = 216 rows
constexpr int kRows0 = 224;     // contiguous rows: elements loaded per row

template <typename T>
using GlobalIn = __attribute__((address_space(1))) const T;
template <typename T>
using GlobalOut = __attribute__((address_space(1))) T;
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// State storage types.  The tile is held in LDS as FP64 whatever the state
// type (the producer converts as it commits), so the consumers run the same
// FP64 arithmetic for every storage type; only the face type changes.
// Contiguous rows are loaded in pieces of two elements (rows of v_vf start
// at even elements only), LPR lanes per row:
//   face : vL/vR element type
//   piece: one load of two state elements (PB bytes)
//   raw  : element as loaded from a strided pencil
//   hp   : two adjacent faces (halo cells are loaded in such pieces)
template <typename TS>
struct Store5;
template <>
struct Store5<double> {
    using face = double;
    using piece = u32x4;
    using raw = double;
    using hp = f64x2;
    static constexpr int PB = 16, LPR = 8;
};
template <>
struct Store5<float> {
    using face = float;
    using piece = u32x2;
    using raw = float;
    using hp = f32x2;
    static constexpr int PB = 8, LPR = 16;
};
template <>
struct Store5<bf16_t> {
    using face = float;
    using piece = unsigned;
    using raw = unsigned short;
    using hp = f32x2;
    static constexpr int PB = 4, LPR = 16;
};

__device__ __forceinline__ double widen(double x) { return x; }
__device__ __forceinline__ double widen(float x) { return static_cast<double>(x); }
__device__ __forceinline__ double widen(unsigned short x)
{
    return static_cast<double>(__uint_as_float(static_cast<unsigned>(x) << 16));
}

// One load piece of a contiguous row -> two consecutive LDS doubles.
__device__ __forceinline__ void put_piece(double* dst, u32x4 p, double)
{
    *reinterpret_cast<f64x2*>(dst) = __builtin_bit_cast(f64x2, p);
}
__device__ __forceinline__ void put_piece(double* dst, u32x2 p, float)
{
    *reinterpret_cast<f64x2*>(dst) =
        f64x2{static_cast<double>(__uint_as_float(p.x)), static_cast<double>(__uint_as_float(p.y))};
}
__device__ __forceinline__ void put_piece(double* dst, unsigned p, bf16_t)
{
    *reinterpret_cast<f64x2*>(dst) = f64x2{static_cast<double>(__uint_as_float(p << 16)),
                                           static_cast<double>(__uint_as_float(p & 0xffff0000u))};
}

__device__ __forceinline__ u32x4 load_piece(__amdgpu_buffer_rsrc_t r, unsigned off, u32x4)
{
    return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ u32x2 load_piece(__amdgpu_buffer_rsrc_t r, unsigned off, u32x2)
{
    return __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ unsigned load_piece(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// Non-temporal face store at a 32-bit byte offset (cache policy bit 1: nt).
__device__ __forceinline__ void store_face(double x, __amdgpu_buffer_rsrc_t r, unsigned off)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, off, 0, 2);
}
__device__ __forceinline__ void store_face(float x, __amdgpu_buffer_rsrc_t r, unsigned off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, off, 0, 2);
}

// Tile t of the launch -> variable and tile index within it.
__device__ __forceinline__ void tile_origin(const TileArgs& a, int t, int& iv, int& tv)
{
    iv = t / a.tiles_per_var;
    tv = t - iv * a.tiles_per_var;
}

// Pencil q of tile tv -> transverse coordinates (a, b): tiles are 16
// consecutive pencils in (a + na*b) order.  Pencils past the end repeat the
// last valid one (same data, same results, same addresses).
__device__ __forceinline__ void tile_pencil(const TileArgs& a, int tv, int q, int& pa, int& pb)
{
    const int P = min(tv * kNP + q, a.na * a.nb - 1);
    pb = P / a.na;
    pa = P - pb * a.na;
}

// Input element offset of pencil (pa, pb) of variable iv at J = lo.
__device__ __forceinline__ long long in_base(const TileArgs& a, int iv, int pa, int pb, int lo)
{
    return a.src_off + pa * a.s_a + pb * a.s_b + iv * a.s_i + static_cast<long long>(lo) * a.s_j;
}

__device__ __forceinline__ long long out_base(const TileArgs& a, int iv, int pa, int pb)
{
    const int k = a.a_is_k ? pa : pb, l = a.a_is_k ? pb : pa;
    return k * a.o_k + l * a.o_l + iv * a.o_i;
}

}  // namespace

// LAY 0: the sweep runs along the contiguous axis (x of v_vf, or a reshaped
//        array): each pencil row is loaded with 16-byte buffer loads, 8 lanes
//        per row.
// LAY 1: the sweep runs along a strided axis (y or z of v_vf): the 16 pencils
//        of a tile are consecutive along x, loaded as 128-byte row segments
//        (one pencil per lane, 4 rows per wave instruction) and transposed
//        into LDS.
// Weight branches other than JS (TENO, WENO-Z, mapped): the raw candidate
// and smoothness coefficients and linear weights of the lane, in registers,
// for the general per-family arithmetic (weno5 / weno7 in weno_device.h).
struct Gen5 {
    Cf<5, double> f;
    BetaRegs<5> b;
};
struct Gen7 {
    Cf<7, double> f;
    BetaRegs<7> b;
};

template <int ORDER, int MODE, typename TS, int LAY, bool COPY, bool MP>
__device__ __forceinline__ void sweep_body(const TileArgs& a)
{
    constexpr int RS = ORDER / 2;   // stencil radius
    using TO = typename Store5<TS>::face;
    using Piece = typename Store5<TS>::piece;
    using Raw = typename Store5<TS>::raw;
    using HP = typename Store5<TS>::hp;
    constexpr int PB = Store5<TS>::PB, LPR = Store5<TS>::LPR;
    constexpr int RPW = 64 / LPR;             // contiguous rows per wave instruction
    constexpr int NH = kNP / RPW;             // instructions to cover the tile's rows
    constexpr int NL0 = kRows0 / (2 * LPR);   // pieces per lane and row

    extern __shared__ __align__(16) double lds[];
    // per tile and pencil: byte offset of the pencil's row in vL and vR
    __shared__ unsigned s_off[2][kNP];
    // per tile, pencil, halo position and side: the values vL, vR hold there
    __shared__ TO s_halo[2][kNP][6][2];

    const int tid = threadIdx.x;
    const int nwave = blockDim.x >> 6;
    // wave-uniform role (a scalar branch: each wave runs only its own
    // path's barriers)
    const bool producer = __builtin_amdgcn_readfirstlane(tid >> 6) == nwave - 1;
    const int ldj = a.ldj;
    double* const buf0 = lds;
    double* const buf1 = lds + kNP * ldj;

    // ---- chunk and tile range of this workgroup ----
    const int bpc = a.bpc;
    const int ch = blockIdx.x / bpc;
    const int bin = blockIdx.x - ch * bpc;
    const int ntile = a.tiles_per_var * a.nv;
    const int xcd = bin & 7, rank = bin >> 3, nbx = bpc >> 3;
    // workgroup bin's tiles: t_first, t_first + bpc, ... < ntile.  Strided
    // layouts: each round of bpc tiles gives groups of G neighbouring tiles
    // (which share input lines: 16 x-values start 24 bytes into a 128-byte
    // line) to the workgroups of one XCD, so a shared line is fetched into
    // one L2 (y, groups of 1 / 16 / 32: L2 fetch 1370 / 805 / 775 MB per
    // 200^3 x 8 sweep)
    constexpr int G = LAY == 1 ? 32 : 1;
    const int t_first = (G > 1 && nbx % G == 0) ? G * (xcd + 8 * (rank / G)) + rank % G : bin;
    if (t_first >= ntile) return;

    const int jo0 = ch * a.chunk;
    const int jo1 = min(a.out_len, jo0 + a.chunk);
    const int c0 = max(jo0, a.j_beg), c1 = min(jo1 - 1, a.j_end);
    if (!COPY && c0 > c1) return;   // nothing to reconstruct in this chunk
    int lo = c0 + a.J0f - RS, hi = c1 + a.J0f + RS;
    if (COPY) {
        lo = c0 <= c1 ? min(jo0 + a.J0f, lo) : jo0 + a.J0f;
        hi = c0 <= c1 ? max(jo1 - 1 + a.J0f, hi) : jo1 - 1 + a.J0f;
    }
    // Rows [Lb, Lb + nrow) of every pencil go to LDS.  LAY 1 loads rows in
    // groups of kRPI: the window is shifted down (never below 0) so that
    // whole groups stay inside the array.  LAY 0 loads 16-byte pieces, so Lb
    // is even.
    int Lb, nrow;
    if (LAY == 0) {
        Lb = lo & ~1;
        nrow = (hi - Lb + 2) & ~1;
    } else {
        nrow = ((hi - lo + kRPI) / kRPI) * kRPI;
        Lb = min(lo, a.src_len - nrow);
    }

    if (producer) {
        const int l = tid & 63;
        const char* const src = static_cast<const char*>(a.src);
        Piece r2[LAY == 0 ? NH * NL0 : 1];
        Raw r1[LAY == 1 ? kMaxLoads : 1];
        // halo cells of vL and vR, one piece of two faces each
        HP rh[2] = {{0, 0}, {0, 0}};
        auto tile_in = [&](int t) {
            int iv, tv;
            tile_origin(a, t, iv, tv);
            // halo cells: lane -> pencil l/4, piece l&3 = outputs {0,1}, {2,3},
            // {ns-4,ns-3}, {ns-2,ns-1}; byte offset into vL and vR
            unsigned hoff = 0;
            if constexpr (!COPY) {
                const int hq = l >> 2, hpc = l & 3;
                const int hj = hpc < 2 ? 2 * hpc : a.out_len - 4 + 2 * (hpc - 2);
                const bool hin = hj >= jo0 && hj < jo1;
                int pa, pb;
                tile_pencil(a, tv, hq, pa, pb);
                hoff = static_cast<unsigned>((out_base(a, iv, pa, pb) + (hin ? hj : jo0)) * sizeof(TO));
            }
            if constexpr (LAY == 0) {
                // LPR lanes per row (LPR consecutive pieces), rows
                // l/LPR + RPW*h, NL0 pieces each (224 elements >= nrow):
                // buffer loads with the array size as the range, so pieces
                // past the row end read the next row or, at the very end of
                // the array, zero; they land in the LDS row padding (ldj =
                // 224) and are never read
                const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
                    const_cast<char*>(src), 0, static_cast<int>(a.src_bytes), 0x00020000);
#pragma unroll
                for (int h = 0; h < NH; ++h) {
                    int pa, pb;
                    tile_pencil(a, tv, l / LPR + RPW * h, pa, pb);
                    const unsigned voff = static_cast<unsigned>(in_base(a, iv, pa, pb, Lb) * sizeof(TS)) +
                                          unsigned(PB) * (l % LPR);
#pragma unroll
                    for (int m = 0; m < NL0; ++m)
                        r2[NL0 * h + m] = load_piece(rsrc, voff + unsigned(LPR * PB) * m, Piece{});
                }
            } else {
                const int q = l & (kNP - 1), J0 = l / kNP;
                const long long step = kRPI * a.s_j * static_cast<long long>(sizeof(TS));
                const int nm = nrow / kRPI;
                int pa0, pb0, pa, pb;
                tile_pencil(a, tv, 0, pa0, pb0);
                tile_pencil(a, tv, q, pa, pb);
                const long long b0 = in_base(a, iv, pa0, pb0, Lb);
                const unsigned lane_off = static_cast<unsigned>(
                    (in_base(a, iv, pa, pb, Lb) - b0 + J0 * a.s_j) * sizeof(TS));
                // uniform row-group address on the scalar unit; groups past
                // the last one re-read it.  Per-tile copies of the counts keep
                // the per-row selects inside the tile loop.
                unsigned long long bm = reinterpret_cast<unsigned long long>(src) + b0 * sizeof(TS);
                int n = nm;
                long long st = step;
                asm volatile("" : "+s"(n), "+s"(st));
#pragma unroll
                for (int m = 0; m < kMaxLoads; ++m) {
                    r1[m] = *reinterpret_cast<GlobalIn<Raw>*>(bm + lane_off);
                    bm += (m + 1 < n) ? st : 0;
                    asm volatile("" : "+s"(bm));
                }
            }
            // then the halo cells of the tile's output rows, vL and vR at the
            // same 32-bit offset (issued last, so the next tile's address
            // arithmetic never waits on them)
            if constexpr (!COPY) {
                const unsigned long long bl = reinterpret_cast<unsigned long long>(a.outL);
                const unsigned long long br = reinterpret_cast<unsigned long long>(a.outR);
                rh[0] = *reinterpret_cast<GlobalIn<HP>*>(bl + hoff);
                rh[1] = *reinterpret_cast<GlobalIn<HP>*>(br + hoff);
            }
        };
        auto commit = [&](int t, double* buf, int slot) {
            int iv, tv;
            tile_origin(a, t, iv, tv);
            if constexpr (LAY == 0) {
#pragma unroll
                for (int h = 0; h < NH; ++h) {
                    double* row = buf + (l / LPR + RPW * h) * ldj + 2 * (l % LPR);
#pragma unroll
                    for (int m = 0; m < NL0; ++m) put_piece(row + 2 * LPR * m, r2[NL0 * h + m], TS{});
                }
            } else {
                double* row = buf + (l & (kNP - 1)) * ldj + l / kNP;
                int n = nrow / kRPI;
                asm volatile("" : "+s"(n));
#pragma unroll
                for (int m = 0; m < kMaxLoads; ++m) {
                    *row = widen(r1[m]);
                    row += (m + 1 < n) ? kRPI : 0;
                }
            }
            if constexpr (!COPY) {
                const int hq = l >> 2, hpc = l & 3;
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    // positions 0-2: outputs 0-2, 3-5: ns-3..ns-1; pieces {2,3}
                    // and {ns-4,ns-3} carry one reconstructed output each
                    if (hpc != 2) s_halo[slot][hq][hpc < 2 ? 2 * hpc : 2 * hpc - 2][side] = rh[side].x;
                    if (hpc != 1) s_halo[slot][hq][hpc < 2 ? 2 * hpc + 1 : 2 * hpc - 1][side] = rh[side].y;
                }
            }
            if (l < kNP) {
                int pa, pb;
                tile_pencil(a, tv, l, pa, pb);
                s_off[slot][l] = static_cast<unsigned>(out_base(a, iv, pa, pb) * sizeof(TO));
            }
        };
        tile_in(t_first);
        commit(t_first, buf0, 0);
        __syncthreads();
        int k = 0;
        for (int t = t_first; t < ntile; t += bpc, ++k) {
            const int tn = t + bpc;
            if (tn < ntile) {
                tile_in(tn);
                commit(tn, (k & 1) ? buf0 : buf1, (k + 1) & 1);
            }
            __syncthreads();
        }
        return;
    }

    // ---- consumer lanes ----
    const int gq = tid / a.jn;
    const int j = jo0 + (tid - gq * a.jn);
    const bool active = gq < 2 && j < jo1;
    const bool core = active && j >= c0 && j <= c1;
    const bool halo = active && !core;
    const int col = j + a.J0f - Lb;
    // halo lanes run the stencil on a clamped window and keep their value
    const int colc = ORDER == 5 ? min(max(col, 2), nrow - 3) : min(max(col, RS), nrow - 1 - RS);
    const int hpos = j < a.j_beg ? j - (a.j_beg - 3) : j - (a.j_end + 1) + 3;
    std::conditional_t<ORDER == 7, std::conditional_t<MODE == MODE_JS, Coef7, Gen7>,
                       std::conditional_t<MODE == MODE_JS, Coef5, Gen5>>
        k5;
    {
        int jc = j - 3;
        jc = jc < 0 ? 0 : (jc > a.c.nc - 1 ? a.c.nc - 1 : jc);
        if constexpr (ORDER == 5 && MODE == MODE_JS) {
            load_coef5(a.c, jc, k5);
        } else if constexpr (ORDER == 7 && MODE == MODE_JS) {
            load_coef7(a.c, jc, k5);
        } else {
            load_cf(a.c, jc, k5.f);
            k5.b.load(a.c, jc);
        }
    }
    const double eps = a.eps;
    const unsigned joff = static_cast<unsigned>(j * sizeof(TO));
    // vL and vR as buffers (offsets below 4 GiB, checked by the planner)
    const int out_bytes = static_cast<int>(static_cast<unsigned long long>(a.o_i) * a.nv * sizeof(TO));
    const __amdgpu_buffer_rsrc_t rL = __builtin_amdgcn_make_buffer_rsrc(a.outL, 0, out_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rR = __builtin_amdgcn_make_buffer_rsrc(a.outR, 0, out_bytes, 0x00020000);
    constexpr int NEWTON = sizeof(TO) == sizeof(double) ? 2 : 1;

    __syncthreads();
    int k = 0;
    for (int t = t_first; t < ntile; t += bpc, ++k) {
        const int slot = k & 1;
        // group g: pencils g, g+2, g+4, ...
        const double* const cur = ((k & 1) ? buf1 : buf0) + gq * ldj;
        if (active) {
            bool allok = true;
            if constexpr (ORDER == 5) {
#pragma unroll 1
                for (int qq = 0; qq < kNPG; ++qq) {
                    const int q = 2 * qq + gq;
                    const double* row = cur + 2 * qq * ldj + colc;
                    const double v0 = row[-2], v1 = row[-1], v2 = row[0], v3 = row[1], v4 = row[2];
                    double oL, oR;
                    if constexpr (MODE == MODE_JS) {
                        allok &= cell5<NEWTON>(v0, v1, v2, v3, v4, k5, eps, oL, oR);
                    } else {
                        // the general kernel's per-family arithmetic, which
                        // carries every input range itself
                        const double w[5] = {v0, v1, v2, v3, v4};
                        weno5<MODE, double>(w, k5.f, k5.b, eps, a.teno_ct, oL, oR);
                    }
                    if constexpr (MP) {
                        // every lane limits (halo lanes replace their value below)
                        const double w[5] = {v0, v1, v2, v3, v4};
                        mp_limit<double>(w, oL, oR);
                    }
                    TO wL = static_cast<TO>(oL), wR = static_cast<TO>(oR);
                    if (halo) {
                        if constexpr (COPY) {
                            wL = wR = static_cast<TO>(row[col - colc]);
                        } else {
                            wL = s_halo[slot][q][hpos][0];
                            wR = s_halo[slot][q][hpos][1];
                        }
                    }
                    const unsigned off = s_off[slot][q] + joff;
                    store_face(wL, rL, off);
                    store_face(wR, rR, off);
                }
            } else {
                // WENO7 lanes take two pencils per step: two independent
                // dependency chains at 2 waves per SIMD
#pragma unroll 1
                for (int qq0 = 0; qq0 < kNPG; qq0 += 2) {
                    double oL[2], oR[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const double* row = cur + 2 * (qq0 + e) * ldj + colc;
                        double w[7];
#pragma unroll
                        for (int u = 0; u < 7; ++u) w[u] = row[u - 3];
                        if constexpr (MODE == MODE_JS)
                            allok &= cell7<NEWTON>(w, k5, eps, oL[e], oR[e]);
                        else
                            weno7<MODE, double>(w, k5.f, k5.b, eps, a.teno_ct, oL[e], oR[e]);
                        if constexpr (MP) {
                            const double wm[5] = {w[1], w[2], w[3], w[4], w[5]};
                            mp_limit<double>(wm, oL[e], oR[e]);
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int q = 2 * (qq0 + e) + gq;
                        const double* row = cur + 2 * (qq0 + e) * ldj + colc;
                        TO wL = static_cast<TO>(oL[e]), wR = static_cast<TO>(oR[e]);
                        if (halo) {
                            if constexpr (COPY) {
                                wL = wR = static_cast<TO>(row[col - colc]);
                            } else {
                                wL = s_halo[slot][q][hpos][0];
                                wR = s_halo[slot][q][hpos][1];
                            }
                        }
                        const unsigned off = s_off[slot][q] + joff;
                        store_face(wL, rL, off);
                        store_face(wR, rR, off);
                    }
                }
            }
            if (!allok && core) {
#pragma unroll 1
                for (int qq = 0; qq < kNPG; ++qq) {
                    const int q = 2 * qq + gq;
                    const double* row = cur + 2 * qq * ldj + colc;
                    double oL, oR;
                    // inline: an out-of-line call takes the coefficient
                    // struct's address, and the compiler then keeps the
                    // lane's 27 (56) coefficients in scratch: 240 bytes per
                    // lane written at the start of every launch (63 MB of
                    // HBM writes per WENO5 sweep at 200^3 x 8), and for
                    // WENO7 reloaded every tile
                    if constexpr (ORDER == 7) {
                        double w[7];
#pragma unroll
                        for (int u = 0; u < 7; ++u) w[u] = row[u - 3];
                        if constexpr (MODE == MODE_JS) cell7_slow_inl(w, k5, eps, oL, oR);
                    } else if constexpr (MODE == MODE_JS) {
                        cell5_slow_inl(row[-2], row[-1], row[0], row[1], row[2], k5, eps, oL, oR);
                    }
                    if constexpr (MP) {
                        const double w[5] = {row[-2], row[-1], row[0], row[1], row[2]};
                        mp_limit<double>(w, oL, oR);
                    }
                    const unsigned off = s_off[slot][q] + joff;
                    store_face(static_cast<TO>(oL), rL, off);
                    store_face(static_cast<TO>(oR), rR, off);
                }
            }
        }
        __syncthreads();
    }
}

// The WENO5-JS kernel: 128 VGPRs, 4 waves per SIMD (two workgroups per CU).
template <typename TS, int LAY, bool COPY, bool MP>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, kThreads5), amdgpu_waves_per_eu(4)))
sweep5_kernel(const TileArgs a)
{
    sweep_body<5, MODE_JS, TS, LAY, COPY, MP>(a);
}

// WENO5's other weight branches (TENO, WENO-Z, mapped) for FP64 state
// without the fused limiter, at the WENO5-JS kernel's occupancy.
template <int MODE, int LAY, bool COPY>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, kThreads5), amdgpu_waves_per_eu(4)))
sweep5_branch_kernel(const TileArgs a)
{
    sweep_body<5, MODE, double, LAY, COPY, false>(a);
}

// The WENO7 kernel: the 56 coefficients of a lane take 112 VGPRs, so it
// runs at 2 waves per SIMD (one workgroup per CU).  JS for every storage
// type, with or without the fused limiter; the other weight branches for
// FP64 state without it.
template <int MODE, typename TS, int LAY, bool COPY, bool MP>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, kThreads5), amdgpu_waves_per_eu(2)))
sweep7_kernel(const TileArgs a)
{
    sweep_body<7, MODE, TS, LAY, COPY, MP>(a);
}

// Host side: decides whether a WENO5 or WENO7 launch (JS for every storage
// type; TENO, WENO-Z and mapped for FP64 state without the fused limiter)
// runs on these kernels and sets its tile shape.  Returns the layout (0 or 1), or -1 when the general sweep
// kernel must be used (other orders/modes, a fused limiter range other than
// the reconstructed range, one-sided output, rows shorter than 96 outputs,
// odd row lengths or misaligned arrays, arrays too large for 32-bit
// offsets).
int plan_sweep5(TileArgs& a, Launch& l, const void* src)
{
    if (l.op != OP_WENO || (l.order != 5 && l.order != 7) || a.sides != 3) return -1;
    if (l.mode != MODE_JS) {
        // the other weight branches: FP64 state, no fused limiter
        if (l.dtype != DT_F64 || l.mp ||
            (l.mode != MODE_TENO && l.mode != MODE_WENOZ && l.mode != MODE_MAPPED))
            return -1;
    }
    const int RS = l.order / 2;
    if (l.mp && (a.mp_beg != a.j_beg || a.mp_end != a.j_end)) return -1;
    int in_b, out_b;
    switch (l.dtype) {
    case DT_F64: in_b = 8, out_b = 8; break;
    case DT_F32: in_b = 4, out_b = 4; break;
    case DT_BF16: in_b = 2, out_b = 4; break;
    default: return -1;
    }
    const int ns = a.out_len;
    if (a.j_beg != 3 || a.j_end != ns - 4 || (ns & 1) || ((a.o_k | a.o_l | a.o_i) & 1)) return -1;
    // halo cells are read in pieces of two faces
    if ((reinterpret_cast<uintptr_t>(a.outL) | reinterpret_cast<uintptr_t>(a.outR)) & (2 * out_b - 1)) return -1;
    if (static_cast<unsigned long long>(a.o_i) * a.nv * out_b > 0xffffffffull) return -1;
    int lay;
    if (a.s_j == 1) {
        // pieces of two elements
        if ((reinterpret_cast<uintptr_t>(src) & (2 * in_b - 1)) || ((a.src_off | a.s_a | a.s_b | a.s_i) & 1))
            return -1;
        // 32-bit buffer offsets
        a.src_bytes = static_cast<unsigned long long>(a.s_i) * a.nv * in_b;
        if (a.src_bytes > 0xffffffffull) return -1;
        lay = 0;
    } else if (a.s_a == 1) {
        if ((a.s_b + kNP + kRPI * a.s_j) * static_cast<long long>(in_b) >= (1LL << 31)) return -1;
        lay = 1;
    } else {
        return -1;
    }
    for (int nch = (ns + kMaxOut - 1) / kMaxOut; nch <= 64; ++nch) {
        const int chunk = ((ns + nch - 1) / nch + 1) & ~1;   // even: 16-byte output pieces
        int rows = 0;
        bool fits = true;
        for (int c = 0; c < nch && fits; ++c) {
            const int jo0 = c * chunk, jo1 = std::min(ns, jo0 + chunk);
            if (jo0 >= jo1) break;
            const int c0 = std::max(jo0, a.j_beg), c1 = std::min(jo1 - 1, a.j_end);
            if (c0 > c1 && !a.copy_halo) continue;
            int lo = c0 + a.J0f - RS, hi = c1 + a.J0f + RS;
            if (a.copy_halo) {
                lo = c0 <= c1 ? std::min(jo0 + a.J0f, lo) : jo0 + a.J0f;
                hi = c0 <= c1 ? std::max(jo1 - 1 + a.J0f, hi) : jo1 - 1 + a.J0f;
            }
            int nrow;
            if (lay == 0) {
                nrow = (hi - (lo & ~1) + 2) & ~1;
                if (nrow > kRows0) fits = false;
            } else {
                nrow = ((hi - lo + kRPI) / kRPI) * kRPI;
                if (nrow > a.src_len || nrow > kRPI * kMaxLoads) fits = false;
            }
            rows = std::max(rows, nrow);
        }
        if (chunk > kMaxOut) fits = false;
        if (!fits) continue;
        if (chunk < 96) return -1;   // short rows: the general kernel uses its lanes better
        a.chunk = chunk;
        a.chunks = nch;
        a.jn = chunk;
        a.np = kNP;
        a.groups = 2;
        a.npg = kNPG;
        a.ldj = lay == 0 ? kRows0 : (rows | 1);
        a.threads = kThreads5;
        const long long pen = static_cast<long long>(a.na) * a.nb;
        a.tiles_per_var = static_cast<int>((pen + kNP - 1) / kNP);
        return lay;
    }
    return -1;
}

namespace {
typedef void (*Sweep5Fn)(const TileArgs);
template <int ORDER, int MODE, typename TS, int LAY, bool COPY, bool MP>
Sweep5Fn kernel_of()
{
    if constexpr (ORDER == 7)
        return sweep7_kernel<MODE, TS, LAY, COPY, MP>;
    else if constexpr (MODE != MODE_JS)
        return sweep5_branch_kernel<MODE, LAY, COPY>;
    else
        return sweep5_kernel<TS, LAY, COPY, MP>;
}
template <int ORDER, int MODE, typename TS, bool MP>
Sweep5Fn pick5(int lay, bool copy)
{
    return lay == 0 ? (copy ? kernel_of<ORDER, MODE, TS, 0, true, MP>() : kernel_of<ORDER, MODE, TS, 0, false, MP>())
                    : (copy ? kernel_of<ORDER, MODE, TS, 1, true, MP>() : kernel_of<ORDER, MODE, TS, 1, false, MP>());
}
template <int ORDER>
Sweep5Fn pick5(int dtype, int lay, bool copy, bool mp)
{
    auto by_type = [&](auto ts) {
        using TS = decltype(ts);
        return mp ? pick5<ORDER, MODE_JS, TS, true>(lay, copy) : pick5<ORDER, MODE_JS, TS, false>(lay, copy);
    };
    return dtype == DT_F64 ? by_type(double{}) : dtype == DT_F32 ? by_type(float{}) : by_type(bf16_t{});
}
// the other weight branches (FP64, no fused limiter)
template <int ORDER>
Sweep5Fn pick_mode(int mode, int lay, bool copy)
{
    switch (mode) {
    case MODE_TENO: return pick5<ORDER, MODE_TENO, double, false>(lay, copy);
    case MODE_WENOZ: return pick5<ORDER, MODE_WENOZ, double, false>(lay, copy);
    default: return pick5<ORDER, MODE_MAPPED, double, false>(lay, copy);
    }
}
}  // namespace

hipError_t launch_sweep5(TileArgs a, int order, int mode, int lay, int dtype, int mp, hipStream_t st)
{
    if (a.threads != kThreads5 || 2 * a.jn > kConsumers || a.ldj > 0xffff) return hipErrorInvalidConfiguration;
    if (mode != MODE_JS && (dtype != DT_F64 || mp)) return hipErrorInvalidConfiguration;
    const size_t lds = 2 * static_cast<size_t>(kNP) * a.ldj * sizeof(double);
    const bool copy = a.copy_halo != 0;
    const Sweep5Fn kern = mode != MODE_JS ? (order == 7 ? pick_mode<7>(mode, lay, copy) : pick_mode<5>(mode, lay, copy))
                          : order == 7    ? pick5<7>(dtype, lay, copy, mp != 0)
                                          : pick5<5>(dtype, lay, copy, mp != 0);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kThreads5, lds) != hipSuccess || occ < 1) occ = 1;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int ntile = a.tiles_per_var * a.nv;
    int bpc = cus * occ;
    const int need = (ntile + 1) / 2;
    if (bpc > need) bpc = need;
    bpc = (bpc + 7) & ~7;
    a.bpc = bpc;
    hipLaunchKernelGGL(kern, dim3(bpc * a.chunks), dim3(kThreads5), lds, st, a);
    return hipGetLastError();
}

}  // namespace wm
