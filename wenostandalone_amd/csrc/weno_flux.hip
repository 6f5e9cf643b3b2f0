// Consumer step of one sweep: the flux divergence over the reconstructed
// faces (no reference counterpart -- the harness stops at vL/vR,
// inputs/weno_standalone.f90:343,363,383; the pairing of faces at an
// interface follows MFC's Riemann solvers, qL(j) with qR(j+1) at j+1/2).
//
// Per variable the scalar flux f(q) = a q + b q^2/2 (linear advection,
// Burgers or both) goes through a Rusanov (local Lax-Friedrichs) interface
// flux,
//   F(qL, qR) = (f(qL) + f(qR))/2 - max(|f'(qL)|, |f'(qR)|) (qR - qL)/2,
// and every cell whose two interfaces lie in the reconstructed core
// (j = 4 .. ns-5) gets
//   rhs(j) = -(F(vL(j), vR(j+1)) - F(vL(j-1), vR(j))) * inv_dx
// (or rhs += that with WENO_ACCUMULATE), in the face layout of the sweep.
//
// weno_reconstruct_flux fuses the reconstruction into it (fused_flux_kernel
// below): each lane reconstructs the faces of one cell, takes its right
// neighbour's vR by a lane shift, forms F(j+1/2), and takes F(j-1/2) from its
// left neighbour by another; the faces never reach HBM.  The sweep reads the
// state once and writes rhs once, where reconstruction followed by a
// separate flux pass reads the state, writes both face arrays, reads them
// back and writes rhs.  weno_flux_divergence is that separate pass, for
// faces that already exist (faces_flux_kernel).
#include "weno_device.h"
#include "weno_cells.h"
#include "weno_tiles.h"
#include "weno_capi_internal.h"

#include <algorithm>
#include <cstring>

namespace wm {
namespace {

struct FluxArgs {
    TileArgs a;       // input/output maps, coefficients, eps, teno_CT, tiles
    const double* inL;   // standalone pass: faces to read
    const double* inR;
    double fa, fb, inv_dx;
    int accumulate;
    int lo;           // first input J of a tile row (fused)
    int nJ;           // tile row length (fused)
    int nf;           // faces per row: cells 3 .. ns-4
    int rows;         // fused: rows whose face cells are laid end to end
};

__device__ __forceinline__ double rusanov(double qL, double qR, double a, double b)
{
    const double fL = qL * fma(0.5 * b, qL, a), fR = qR * fma(0.5 * b, qR, a);
    const double s = fmax(fabs(fma(b, qL, a)), fabs(fma(b, qR, a)));
    return 0.5 * (fL + fR) - 0.5 * s * (qR - qL);
}

template <int ORDER, int MODE>
__device__ __forceinline__ void cell_faces(const double* v, const Cf<ORDER, double>& f, const BetaRegs<ORDER>& B,
                                           double eps, double ct, double& oL, double& oR)
{
    if constexpr (ORDER == 3)
        weno3<MODE, double>(v, f, B, eps, ct, oL, oR);
    else if constexpr (ORDER == 5)
        weno5<MODE, double>(v, f, B, eps, ct, oL, oR);
    else
        weno7<MODE, double>(v, f, B, eps, ct, oL, oR);
}

// Interface fluxes from neighbouring lanes.  Lane i of wave w holds cell
// j = 3 + 62 w + i: lanes 1..62 own the rhs of their cells; lanes 0 and 63
// only supply the faces of the cells on either side of the wave's range, so
// both interfaces of every owned cell are formed inside one wave with two
// shuffles and no LDS round trip.
constexpr int kCellsPerWave = 62;

// Every lane forms F(j+1/2) = F(vL(j), vR(j+1)) with its right neighbour's
// vR, and takes F(j-1/2) from its left neighbour: one flux evaluation and
// two lane shifts per cell.
__device__ __forceinline__ void emit_rhs(const FluxArgs& fx, double* out, double oL, double oR, bool own)
{
    const double Fp = rusanov(oL, from_next_lane(oR), fx.fa, fx.fb);
    const double Fm = from_prev_lane(Fp);
    if (own) {
        const double d = -(Fp - Fm) * fx.inv_dx;
        *out = fx.accumulate ? *out + d : d;
    }
}

#ifndef WENO_FLUX_SLOTS
#define WENO_FLUX_SLOTS 28
#endif
#ifndef WENO_FLUX_MAXNP
#define WENO_FLUX_MAXNP 8
#endif
constexpr int kFluxSlots = WENO_FLUX_SLOTS;   // producer registers (doubles per lane)
#ifndef WENO_FLUX_CONSUMERS
#define WENO_FLUX_CONSUMERS 7
#endif
constexpr int kFluxConsumers = WENO_FLUX_CONSUMERS;   // consumer waves per workgroup (at most)
constexpr int kFluxMaxNP = WENO_FLUX_MAXNP;   // pencils per tile

// Fused reconstruction + flux divergence.  One workgroup = up to 7 consumer
// waves + 1 producer wave, persistent over tiles of np pencils (of one
// plane) x whole rows (input J = lo .. lo+nJ-1), double-buffered in LDS as
// [pencil][J] with an odd row stride (ldj).
//   producer : loads tile k+1 into registers, then into LDS; it never
//              stores to global memory.  Contiguous rows (x sweeps and
//              reshaped sources): the tile's span, 64 consecutive elements
//              per instruction.  Strided rows (y, z of the field): 64/np rows
//              per instruction, np consecutive pencils each.
//   consumers: the face cells of R rows, end to end, in segments of 62 (one
//              per wave, or several per wave for long rows); a tile of np
//              pencils takes np / R passes.  With one segment per wave the
//              per-cell coefficients are loaded once for the whole launch,
//              otherwise once per segment and tile.  At 200^3 (206 face
//              cells per row) R = 2 rows fill 7 waves to 95 %.
// One barrier per tile hands the buffers over.  The register allocation is
// the larger of the two roles, not their sum.
template <int ORDER, int MODE, bool MP, int WPE>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, 64 * (kFluxConsumers + 1)), amdgpu_waves_per_eu(WPE)))
fused_flux_kernel(const FluxArgs fx)
{
    const TileArgs& a = fx.a;
    constexpr int RW = ORDER / 2;
    constexpr int R5 = MP ? (RW > 2 ? RW : 2) : RW;   // stencil radius
    constexpr int W = 2 * R5 + 1;
    extern __shared__ __align__(16) double lds[];
    const int np = a.np, nJ = fx.nJ, ld = a.ldj, ns = a.out_len;
    double* const buf0 = lds;
    double* const buf1 = lds + np * ld;
    const int tid = threadIdx.x;
    const int nwave = blockDim.x >> 6;
    const bool producer = __builtin_amdgcn_readfirstlane(tid >> 6) == nwave - 1;
    const bool rows = a.s_j == 1;

    // tiles of this workgroup: t_first, t_first + tstep, ... < t_end.
    // Strided rows share input lines with the neighbouring tiles: one
    // contiguous range per XCD keeps them in one L2.
    const int ntile = a.tiles_per_var * a.nv;
    int t_first = blockIdx.x, t_end = ntile, tstep = gridDim.x;
    if (!rows && (gridDim.x & 7) == 0) {
        const int xcd = blockIdx.x & 7, rank = blockIdx.x >> 3;
        const int per = ntile >> 3, rem = ntile & 7;
        t_first = xcd * per + min(xcd, rem) + rank;
        t_end = (xcd + 1) * per + min(xcd + 1, rem);
        tstep = gridDim.x >> 3;
    }
    if (t_first >= t_end) return;

    if (producer) {
        double r[kFluxSlots];
        tile_commit(a, nJ, ld, tile_issue(a, fx.lo, nJ, t_first, r), r, buf0);
        __syncthreads();
        int k = 0;
        for (int t = t_first; t < t_end; t += tstep, ++k) {
            const int tn = t + tstep;
            if (tn < t_end) tile_commit(a, nJ, ld, tile_issue(a, fx.lo, nJ, tn, r), r, (k & 1) ? buf0 : buf1);
            __syncthreads();
        }
        return;
    }

    // Consumers.  The face cells j = 3 .. ns-4 of R rows are laid end to
    // end as positions p = r nf + (j - 3); wave w, lane i holds position
    // 62 s + i - 1 of its segment s.  Lanes 1..62 own their position; lanes
    // 0 and 63 only supply faces to their neighbours.  The face-only cells
    // j = 3 and j = ns-4 separate the rows, so every owned cell's two
    // neighbours are in its own row.
    const int i = tid & 63, cw = tid >> 6, nc = nwave - 1;
    const int nf = ns - 6, R = fx.rows, npos = R * nf;
    const int nseg = (npos + kCellsPerWave - 1) / kCellsPerWave;
    const bool one = nseg <= nc;
    double* const rhs = static_cast<double*>(a.outL);
    const long long o_a = a.a_is_k ? a.o_k : a.o_l, o_b = a.a_is_k ? a.o_l : a.o_k;
    // WENO5-JS takes the shared-reciprocal cell5; the other families the
    // general per-family arithmetic
    constexpr bool kJS5 = ORDER == 5 && MODE == MODE_JS;
    Cf<kJS5 ? 3 : ORDER, double> f;
    BetaRegs<kJS5 ? 3 : ORDER> breg;
    Coef5 k5;
    int r = 0, j = 3;
    bool own = false;
    auto locate = [&](int seg) {
        const int pos = kCellsPerWave * seg + i - 1;
        const int pc = min(max(pos, 0), npos - 1);
        r = pc / nf;
        j = 3 + pc - r * nf;
        own = i >= 1 && i <= kCellsPerWave && pos < npos && j >= 4 && j <= ns - 5;
        if constexpr (kJS5) {
            load_coef5(a.c, j - 3, k5);
        } else {
            load_cf(a.c, j - 3, f);
            breg.load(a.c, j - 3);
        }
    };
    if (one) locate(cw);
    __syncthreads();
    int k = 0;
    for (int t = t_first; t < t_end; t += tstep, ++k) {
        const double* const cur = (k & 1) ? buf1 : buf0;
        int iv, pa0, pb, cnt;
        plane_tile(a, t, iv, pa0, pb, cnt);
        const long long obase = pa0 * o_a + pb * o_b + iv * a.o_i;
        for (int seg = cw; seg < nseg; seg += nc) {
            if (!one) locate(seg);
            const int col = j + a.J0f - fx.lo;
            for (int q0 = 0; q0 < cnt; q0 += R) {
                const int q = q0 + r;
                const double* row = cur + min(q, cnt - 1) * ld + col;
                double v[W];
#pragma unroll
                for (int u = 0; u < W; ++u) v[u] = row[u - R5];
                double oL, oR;
                if constexpr (kJS5) {
                    const double* w = v + (R5 - 2);
                    const bool ok = cell5<2>(w[0], w[1], w[2], w[3], w[4], k5, a.eps, oL, oR);
                    // a wave-wide (scalar) test keeps the rare slow path a
                    // branch instead of arithmetic every lane runs
                    if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
                        // opaque copies keep the divisions inside the branch
                        double x0 = w[0], x1 = w[1], x2 = w[2], x3 = w[3], x4 = w[4];
                        asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4));
                        if (!ok) cell5_slow_inl(x0, x1, x2, x3, x4, k5, a.eps, oL, oR);
                    }
                } else {
                    cell_faces<ORDER, MODE>(v + (R5 - RW), f, breg, a.eps, a.teno_ct, oL, oR);
                }
                if constexpr (MP) mp_limit<double>(v + (R5 - 2), oL, oR);
                emit_rhs(fx, rhs + obase + q * o_a + j, oL, oR, own && q < cnt);
            }
        }
        __syncthreads();
    }
}

// Separate pass over existing faces: one wave per 62 cells of one pencil
// row, faces read straight from HBM (contiguous along the row).
__global__ void __launch_bounds__(256) faces_flux_kernel(const FluxArgs fx)
{
    const TileArgs& a = fx.a;
    const int ns = a.out_len;
    const int wpr = (ns - 8 + kCellsPerWave - 1) / kCellsPerWave;   // waves per row
    const long long nrow = static_cast<long long>(a.na) * a.nb * a.nv;
    const long long gw = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int i = threadIdx.x & 63;
    if (gw >= nrow * wpr) return;
    const long long row = gw / wpr;
    const int w = static_cast<int>(gw - row * wpr);
    const int iv = static_cast<int>(row / (static_cast<long long>(a.na) * a.nb));
    const long long P = row - static_cast<long long>(iv) * a.na * a.nb;
    const int pb = static_cast<int>(P / a.na), pa = static_cast<int>(P - static_cast<long long>(pb) * a.na);
    const int kk = a.a_is_k ? pa : pb, ll = a.a_is_k ? pb : pa;
    const long long o = kk * a.o_k + ll * a.o_l + iv * a.o_i;
    const int j = 3 + kCellsPerWave * w + i;
    const bool face = j <= ns - 4;
    const double oL = face ? fx.inL[o + j] : 0.0;
    const double oR = face ? fx.inR[o + j] : 0.0;
    const bool own = face && i >= 1 && i <= kCellsPerWave && j >= 4 && j <= ns - 5;
    emit_rhs(fx, static_cast<double*>(a.outL) + o + j, oL, oR, own);
}

typedef void (*FluxFn)(const FluxArgs);

// Longest fused row: a tile row of nJ <= 64 * kFluxSlots inputs fits one
// pencil per tile (nJ = core + 2R <= core + 6).
constexpr int kFluxMaxCore = 1024;
static_assert(kFluxMaxCore + 6 <= 64 * kFluxSlots, "a fused row must fit the producer's registers");
static_assert(64 * kFluxSlots <= 2048, "span element indices must stay exact in FP32");

template <int ORDER, bool MP>
FluxFn pick_fused(int mode)
{
    constexpr int WPE = ORDER == 7 ? 2 : 4;   // 256 / 128 VGPRs
    switch (mode) {
    case MODE_JS: return fused_flux_kernel<ORDER, MODE_JS, MP, WPE>;
    case MODE_MAPPED: return fused_flux_kernel<ORDER, MODE_MAPPED, MP, WPE>;
    case MODE_WENOZ: return fused_flux_kernel<ORDER, MODE_WENOZ, MP, WPE>;
    case MODE_TENO:
        if constexpr (ORDER != 3) return fused_flux_kernel<ORDER, MODE_TENO, MP, WPE>;
        return fused_flux_kernel<ORDER, MODE_NONE, MP, WPE>;
    default: return fused_flux_kernel<ORDER, MODE_NONE, MP, WPE>;
    }
}

FluxFn pick_fused(int order, int mode, bool mp)
{
    if (order == 3) return mp ? pick_fused<3, true>(mode) : pick_fused<3, false>(mode);
    if (order == 5) return mp ? pick_fused<5, true>(mode) : pick_fused<5, false>(mode);
    return mp ? pick_fused<7, true>(mode) : pick_fused<7, false>(mode);
}

int cu_count()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus;
}

// waves of 62 cells covering the rows' rhs range 4 .. ns-5
int waves_per_row(int ns) { return (ns - 8 + kCellsPerWave - 1) / kCellsPerWave; }

hipError_t launch_fused(int order, int mode, bool mp, FluxArgs fx, hipStream_t st)
{
    TileArgs& a = fx.a;
    const FluxFn kern = pick_fused(order, mode, mp);
    a.np = tile_pencils(fx.nJ, a.s_a, a.s_j == 1, kFluxSlots, kFluxMaxNP);
    if (a.np < 1) return hipErrorInvalidConfiguration;
    // rows laid end to end: as many as kFluxConsumers waves of 62 cells cover
    const int nf = a.out_len - 6;
    fx.rows = std::max(1, std::min(a.np, kFluxConsumers * kCellsPerWave / nf));
    const int nseg = (fx.rows * nf + kCellsPerWave - 1) / kCellsPerWave;
    a.threads = 64 * (std::min(nseg, kFluxConsumers) + 1);
    a.ldj = fx.nJ | 1;
    a.chunks = (a.na + a.np - 1) / a.np;   // tiles per plane
    a.tiles_per_var = a.chunks * a.nb;
    const size_t lds = 2 * static_cast<size_t>(a.np) * a.ldj * sizeof(double);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    int occ = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, a.threads, lds) != hipSuccess || occ < 1) occ = 1;
    const long long ntile = static_cast<long long>(a.tiles_per_var) * a.nv;
    long long grid = std::min<long long>(ntile, 1LL * cu_count() * occ);
    if (a.s_j != 1 && grid >= 8) grid &= ~7LL;   // whole XCD ranges
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(std::max<long long>(1, grid))), dim3(a.threads), lds, st, fx);
    return hipGetLastError();
}

hipError_t launch_faces(FluxArgs fx, hipStream_t st)
{
    const TileArgs& a = fx.a;
    const long long waves = static_cast<long long>(a.na) * a.nb * a.nv * waves_per_row(a.out_len);
    const long long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(faces_flux_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, fx);
    return hipGetLastError();
}

int check_flux(const weno_grid* g, int dir, const weno_flux* fx, int flags, int allowed)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir))) return rc;
    if (!fx) return fail(WENO_E_INVALID, "flux parameters are NULL");
    if (flags & ~allowed) return fail(WENO_E_INVALID, "unknown flag bits 0x%x", flags & ~allowed);
    const Dims d = dims_of(g, dir);
    if (d.ns - 6 < 3)
        return fail(WENO_E_INVALID, "sweep of %d points has no cell with both interfaces in the core", d.ns);
    return WENO_OK;
}

}  // namespace
}  // namespace wm

using namespace wm;

extern "C" {

int weno_reconstruct_flux(const weno_grid* g, int order, int dir, int src_kind, const void* src,
                          const weno_coeffs* c, const weno_flux* fx, void* rhs, int flags)
{
    int rc;
    if ((rc = check_flux(g, dir, fx, flags, WENO_FUSE_MP | WENO_ACCUMULATE))) return rc;
    if (order != 3 && order != 5 && order != 7)
        return fail(WENO_E_INVALID, "order must be 3, 5 or 7 (got %d)", order);
    if (src_kind != WENO_SRC_FIELD && src_kind != WENO_SRC_RESHAPED)
        return fail(WENO_E_INVALID, "src_kind must be WENO_SRC_FIELD or WENO_SRC_RESHAPED");
    const Dims d = dims_of(g, dir);
    const int nf = d.ns - 6;
    if (nf > kFluxMaxCore)
        return fail(WENO_E_INVALID, "fused flux: sweeps of at most %d core points (got %d)", kFluxMaxCore, nf);
    if ((rc = ensure_init())) return rc;
    FluxArgs F;
    std::memset(&F, 0, sizeof F);
    TileArgs& a = F.a;
    const void* s;
    void* o;
    if ((rc = resolve(src, &s, "src", src_elems(d, src_kind) * sizeof(double))) ||
        (rc = resolve_mut(rhs, &o, "rhs", face_elems(d) * sizeof(double))))
        return rc;
    if ((rc = resolve_coeffs(c, a.c, nf))) return rc;
    input_map(d, dir, src_kind, a);
    output_map(d, d.ns, a);
    a.src = s;
    a.outL = o;
    a.J0f = 3;
    a.eps = g->weno_eps;
    a.teno_ct = g->teno_CT;
    const bool mp = (flags & WENO_FUSE_MP) != 0;
    const int R = mp ? std::max(order / 2, 2) : order / 2;
    // faces of cells 3 .. ns-4 need inputs J = 6-R .. ns-1+R
    F.lo = 6 - R;
    F.nJ = (d.ns - 1 + R) - F.lo + 1;
    F.nf = nf;
    F.fa = fx->a;
    F.fb = fx->b;
    F.inv_dx = fx->inv_dx;
    F.accumulate = (flags & WENO_ACCUMULATE) ? 1 : 0;
    const int mode = mode_for(g, order);
    return finish(launch_fused(order, mode, mp, F, stream()), "reconstruct_flux");
}

int weno_flux_divergence(const weno_grid* g, int dir, const void* vL, const void* vR, const weno_flux* fx,
                         void* rhs, int flags)
{
    int rc;
    if ((rc = check_flux(g, dir, fx, flags, WENO_ACCUMULATE))) return rc;
    const Dims d = dims_of(g, dir);
    const int nf = d.ns - 6;
    if ((rc = ensure_init())) return rc;
    FluxArgs F;
    std::memset(&F, 0, sizeof F);
    TileArgs& a = F.a;
    const void *l, *r;
    void* o;
    const size_t fb = face_elems(d) * sizeof(double);
    if ((rc = resolve(vL, &l, "vL", fb)) || (rc = resolve(vR, &r, "vR", fb)) || (rc = resolve_mut(rhs, &o, "rhs", fb)))
        return rc;
    input_map(d, dir, WENO_SRC_FIELD, a);
    output_map(d, d.ns, a);
    a.outL = o;
    F.inL = static_cast<const double*>(l);
    F.inR = static_cast<const double*>(r);
    F.nf = nf;
    F.fa = fx->a;
    F.fb = fx->b;
    F.inv_dx = fx->inv_dx;
    F.accumulate = (flags & WENO_ACCUMULATE) ? 1 : 0;
    return finish(launch_faces(F, stream()), "flux_divergence");
}

}  // extern "C"
