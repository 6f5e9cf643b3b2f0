// extern "C" entry point of the WENO5/7-JS adjoint (weno_reconstruct_vjp).
#include "weno_capi_internal.h"

#include <cstring>

namespace wm {
hipError_t launch_vjp(TileArgs a, int order, double* grad, hipStream_t st);
hipError_t launch_vjp5_tiles(TileArgs a, double* grad, hipStream_t st);
int vjp_max_threads();
int vjp_max_slots();
int vjp_max_np();
}  // namespace wm

using namespace wm;

extern "C" int weno_reconstruct_vjp(const weno_grid* g, int order, int dir, int src_kind, const void* src,
                                    const weno_coeffs* c, const void* gL, const void* gR, void* gv)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir))) return rc;
    if (order != 5 && order != 7) return fail(WENO_E_INVALID, "the adjoint covers WENO5 and WENO7 (got order %d)", order);
    if (!g->wenojs) return fail(WENO_E_INVALID, "the adjoint covers the JS weights only (wenojs must be set)");
    if (src_kind != WENO_SRC_FIELD && src_kind != WENO_SRC_RESHAPED)
        return fail(WENO_E_INVALID, "src_kind must be WENO_SRC_FIELD or WENO_SRC_RESHAPED");
    if ((rc = ensure_init())) return rc;
    const Dims d = dims_of(g, dir);
    TileArgs a;
    std::memset(&a, 0, sizeof a);
    {
        // the longest sweeps whose one-pencil tile fits LDS: WENO5 the
        // workgroup limit (512), WENO7 479 core points (its 33-entry
        // coefficient table per lane takes most of the 158 KB)
        const int max_core = order == 5 ? vjp_max_threads() : 479;
        if (d.ns - 6 > max_core)
            return fail(WENO_E_INVALID, "WENO%d adjoint sweep length %d exceeds %d core points", order, d.ns - 6,
                        max_core);
    }
    const void *s, *l, *r;
    void* o;
    const size_t sb = src_elems(d, src_kind) * sizeof(double), fb = face_elems(d) * sizeof(double);
    if ((rc = resolve(src, &s, "src", sb)) || (rc = resolve(gL, &l, "gL", fb)) || (rc = resolve(gR, &r, "gR", fb)) ||
        (rc = resolve_mut(gv, &o, "gv", sb)))
        return rc;
    if ((rc = resolve_coeffs(c, a.c, d.ns - 6))) return rc;
    input_map(d, dir, src_kind, a);
    output_map(d, d.ns, a);
    a.src = s;
    a.outL = const_cast<void*>(l);
    a.outR = const_cast<void*>(r);
    a.J0f = 3;
    a.j_beg = 3;
    a.j_end = d.ns - 4;
    a.chunk = d.ns;
    a.chunks = 1;
    a.eps = g->weno_eps;
    // WENO5 rows of up to 420 inputs: whole-row tiles with the
    // shared-reciprocal adjoint (vjp5_tile_kernel)
    if (order == 5) {
        const hipError_t e = launch_vjp5_tiles(a, static_cast<double*>(o), stream());
        if (e != hipErrorNotSupported) return finish(e, "reconstruct_vjp");
    }
    const int R = order / 2, W = order;
    const int jn = a.j_end - a.j_beg + 1;
    const int nJ = jn + 2 * R;
    const int tmax = vjp_max_threads();
    // G lane groups of jn lanes.  The kernel's register use allows 8 waves
    // per CU: workgroups of at most 4 waves run two per CU, so that one
    // loads while the other computes; take that shape when it keeps at
    // least 80% of the lanes busy, otherwise the best fill up to tmax.
    int G = 1, T = 64 * ((jn + 63) / 64);
    double best = double(jn) / T;
    auto search = [&](int cap) {
        G = 1;
        T = 64 * ((jn + 63) / 64);
        best = double(jn) / T;
        for (int gg = 2; gg <= 16; ++gg) {
            const int t = 64 * ((gg * jn + 63) / 64);
            if (t > cap) break;
            if (double(gg * jn) / t > best + 1e-9) {
                best = double(gg * jn) / t;
                G = gg;
                T = t;
            }
        }
    };
    search(256);
    if (T > 256 || best < 0.8) search(tmax);
    const int entries = (order == 5) ? 15 : 33;   // coefficient table row stride
    const int ldj = nJ | 1;
    // LDS per workgroup: input tiles (two for WENO5, which loads the next
    // tile during the arithmetic) and the gradient tile, the contributions
    // of the outputs (one pencil per lane group, 2(W-1) zero rows each), the
    // coefficient table.  Workgroups of at most 4 waves run two per CU (78 KB
    // each); larger ones run alone and may take 158 KB.
    auto lds_of = [&](int np) {
        return ((order == 5 ? 3 : 2) * size_t(np) * ldj + size_t(G) * W * (jn + 2 * (W - 1)) +
                size_t(entries) * jn) *
               sizeof(double);
    };
    auto fits = [&](int npg, size_t cap) {
        return lds_of(G * npg) <= cap && (size_t(G) * npg * ldj + T - 1) / T <= size_t(vjp_max_slots()) &&
               G * npg <= vjp_max_np();
    };
    const size_t cap = T <= 256 ? 78 * 1024 : 158 * 1024;   // + 1.5 KB of static LDS <= 160 KB
    // pencils per tile: the contributions are staged one pencil per group at
    // a time, so the tiles can be large, which spreads the per-tile loads and
    // barriers over more cells (128^3 x 8: WENO5 15 -> 24 pencils, 1.24 ->
    // 1.19 ms per x+y+z; WENO7 6 -> 24 pencils, 1.90 -> 1.63 ms)
    int npg = std::max(1, 24 / G);
    while (npg > 1 && !fits(npg, cap)) --npg;
    if (!fits(npg, cap) && !fits(npg, 158 * 1024))
        return fail(WENO_E_INVALID, "adjoint tile does not fit LDS for sweep length %d", d.ns);
    a.groups = G;
    a.npg = npg;
    a.np = G * npg;
    a.jn = jn;
    a.threads = T;
    a.ldj = ldj;
    a.slots = static_cast<int>((size_t(a.np) * ldj + T - 1) / T);
    a.tiles_per_var = static_cast<int>((static_cast<long long>(a.na) * a.nb + a.np - 1) / a.np);
    return finish(launch_vjp(a, order, static_cast<double*>(o), stream()), "reconstruct_vjp");
}
