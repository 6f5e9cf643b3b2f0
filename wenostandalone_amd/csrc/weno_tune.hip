// Tile-shape table consulted by the launch set-up.  Entries are filled by the
// autotuner (weno_autotune) and can be set explicitly (weno_set_tile).
#include "weno_capi_internal.h"
#include <mutex>

namespace wm {

namespace {
struct TileChoice {
    int np = 0;
    int chunk = 0;
    int threads = 0;
    int prefetch = -1;
};
std::mutex g_tmu;
TileChoice g_table[8][3][3];  // [order][dtype][dir]
}  // namespace

int tuned_np(int order, int dtype, int dir)
{
    if (order < 0 || order > 7 || dtype < 0 || dtype > 2 || dir < 0 || dir > 2) return 0;
    std::lock_guard<std::mutex> lk(g_tmu);
    return g_table[order][dtype][dir].np;
}

int tuned_chunk(int order, int dtype, int dir)
{
    if (order < 0 || order > 7 || dtype < 0 || dtype > 2 || dir < 0 || dir > 2) return 0;
    std::lock_guard<std::mutex> lk(g_tmu);
    return g_table[order][dtype][dir].chunk;
}

int tuned_threads(int order, int dtype, int dir)
{
    if (order < 0 || order > 7 || dtype < 0 || dtype > 2 || dir < 0 || dir > 2) return 0;
    std::lock_guard<std::mutex> lk(g_tmu);
    return g_table[order][dtype][dir].threads;
}

int tuned_prefetch(int order, int dtype, int dir)
{
    if (order < 0 || order > 7 || dtype < 0 || dtype > 2 || dir < 0 || dir > 2) return -1;
    std::lock_guard<std::mutex> lk(g_tmu);
    return g_table[order][dtype][dir].prefetch;
}

void set_tuned(int order, int dtype, int dir, int np, int chunk, int threads, int prefetch)
{
    std::lock_guard<std::mutex> lk(g_tmu);
    g_table[order][dtype][dir] = TileChoice{np, chunk, threads, prefetch};
}

}  // namespace wm

extern "C" int weno_set_tile(int order, int dtype, int dir, int pencils, int chunk, int threads, int prefetch)
{
    using namespace wm;
    if (order != 3 && order != 5 && order != 7) return fail(WENO_E_INVALID, "order must be 3, 5 or 7");
    if (dtype < 0 || dtype > 2) return fail(WENO_E_INVALID, "bad dtype %d", dtype);
    if (dir < 0 || dir > 2) return fail(WENO_E_INVALID, "bad dir %d", dir);
    if (pencils < 0 || pencils > 64) return fail(WENO_E_INVALID, "pencils per tile must be in [0,64] (got %d)", pencils);
    if (chunk < 0) return fail(WENO_E_INVALID, "chunk must be >= 0");
    if (threads != 0 && (threads < 64 || threads > 1024 || threads % 64))
        return fail(WENO_E_INVALID, "threads must be 0 or a multiple of 64 in [64,1024] (got %d)", threads);
    if (prefetch < -1 || prefetch > 15) return fail(WENO_E_INVALID, "variant must be -1 or a bit set in [0,15]");
    set_tuned(order, dtype, dir, pencils, chunk, threads, prefetch);
    return WENO_OK;
}
