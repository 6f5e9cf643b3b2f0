// Tile-shape table consulted by the launch set-up.  Entries are filled by the
// autotuner (weno_autotune) and can be set explicitly (weno_set_tile).
#include "weno_capi_internal.h"
#include <mutex>

namespace wm {

namespace {
struct TileChoice {
    int np = 0;
    int chunk = 0;
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

void set_tuned(int order, int dtype, int dir, int np, int chunk)
{
    std::lock_guard<std::mutex> lk(g_tmu);
    g_table[order][dtype][dir].np = np;
    g_table[order][dtype][dir].chunk = chunk;
}

}  // namespace wm

extern "C" int weno_set_tile(int order, int dtype, int dir, int pencils, int chunk)
{
    using namespace wm;
    if (order != 3 && order != 5 && order != 7) return fail(WENO_E_INVALID, "order must be 3, 5 or 7");
    if (dtype < 0 || dtype > 2) return fail(WENO_E_INVALID, "bad dtype %d", dtype);
    if (dir < 0 || dir > 2) return fail(WENO_E_INVALID, "bad dir %d", dir);
    if (pencils != 0 && (pencils < 1 || pencils > 32 || (pencils & (pencils - 1))))
        return fail(WENO_E_INVALID, "pencils per tile must be 0 or a power of two in [1,32] (got %d)", pencils);
    if (chunk < 0) return fail(WENO_E_INVALID, "chunk must be >= 0");
    set_tuned(order, dtype, dir, pencils, chunk);
    return WENO_OK;
}
