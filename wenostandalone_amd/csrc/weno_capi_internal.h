// Host-side state and helpers shared by the C-ABI translation units.
#pragma once
#include "weno_internal.h"
#include "../../include/weno_c.h"
#include <algorithm>
#include <map>

namespace wm {

struct Present {
    size_t bytes;
    void* dev;
    int refcount;
};

struct State {
    bool initialised = false;
    bool async = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t user_stream = nullptr;
    std::map<uintptr_t, Present> present;
};
extern State g_state;

struct Dims {
    int nx, ny, nz, nv;
    int ns, nk, nl;
};

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int ensure_init();
hipStream_t stream();
int finish(hipError_t e, const char* what);
// need: bytes from p that the launch will touch (0: not checked)
int resolve(const void* p, const void** out, const char* name, size_t need = 0);
int resolve_mut(void* p, void** out, const char* name, size_t need = 0);
// element counts of the reference's arrays (weno_standalone.f90:163-187)
size_t field_elems(const Dims& d);
size_t reshaped_elems(const Dims& d);
size_t face_elems(const Dims& d);
size_t src_elems(const Dims& d, int src_kind);
int state_bytes(int dtype);   // v_vf / v_rs_ws element size
int face_bytes(int dtype);    // vL / vR element size
int check_grid(const weno_grid* g);
int check_dir(int dir);
int check_dtype(int dtype);
Dims dims_of(const weno_grid* g, int dir);
void input_map(const Dims& d, int dir, int src_kind, TileArgs& a);
void output_map(const Dims& d, int row_len, TileArgs& a);
void plan_tiles(TileArgs& a, Launch& l, int np_override, int chunk_override, int threads_cap = 0,
                int prefetch = -1);
int mode_for(const weno_grid* g, int order);
int resolve_coeffs(const weno_coeffs* c, Coefs& out, int nc);
int validate_reconstruct(const weno_grid* g, int order, int dir, int src_kind, int dtype, int flags);
int build_reconstruct(const weno_grid* g, int order, int dir, int src_kind, int dtype, const void* src,
                      const weno_coeffs* c, void* vL, void* vR, int flags, TileArgs& a, Launch& l);

// tile shapes chosen by the autotuner (0 = default)
int tuned_np(int order, int dtype, int dir);
int tuned_chunk(int order, int dtype, int dir);
int tuned_threads(int order, int dtype, int dir);
int tuned_prefetch(int order, int dtype, int dir);

}  // namespace wm
