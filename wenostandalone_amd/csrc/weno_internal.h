// Internal declarations shared by the HIP translation units of libweno_mi355.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace wm {

enum Op { OP_RESHAPE = 0, OP_COPY = 1, OP_WENO = 2, OP_MP = 3 };
// Weight branches of the reference (weno_standalone.f90:552-606), in its
// order of precedence: wenojs, mapped_weno, wenoz, teno; none -> NaN.
enum Mode { MODE_NONE = 0, MODE_JS = 1, MODE_TENO = 2, MODE_WENOZ = 3, MODE_MAPPED = 4 };
enum Dtype { DT_F64 = 0, DT_F32 = 1, DT_BF16 = 2 };

// Per-direction stencil coefficients, in the reference's layouts
// (weno_standalone.f90:171-187): poly (nc,5,3), d (5,nc), beta (nc,5,6).
struct Coefs {
    const double* polyL;
    const double* polyR;
    const double* dL;
    const double* dR;
    const double* beta;
    int nc;
};

// One sweep over a set of pencils.  A pencil is one line along the sweep
// direction at fixed transverse position (a,b) and variable i.  (a,b) are
// the transverse coordinates in load order: `a` is the one whose input
// stride is smallest, so that consecutive pencils of a tile sit next to each
// other in memory.
struct TileArgs {
    const void* src;
    long long src_off;            // element offset of (J=0, a=0, b=0, i=0)
    long long s_j, s_a, s_b, s_i; // input strides (elements)
    int src_len;                  // input extent along the sweep (ns+6)
    int na, nb, nv;
    int a_is_k;                   // pencil (a,b) -> output (k,l) = a_is_k ? (a,b) : (b,a)
    void* outL;
    void* outR;
    long long o_k, o_l, o_i;      // output strides; stride along j is 1
    int out_len;                  // output row length (ns; ns+6 for reshape)
    int J0f;                      // input J of output j (J = j + J0f)
    int j_beg, j_end;             // reconstructed (core) range, inclusive; empty if j_beg>j_end
    int copy_halo;                // write order-1 values into the rest of [0,out_len)
    int sides;                    // bit 0: write outL, bit 1: write outR
    int mp_beg, mp_end;           // range the fused limiter applies to (inclusive)
    int chunk;                    // outputs per tile along j
    int chunks;                   // tiles along j
    int np;                       // pencils per tile (= groups * npg)
    int groups;                   // lane groups per block
    int npg;                      // pencils per lane group
    int jn;                       // lanes per group (outputs per chunk row)
    int threads;                  // block size
    int slots;                    // tile elements loaded per thread
    int prefetch;                 // load tile k+1 while computing tile k
    int tiles_per_var;            // pencil groups per variable
    int bpc;                      // blocks per chunk (set at launch)
    int ldj;                      // LDS row stride (elements)
    unsigned long long src_bytes; // size of the input array (all variables)
    Coefs c;
    double eps, teno_ct;
};

struct Launch {
    int op, order, mode, mp, dtype;
    int nontemporal;   // streaming stores
    int blocks_per_cu; // 0 = occupancy limit
    int creg;          // WENO5: beta/d coefficients in registers instead of LDS
    int one_tile;      // one tile per workgroup (no persistent loop)
    int fast5;         // 0: general kernel; 1 + layout: the WENO5/7-JS producer/consumer kernel
};

int kernel_max_slots();
int kernel_max_np();
int max_threads(int op, int order);

// Launches one tile-kernel sweep; returns hipSuccess or an error.
hipError_t launch_tile(const TileArgs& a, const Launch& l, hipStream_t st);

// Order-1 copy and reshape (rows of the field permuted along the sweep; no
// tile planning): streaming rows for x sweeps, LDS-transposed tiles for y/z.
hipError_t launch_permute(const TileArgs& a, int dtype, hipStream_t st);

// WENO5-JS producer/consumer kernel: planning (returns the layout, or -1 if
// the launch stays on the general kernel) and launch.
int plan_sweep5(TileArgs& a, Launch& l, const void* src);
hipError_t launch_sweep5(TileArgs a, int order, int mode, int lay, int dtype, int mp, hipStream_t st);

// LDS element size of the compute type for a storage dtype
inline int lds_elem_bytes(int dtype) { return dtype == DT_F64 ? 8 : 4; }

}  // namespace wm
