/*
 * libweno_mi355 -- C ABI of the MI355X (gfx950) WENO reconstruction path.
 *
 * This is the boundary a Fortran host binds with iso_c_binding (see
 * wenostandalone_amd/fortran/m_weno_standalone.f90) and Python binds with
 * ctypes.  Plain C types only: pointers, sizes, ints and doubles.
 *
 * The reference has no C boundary: its "API" is the Fortran module
 * m_weno_standalone of inputs/weno_standalone.f90, whose procedures run
 * `!$omp target` regions on arrays mapped with `target enter data`.  Each
 * entry point below names the reference procedure it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on error;
 *     weno_last_error() gives the message.  Nothing is launched after a
 *     validation failure.
 *   - Array arguments are device addresses, or host addresses that are
 *     present in the data environment (weno_target_enter_data), in which case
 *     the mapped device copy is used -- the semantics of OpenMP's implicit
 *     map of a present array.  A host address that is not present is an
 *     error (never dereferenced on the device).
 *   - Calls are synchronous (like the reference's target regions, which
 *     carry no `nowait`) unless weno_set_async(1) was called; then they are
 *     stream-ordered on the library stream and weno_sync() waits.
 *   - Layouts are the reference's column-major arrays:
 *       v_vf       (nx+6, ny+6, nz+6, v_size)    weno_standalone.f90:163-165
 *       v_rs_ws_d  (ns+6, nk, nl, v_size)        :167-169, :281-330
 *       vL_d, vR_d (ns, nk, nl, v_size)          :343, :363, :383
 *     with nx = m_cells+1+2*buff_size (likewise ny, nz) and, per direction,
 *     (ns,nk,nl) = x:(nx,ny,nz)  y:(ny,nx,nz)  z:(nz,ny,nx).
 *     Coefficients (nc = ns-6 core points):
 *       polyL/polyR (nc,5,3), dL/dR (5,nc), beta (nc,5,6)   :171-187
 */
#ifndef WENO_C_H
#define WENO_C_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define WENO_OK 0
#define WENO_E_INVALID (-1)   /* argument validation failed */
#define WENO_E_HIP (-2)       /* HIP runtime error */
#define WENO_E_NOT_PRESENT (-3) /* host address not in the data environment */
#define WENO_E_STATE (-4)     /* library not initialised / wrong state */

enum { WENO_DIR_X = 0, WENO_DIR_Y = 1, WENO_DIR_Z = 2 };
enum { WENO_SRC_FIELD = 0, WENO_SRC_RESHAPED = 1 };
enum { WENO_F64 = 0, WENO_F32 = 1, WENO_BF16 = 2 };

/* reconstruct flags */
#define WENO_FUSE_MP 1     /* apply s_preserve_monotonicity in the same pass */
#define WENO_FUSE_COPY 2   /* write order-1 values (s_order1_copy) into j outside the core */
#define WENO_STREAMING 4   /* non-temporal stores */
#define WENO_ONLY_L 8      /* write vL only (s_weno5_left_kernel, :1145) */
#define WENO_ONLY_R 16     /* write vR only (s_weno5_right_kernel, :1272) */
#define WENO_ACCUMULATE 32 /* flux divergence: rhs += contribution instead of = */

/* Grid and scheme parameters: namelist &grid_params (weno_standalone.f90:
 * 57-66, :117-118) plus `teno`, `mapped_weno` and `wenoz`, compile-time
 * parameters in the reference (:72-74) and runtime switches here.  The
 * weight branch follows the reference's precedence: wenojs, then
 * mapped_weno, then wenoz, then teno (:552-606, :1475-1503); WENO3 has no
 * TENO branch; with no branch taken the outputs are NaN exactly as in the
 * reference.  WENO-Z uses (tau/beta) to the power 1 for WENO3/5 and
 * wenoz_q = 2 for WENO7 (:75), as the reference writes them. */
typedef struct weno_grid {
    int m_cells, n_cells, p_cells;
    int v_size;
    int buff_size;
    double weno_eps;
    int wenojs;
    int teno;
    double teno_CT;
    int mapped_weno;
    int wenoz;
} weno_grid;

/* One direction's coefficient set (device or present host addresses). */
typedef struct weno_coeffs {
    const double* polyL;
    const double* polyR;
    const double* dL;
    const double* dR;
    const double* beta;
} weno_coeffs;

/* ---- runtime ------------------------------------------------------------ */
/* Selects the device and creates the library stream.  Replaces the implicit
 * OpenMP offload runtime start-up (omp_get_num_devices, :1755-1756). */
int weno_init(int device);
int weno_finalize(void);
const char* weno_last_error(void);
int weno_sync(void);
int weno_set_async(int async_mode);
/* Use a caller's hipStream_t (NULL = library stream). */
int weno_set_stream(void* hip_stream);
void* weno_get_stream(void);
int weno_device_count(void);

/* ---- data environment: `!$omp target enter/update/exit data` ------------
 * :194-203, :209-212, :1790-1791, :1831, :1856 */
int weno_target_enter_data(const void* host, size_t bytes); /* map(alloc:) */
int weno_target_update_to(const void* host, size_t bytes);
int weno_target_update_from(void* host, size_t bytes);
int weno_target_exit_data(const void* host);                /* map(release:) */
void* weno_device_address(const void* host);                /* NULL if absent */

/* plain device memory for callers without an allocator of their own */
void* weno_device_alloc(size_t bytes);
int weno_device_free(void* dev);
int weno_memcpy_to_device(void* dev, const void* host, size_t bytes);
int weno_memcpy_to_host(void* host, const void* dev, size_t bytes);

/* ---- shapes ------------------------------------------------------------- */
size_t weno_field_elems(const weno_grid* g);             /* v_vf          */
size_t weno_reshaped_elems(const weno_grid* g, int dir); /* v_rs_ws_dir   */
size_t weno_face_elems(const weno_grid* g, int dir);     /* vL_dir, vR_dir */
int weno_core_count(const weno_grid* g, int dir);        /* nc            */

/* ---- kernels ----------------------------------------------------------- */
/* s_initialize_weno_{x,y,z} (:281-330): v_rs_ws_dir = permuted v_vf.
 * dtype WENO_F64 or WENO_F32. */
int weno_reshape(const weno_grid* g, int dir, int dtype, const void* v_vf, void* v_rs);

/* s_order1_copy_{x,y,z} (:332-390): vL = vR = v_vf in reshaped coordinates.
 * dtype of the field; F64 -> double faces, F32 and BF16 -> float faces. */
int weno_order1_copy(const weno_grid* g, int dir, int dtype, const void* v_vf, void* vL, void* vR);

/* s_weno3_kernel (:392), s_weno5_kernel / s_weno5_kernel_dispatch
 * (:477, :1399), s_weno5_kernel_{x,y,z} (:671, :829, :987),
 * s_weno5_{left,right}_kernel (:1145, :1272 -- same outputs), s_weno7_kernel
 * (:1419).  order in {3,5,7}.  src_kind WENO_SRC_RESHAPED takes v_rs_ws_dir
 * as the reference does; WENO_SRC_FIELD reads v_vf directly with
 * direction-aware loads (no reshape pass).  Writes the core range
 * (order 5/7) or the whole `is` range (order 3); with WENO_FUSE_COPY the
 * rest of vL/vR gets the order-1 values; with WENO_FUSE_MP the result is
 * additionally limited by s_preserve_monotonicity.
 * dtype: WENO_F64 (double field, double faces), WENO_F32 (float field,
 * float faces), WENO_BF16 (bf16 field, float faces).  WENO5-JS launches
 * that write both sides over the core range (rows of at least 96 faces)
 * widen FP32/BF16 state to FP64 and do all of the stencil arithmetic in
 * FP64, rounding each face once; every other FP32/BF16 launch does the
 * candidate polynomials in FP32 and beta and the weight normalisation in
 * FP64. */
int weno_reconstruct(const weno_grid* g, int order, int dir, int src_kind, int dtype,
                     const void* src, const weno_coeffs* c, void* vL, void* vR, int flags);

/* s_preserve_monotonicity (:1544-1615): limits vL/vR in place on the core. */
int weno_mp_limit(const weno_grid* g, int dir, int src_kind, int dtype, const void* src,
                  void* vL, void* vR);

/* Adjoint of weno_reconstruct for order 5/7 with JS weights (wenojs set),
 * FP64 (no reference counterpart: the reference has no derivative).  For
 * cotangents gL, gR on the core outputs of one sweep (face layout), adds
 *   (dvL/dsrc)^T gL + (dvR/dsrc)^T gR
 * to gv, which has src's layout (v_vf or v_rs_ws_dir) and must be zeroed by
 * the caller before the first sweep; beta and the weights are recomputed,
 * nothing is stored by the forward pass.  Sweeps of one field accumulate.
 * Sweeps of up to 512 core points (ns <= 518) for WENO5 and 479 core
 * points (ns <= 485) for WENO7 are supported; longer ones return
 * WENO_E_INVALID. */
int weno_reconstruct_vjp(const weno_grid* g, int order, int dir, int src_kind, const void* src,
                         const weno_coeffs* c, const void* gL, const void* gR, void* gv);

/* ---- consumer step (no reference counterpart) ---------------------------
 * The harness stops at vL/vR (:343, :363, :383); MFC's next step, its
 * Riemann solver and flux divergence, pairs qL(j) with qR(j+1) at interface
 * j+1/2.  Here, per variable, the scalar flux f(q) = a q + b q^2 / 2 with a
 * Rusanov (local Lax-Friedrichs) interface flux
 *   F(qL, qR) = (f(qL) + f(qR))/2 - max(|a + b qL|, |a + b qR|) (qR - qL)/2
 * gives, for the cells j = 4 .. ns-5 whose two interfaces lie in the
 * reconstructed core,
 *   rhs(j) = -(F(vL(j), vR(j+1)) - F(vL(j-1), vR(j))) * inv_dx
 * in the sweep's face layout (ns, nk, nl, v_size), FP64; other cells of rhs
 * keep their values.  WENO_ACCUMULATE adds the contribution to rhs. */
typedef struct weno_flux {
    double a;       /* linear part of f */
    double b;       /* quadratic part of f (Burgers: a = 0, b = 1) */
    double inv_dx;  /* 1 / cell width along the sweep */
} weno_flux;

/* Reconstruction (weno_reconstruct's order, weight branch and WENO_FUSE_MP)
 * and flux divergence in one pass: the faces stay on chip, so the sweep
 * reads the state once and writes rhs once.  FP64 state; sweeps of up to
 * 1024 core points.  flags: WENO_FUSE_MP, WENO_ACCUMULATE. */
int weno_reconstruct_flux(const weno_grid* g, int order, int dir, int src_kind, const void* src,
                          const weno_coeffs* c, const weno_flux* fx, void* rhs, int flags);

/* The same flux divergence over existing faces vL, vR (FP64, face layout):
 * the second pass of an unfused reconstruction.  flags: WENO_ACCUMULATE. */
int weno_flux_divergence(const weno_grid* g, int dir, const void* vL, const void* vR, const weno_flux* fx,
                         void* rhs, int flags);

/* ---- batched blocks, graphs, autotuning (no reference counterpart) ------ */
/* One independent block (sub-domain): its own extents, variable count,
 * order, field, per-direction coefficients and face arrays. */
typedef struct weno_block {
    weno_grid grid;
    int order;                 /* 3, 5 or 7 */
    const void* v_vf;          /* padded field, FP64 */
    weno_coeffs coeffs[3];     /* x, y, z */
    void* vL[3];
    void* vR[3];
} weno_block;

/* All blocks' sweeps along dir in one persistent launch (FP64; flags:
 * WENO_FUSE_MP, WENO_FUSE_COPY).  Work items (block, tile) are pulled from
 * an atomic counter, largest blocks first. */
int weno_batched_reconstruct(const weno_block* blocks, int nblocks, int dir, int flags);

/* weno_batched_reconstruct that allocates nothing.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The work-item counter and the block descriptors live in a library-owned
 * static device array (a 256-byte counter slot and WENO_NOALLOC_MAX_BLOCKS
 * descriptors of 280 bytes: 71,936 bytes per GPU, held until the process
 * ends) instead of memory the library allocates; more than
 * WENO_NOALLOC_MAX_BLOCKS blocks is WENO_E_INVALID.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other host threads, other streams) race.  Same kernels, same results as
 * weno_batched_reconstruct. */
#define WENO_NOALLOC_MAX_BLOCKS 256
int weno_batched_reconstruct_noalloc(const weno_block* blocks, int nblocks, int dir, int flags);

/* Captures the x, y and z batched launches into one hipGraph (the blocks'
 * arrays must stay allocated while the graph lives). */
int weno_batched_graph_create(const weno_block* blocks, int nblocks, int flags, void** graph);

/* Captures the x, y and z sweeps of weno_reconstruct for one box into one
 * hipGraph: coeffs3, vL3, vR3 hold the three directions' arguments. */
int weno_sweep_graph_create(const weno_grid* g, int order, int src_kind, int dtype, const void* src,
                            const weno_coeffs* coeffs3, void* const* vL3, void* const* vR3, int flags,
                            void** graph);
int weno_graph_launch(void* graph);
int weno_graph_destroy(void* graph);

/* Times tile-shape candidates (pencils per tile x threads x kernel variant)
 * for the three sweeps of one box and keeps the fastest in the table that
 * weno_reconstruct consults.  best (optional) receives per direction
 * {pencils, threads, variant, microseconds}.  Runs the sweeps repeatedly. */
int weno_autotune_sweep(const weno_grid* g, int order, int dtype, int src_kind, const void* src,
                        const weno_coeffs* coeffs3, void* const* vL3, void* const* vR3, int flags, int* best);

/* Same for the batched launch: chooses workgroup size and pencils per tile;
 * best (optional) receives {threads, pencils, microseconds for x+y+z}. */
int weno_batched_autotune(const weno_block* blocks, int nblocks, int flags, int* best);

/* Plain device-to-device streaming copy (16-byte multiples): the bandwidth
 * yardstick for the sweeps' HBM traffic counters.  No reference counterpart. */
int weno_stream_copy(void* dst, const void* src, size_t bytes);

/* Tile shape of the sweep kernel for (order, dtype, dir): pencils per tile,
 * outputs per tile along the sweep, threads per workgroup (0 = automatic
 * for each) and the kernel variant (-1 = automatic; otherwise bits: 1 load
 * the next tile during compute, 2 WENO5 coefficients in registers, 4 one
 * tile per workgroup, 8 keep the launch on the general kernel instead of the
 * producer/consumer kernel, whose tile shape is fixed; the general kernel's
 * FP32/BF16-state form does its stencil arithmetic in FP32).  Used by the
 * autotuner; no reference counterpart. */
int weno_set_tile(int order, int dtype, int dir, int pencils, int chunk, int threads, int variant);

#ifdef __cplusplus
}
#endif
#endif
