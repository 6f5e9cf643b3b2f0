// extern "C" entry points of libweno_mi355 (declared in include/weno_c.h):
// runtime state, the host-address -> device-buffer data environment, argument
// validation and the set-up of the sweep kernel for every reference kernel
// family.
#include "weno_internal.h"
#include "weno_capi_internal.h"
#include "../../include/weno_c.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

namespace wm {

static thread_local std::string g_err;
static std::mutex g_mu;
State g_state;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(WENO_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int ensure_init()
{
    if (g_state.initialised) return WENO_OK;
    return weno_init(-1);
}

hipStream_t stream() { return g_state.user_stream ? g_state.user_stream : g_state.stream; }

int finish(hipError_t e, const char* what)
{
    if (e != hipSuccess) return hip_fail(e, what);
    if (!g_state.async) {
        e = hipStreamSynchronize(stream());
        if (e != hipSuccess) return hip_fail(e, what);
    }
    return WENO_OK;
}

// Translate a user pointer into a device address: a host address present in
// the data environment maps to its device copy; otherwise the pointer must be
// device memory.  With need > 0 the `need` bytes from p must lie inside the
// mapping (OpenMP's present check rejects a partly present array) or inside
// the device allocation, so that no launch reaches past either.
int resolve(const void* p, const void** out, const char* name, size_t need)
{
    if (p == nullptr) return fail(WENO_E_INVALID, "%s is NULL", name);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_state.present.upper_bound(reinterpret_cast<uintptr_t>(p));
        if (it != g_state.present.begin()) {
            --it;
            const uintptr_t off = reinterpret_cast<uintptr_t>(p) - it->first;
            if (off < it->second.bytes) {
                if (need > it->second.bytes - off)
                    return fail(WENO_E_INVALID,
                                "%s (%p) is present with %zu bytes from that address; the call needs %zu", name, p,
                                it->second.bytes - off, need);
                *out = static_cast<const char*>(it->second.dev) + off;
                return WENO_OK;
            }
        }
    }
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e == hipSuccess && (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged)) {
        if (need > 0) {
            hipDeviceptr_t base = nullptr;
            size_t size = 0;
            if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess || !base) {
                (void)hipGetLastError();
                return fail(WENO_E_INVALID, "%s (%p): the extent of this device allocation cannot be determined",
                            name, p);
            }
            const size_t off = static_cast<size_t>(static_cast<const char*>(p) - static_cast<const char*>(base));
            if (off > size || need > size - off)
                return fail(WENO_E_INVALID,
                            "%s (%p) has %zu bytes of device memory from that address; the call needs %zu", name, p,
                            off > size ? size_t(0) : size - off, need);
        }
        *out = p;
        return WENO_OK;
    }
    (void)hipGetLastError();
    return fail(WENO_E_NOT_PRESENT,
                "%s (%p) is neither device memory nor present in the data environment", name, p);
}

int resolve_mut(void* p, void** out, const char* name, size_t need)
{
    const void* q = nullptr;
    int rc = resolve(p, &q, name, need);
    *out = const_cast<void*>(q);
    return rc;
}

size_t field_elems(const Dims& d) { return size_t(d.nx + 6) * size_t(d.ny + 6) * size_t(d.nz + 6) * size_t(d.nv); }
size_t reshaped_elems(const Dims& d) { return size_t(d.ns + 6) * size_t(d.nk) * size_t(d.nl) * size_t(d.nv); }
size_t face_elems(const Dims& d) { return size_t(d.ns) * size_t(d.nk) * size_t(d.nl) * size_t(d.nv); }
size_t src_elems(const Dims& d, int src_kind)
{
    return src_kind == WENO_SRC_RESHAPED ? reshaped_elems(d) : field_elems(d);
}
int state_bytes(int dtype) { return dtype == WENO_F64 ? 8 : (dtype == WENO_F32 ? 4 : 2); }
int face_bytes(int dtype) { return dtype == WENO_F64 ? 8 : 4; }

int check_grid(const weno_grid* g)
{
    if (!g) return fail(WENO_E_INVALID, "grid is NULL");
    if (g->m_cells < 0 || g->n_cells < 0 || g->p_cells < 0)
        return fail(WENO_E_INVALID, "m/n/p_cells must be >= 0 (got %d %d %d)", g->m_cells, g->n_cells,
                    g->p_cells);
    if (g->v_size < 1) return fail(WENO_E_INVALID, "v_size must be >= 1 (got %d)", g->v_size);
    if (g->buff_size < 0) return fail(WENO_E_INVALID, "buff_size must be >= 0 (got %d)", g->buff_size);
    const int nx = g->m_cells + 1 + 2 * g->buff_size, ny = g->n_cells + 1 + 2 * g->buff_size,
              nz = g->p_cells + 1 + 2 * g->buff_size;
    if (nx < 7 || ny < 7 || nz < 7)
        return fail(WENO_E_INVALID,
                    "each `is` extent (cells+1+2*buff_size) must be >= 7 so the core range is non-empty "
                    "(got %d %d %d)",
                    nx, ny, nz);
    const double elems = double(nx + 6) * double(ny + 6) * double(nz + 6) * double(g->v_size);
    if (elems > 4.0e9) return fail(WENO_E_INVALID, "field too large (%.3g elements)", elems);
    if (!(g->weno_eps > 0.0)) return fail(WENO_E_INVALID, "weno_eps must be > 0");
    return WENO_OK;
}

Dims dims_of(const weno_grid* g, int dir)
{
    Dims d;
    d.nx = g->m_cells + 1 + 2 * g->buff_size;
    d.ny = g->n_cells + 1 + 2 * g->buff_size;
    d.nz = g->p_cells + 1 + 2 * g->buff_size;
    d.nv = g->v_size;
    if (dir == 0) { d.ns = d.nx; d.nk = d.ny; d.nl = d.nz; }
    else if (dir == 1) { d.ns = d.ny; d.nk = d.nx; d.nl = d.nz; }
    else { d.ns = d.nz; d.nk = d.ny; d.nl = d.nx; }
    return d;
}

// Input addressing of one direction.  Field: v_vf with 3 padding cells on
// every axis; reshaped: (ns+6, nk, nl, v).
void input_map(const Dims& d, int dir, int src_kind, TileArgs& a)
{
    const long long NX = d.nx + 6, NY = d.ny + 6, NZ = d.nz + 6;
    a.src_len = d.ns + 6;
    a.nv = d.nv;
    if (src_kind == WENO_SRC_RESHAPED) {
        const long long S = d.ns + 6;
        a.s_j = 1;
        a.s_a = S;
        a.s_b = S * d.nk;
        a.s_i = S * d.nk * d.nl;
        a.src_off = 0;
        a.na = d.nk;
        a.nb = d.nl;
        a.a_is_k = 1;
        return;
    }
    a.s_i = NX * NY * NZ;
    if (dir == 0) {          // J=x, k=y, l=z
        a.s_j = 1;
        a.s_a = NX;
        a.s_b = NX * NY;
        a.src_off = 3 * NX + 3 * NX * NY;
        a.na = d.nk;
        a.nb = d.nl;
        a.a_is_k = 1;
    } else if (dir == 1) {   // J=y, k=x, l=z
        a.s_j = NX;
        a.s_a = 1;
        a.s_b = NX * NY;
        a.src_off = 3 + 3 * NX * NY;
        a.na = d.nk;
        a.nb = d.nl;
        a.a_is_k = 1;
    } else {                 // J=z, k=y, l=x
        a.s_j = NX * NY;
        a.s_a = 1;
        a.s_b = NX;
        a.src_off = 3 + 3 * NX;
        a.na = d.nl;
        a.nb = d.nk;
        a.a_is_k = 0;
    }
}

void output_map(const Dims& d, int row_len, TileArgs& a)
{
    a.out_len = row_len;
    a.o_k = row_len;
    a.o_l = static_cast<long long>(row_len) * d.nk;
    a.o_i = static_cast<long long>(row_len) * d.nk * d.nl;
}

// Defaults from tile-shape sweeps on MI355X at 200^3 x 8 (FP64 WENO5): the
// persistent loop with the coefficients in registers and no prefetch, in
// 448-thread workgroups (two lane groups of 206 + 2 workgroups per CU).
constexpr int kDefaultVariant = 2;
constexpr int kDefaultWeno7Threads = 256;
constexpr int kDefaultWeno3Threads = 512;
constexpr int kDefaultWeno5Threads = 448;
constexpr int kDefaultMpThreads = 512;
constexpr int kDefaultMappedWeno5Threads = 256;

// Pencils per tile, j-chunking and LDS row stride for one launch.
void plan_tiles(TileArgs& a, Launch& l, int np_override, int chunk_override, int threads_cap, int prefetch)
{
    const int elem = lds_elem_bytes(l.dtype);
    const int np_target = np_override > 0 ? np_override : (l.dtype == DT_F64 ? 16 : 32);
    const int halo = 8;                      // 2*R_max + 1, rounded up
    const long long lds_budget = 96 * 1024;  // both tile buffers
    const int max_np = kernel_max_np(), max_slots = kernel_max_slots();
    // FP64 mapped WENO5 (heavier arithmetic per face): 256-thread tiles, two
    // or more workgroups per CU (measured 1028 -> 984 us per 200^3 x 8 sweep)
    if (threads_cap <= 0 && l.op == OP_WENO && l.order == 5 && l.mode == MODE_MAPPED && l.dtype == DT_F64)
        threads_cap = kDefaultMappedWeno5Threads;
    if (threads_cap <= 0 && l.op == OP_WENO && l.order == 5) threads_cap = kDefaultWeno5Threads;
    // FP64 WENO7 with its coefficients in registers (~200-250 VGPRs): two
    // 256-thread workgroups per CU hold 8 waves where one 448-thread
    // workgroup holds 7, and need no LDS coefficient table (measured 787 ->
    // 755 us per 200^3 x 8 sweep)
    if (threads_cap <= 0 && l.op == OP_WENO && l.order == 7 && l.dtype == DT_F64 &&
        (((prefetch < 0 ? kDefaultVariant : prefetch) >> 1) & 1))
        threads_cap = kDefaultWeno7Threads;
    // FP64 WENO3: two lane groups (448 threads) rather than the three (640)
    // the lane-utilisation search would pick (measured 497 -> 430 us per
    // 200^3 x 8 sweep)
    if (threads_cap <= 0 && l.op == OP_WENO && l.order == 3 && l.dtype == DT_F64) threads_cap = kDefaultWeno3Threads;
    // FP64 standalone limiter: two lane groups (448 threads) rather than the
    // four (832) the lane-utilisation search would pick (measured 624 -> 609
    // us per 200^3 x 8 sweep)
    if (threads_cap <= 0 && l.op == OP_MP && l.dtype == DT_F64) threads_cap = kDefaultMpThreads;
    // a whole number of waves, at most the kernel's workgroup limit
    const int tmax = std::max(64, (threads_cap > 0 ? std::min(threads_cap, max_threads(l.op, l.order))
                                                   : max_threads(l.op, l.order)) & ~63);
    int chunk = chunk_override > 0 ? std::min(chunk_override, a.out_len) : a.out_len;
    const int min_np = std::min(np_target, 8);
    if (2LL * min_np * (chunk + halo) * elem > lds_budget)
        chunk = static_cast<int>(lds_budget / (2LL * min_np * elem)) - halo;
    // one lane per output of a chunk row: rows longer than the workgroup
    // limit are split into more chunks
    chunk = std::min(chunk, tmax);
    const int chunks = (a.out_len + chunk - 1) / chunk;
    chunk = (a.out_len + chunks - 1) / chunks;
    // lanes per group: the most outputs any chunk row has
    int jn = 1;
    for (int c = 0; c < chunks; ++c) {
        const int jo0 = c * chunk, jo1 = std::min(a.out_len, jo0 + chunk);
        const int c0 = std::max(jo0, a.j_beg), c1 = std::min(jo1 - 1, a.j_end);
        const int n = a.copy_halo ? jo1 - jo0 : std::max(0, c1 - c0 + 1);
        jn = std::max(jn, n);
    }
    // lane groups per block: best lane utilisation within 1024 threads
    int G = 1, T = 64 * ((jn + 63) / 64);
    double best = double(jn) / T;
    auto search = [&](int cap) {
        G = 1;
        T = 64 * ((jn + 63) / 64);
        best = double(jn) / T;
        for (int g = 2; g <= 16; ++g) {
            const int t = 64 * ((g * jn + 63) / 64);
            if (t > cap) break;
            const double u = double(g * jn) / t;
            if (u > best + 1e-9) {
                best = u;
                G = g;
                T = t;
            }
        }
    };
    search(tmax);
    const int ldj = std::min(chunk + halo - 1, a.src_len) | 1;
    int npg = std::max(1, (np_target + G - 1) / G);
    npg = std::min(npg, std::max(1, max_np / G));
    while (npg > 1 && (2LL * G * npg * ldj * elem > lds_budget ||
                       (static_cast<long long>(G) * npg * ldj + T - 1) / T > max_slots))
        --npg;
    a.ldj = ldj;
    a.chunk = chunk;
    a.chunks = chunks;
    a.groups = G;
    a.npg = npg;
    a.np = G * npg;
    a.jn = jn;
    a.threads = T;
    a.slots = static_cast<int>((static_cast<long long>(a.np) * ldj + T - 1) / T);
    // variant bits: 1 prefetch, 2 coefficients in registers (WENO5; WENO7
    // with FP64 state: the FP32/BF16 WENO7 kernels are held to 128 VGPRs),
    // 4 one tile per workgroup; -1 = automatic
    const int variant = prefetch < 0 ? kDefaultVariant : prefetch;
    a.prefetch = variant & 1;
    l.creg = (variant >> 1) & 1;
    l.one_tile = (variant >> 2) & 1;
    l.fast5 = 0;
    const long long pen = static_cast<long long>(a.na) * a.nb;
    a.tiles_per_var = static_cast<int>((pen + a.np - 1) / a.np);
    l.blocks_per_cu = 0;
}

int mode_for(const weno_grid* g, int order)
{
    // precedence of the reference's if/elseif chain (:552, :557, :572, :578)
    if (g->wenojs) return MODE_JS;
    if (g->mapped_weno) return MODE_MAPPED;
    if (g->wenoz) return MODE_WENOZ;
    if (g->teno && order != 3) return MODE_TENO;
    return MODE_NONE;
}

int check_dtype(int dtype)
{
    if (dtype != WENO_F64 && dtype != WENO_F32 && dtype != WENO_BF16)
        return fail(WENO_E_INVALID, "dtype must be WENO_F64, WENO_F32 or WENO_BF16 (got %d)", dtype);
    return WENO_OK;
}

int check_dir(int dir)
{
    if (dir < 0 || dir > 2) return fail(WENO_E_INVALID, "dir must be 0 (x), 1 (y) or 2 (z) (got %d)", dir);
    return WENO_OK;
}

int resolve_coeffs(const weno_coeffs* c, Coefs& out, int nc)
{
    if (!c) return fail(WENO_E_INVALID, "coefficients are NULL");
    const void* p;
    int rc;
    // poly (nc,5,3), d (5,nc), beta (nc,5,6) doubles (:171-187)
    const size_t n = static_cast<size_t>(std::max(nc, 0)) * sizeof(double);
    if ((rc = resolve(c->polyL, &p, "polyL", 15 * n))) return rc;
    out.polyL = static_cast<const double*>(p);
    if ((rc = resolve(c->polyR, &p, "polyR", 15 * n))) return rc;
    out.polyR = static_cast<const double*>(p);
    if ((rc = resolve(c->dL, &p, "dL", 5 * n))) return rc;
    out.dL = static_cast<const double*>(p);
    if ((rc = resolve(c->dR, &p, "dR", 5 * n))) return rc;
    out.dR = static_cast<const double*>(p);
    if ((rc = resolve(c->beta, &p, "beta_coef", 30 * n))) return rc;
    out.beta = static_cast<const double*>(p);
    out.nc = nc;
    return WENO_OK;
}

int validate_reconstruct(const weno_grid* g, int order, int dir, int src_kind, int dtype, int flags)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir)) || (rc = check_dtype(dtype))) return rc;
    if (order != 3 && order != 5 && order != 7)
        return fail(WENO_E_INVALID, "order must be 3, 5 or 7 (got %d)", order);
    if (src_kind != WENO_SRC_FIELD && src_kind != WENO_SRC_RESHAPED)
        return fail(WENO_E_INVALID, "src_kind must be WENO_SRC_FIELD or WENO_SRC_RESHAPED");
    if (flags & ~(WENO_FUSE_MP | WENO_FUSE_COPY | WENO_STREAMING | WENO_ONLY_L | WENO_ONLY_R))
        return fail(WENO_E_INVALID, "unknown flag bits 0x%x", flags);
    if ((flags & WENO_ONLY_L) && (flags & WENO_ONLY_R))
        return fail(WENO_E_INVALID, "WENO_ONLY_L and WENO_ONLY_R are exclusive (flag bits 0x%x)", flags);
    return WENO_OK;
}

int build_reconstruct(const weno_grid* g, int order, int dir, int src_kind, int dtype, const void* src,
                      const weno_coeffs* c, void* vL, void* vR, int flags, TileArgs& a, Launch& l)
{
    int rc;
    if ((rc = validate_reconstruct(g, order, dir, src_kind, dtype, flags))) return rc;
    if ((rc = ensure_init())) return rc;
    std::memset(&a, 0, sizeof a);
    const Dims d = dims_of(g, dir);
    const void* s;
    void *oL, *oR;
    const size_t fb = face_elems(d) * face_bytes(dtype);
    if ((rc = resolve(src, &s, "src", src_elems(d, src_kind) * state_bytes(dtype))) ||
        (rc = resolve_mut(vL, &oL, "vL", fb)) || (rc = resolve_mut(vR, &oR, "vR", fb)))
        return rc;
    if ((rc = resolve_coeffs(c, a.c, d.ns - 6))) return rc;
    input_map(d, dir, src_kind, a);
    output_map(d, d.ns, a);
    a.src = s;
    a.outL = oL;
    a.outR = oR;
    a.J0f = 3;
    if (order == 3) {
        a.j_beg = 0;
        a.j_end = d.ns - 1;
    } else {
        a.j_beg = 3;
        a.j_end = d.ns - 4;
    }
    // s_preserve_monotonicity works on the core range only (:1562)
    a.mp_beg = 3;
    a.mp_end = d.ns - 4;
    a.copy_halo = (flags & WENO_FUSE_COPY) ? 1 : 0;
    a.sides = (flags & WENO_ONLY_L) ? 1 : ((flags & WENO_ONLY_R) ? 2 : 3);
    a.eps = g->weno_eps;
    a.teno_ct = g->teno_CT;
    l.op = OP_WENO;
    l.order = order;
    l.mode = mode_for(g, order);
    l.mp = (flags & WENO_FUSE_MP) ? 1 : 0;
    l.dtype = dtype;
    l.nontemporal = (flags & WENO_STREAMING) ? 1 : 0;
    const int variant = tuned_prefetch(order, dtype, dir);
    plan_tiles(a, l, tuned_np(order, dtype, dir), tuned_chunk(order, dtype, dir), tuned_threads(order, dtype, dir),
               variant);
    // variant bit 8 keeps the launch on the general kernel
    if (variant < 0 || !(variant & 8)) l.fast5 = plan_sweep5(a, l, s) + 1;
    return WENO_OK;
}

}  // namespace wm

using namespace wm;

extern "C" {

int weno_init(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(WENO_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    }
    if (device < 0) device = g_state.initialised ? g_state.device : 0;
    if (device >= n) return fail(WENO_E_INVALID, "device %d out of range (%d devices)", device, n);
    if (g_state.initialised && g_state.device == device) return WENO_OK;
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (!g_state.stream && (e = hipStreamCreateWithFlags(&g_state.stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
    g_state.device = device;
    g_state.initialised = true;
    return WENO_OK;
}

int weno_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_state.initialised) return WENO_OK;
    (void)hipDeviceSynchronize();
    for (auto& kv : g_state.present) (void)hipFree(kv.second.dev);
    g_state.present.clear();
    if (g_state.stream) (void)hipStreamDestroy(g_state.stream);
    g_state.stream = nullptr;
    g_state.user_stream = nullptr;
    g_state.initialised = false;
    return WENO_OK;
}

const char* weno_last_error(void) { return g_err.c_str(); }

int weno_sync(void)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    hipError_t e = hipStreamSynchronize(stream());
    return e == hipSuccess ? WENO_OK : hip_fail(e, "hipStreamSynchronize");
}

int weno_set_async(int async_mode)
{
    g_state.async = async_mode != 0;
    return WENO_OK;
}

int weno_set_stream(void* s)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    g_state.user_stream = static_cast<hipStream_t>(s);
    return WENO_OK;
}

void* weno_get_stream(void)
{
    if (ensure_init()) return nullptr;
    return stream();
}

int weno_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int weno_target_enter_data(const void* host, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    if (!host || bytes == 0) return fail(WENO_E_INVALID, "enter data: NULL host address or zero size");
    std::lock_guard<std::mutex> lk(g_mu);
    const uintptr_t h = reinterpret_cast<uintptr_t>(host);
    auto it = g_state.present.find(h);
    if (it != g_state.present.end()) {
        if (it->second.bytes != bytes)
            return fail(WENO_E_INVALID, "enter data: %p already present with a different size", host);
        ++it->second.refcount;
        return WENO_OK;
    }
    // reject overlap with another mapping
    auto nx = g_state.present.lower_bound(h);
    if (nx != g_state.present.end() && nx->first < h + bytes)
        return fail(WENO_E_INVALID, "enter data: [%p,+%zu) overlaps a present array", host, bytes);
    if (nx != g_state.present.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second.bytes > h)
            return fail(WENO_E_INVALID, "enter data: [%p,+%zu) overlaps a present array", host, bytes);
    }
    void* dev = nullptr;
    hipError_t e = hipMalloc(&dev, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    g_state.present[h] = Present{bytes, dev, 1};
    return WENO_OK;
}

static int find_present(const void* host, Present** out)
{
    auto it = g_state.present.find(reinterpret_cast<uintptr_t>(host));
    if (it == g_state.present.end())
        return fail(WENO_E_NOT_PRESENT, "%p is not present in the data environment", host);
    *out = &it->second;
    return WENO_OK;
}

int weno_target_update_to(const void* host, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    Present* p;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if ((rc = find_present(host, &p))) return rc;
    }
    if (bytes > p->bytes) return fail(WENO_E_INVALID, "update to: %zu bytes > mapped %zu", bytes, p->bytes);
    hipError_t e = hipMemcpyAsync(p->dev, host, bytes, hipMemcpyHostToDevice, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    return e == hipSuccess ? WENO_OK : hip_fail(e, "update to");
}

int weno_target_update_from(void* host, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    Present* p;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if ((rc = find_present(host, &p))) return rc;
    }
    if (bytes > p->bytes) return fail(WENO_E_INVALID, "update from: %zu bytes > mapped %zu", bytes, p->bytes);
    hipError_t e = hipMemcpyAsync(host, p->dev, bytes, hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    return e == hipSuccess ? WENO_OK : hip_fail(e, "update from");
}

int weno_target_exit_data(const void* host)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    Present* p;
    if ((rc = find_present(host, &p))) return rc;
    if (--p->refcount > 0) return WENO_OK;
    (void)hipStreamSynchronize(stream());
    hipError_t e = hipFree(p->dev);
    g_state.present.erase(reinterpret_cast<uintptr_t>(host));
    return e == hipSuccess ? WENO_OK : hip_fail(e, "hipFree");
}

void* weno_device_address(const void* host)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_state.present.upper_bound(reinterpret_cast<uintptr_t>(host));
    if (it == g_state.present.begin()) return nullptr;
    --it;
    const uintptr_t off = reinterpret_cast<uintptr_t>(host) - it->first;
    if (off >= it->second.bytes) return nullptr;
    return static_cast<char*>(it->second.dev) + off;
}

void* weno_device_alloc(size_t bytes)
{
    if (ensure_init()) return nullptr;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        hip_fail(e, "hipMalloc");
        return nullptr;
    }
    return p;
}

int weno_device_free(void* dev)
{
    if (!dev) return WENO_OK;
    hipError_t e = hipFree(dev);
    return e == hipSuccess ? WENO_OK : hip_fail(e, "hipFree");
}

int weno_memcpy_to_device(void* dev, const void* host, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    hipError_t e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    return e == hipSuccess ? WENO_OK : hip_fail(e, "memcpy to device");
}

int weno_memcpy_to_host(void* host, const void* dev, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    return e == hipSuccess ? WENO_OK : hip_fail(e, "memcpy to host");
}

size_t weno_field_elems(const weno_grid* g)
{
    if (check_grid(g)) return 0;
    const Dims d = dims_of(g, 0);
    return size_t(d.nx + 6) * size_t(d.ny + 6) * size_t(d.nz + 6) * size_t(d.nv);
}

size_t weno_reshaped_elems(const weno_grid* g, int dir)
{
    if (check_grid(g) || check_dir(dir)) return 0;
    const Dims d = dims_of(g, dir);
    return size_t(d.ns + 6) * size_t(d.nk) * size_t(d.nl) * size_t(d.nv);
}

size_t weno_face_elems(const weno_grid* g, int dir)
{
    if (check_grid(g) || check_dir(dir)) return 0;
    const Dims d = dims_of(g, dir);
    return size_t(d.ns) * size_t(d.nk) * size_t(d.nl) * size_t(d.nv);
}

int weno_core_count(const weno_grid* g, int dir)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir))) return rc;
    return dims_of(g, dir).ns - 6;
}

int weno_reshape(const weno_grid* g, int dir, int dtype, const void* v_vf, void* v_rs)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir))) return rc;
    if (dtype != WENO_F64 && dtype != WENO_F32)
        return fail(WENO_E_INVALID, "reshape supports WENO_F64 and WENO_F32 (got %d)", dtype);
    if ((rc = ensure_init())) return rc;
    TileArgs a;
    std::memset(&a, 0, sizeof a);
    const Dims d = dims_of(g, dir);
    const void* s;
    void* o;
    if ((rc = resolve(v_vf, &s, "v_vf", field_elems(d) * state_bytes(dtype))) ||
        (rc = resolve_mut(v_rs, &o, "v_rs_ws", reshaped_elems(d) * state_bytes(dtype))))
        return rc;
    input_map(d, dir, WENO_SRC_FIELD, a);
    output_map(d, d.ns + 6, a);
    a.src = s;
    a.outL = o;
    a.outR = o;
    a.J0f = 0;
    a.j_beg = 1;
    a.j_end = 0;
    a.copy_halo = 1;
    a.sides = 1;
    return finish(launch_permute(a, dtype, stream()), "reshape");
}

int weno_order1_copy(const weno_grid* g, int dir, int dtype, const void* v_vf, void* vL, void* vR)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir)) || (rc = check_dtype(dtype)) || (rc = ensure_init()))
        return rc;
    TileArgs a;
    std::memset(&a, 0, sizeof a);
    const Dims d = dims_of(g, dir);
    const void* s;
    void *oL, *oR;
    const size_t fb = face_elems(d) * face_bytes(dtype);
    if ((rc = resolve(v_vf, &s, "v_vf", field_elems(d) * state_bytes(dtype))) ||
        (rc = resolve_mut(vL, &oL, "vL", fb)) || (rc = resolve_mut(vR, &oR, "vR", fb)))
        return rc;
    input_map(d, dir, WENO_SRC_FIELD, a);
    output_map(d, d.ns, a);
    a.src = s;
    a.outL = oL;
    a.outR = oR;
    a.J0f = 3;
    a.j_beg = 1;
    a.j_end = 0;
    a.copy_halo = 1;
    a.sides = 3;
    return finish(launch_permute(a, dtype, stream()), "order1_copy");
}

int weno_reconstruct(const weno_grid* g, int order, int dir, int src_kind, int dtype, const void* src,
                     const weno_coeffs* c, void* vL, void* vR, int flags)
{
    int rc;
    TileArgs a;
    Launch l{};
    if ((rc = build_reconstruct(g, order, dir, src_kind, dtype, src, c, vL, vR, flags, a, l))) return rc;
    return finish(launch_tile(a, l, stream()), "reconstruct");
}

int weno_mp_limit(const weno_grid* g, int dir, int src_kind, int dtype, const void* src, void* vL, void* vR)
{
    int rc;
    if ((rc = check_grid(g)) || (rc = check_dir(dir)) || (rc = check_dtype(dtype)) || (rc = ensure_init()))
        return rc;
    if (src_kind != WENO_SRC_FIELD && src_kind != WENO_SRC_RESHAPED)
        return fail(WENO_E_INVALID, "src_kind must be WENO_SRC_FIELD or WENO_SRC_RESHAPED");
    if ((rc = ensure_init())) return rc;
    TileArgs a;
    std::memset(&a, 0, sizeof a);
    Launch l{};
    const Dims d = dims_of(g, dir);
    const void* s;
    void *oL, *oR;
    const size_t fb = face_elems(d) * face_bytes(dtype);
    if ((rc = resolve(src, &s, "src", src_elems(d, src_kind) * state_bytes(dtype))) ||
        (rc = resolve_mut(vL, &oL, "vL", fb)) || (rc = resolve_mut(vR, &oR, "vR", fb)))
        return rc;
    input_map(d, dir, src_kind, a);
    output_map(d, d.ns, a);
    a.src = s;
    a.outL = oL;
    a.outR = oR;
    a.J0f = 3;
    a.j_beg = 3;
    a.j_end = d.ns - 4;
    a.mp_beg = 3;
    a.mp_end = d.ns - 4;
    a.copy_halo = 0;
    a.sides = 3;
    l.op = OP_MP;
    l.dtype = dtype;
    plan_tiles(a, l, 0, 0);
    return finish(launch_tile(a, l, stream()), "mp_limit");
}

}  // extern "C"
