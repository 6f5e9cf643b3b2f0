// extern "C" entry points for batched blocks, hipGraph capture of a whole
// x/y/z sweep, and the tile-shape autotuner (none has a reference
// counterpart: the harness reconstructs one box, direction by direction).
#include "weno_capi_internal.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

namespace wm {

struct BatchDesc {
    TileArgs a;
    int order;
    int mode;
    int item0;
    int pad;
};
hipError_t launch_batched(const BatchDesc* descs_dev, int ndesc, int total, int* counter, size_t lds, int threads,
                          int mp, hipStream_t st);
int batch_max_threads();
int batch_max_np();
void set_tuned(int order, int dtype, int dir, int np, int chunk, int threads, int prefetch);

// Allocation-free batched launches keep the work-item counter and the block
// descriptors in this file-scope device array instead of library-allocated
// memory: the counter in the first 256 bytes, then up to kNoallocMaxBlocks
// descriptors.  Each GPU holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: the
// device address comes from hipGetSymbolAddress, once per device.
constexpr int kNoallocMaxBlocks = WENO_NOALLOC_MAX_BLOCKS;
constexpr size_t kNoallocCounterBytes = 256;
constexpr size_t kNoallocScratchBytes = kNoallocCounterBytes + size_t(kNoallocMaxBlocks) * sizeof(BatchDesc);
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
static_assert(sizeof(BatchDesc) == 280, "the header states the scratch size from a 280-byte descriptor");

namespace {

struct BatchTune {
    int threads = 512;
    int np_target = 16;
};
std::mutex g_bmu;
BatchTune g_btune;

// one direction of a batch, ready to launch
struct BatchPlan {
    std::vector<BatchDesc> descs;
    int total = 0;
    size_t lds = 0;
    int threads = 0;
    int mp = 0;
};

struct GraphHandle {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void*> buffers;   // device memory owned by the graph
};

// coefficient table row stride (BetaTraits::LD) per order
int entries_of(int order) { return order == 3 ? 7 : (order == 5 ? 15 : 33); }

// One block's launch description along dir; cost orders the work queue
// (largest blocks first) and lds is the dynamic LDS the block's tiles need.
int describe_block(const weno_block& B, int b, int dir, int flags, int threads, int np_target, BatchDesc& D,
                   double& cost, size_t& lds)
{
    int rc;
    const weno_grid* g = &B.grid;
    if ((rc = check_grid(g))) return fail(rc, "block %d: %s", b, weno_last_error());
    if (B.order != 3 && B.order != 5 && B.order != 7)
        return fail(WENO_E_INVALID, "block %d: order must be 3, 5 or 7 (got %d)", b, B.order);
    const int mp = (flags & WENO_FUSE_MP) ? 1 : 0;
    const Dims d = dims_of(g, dir);
    std::memset(&D, 0, sizeof D);
    TileArgs& a = D.a;
    const void* s;
    void *oL, *oR;
    const size_t fb = face_elems(d) * sizeof(double);
    if ((rc = resolve(B.v_vf, &s, "v_vf", field_elems(d) * sizeof(double))) ||
        (rc = resolve_mut(B.vL[dir], &oL, "vL", fb)) || (rc = resolve_mut(B.vR[dir], &oR, "vR", fb)) ||
        (rc = resolve_coeffs(&B.coeffs[dir], a.c, d.ns - 6)))
        return fail(rc, "block %d: %s", b, weno_last_error());
    input_map(d, dir, WENO_SRC_FIELD, a);
    output_map(d, d.ns, a);
    a.src = s;
    a.outL = oL;
    a.outR = oR;
    a.J0f = 3;
    a.j_beg = B.order == 3 ? 0 : 3;
    a.j_end = B.order == 3 ? d.ns - 1 : d.ns - 4;
    a.mp_beg = 3;
    a.mp_end = d.ns - 4;
    a.copy_halo = (flags & WENO_FUSE_COPY) ? 1 : 0;
    a.sides = 3;
    a.eps = g->weno_eps;
    a.teno_ct = g->teno_CT;
    a.chunk = d.ns;
    a.chunks = 1;
    const int R = std::max(B.order / 2, mp ? 2 : 0);
    const int jn = a.copy_halo ? d.ns : a.j_end - a.j_beg + 1;
    if (jn > threads)
        return fail(WENO_E_INVALID, "block %d: sweep length %d exceeds the batched workgroup (%d threads)", b, jn,
                    threads);
    const int lo = a.copy_halo ? std::min(3, a.j_beg + 3 - R) : a.j_beg + 3 - R;
    const int hi = a.copy_halo ? std::max(d.ns + 2, a.j_end + 3 + R) : a.j_end + 3 + R;
    const int ldj = (hi - lo + 1) | 1;
    const int G = std::max(1, threads / jn);
    int npg = std::max(1, np_target / G);
    npg = std::min(npg, std::max(1, batch_max_np() / G));
    auto lds_of = [&](int np) { return (size_t(np) * ldj + size_t(entries_of(B.order)) * jn) * sizeof(double); };
    while (npg > 1 && lds_of(G * npg) > 150 * 1024) --npg;
    a.groups = std::min(G, batch_max_np());
    a.npg = npg;
    a.np = a.groups * npg;
    a.jn = jn;
    a.ldj = ldj;
    a.threads = threads;
    a.tiles_per_var = static_cast<int>((static_cast<long long>(a.na) * a.nb + a.np - 1) / a.np);
    D.order = B.order;
    D.mode = mode_for(g, B.order);
    lds = lds_of(a.np);
    cost = double(a.tiles_per_var) * a.nv * a.np * jn;
    return WENO_OK;
}

int check_batch_args(const weno_block* blocks, int nblocks, int flags)
{
    if (!blocks || nblocks <= 0) return fail(WENO_E_INVALID, "no blocks");
    if (flags & ~(WENO_FUSE_MP | WENO_FUSE_COPY))
        return fail(WENO_E_INVALID, "batched flags: only WENO_FUSE_MP and WENO_FUSE_COPY (got 0x%x)", flags);
    return WENO_OK;
}

int plan_batch(const weno_block* blocks, int nblocks, int dir, int flags, int threads, int np_target,
               BatchPlan& plan)
{
    int rc;
    if ((rc = check_batch_args(blocks, nblocks, flags))) return rc;
    plan.descs.clear();
    plan.threads = threads;
    plan.mp = (flags & WENO_FUSE_MP) ? 1 : 0;
    plan.lds = 0;
    std::vector<std::pair<double, BatchDesc>> work(nblocks);
    for (int b = 0; b < nblocks; ++b) {
        size_t lds = 0;
        if ((rc = describe_block(blocks[b], b, dir, flags, threads, np_target, work[b].second, work[b].first, lds)))
            return rc;
        plan.lds = std::max(plan.lds, lds);
    }
    // largest blocks first: their items are pulled first and finish last
    std::stable_sort(work.begin(), work.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    plan.total = 0;
    for (auto& w : work) {
        w.second.item0 = plan.total;
        plan.total += w.second.a.tiles_per_var * w.second.a.nv;
        plan.descs.push_back(w.second);
    }
    return WENO_OK;
}

int upload(const void* host, size_t bytes, void** dev, std::vector<void*>& owned)
{
    hipError_t e = hipMalloc(dev, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    owned.push_back(*dev);
    e = hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? WENO_OK : hip_fail(e, "descriptor upload");
}

// library-owned buffers for the un-captured batched call
std::vector<void*> g_batch_bufs;
size_t g_batch_cap = 0;
void* g_batch_desc = nullptr;
int* g_batch_counter = nullptr;

int batch_once(const BatchPlan& plan)
{
    const size_t bytes = plan.descs.size() * sizeof(BatchDesc);
    hipError_t e;
    if (bytes > g_batch_cap) {
        if (g_batch_desc) (void)hipFree(g_batch_desc);
        g_batch_desc = nullptr;
        if ((e = hipMalloc(&g_batch_desc, bytes)) != hipSuccess) return hip_fail(e, "hipMalloc");
        g_batch_cap = bytes;
    }
    if (!g_batch_counter && (e = hipMalloc(reinterpret_cast<void**>(&g_batch_counter), 64)) != hipSuccess)
        return hip_fail(e, "hipMalloc");
    if ((e = hipMemcpyAsync(g_batch_desc, plan.descs.data(), bytes, hipMemcpyHostToDevice, stream())) != hipSuccess)
        return hip_fail(e, "descriptor upload");
    return finish(launch_batched(static_cast<const BatchDesc*>(g_batch_desc), static_cast<int>(plan.descs.size()),
                                 plan.total, g_batch_counter, plan.lds, plan.threads, plan.mp, stream()),
                  "batched");
}

float time_ms(const std::function<int()>& fn, int reps)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.f;
    float ms = -1.f;
    if (fn() == WENO_OK && hipEventRecord(a, stream()) == hipSuccess) {
        bool ok = true;
        for (int r = 0; r < reps && ok; ++r) ok = fn() == WENO_OK;
        if (ok && hipEventRecord(b, stream()) == hipSuccess && hipEventSynchronize(b) == hipSuccess)
            (void)hipEventElapsedTime(&ms, a, b);
        if (ms >= 0) ms /= reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace
}  // namespace wm

using namespace wm;

extern "C" {

int weno_batched_reconstruct(const weno_block* blocks, int nblocks, int dir, int flags)
{
    int rc;
    if ((rc = check_dir(dir)) || (rc = check_batch_args(blocks, nblocks, flags))) return rc;
    if ((rc = ensure_init())) return rc;
    BatchTune t;
    {
        std::lock_guard<std::mutex> lk(g_bmu);
        t = g_btune;
    }
    BatchPlan plan;
    if ((rc = plan_batch(blocks, nblocks, dir, flags, t.threads, t.np_target, plan))) return rc;
    return batch_once(plan);
}

int weno_batched_graph_create(const weno_block* blocks, int nblocks, int flags, void** handle)
{
    int rc;
    if (!handle) return fail(WENO_E_INVALID, "handle is NULL");
    *handle = nullptr;
    if ((rc = ensure_init())) return rc;
    BatchTune t;
    {
        std::lock_guard<std::mutex> lk(g_bmu);
        t = g_btune;
    }
    auto* G = new GraphHandle();
    BatchPlan plans[3];
    void* descs[3] = {nullptr, nullptr, nullptr};
    void* counters = nullptr;
    for (int d = 0; d < 3 && rc == WENO_OK; ++d) {
        if ((rc = plan_batch(blocks, nblocks, d, flags, t.threads, t.np_target, plans[d]))) break;
        rc = upload(plans[d].descs.data(), plans[d].descs.size() * sizeof(BatchDesc), &descs[d], G->buffers);
    }
    if (rc == WENO_OK) {
        hipError_t e = hipMalloc(&counters, 3 * 64);
        if (e != hipSuccess) rc = hip_fail(e, "hipMalloc");
        else G->buffers.push_back(counters);
    }
    if (rc == WENO_OK) {
        hipError_t e = hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) rc = hip_fail(e, "hipStreamBeginCapture");
        for (int d = 0; d < 3 && e == hipSuccess; ++d)
            e = launch_batched(static_cast<const BatchDesc*>(descs[d]), static_cast<int>(plans[d].descs.size()),
                               plans[d].total, reinterpret_cast<int*>(static_cast<char*>(counters) + 64 * d),
                               plans[d].lds, plans[d].threads, plans[d].mp, stream());
        hipError_t e2 = hipStreamEndCapture(stream(), &G->graph);
        if (rc == WENO_OK && e != hipSuccess) rc = hip_fail(e, "batched launch under capture");
        if (rc == WENO_OK && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
        if (rc == WENO_OK && (e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0)) != hipSuccess)
            rc = hip_fail(e, "hipGraphInstantiate");
    }
    if (rc != WENO_OK) {
        weno_graph_destroy(G);
        return rc;
    }
    *handle = G;
    return WENO_OK;
}

int weno_sweep_graph_create(const weno_grid* g, int order, int src_kind, int dtype, const void* src,
                            const weno_coeffs* coeffs3, void* const* vL3, void* const* vR3, int flags, void** handle)
{
    int rc;
    if (!handle || !coeffs3 || !vL3 || !vR3) return fail(WENO_E_INVALID, "NULL argument");
    *handle = nullptr;
    for (int d = 0; d < 3; ++d)
        if ((rc = validate_reconstruct(g, order, d, src_kind, dtype, flags))) return rc;
    if ((rc = ensure_init())) return rc;
    // plan (and resolve) every launch before capturing
    TileArgs a[3];
    Launch l[3];
    for (int d = 0; d < 3; ++d)
        if ((rc = build_reconstruct(g, order, d, src_kind, dtype, src, &coeffs3[d], vL3[d], vR3[d], flags, a[d],
                                    l[d])))
            return rc;
    auto* G = new GraphHandle();
    hipError_t e = hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete G;
        return hip_fail(e, "hipStreamBeginCapture");
    }
    for (int d = 0; d < 3 && e == hipSuccess; ++d) e = launch_tile(a[d], l[d], stream());
    hipError_t e2 = hipStreamEndCapture(stream(), &G->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        weno_graph_destroy(G);
        return hip_fail(e, "sweep graph");
    }
    *handle = G;
    return WENO_OK;
}

int weno_graph_launch(void* handle)
{
    int rc;
    if (!handle) return fail(WENO_E_INVALID, "graph handle is NULL");
    if ((rc = ensure_init())) return rc;
    auto* G = static_cast<GraphHandle*>(handle);
    return finish(hipGraphLaunch(G->exec, stream()), "hipGraphLaunch");
}

int weno_graph_destroy(void* handle)
{
    if (!handle) return WENO_OK;
    auto* G = static_cast<GraphHandle*>(handle);
    (void)hipStreamSynchronize(stream());
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    for (void* p : G->buffers) (void)hipFree(p);
    delete G;
    return WENO_OK;
}

int weno_autotune_sweep(const weno_grid* g, int order, int dtype, int src_kind, const void* src,
                        const weno_coeffs* coeffs3, void* const* vL3, void* const* vR3, int flags, int* best)
{
    int rc;
    if (!coeffs3 || !vL3 || !vR3) return fail(WENO_E_INVALID, "NULL argument");
    for (int d = 0; d < 3; ++d)
        if ((rc = validate_reconstruct(g, order, d, src_kind, dtype, flags))) return rc;
    if ((rc = ensure_init())) return rc;
    const bool was_async = g_state.async;
    g_state.async = true;
    static const int kThreads[] = {256, 384, 448, 512, 640, 832};
    // general-kernel variants (bit 8 set: WENO5-JS FP64 stays off the
    // producer/consumer kernel, which has a fixed shape and is timed once)
    static const int kVariants[] = {10, 11, 14, 8};
    static const int kPencils[] = {8, 12, 16, 24, 32};
    for (int d = 0; d < 3 && rc == WENO_OK; ++d) {
        float best_ms = 1e30f;
        int bt = 0, bv = -1, bp = 0;
        set_tuned(order, dtype, d, 0, 0, 0, -1);
        {
            const float ms = time_ms([&] {
                return weno_reconstruct(g, order, d, src_kind, dtype, src, &coeffs3[d], vL3[d], vR3[d], flags);
            }, 3);
            if (ms > 0) {
                best_ms = ms;
                bt = 512;
                bv = -1;
                bp = 16;
            }
        }
        bool ran = best_ms < 1e29f;
        for (int th : kThreads)
            for (int var : kVariants)
                for (int np : kPencils) {
                    if (order == 7 && th > 512) continue;
                    set_tuned(order, dtype, d, np, 0, th, var);
                    const float ms = time_ms([&] {
                        return weno_reconstruct(g, order, d, src_kind, dtype, src, &coeffs3[d], vL3[d], vR3[d],
                                                flags);
                    }, 3);
                    if (ms > 0) ran = true;
                    if (ms > 0 && ms < best_ms) {
                        best_ms = ms;
                        bt = th;
                        bv = var;
                        bp = np;
                    }
                }
        if (!ran) {
            rc = fail(WENO_E_HIP, "autotune: no candidate ran for direction %d", d);
            set_tuned(order, dtype, d, 0, 0, 0, -1);
            break;
        }
        if (bv < 0)
            set_tuned(order, dtype, d, 0, 0, 0, -1);
        else
            set_tuned(order, dtype, d, bp, 0, bt, bv);
        if (best) {
            best[4 * d + 0] = bp;
            best[4 * d + 1] = bt;
            best[4 * d + 2] = bv;
            best[4 * d + 3] = static_cast<int>(best_ms * 1000.f);
        }
    }
    (void)hipStreamSynchronize(stream());
    g_state.async = was_async;
    return rc;
}

int weno_batched_autotune(const weno_block* blocks, int nblocks, int flags, int* best)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    const bool was_async = g_state.async;
    g_state.async = true;
    float best_ms = 1e30f;
    BatchTune bt;
    for (int th : {256, 512})
        for (int np : {8, 16, 32, 64, 128, 256}) {
            BatchPlan plans[3];
            bool ok = true;
            for (int d = 0; d < 3 && ok; ++d) ok = plan_batch(blocks, nblocks, d, flags, th, np, plans[d]) == WENO_OK;
            if (!ok) continue;
            const float ms = time_ms([&] {
                int r = WENO_OK;
                for (int d = 0; d < 3 && r == WENO_OK; ++d) r = batch_once(plans[d]);
                return r;
            }, 3);
            if (ms > 0 && ms < best_ms) {
                best_ms = ms;
                bt.threads = th;
                bt.np_target = np;
            }
        }
    (void)hipStreamSynchronize(stream());
    g_state.async = was_async;
    if (best_ms >= 1e30f) return fail(WENO_E_INVALID, "batched autotune: no candidate fits (%s)", weno_last_error());
    {
        std::lock_guard<std::mutex> lk(g_bmu);
        g_btune = bt;
    }
    if (best) {
        best[0] = bt.threads;
        best[1] = bt.np_target;
        best[2] = static_cast<int>(best_ms * 1000.f);
    }
    return WENO_OK;
}

int weno_batched_reconstruct_noalloc(const weno_block* blocks, int nblocks, int dir, int flags)
{
    int rc;
    if ((rc = check_dir(dir))) return rc;
    if ((rc = check_batch_args(blocks, nblocks, flags))) return rc;
    if (nblocks > kNoallocMaxBlocks)
        return fail(WENO_E_INVALID, "allocation-free batched launch: %d blocks, at most %d", nblocks,
                    kNoallocMaxBlocks);
    if ((rc = ensure_init())) return rc;
    // one call at a time per process: the host staging below and each GPU's
    // scratch array are shared by every call
    static std::mutex mu;
    static void* dev_scratch[64] = {};
    static BatchDesc described[kNoallocMaxBlocks], queued[kNoallocMaxBlocks];
    static double cost[kNoallocMaxBlocks];
    static int order_of[kNoallocMaxBlocks];
    std::lock_guard<std::mutex> lk(mu);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(WENO_E_INVALID, "device %d out of range for the scratch table", dev);
    if (!dev_scratch[dev]) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess)
            return hip_fail(e, "hipGetSymbolAddress(scratch)");
        dev_scratch[dev] = p;
    }
    BatchTune t;
    {
        std::lock_guard<std::mutex> tl(g_bmu);
        t = g_btune;
    }
    size_t lds = 0;
    for (int b = 0; b < nblocks; ++b) {
        size_t need = 0;
        if ((rc = describe_block(blocks[b], b, dir, flags, t.threads, t.np_target, described[b], cost[b], need)))
            return rc;
        lds = std::max(lds, need);
    }
    // largest blocks first (stable for equal costs), as weno_batched_reconstruct
    for (int b = 0; b < nblocks; ++b) {
        int j = b;
        while (j > 0 && cost[order_of[j - 1]] < cost[b]) {
            order_of[j] = order_of[j - 1];
            --j;
        }
        order_of[j] = b;
    }
    int total = 0;
    for (int k = 0; k < nblocks; ++k) {
        queued[k] = described[order_of[k]];
        queued[k].item0 = total;
        total += queued[k].a.tiles_per_var * queued[k].a.nv;
    }
    char* const base = static_cast<char*>(dev_scratch[dev]);
    BatchDesc* const descs = reinterpret_cast<BatchDesc*>(base + kNoallocCounterBytes);
    int* const counter = reinterpret_cast<int*>(base);
    if ((e = hipMemcpyAsync(descs, queued, size_t(nblocks) * sizeof(BatchDesc), hipMemcpyHostToDevice, stream())) !=
        hipSuccess)
        return hip_fail(e, "descriptor upload");
    return finish(launch_batched(descs, nblocks, total, counter, lds, t.threads, (flags & WENO_FUSE_MP) ? 1 : 0,
                                 stream()),
                  "batched (allocation-free)");
}

}  // extern "C"
