// A plain streaming copy: the yardstick the sweep kernels' HBM traffic and
// bandwidth are calibrated against (no reference counterpart).
#include "weno_capi_internal.h"

namespace wm {

// each workgroup moves one contiguous 16 KiB chunk: four 16-byte loads per
// lane in flight, then four stores
__global__ void __launch_bounds__(256) stream_copy_kernel(const double2* __restrict__ src,
                                                          double2* __restrict__ dst, size_t n)
{
    const size_t base = blockIdx.x * size_t(1024) + threadIdx.x;
    double2 r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const size_t i = base + size_t(u) * 256;
        if (i < n) r[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const size_t i = base + size_t(u) * 256;
        if (i < n) dst[i] = r[u];
    }
}

}  // namespace wm

using namespace wm;

extern "C" int weno_stream_copy(void* dst, const void* src, size_t bytes)
{
    int rc;
    if ((rc = ensure_init())) return rc;
    if (bytes % 16) return fail(WENO_E_INVALID, "stream copy size must be a multiple of 16 bytes");
    const void* s;
    void* d;
    if ((rc = resolve(src, &s, "src", bytes)) || (rc = resolve_mut(dst, &d, "dst", bytes))) return rc;
    if ((reinterpret_cast<uintptr_t>(s) | reinterpret_cast<uintptr_t>(d)) & 15)
        return fail(WENO_E_INVALID, "stream copy needs 16-byte aligned buffers");
    const size_t n = bytes / 16;
    const size_t blocks = (n + 1023) / 1024;
    if (blocks == 0) return WENO_OK;
    if (blocks > 0x7fffffffu) return fail(WENO_E_INVALID, "stream copy too large");
    hipLaunchKernelGGL(stream_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream(),
                       static_cast<const double2*>(s), static_cast<double2*>(d), n);
    return finish(hipGetLastError(), "stream_copy");
}
