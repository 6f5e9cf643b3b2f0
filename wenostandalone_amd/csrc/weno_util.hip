// A plain streaming copy: the yardstick the sweep kernels' HBM traffic and
// bandwidth are calibrated against (no reference counterpart).
#include "weno_capi_internal.h"

namespace wm {

__global__ void __launch_bounds__(256) stream_copy_kernel(const double2* __restrict__ src,
                                                          double2* __restrict__ dst, size_t n)
{
    size_t i = blockIdx.x * size_t(256) + threadIdx.x;
    const size_t stride = size_t(gridDim.x) * 256;
    for (; i < n; i += stride) dst[i] = src[i];
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
    if ((rc = resolve(src, &s, "src")) || (rc = resolve_mut(dst, &d, "dst"))) return rc;
    if ((reinterpret_cast<uintptr_t>(s) | reinterpret_cast<uintptr_t>(d)) & 15)
        return fail(WENO_E_INVALID, "stream copy needs 16-byte aligned buffers");
    const size_t n = bytes / 16;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    size_t blocks = (n + 255) / 256;
    if (blocks > size_t(cus) * 16) blocks = size_t(cus) * 16;
    if (blocks == 0) return WENO_OK;
    hipLaunchKernelGGL(stream_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream(),
                       static_cast<const double2*>(s), static_cast<double2*>(d), n);
    return finish(hipGetLastError(), "stream_copy");
}
