// traceRayRef on the GPU (src/library/raytracing.cpp:480-535): the reference's brute-force
// raycast, one ray per lane; the per-ray walk is trace_ref_core.h.
#include "device_internal.h"
#include "trace_ref_core.h"

using cqx::fail;

namespace {

constexpr uint32_t kRows = 33; // heights 0..32
constexpr uint32_t kBlock = 128;

__global__ void __launch_bounds__(kBlock) k_trace_ref(const uint32_t* __restrict__ nodes, uint32_t root,
                                                      const cqx_ray* __restrict__ rays, uint64_t n,
                                                      cqx_intersection* __restrict__ out)
{
    __shared__ uint32_t stk_node[kRows * kBlock];
    __shared__ uint8_t stk_next[kRows * kBlock];
    const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = cqxd::trace_ref_ray(nodes, root, rays[i], stk_node + threadIdx.x, stk_next + threadIdx.x, kBlock);
}

} // namespace

extern "C" {

cqx_status cqx_trace_ref(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                         cqx_intersection* out, void* stream)
{
    if (!ctx || !vol) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (n == 0) return CQX_OK;
    if (!rays || !out) return fail(CQX_ERR_INVALID_ARG, "null ray or result buffer");
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 0x7fffffffull) return fail(CQX_ERR_LIMIT, "too many rays for one launch");
    HIP_TRY(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_trace_ref, dim3(uint32_t(blocks)), dim3(kBlock), 0, hipStream_t(stream), vol->nodes,
                       vol->root, rays, n, out);
    HIP_TRY(hipGetLastError());
    return CQX_OK;
}

} // extern "C"
