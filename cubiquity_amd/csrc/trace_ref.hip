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

cqx_status cqx_trace_ref_host(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                              cqx_intersection* out)
{
    if (!ctx || !vol) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (n == 0) return CQX_OK;
    if (!rays || !out) return fail(CQX_ERR_INVALID_ARG, "null ray or result buffer");
    HIP_TRY(hipSetDevice(ctx->device));
    cqx_status st = cqx::ensure_scratch(ctx, size_t(n) * (sizeof(cqx_ray) + sizeof(cqx_intersection)));
    if (st != CQX_OK) return st;
    cqx_ray* d_rays = static_cast<cqx_ray*>(ctx->scratch);
    auto* d_out = reinterpret_cast<cqx_intersection*>(static_cast<char*>(ctx->scratch) + size_t(n) * sizeof(cqx_ray));
    HIP_TRY(hipMemcpy(d_rays, rays, size_t(n) * sizeof(cqx_ray), hipMemcpyHostToDevice));
    st = cqx_trace_ref(ctx, vol, d_rays, n, d_out, nullptr);
    if (st != CQX_OK) return st;
    HIP_TRY(hipMemcpy(out, d_out, size_t(n) * sizeof(cqx_intersection), hipMemcpyDeviceToHost));
    return CQX_OK;
}

} // extern "C"
