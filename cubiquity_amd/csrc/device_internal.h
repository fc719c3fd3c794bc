// Internal declarations shared by the HIP sources of libcqx.so.
#pragma once

#include "cqx_internal.h"
#include "trace_ranges.h"
#include "traverse_core.h"

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#define HIP_TRY(expr)                                                                                            \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return cqx::fail(CQX_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

namespace cqx {

struct PtPlan; // path-tracer buffers and captured graph (pathtrace.hip)
void destroy_pt_plan(PtPlan* plan);
struct DemoPlan; // buffers of the reference-shading mode (demo.hip)
void destroy_demo_plan(DemoPlan* plan);
struct DeviceEdits; // a volume's on-device edit state (edits_gpu.hip)
void destroy_device_edits(DeviceEdits* e);

// Best-effort release paths: an error here has nowhere useful to go.
inline void release(void* p) { (void)hipFree(p); }

constexpr uint32_t kDefaultMaxIters = 1u << 22;

} // namespace cqx

struct cqx_ctx {
    int device = 0;
    int cu_count = 0;
    int lds_per_block = 0;
    uint64_t* d_counter = nullptr; // work counters of the persistent kernel (8 queues, 128 B apart)
    void* scratch = nullptr;       // staging for cqx_intersect_host
    size_t scratch_bytes = 0;
    void* bins = nullptr; // ray binning: cell counts/offsets, then the order (u32) and cells (u16) of n rays
    size_t bins_bytes = 0;
    hipStream_t capture_stream = nullptr; // private stream for graph capture
    cqx::PtPlan* pt_plan = nullptr;
    cqx::DemoPlan* demo_plan = nullptr;
    void* batch_desc = nullptr; // descriptor table of the batched launch
    size_t batch_bytes = 0;
    void* noalloc_scratch = nullptr; // this device's static scratch (cqx_intersect_noalloc), not owned
    struct Tuned {
        int height; // SubDAG height of the volume tuned on
        int cls;    // batch-size class: floor(log2 n)
        cqx_launch_cfg cfg;
        float ms;
    };
    std::vector<Tuned> tuned; // autotuner cache
};

struct cqx_dvolume {
    uint32_t* nodes = nullptr;
    cqx_subdag* sub = nullptr;
    uint64_t node_count = 0;
    uint64_t capacity = 0; // nodes the device buffer holds (headroom for appended edits)
    uint32_t root = 0;
    cqx_subdag hsub[8];
    int max_height = 0; // largest SubDAG start height (stack rows needed: max_height + 1)
    cqx::DeviceEdits* edits = nullptr; // index and buffers of on-device edits, made on first use
};

namespace cqx {

// Autotuner cache lookup: the tuned configuration for a volume of this SubDAG height and a batch of
// about n rays (or of the nearest tuned height and batch size, each within 2), or null.
const cqx_launch_cfg* tuned_cfg(const cqx_ctx* ctx, int height, uint64_t n);

// Device scratch a launch uses instead of the context's buffers: the persistent kernels' work
// counters (1 KiB) and the ray-binning buffers (bins_bytes long).
struct ScratchView {
    void* counters = nullptr;
    void* bins = nullptr;
    size_t bins_bytes = 0;
};

// Grows the context's staging buffer (host-buffer entry points) to at least `bytes`.
cqx_status ensure_scratch(cqx_ctx* ctx, size_t bytes);
// Launches the traversal of `n` rays (or of *n_dev rays, at most n, when n_dev is not null).
// Counters and binning buffers are the context's, or `sv`'s when it is given (then nothing is
// allocated).
cqx_status launch_intersect(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                            const uint32_t* n_dev, uint32_t flags, float max_footprint, cqx_hit* hits,
                            const cqx_launch_cfg* cfg, hipStream_t stream, const ScratchView* sv = nullptr);

// Threads per block of the kernels that compact lists with block_slots.
constexpr uint32_t kCompactBlock = 1024;

// Stream compaction into N lists at once: a thread with take[l] set gets slot[l], its entry's
// index in list l. Each list's slots for the whole block are claimed with one atomicAdd on
// count[l] (counts may be shared: every list claims its own range). One atomic per wave on a
// counter that every block uses serialises at the memory side (several nanoseconds each), which
// made it the largest cost of a shading pass over millions of rays; per block it is 16 times
// rarer. Entries keep no order across blocks. Call once per kernel, from every thread of a
// kBlock-thread block (it synchronises the block).
template <uint32_t kBlock, int N>
__device__ __forceinline__ void block_slots(const bool (&take)[N], uint32_t* const (&count)[N], uint32_t (&slot)[N])
{
    static_assert(kBlock % 64 == 0 && kBlock <= 1024, "whole waves, at most 16");
    constexpr uint32_t kWaves = kBlock / 64;
    __shared__ uint32_t part[N][kWaves + 1]; // per-wave counts, then their prefix; [kWaves]: the block's base
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t mask[N];
#pragma unroll
    for (int l = 0; l < N; l++) {
        mask[l] = __ballot(take[l]);
        if (lane == 0) part[l][wave] = uint32_t(__popcll(mask[l]));
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < N; l++) {
        if (threadIdx.x == uint32_t(l)) { // list l's prefix and claim, on thread l
            uint32_t sum = 0;
            for (uint32_t w = 0; w < kWaves; w++) {
                const uint32_t c = part[l][w];
                part[l][w] = sum;
                sum += c;
            }
            part[l][kWaves] = sum ? atomicAdd(count[l], sum) : 0u;
        }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int l = 0; l < N; l++) slot[l] = part[l][kWaves] + part[l][wave] + uint32_t(__popcll(mask[l] & below));
}

} // namespace cqx
