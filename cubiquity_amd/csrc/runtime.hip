// libcqx device runtime: contexts, device node stores, and the traversal kernels for gfx950.
//
// Kernels
//   k_intersect_lane        one ray per lane, one pass over the ray array (variant 1)
//   k_intersect_persistent  persistent waves whose traversal steps run SubDAG walks only: a lane
//                           whose walk ends waits; once enough lanes of the wave wait, a service
//                           pass moves them on to their next SubDAG, writes the records of the
//                           finished rays, and claims a contiguous block of new rays from a
//                           global counter with one atomic, handed out by ballot prefix
//                           (variant 2); variant 3 also keeps the current node's 8 children in
//                           registers
//   k_ray_bounds / k_bin_grid / k_bin_count / k_bin_columns / k_bin_scan / k_bin_scatter
//                           optional prepass of the persistent variants: rays binned by origin
//                           into a coarse Morton grid over the origins' bounds, handed out cell
//                           by cell
// All keep the traversal stack in LDS (one column per lane, one row per node height).
#include "device_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

using namespace cqxd;
using cqx::fail;
using cqx::release;

namespace {

// ---- variant 1: one ray per lane --------------------------------------------------------------
__global__ void __launch_bounds__(256) k_intersect_lane(TraceParams P, const cqx_ray* __restrict__ rays, uint64_t n,
                                                        const uint32_t* __restrict__ n_dev, cqx_hit* __restrict__ hits)
{
    extern __shared__ uint32_t lds_stack[];
    if (n_dev) {
        const uint64_t nd = *n_dev;
        n = nd < n ? nd : n;
    }
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // rays are read once and records written once: streaming accesses keep node lines in L2
    const float* r = reinterpret_cast<const float*>(rays + i);
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = __builtin_nontemporal_load(r + k);
    cqx_hit h;
    trace_ray(P, v[0], v[1], v[2], v[3], v[4], v[5], lds_stack + threadIdx.x, blockDim.x, h, occupied_octants(P.sub));
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&h);
    uint32_t* dst = reinterpret_cast<uint32_t*>(hits + i);
#pragma unroll
    for (int k = 0; k < 8; k++) __builtin_nontemporal_store(src[k], dst + k);
}

// ---- variant 2: persistent waves with ballot refill ------------------------------------------
// Every wave owns 64 lanes for the whole launch. Between bursts of `regroup` traversal steps the
// wave ballots its idle lanes; once at least `refill` lanes are idle (or all are), one lane claims
// that many consecutive ray ids from the wave's queue with a single atomic and the idle lanes take
// them in lane order (ballot prefix). The ray array is split into one queue per XCD
// (blockIdx.x % 8, the order in which workgroups are spread over the XCDs), so waves of one XCD
// draw from their own counter; a wave whose queue is drained moves on to the next queue. A wave
// exits when every queue is drained and all its lanes are idle.
constexpr int kQueues = 8;
constexpr int kCounterStride = 16; // queue counters 128 bytes apart: one cache line each

__global__ void k_reset_counters(unsigned long long* counters)
{
    if (threadIdx.x < kQueues) counters[threadIdx.x * kCounterStride] = 0;
}

// Rays are read once and records written once: streaming (non-temporal) accesses keep them from
// displacing node lines in L2.
__device__ __forceinline__ void setup_from(RayState& S, const float* r)
{
    float v[6];
#pragma unroll
    for (int i = 0; i < 6; i++) v[i] = __builtin_nontemporal_load(r + i);
    ray_setup(S, v[0], v[1], v[2], v[3], v[4], v[5]);
}

__device__ __forceinline__ void store_record(cqx_hit* dst, const cqx_hit& h)
{
    const uint32_t* s = reinterpret_cast<const uint32_t*>(&h);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
#pragma unroll
    for (int i = 0; i < 8; i++) __builtin_nontemporal_store(s[i], d + i);
}

// Where a persistent wave's ray ids come from. SingleSource: one volume, one ray array; its
// parameters stay wave-uniform. BatchSource: many (volume, ray array) batches laid end to end in
// one id space; a lane finds its batch by binary search over the batches' first ids.
struct SingleSource {
    static constexpr bool kUniform = true; // P is the same for every lane (kernel argument)
    TraceParams P;
    const cqx_ray* __restrict__ rays;
    cqx_hit* __restrict__ hits;
    const uint32_t* __restrict__ order = nullptr; // binned launch: the ray behind each id
    __device__ __forceinline__ void fetch(uint64_t id, TraceParams&, const float*& ray, cqx_hit*& out) const
    {
        const uint64_t r = order ? uint64_t(order[id]) : id;
        ray = reinterpret_cast<const float*>(rays + r);
        out = hits + r;
    }
};

struct BatchDesc {
    const uint32_t* nodes;
    const cqx_subdag* sub;
    const cqx_ray* rays;
    cqx_hit* hits;
    uint64_t first; // first id of this batch in the launch's id space
    uint32_t flags;
    float max_footprint;
};

struct BatchSource {
    static constexpr bool kUniform = false; // each lane's batch has its own volume and flags
    const BatchDesc* __restrict__ desc;
    uint32_t n_batches;
    uint32_t max_iters;
    __device__ __forceinline__ void fetch(uint64_t id, TraceParams& T, const float*& ray, cqx_hit*& out) const
    {
        uint32_t lo = 0, hi = n_batches; // invariant: desc[lo].first <= id < desc[hi].first
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (desc[mid].first <= id) lo = mid;
            else hi = mid;
        }
        const BatchDesc& b = desc[lo];
        T.nodes = b.nodes;
        T.sub = b.sub;
        T.surf = (b.flags & CQX_SURFACE_PROPERTIES) ? 1u : 0u;
        T.max_footprint = b.max_footprint;
        T.max_iters = max_iters;
        T.fp_disabled = footprint_is_disabled(b.max_footprint) ? 1u : 0u;
        const uint64_t k = id - b.first;
        ray = reinterpret_cast<const float*>(b.rays + k);
        out = b.hits + k;
    }
};

// The parameters the traversal steps use: the kernel argument when they are the same for every
// lane (kept in scalar registers), the lane's own copy otherwise.
__device__ __forceinline__ const TraceParams& uniform_params(const SingleSource& s, const TraceParams&) { return s.P; }
__device__ __forceinline__ const TraceParams& uniform_params(const BatchSource&, const TraceParams& lane) { return lane; }

// Static LDS of k_intersect_persistent besides its dynamic stacks: the copy of the 8 SubDAGs.
constexpr size_t kSubdagCopyBytes = 8 * sizeof(cqx_subdag);

template <typename Source, bool kCache, bool kTop = false, bool kShort = false>
__global__ void __launch_bounds__(256) k_intersect_persistent(Source src, uint64_t n, const uint32_t* __restrict__ n_dev,
                                                              unsigned long long* __restrict__ counters,
                                                              uint32_t regroup, uint32_t refill, uint32_t rows,
                                                              uint32_t top_nodes, uint32_t chunk)
{
    // kShort: `rows` (a power of two) is the short stack's depth; rows are reused modulo it
    const uint32_t smask = kShort ? rows - 1u : 0u;
    extern __shared__ uint32_t lds_stack[];
    // kTop: after the stacks, a copy of internal nodes [256, 256 + top_nodes) (the top levels of
    // a breadth-first store, read by most rays) so that those child reads stay out of L2.
    uint32_t* const top = lds_stack + rows * blockDim.x;
    if constexpr (kTop) { // single-volume launches only
        const uint32_t* g = src.P.nodes + kMaterialCount * 8;
        for (uint32_t i = threadIdx.x; i < top_nodes * 8; i += blockDim.x) top[i] = g[i];
        __syncthreads();
    }
    if (n_dev) {
        // ray count produced on the device by an earlier kernel (path-tracer bounce loop)
        const uint64_t nd = *n_dev;
        n = nd < n ? nd : n;
    }
    TraceParams P;
    cqx_hit* dst = nullptr;
    // Each wave owns `rows` x 64 words of the block's stack memory; slot h of a lane is
    // stack[h * 64] (a shift, not a multiply, in the address).
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t* stack = lds_stack + (threadIdx.x >> 6) * rows * 64u + lane;
    constexpr uint32_t stride = 64u;
    const uint64_t per_q = (n + kQueues - 1) / kQueues;
    uint32_t qc = blockIdx.x % kQueues; // current queue
    uint32_t drained = 0;               // queues found empty so far
    // Ray ids [c_next, c_end) claimed by this wave: one atomic claims `chunk` ids, and service
    // passes hand them to idle lanes without further atomics.
    uint64_t c_next = 0, c_end = 0;

    RayState S;
    S.phase = kIdle;
    // single-volume launches pass over the empty octants without reading their SubDAGs, and read
    // the others from an LDS copy
    uint32_t occ = 0xffu;
    __shared__ cqx_subdag lds_sub[8];
    if constexpr (Source::kUniform) {
        occ = occupied_octants(src.P.sub);
        if (threadIdx.x < 8) lds_sub[threadIdx.x] = src.P.sub[threadIdx.x];
        __syncthreads();
    }
    const auto enter = [&](RayState& R) {
        if constexpr (Source::kUniform) enter_octants_from<kCache, kShort>(lds_sub, src.P, R, stack, stride, smask, occ);
        else enter_octants<kCache, kShort>(uniform_params(src, P), R, stack, stride, smask);
    };

    for (;;) {
        // Lanes are busy while they walk a SubDAG. Between SubDAGs, finished or idle, they wait
        // for the next service pass, which runs every `regroup` steps once `refill` lanes wait.
        const uint64_t walking = __ballot(S.phase == kWalk);
        const uint32_t nwait = 64u - uint32_t(__popcll(walking));
        if (nwait >= refill || walking == 0) {
            enter(S);
            if (is_finished(S.phase)) {
                cqx_hit out;
                emit_record(uniform_params(src, P), S, out);
                store_record(dst, out);
                S.phase = kIdle;
            }
            // Claim a new chunk once the current one is used up (at most one claim per queue
            // finds it empty, so this ends).
            while (c_next == c_end && drained < kQueues) {
                const uint64_t q_begin = per_q * qc;
                const uint64_t q_end = q_begin + per_q < n ? q_begin + per_q : n;
                const uint64_t q_len = q_end > q_begin ? q_end - q_begin : 0;
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&counters[qc * kCounterStride], (unsigned long long)chunk);
                base = __shfl(base, 0);
                if (base < q_len) {
                    c_next = q_begin + base;
                    c_end = q_begin + (base + chunk < q_len ? base + chunk : q_len);
                } else {
                    qc = (qc + 1) % kQueues;
                    drained++;
                }
            }
            const uint64_t idle = __ballot(S.phase == kIdle);
            if (S.phase == kIdle) {
                const uint32_t rank = uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                const uint64_t k = c_next + rank;
                if (k < c_end) {
                    const float* r;
                    src.fetch(k, P, r, dst);
                    setup_from(S, r);
                    enter(S);
                }
            }
            const uint64_t nidle = uint64_t(__popcll(idle));
            c_next = c_next + nidle < c_end ? c_next + nidle : c_end;
            if (drained == kQueues && __ballot(S.phase != kIdle) == 0) break;
        }
        const TraceParams& PP = uniform_params(src, P);
        for (uint32_t s = 0; s < regroup; s++)
            if (S.phase == kWalk) walk_step<kCache, kTop, kShort>(PP, S, stack, stride, top, top_nodes, smask);
    }
}

// Variant 5: the persistent kernel with two rays per lane (slots 0 and 1, each with its own
// state and stack column). Each traversal step issues the child reads of both slots before it
// waits for either, so a wave keeps twice as many dependent loads in flight; service passes and
// chunked claims are those of k_intersect_persistent, counted over both slots.
__global__ void __launch_bounds__(256) k_intersect_pair(SingleSource src, uint64_t n, const uint32_t* __restrict__ n_dev,
                                                        unsigned long long* __restrict__ counters, uint32_t regroup,
                                                        uint32_t refill, uint32_t rows, uint32_t chunk)
{
    extern __shared__ uint32_t lds_stack[];
    if (n_dev) {
        const uint64_t nd = *n_dev;
        n = nd < n ? nd : n;
    }
    const TraceParams& P = src.P;
    const uint32_t lane = threadIdx.x & 63u;
    // Each wave owns 2 x rows x 64 words: slot j's column starts at (j * rows) * 64 + lane.
    uint32_t* const wave_stack = lds_stack + (threadIdx.x >> 6) * 2u * rows * 64u + lane;
    uint32_t* const stk[2] = { wave_stack, wave_stack + rows * 64u };
    constexpr uint32_t stride = 64u;
    const uint64_t per_q = (n + kQueues - 1) / kQueues;
    uint32_t qc = blockIdx.x % kQueues;
    uint32_t drained = 0;
    uint64_t c_next = 0, c_end = 0;

    RayState S[2];
    cqx_hit* dst[2] = { nullptr, nullptr };
    S[0].phase = kIdle;
    S[1].phase = kIdle;
    TraceParams unused;

    for (;;) {
        const uint32_t walking = uint32_t(__popcll(__ballot(S[0].phase == kWalk))) +
                                 uint32_t(__popcll(__ballot(S[1].phase == kWalk)));
        if (128u - walking >= refill || walking == 0) {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                enter_octants(P, S[j], stk[j], stride);
                if (is_finished(S[j].phase)) {
                    cqx_hit out;
                    emit_record(P, S[j], out);
                    store_record(dst[j], out);
                    S[j].phase = kIdle;
                }
            }
            while (c_next == c_end && drained < kQueues) {
                const uint64_t q_begin = per_q * qc;
                const uint64_t q_end = q_begin + per_q < n ? q_begin + per_q : n;
                const uint64_t q_len = q_end > q_begin ? q_end - q_begin : 0;
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&counters[qc * kCounterStride], (unsigned long long)chunk);
                base = __shfl(base, 0);
                if (base < q_len) {
                    c_next = q_begin + base;
                    c_end = q_begin + (base + chunk < q_len ? base + chunk : q_len);
                } else {
                    qc = (qc + 1) % kQueues;
                    drained++;
                }
            }
            const uint64_t idle0 = __ballot(S[0].phase == kIdle);
            const uint64_t idle1 = __ballot(S[1].phase == kIdle);
            const uint64_t below = (1ull << lane) - 1ull;
            const uint32_t n0 = uint32_t(__popcll(idle0));
#pragma unroll
            for (int j = 0; j < 2; j++) {
                if (S[j].phase == kIdle) {
                    const uint32_t rank = j == 0 ? uint32_t(__popcll(idle0 & below))
                                                 : n0 + uint32_t(__popcll(idle1 & below));
                    const uint64_t k = c_next + rank;
                    if (k < c_end) {
                        const float* r;
                        src.fetch(k, unused, r, dst[j]);
                        setup_from(S[j], r);
                        enter_octants(P, S[j], stk[j], stride);
                    }
                }
            }
            const uint64_t nidle = uint64_t(n0) + uint64_t(__popcll(idle1));
            c_next = c_next + nidle < c_end ? c_next + nidle : c_end;
            if (drained == kQueues && __ballot(S[0].phase != kIdle || S[1].phase != kIdle) == 0) break;
        }
        for (uint32_t s = 0; s < regroup; s++) {
            const bool w0 = S[0].phase == kWalk, w1 = S[1].phase == kWalk;
            uint32_t c0 = 0, c1 = 0;
            if (w0) c0 = walk_fetch(P, S[0]);
            if (w1) c1 = walk_fetch(P, S[1]);
            if (w0) walk_step_on(P, S[0], stk[0], stride, c0);
            if (w1) walk_step_on(P, S[1], stk[1], stride, c1);
        }
    }
}

// ---- ray binning -------------------------------------------------------------------------------
// An incoherent batch taken in array order spreads the rays in flight on one XCD over the whole
// volume, so its L2 holds few of the nodes they read. The prepass bins the rays by origin into a
// coarse Morton-ordered grid over the volume's bounds; the persistent waves then take ray ids in
// cell order through `order`, and as each XCD drains its own contiguous part of the id space, its
// rays in flight come from a few neighbouring cells. The order within a cell is whatever the
// atomics give: every ray's record is independent of it.
// A counting sort without global atomics: each of up to kMaxBinBlocks workgroups counts the cells
// of its contiguous range of rays in LDS and writes its histogram; one thread per cell turns the
// histograms into each workgroup's offset within the cell and the cell's total; the totals are
// scanned; each workgroup then places its rays from its own offsets (LDS atomics only).
// The grid covers the rays' origins: the bounds of a sample of them, within the box of the
// volume's SubDAGs. Edits past a volume's bounds grow that box (new SubDAGs in other octants, a
// taller one in the first), and a grid over the whole box would put rays that start inside the old
// bounds into a few of its cells.
constexpr uint32_t kMaxBinBits = 4;
constexpr uint32_t kMaxCells = 1u << (3 * kMaxBinBits); // 4096
constexpr uint32_t kMaxBinBlocks = 256;
constexpr uint32_t kBinBlock = 1024;   // threads per workgroup of the count and scatter kernels
constexpr uint32_t kBinMinRange = 16384; // rays per workgroup at least

constexpr uint32_t kBoundsBlocks = 256;      // workgroups of the origin-bounds pass, at most
constexpr uint32_t kBoundsSample = 1u << 20; // about this many origins are read for the bounds

struct BinGrid {
    float lo[3];
    float scale[3]; // cells per unit length along each axis
    uint32_t bits;  // 2^bits cells per axis
};

// The union of the boxes of the non-empty SubDAGs.
struct BinBox {
    float lo[3], hi[3];
    uint32_t bits;
};

__device__ __forceinline__ uint32_t bin_cell(const BinGrid& g, const float* r)
{
    const uint32_t side = 1u << g.bits;
    uint32_t c = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float f = (r[a] - g.lo[a]) * g.scale[a];
        // clamped to [0, side): origins outside the bounds go to the border cells, NaN to cell 0
        const uint32_t k = f >= 0.0f ? (f < float(side) ? uint32_t(f) : side - 1u) : 0u;
#pragma unroll
        for (uint32_t b = 0; b < kMaxBinBits; b++) c |= ((k >> b) & 1u) << (3 * b + a);
    }
    return c; // < 2^(3 bits)
}

// Minimum and maximum of x, y, z over the workgroup's share of the origins of rays 0, stride,
// 2 * stride, ...: partial[6 * blockIdx.x + (0..2 minima, 3..5 maxima)]. NaN coordinates are
// skipped (fminf and fmaxf return the other operand); a share with none leaves (+inf, -inf).
__global__ void __launch_bounds__(256) k_ray_bounds(const cqx_ray* __restrict__ rays, uint64_t n, uint64_t stride,
                                                    float* __restrict__ partial)
{
    __shared__ float part[4][6];
    float v[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
    const uint64_t step = uint64_t(gridDim.x) * blockDim.x * stride;
    for (uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) * stride; i < n; i += step) {
        const float* r = reinterpret_cast<const float*>(rays + i);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = fminf(v[a], r[a]);
            v[3 + a] = fmaxf(v[3 + a], r[a]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = fminf(v[a], __shfl_xor(v[a], off));
            v[3 + a] = fmaxf(v[3 + a], __shfl_xor(v[3 + a], off));
        }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0)
        for (int k = 0; k < 6; k++) part[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t k = threadIdx.x;
        float x = part[0][k];
        for (uint32_t j = 1; j < blockDim.x / 64; j++) x = k < 3 ? fminf(x, part[j][k]) : fmaxf(x, part[j][k]);
        partial[6 * blockIdx.x + k] = x;
    }
}

// The grid (one wave): per axis, the origins' bounds within the SubDAG box, or the whole box along
// an axis where that is empty or a single point (no sampled origin in the box, or all at one
// coordinate).
__global__ void __launch_bounds__(64) k_bin_grid(const float* __restrict__ partial, uint32_t blocks, BinBox box,
                                                 BinGrid* __restrict__ grid)
{
    float v[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
    for (uint32_t b = threadIdx.x; b < blocks; b += 64)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = fminf(v[a], partial[6 * b + a]);
            v[3 + a] = fmaxf(v[3 + a], partial[6 * b + 3 + a]);
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            v[a] = fminf(v[a], __shfl_xor(v[a], off));
            v[3 + a] = fmaxf(v[3 + a], __shfl_xor(v[3 + a], off));
        }
    if (threadIdx.x != 0) return;
    BinGrid g;
    g.bits = box.bits;
    for (int a = 0; a < 3; a++) {
        float l = fmaxf(v[a], box.lo[a]), h = fminf(v[3 + a], box.hi[a]);
        if (!(l < h)) {
            l = box.lo[a];
            h = box.hi[a];
        }
        g.lo[a] = l;
        g.scale[a] = float(double(1u << box.bits) / (double(h) - double(l)));
    }
    *grid = g;
}

// Cell of every ray of the workgroup's range [blockIdx.x * range, ...), and the range's cell
// histogram, hist[blockIdx.x * ncells + cell].
__global__ void __launch_bounds__(kBinBlock) k_bin_count(const cqx_ray* __restrict__ rays, uint64_t n, uint64_t range,
                                                         const BinGrid* __restrict__ grid, uint32_t ncells,
                                                         uint16_t* __restrict__ cells, uint32_t* __restrict__ hist)
{
    const BinGrid g = *grid;
    __shared__ uint32_t h[kMaxCells];
    for (uint32_t c = threadIdx.x; c < ncells; c += blockDim.x) h[c] = 0u;
    __syncthreads();
    const uint64_t first = uint64_t(blockIdx.x) * range, last = first + range < n ? first + range : n;
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) {
        const uint32_t c = bin_cell(g, reinterpret_cast<const float*>(rays + i));
        cells[i] = uint16_t(c);
        atomicAdd(&h[c], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < ncells; c += blockDim.x) hist[uint64_t(blockIdx.x) * ncells + c] = h[c];
}

// Per cell (one thread each): the workgroups' counts become their offsets within the cell, in
// workgroup order, and totals[cell] the cell's count.
__global__ void __launch_bounds__(256) k_bin_columns(uint32_t* __restrict__ hist, uint32_t groups, uint32_t ncells,
                                                     uint32_t* __restrict__ totals)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    uint32_t run = 0;
    for (uint32_t b = 0; b < groups; b++) {
        const uint32_t v = hist[uint64_t(b) * ncells + c];
        hist[uint64_t(b) * ncells + c] = run;
        run += v;
    }
    totals[c] = run;
}

// Exclusive prefix sum of counts[0, ncells) in place (ncells <= 4096: four per thread).
__global__ void __launch_bounds__(1024) k_bin_scan(uint32_t* __restrict__ counts, uint32_t ncells)
{
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        const uint32_t i = 4 * t + k;
        v[k] = i < ncells ? counts[i] : 0u;
        sum += v[k];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t x = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        const uint32_t i = 4 * t + k;
        if (i < ncells) counts[i] = run;
        run += v[k];
    }
}

// order[...] = the ids of the rays of cell c, from cell c's first slot (base[c]) on, each
// workgroup's rays of the cell from its offset (hist) within the cell on. The offsets of the
// workgroups of cell c add up to its count, so every slot of [0, n) is written once.
__global__ void __launch_bounds__(kBinBlock) k_bin_scatter(const uint16_t* __restrict__ cells, uint64_t n,
                                                           uint64_t range, uint32_t ncells,
                                                           const uint32_t* __restrict__ base,
                                                           const uint32_t* __restrict__ hist,
                                                           uint32_t* __restrict__ order)
{
    __shared__ uint32_t cur[kMaxCells];
    for (uint32_t c = threadIdx.x; c < ncells; c += blockDim.x) cur[c] = base[c] + hist[uint64_t(blockIdx.x) * ncells + c];
    __syncthreads();
    const uint64_t first = uint64_t(blockIdx.x) * range, last = first + range < n ? first + range : n;
    for (uint64_t i = first + threadIdx.x; i < last; i += blockDim.x) order[atomicAdd(&cur[cells[i]], 1u)] = uint32_t(i);
}

// The box the binning grid lies in: the union of the boxes of the non-empty SubDAGs. False when
// every SubDAG is empty (nothing to bin for).
bool bin_box(const cqx_dvolume* vol, uint32_t bits, BinBox& b)
{
    double lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
    bool any = false;
    for (const cqx_subdag& s : vol->hsub) {
        if (s.node == 0) continue;
        const double side = std::ldexp(1.0, s.height);
        for (int a = 0; a < 3; a++) {
            const double l = s.lower[a], h = double(s.lower[a]) + side;
            lo[a] = any ? std::min(lo[a], l) : l;
            hi[a] = any ? std::max(hi[a], h) : h;
        }
        any = true;
    }
    if (!any) return false;
    b.bits = bits;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = float(lo[a]);
        b.hi[a] = float(hi[a]);
    }
    return true;
}

// Binning buffers: kMaxCells counts, kMaxBinBlocks x kMaxCells histogram entries, then n order
// entries (u32) and n cells (u16).
static_assert(sizeof(BinGrid) <= kMaxCells * 4, "the grid fits the counts area");
static_assert(kBoundsBlocks * 6 <= kMaxBinBlocks * kMaxCells, "the bounds partials fit the histogram area");
size_t bin_bytes(uint64_t n);

cqx_status ensure_bins(cqx_ctx* ctx, uint64_t n)
{
    const size_t bytes = bin_bytes(n);
    if (bytes <= ctx->bins_bytes) return CQX_OK;
    HIP_TRY(hipDeviceSynchronize()); // launches in flight may use the old buffer
    release(ctx->bins);
    ctx->bins = nullptr;
    ctx->bins_bytes = 0;
    if (hipMalloc(&ctx->bins, bytes) != hipSuccess) return fail(CQX_ERR_OOM, "ray binning buffer of %zu bytes failed", bytes);
    ctx->bins_bytes = bytes;
    return CQX_OK;
}

// Runs the binning prepass of n rays on `stream`; `order` receives the device pointer of the
// resulting id -> ray map. The buffers are the context's (grown as needed) or, with `sv`, the
// caller-provided scratch, which must hold bin_bytes(n).
size_t bin_bytes(uint64_t n)
{
    return size_t(kMaxCells) * 4 + size_t(kMaxBinBlocks) * kMaxCells * 4 + size_t(n) * 4 + size_t(n) * 2;
}

cqx_status bin_rays(cqx_ctx* ctx, const cqx_ray* rays, uint64_t n, const BinBox& box, hipStream_t stream,
                    const cqx::ScratchView* sv, const uint32_t*& order)
{
    void* bins = nullptr;
    if (sv) {
        if (bin_bytes(n) > sv->bins_bytes) return fail(CQX_ERR_LIMIT, "binning %llu rays needs %zu B of scratch, %zu B held",
                                                      (unsigned long long)n, bin_bytes(n), sv->bins_bytes);
        bins = sv->bins;
    } else {
        cqx_status st = ensure_bins(ctx, n);
        if (st != CQX_OK) return st;
        bins = ctx->bins;
    }
    uint32_t* counts = static_cast<uint32_t*>(bins);
    uint32_t* hist = counts + kMaxCells;
    uint32_t* ord = hist + size_t(kMaxBinBlocks) * kMaxCells;
    uint16_t* cells = reinterpret_cast<uint16_t*>(ord + n);
    const uint32_t ncells = 1u << (3 * box.bits);
    // the bounds pass's partials use the histogram area and the grid the counts area: both are
    // read before the kernels that fill those areas run
    float* partial = reinterpret_cast<float*>(hist);
    BinGrid* grid = reinterpret_cast<BinGrid*>(counts);
    const uint64_t stride = std::max<uint64_t>(1, n / kBoundsSample);
    const uint32_t bblocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(kBoundsBlocks, (n / stride + 255) / 256)));
    hipLaunchKernelGGL(k_ray_bounds, dim3(bblocks), dim3(256), 0, stream, rays, n, stride, partial);
    hipLaunchKernelGGL(k_bin_grid, dim3(1), dim3(64), 0, stream, partial, bblocks, box, grid);
    const uint64_t groups = std::max<uint64_t>(1, std::min<uint64_t>(kMaxBinBlocks, n / kBinMinRange));
    const uint64_t range = (n + groups - 1) / groups;
    hipLaunchKernelGGL(k_bin_count, dim3(uint32_t(groups)), dim3(kBinBlock), 0, stream, rays, n, range, grid, ncells,
                       cells, hist);
    hipLaunchKernelGGL(k_bin_columns, dim3((ncells + 255) / 256), dim3(256), 0, stream, hist, uint32_t(groups), ncells,
                       counts);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, stream, counts, ncells);
    hipLaunchKernelGGL(k_bin_scatter, dim3(uint32_t(groups)), dim3(kBinBlock), 0, stream, cells, n, range, ncells, counts,
                       hist, ord);
    HIP_TRY(hipGetLastError());
    order = ord;
    return CQX_OK;
}

// Launch shape of the persistent kernel: CUs x waves_per_cu waves, at most one lane per ray.
// `chunk`: ray ids a wave claims per atomic, about a 32nd of a wave's share (64 to 4096). Smaller
// claims end the launch with less left over in single waves, and with the 8 per-XCD counters on
// separate cache lines their atomics are cheap: a quarter of the share was 4-9 % slower.
cqx_status persistent_shape(const cqx_ctx* ctx, const cqx_launch_cfg* cfg, uint32_t block, uint64_t n,
                            uint32_t& blocks, uint32_t& regroup, uint32_t& refill, uint32_t& chunk)
{
    // built-in defaults: the best launch of a sweep on the 1024^3 incoherent workload
    const uint32_t waves_per_cu = (cfg && cfg->waves_per_cu) ? cfg->waves_per_cu : 32u;
    regroup = (cfg && cfg->regroup_interval) ? cfg->regroup_interval : 8u;
    refill = (cfg && cfg->refill_lanes) ? cfg->refill_lanes : 16u;
    if (waves_per_cu > 64 || regroup > 4096 || refill > 64)
        return fail(CQX_ERR_INVALID_ARG, "persistent launch parameters out of range");
    const uint64_t want_blocks = (uint64_t(ctx->cu_count) * waves_per_cu * 64 + block - 1) / block;
    const uint64_t need_blocks = (n + block - 1) / block;
    blocks = uint32_t(std::max<uint64_t>(1, std::min(want_blocks, need_blocks)));
    const uint64_t waves = uint64_t(blocks) * (block / 64);
    const uint64_t share = (n / (waves * 32) + 63) / 64 * 64;
    chunk = uint32_t(std::min<uint64_t>(4096, std::max<uint64_t>(64, share)));
    return CQX_OK;
}

cqx_status check_ctx(cqx_ctx* ctx)
{
    if (!ctx) return fail(CQX_ERR_INVALID_ARG, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    return CQX_OK;
}

} // namespace

// Scratch of cqx_intersect_noalloc: the persistent kernels' work counters (8 queues, 128 B apart)
// and the ray-binning buffers (cell counts, then the order and cells of at most
// CQX_NOALLOC_MAX_BINNED_RAYS rays), in one file-scope device array. Every device the library
// runs on holds its own copy until the process ends. Host code uses it only through the address
// hipGetSymbolAddress returns with that device current (kept per device in noalloc_scratch_ptr),
// never through the array's name, which on the host is a stand-in and not a device address.
constexpr size_t kNoallocCounterBytes = 1024;
constexpr size_t kNoallocBinBytes =
    size_t(kMaxCells) * 4 + size_t(kMaxBinBlocks) * kMaxCells * 4 + size_t(CQX_NOALLOC_MAX_BINNED_RAYS) * 6;
static_assert(kNoallocCounterBytes + kNoallocBinBytes == CQX_NOALLOC_SCRATCH_BYTES, "cqx.h states the scratch size");
static_assert(sizeof(unsigned long long) * kQueues * kCounterStride <= kNoallocCounterBytes, "counters fit");
static __device__ __align__(256) unsigned char scratch[kNoallocCounterBytes + kNoallocBinBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex noalloc_mutex;
void* noalloc_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current), looked up once per device.
cqx_status noalloc_scratch_address(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(CQX_ERR_LIMIT, "device %d beyond the %d supported", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CQX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
        noalloc_scratch_ptr[device] = p;
    }
    *out = noalloc_scratch_ptr[device];
    return CQX_OK;
}

} // namespace

namespace cqx {

cqx_status ensure_scratch(cqx_ctx* ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return CQX_OK;
    HIP_TRY(hipDeviceSynchronize()); // launches in flight may use the old buffer
    release(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    if (hipMalloc(&ctx->scratch, bytes) != hipSuccess) return fail(CQX_ERR_OOM, "device staging of %zu bytes failed", bytes);
    ctx->scratch_bytes = bytes;
    return CQX_OK;
}

cqx_status launch_intersect(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                            const uint32_t* n_dev, uint32_t flags, float max_footprint, cqx_hit* hits,
                            const cqx_launch_cfg* cfg, hipStream_t stream, const ScratchView* sv)
{
    unsigned long long* const counters =
        reinterpret_cast<unsigned long long*>(sv ? sv->counters : static_cast<void*>(ctx->d_counter));
    TraceParams P;
    P.nodes = vol->nodes;
    P.sub = vol->sub;
    P.surf = (flags & CQX_SURFACE_PROPERTIES) ? 1u : 0u;
    P.max_footprint = max_footprint;
    P.max_iters = (cfg && cfg->max_iters) ? cfg->max_iters : kDefaultMaxIters;
    P.fp_disabled = footprint_is_disabled(max_footprint) ? 1u : 0u;
    P.offset32 = vol->node_count <= (uint64_t(1) << 27) ? 1u : 0u;

    const uint32_t block = (cfg && cfg->block) ? cfg->block : 256u;
    if (block % 64 != 0 || block == 0 || block > 256)
        return fail(CQX_ERR_INVALID_ARG, "block %u must be a multiple of 64, <= 256", block);
    const uint32_t variant = (cfg && cfg->variant) ? cfg->variant : 2u;
    uint32_t rows = uint32_t(vol->max_height) + 1;
    bool short_stack = false;
    if (cfg && cfg->stack_rows) {
        const uint32_t k = cfg->stack_rows;
        if (k < 2 || (k & (k - 1)) != 0) return fail(CQX_ERR_INVALID_ARG, "stack_rows %u must be a power of two >= 2", k);
        if (variant != 2 && variant != 3)
            return fail(CQX_ERR_INVALID_ARG, "short stacks (stack_rows) need variant 2 or 3");
        if (k < rows) {
            rows = k;
            short_stack = true;
        }
    }
    const size_t lds = size_t(rows) * block * sizeof(uint32_t);
    if (lds > size_t(ctx->lds_per_block)) return fail(CQX_ERR_LIMIT, "stack needs %zu B of LDS per block", lds);

    const uint32_t bin_bits = cfg ? cfg->bin_bits : 0u;
    const uint32_t* order = nullptr;
    if (bin_bits) {
        if (bin_bits > kMaxBinBits) return fail(CQX_ERR_INVALID_ARG, "bin_bits %u outside 0..%u", bin_bits, kMaxBinBits);
        if (variant < 2 || variant > 5) return fail(CQX_ERR_INVALID_ARG, "ray binning needs a persistent variant (2-5)");
        if (n_dev) return fail(CQX_ERR_INVALID_ARG, "ray binning needs the ray count on the host");
        if (n > 0xffffffffull) return fail(CQX_ERR_LIMIT, "ray binning takes fewer than 2^32 rays");
        BinBox box;
        if (bin_box(vol, bin_bits, box)) {
            const cqx_status bst = bin_rays(ctx, rays, n, box, stream, sv, order);
            if (bst != CQX_OK) return bst;
        }
    }
    if (variant == 1) {
        const uint64_t blocks = (n + block - 1) / block;
        if (blocks > 0x7fffffffull) return fail(CQX_ERR_LIMIT, "too many rays for one launch");
        hipLaunchKernelGGL(k_intersect_lane, dim3(uint32_t(blocks)), dim3(block), lds, stream, P, rays, n, n_dev, hits);
    } else if (variant == 5) {
        uint32_t blocks, regroup, refill, chunk;
        cqx_status st = persistent_shape(ctx, cfg, block, (n + 1) / 2, blocks, regroup, refill, chunk);
        if (st != CQX_OK) return st;
        const size_t lds2 = 2 * lds;
        if (lds2 > size_t(ctx->lds_per_block))
            return fail(CQX_ERR_LIMIT, "two stacks per lane need %zu B of LDS per block", lds2);
        SingleSource src{ P, rays, hits, order };
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, counters);
        hipLaunchKernelGGL(k_intersect_pair, dim3(blocks), dim3(block), lds2, stream, src, n, n_dev, counters, regroup,
                           refill, rows, chunk);
    } else if (variant >= 2 && variant <= 4) {
        uint32_t blocks, regroup, refill, chunk;
        cqx_status st = persistent_shape(ctx, cfg, block, n, blocks, regroup, refill, chunk);
        if (st != CQX_OK) return st;
        uint32_t top_nodes = 0;
        size_t lds_total = lds;
        if (variant == 4) {
            const uint64_t internal = vol->node_count - CQX_MATERIAL_COUNT;
            top_nodes = uint32_t(std::min<uint64_t>((cfg && cfg->lds_nodes) ? cfg->lds_nodes : 512u, internal));
            lds_total += size_t(top_nodes) * 32;
            if (lds_total > size_t(ctx->lds_per_block))
                return fail(CQX_ERR_LIMIT, "stack and %u LDS nodes need %zu B of LDS per block", top_nodes, lds_total);
        }
        if (lds_total + kSubdagCopyBytes > size_t(ctx->lds_per_block))
            return fail(CQX_ERR_LIMIT, "stack and SubDAG copy need %zu B of LDS per block", lds_total + kSubdagCopyBytes);
        SingleSource src{ P, rays, hits, order };
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, counters);
        if (variant == 2 && short_stack)
            hipLaunchKernelGGL((k_intersect_persistent<SingleSource, false, false, true>), dim3(blocks), dim3(block), lds,
                               stream, src, n, n_dev, counters, regroup, refill, rows, 0u, chunk);
        else if (variant == 3 && short_stack)
            hipLaunchKernelGGL((k_intersect_persistent<SingleSource, true, false, true>), dim3(blocks), dim3(block), lds,
                               stream, src, n, n_dev, counters, regroup, refill, rows, 0u, chunk);
        else if (variant == 2)
            hipLaunchKernelGGL((k_intersect_persistent<SingleSource, false>), dim3(blocks), dim3(block), lds, stream, src,
                               n, n_dev, counters, regroup, refill, rows, 0u, chunk);
        else if (variant == 3)
            hipLaunchKernelGGL((k_intersect_persistent<SingleSource, true>), dim3(blocks), dim3(block), lds, stream, src,
                               n, n_dev, counters, regroup, refill, rows, 0u, chunk);
        else
            hipLaunchKernelGGL((k_intersect_persistent<SingleSource, false, true>), dim3(blocks), dim3(block), lds_total,
                               stream, src, n, n_dev, counters, regroup, refill, rows, top_nodes, chunk);
    } else {
        return fail(CQX_ERR_INVALID_ARG, "unknown kernel variant %u", variant);
    }
    HIP_TRY(hipGetLastError());
    return CQX_OK;
}

} // namespace cqx

extern "C" {

cqx_status cqx_create(int device, cqx_ctx** out)
{
    if (!out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CQX_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(CQX_ERR_NO_DEVICE, "device %d not present (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    cqx_ctx* c = new (std::nothrow) cqx_ctx;
    if (!c) return fail(CQX_ERR_OOM, "out of host memory");
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    c->lds_per_block = int(prop.sharedMemPerBlock);
    if (hipMalloc(&c->d_counter, sizeof(uint64_t) * kQueues * kCounterStride) != hipSuccess) {
        delete c;
        return fail(CQX_ERR_OOM, "device allocation of the work counters failed");
    }
    if (hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking) != hipSuccess) {
        release(c->d_counter);
        delete c;
        return fail(CQX_ERR_HIP, "stream creation failed");
    }
    const cqx_status st = noalloc_scratch_address(device, &c->noalloc_scratch);
    if (st != CQX_OK) {
        (void)hipStreamDestroy(c->capture_stream);
        release(c->d_counter);
        delete c;
        return st;
    }
    *out = c;
    return CQX_OK;
}

cqx_status cqx_destroy(cqx_ctx* ctx)
{
    if (!ctx) return CQX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    cqx::destroy_pt_plan(ctx->pt_plan);
    cqx::destroy_demo_plan(ctx->demo_plan);
    release(ctx->d_counter);
    release(ctx->scratch);
    release(ctx->bins);
    release(ctx->batch_desc);
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    delete ctx;
    return CQX_OK;
}

cqx_status cqx_volume_upload(cqx_ctx* ctx, const uint32_t* nodes, uint64_t node_count, uint32_t root,
                             const cqx_subdag sub[8], cqx_dvolume** out)
{
    if (!out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    st = cqx::validate_dag(nodes, node_count, root);
    if (st != CQX_OK) return st;
    cqx_subdag s[8];
    if (sub) {
        std::memcpy(s, sub, sizeof(s));
        for (int c = 0; c < 8; c++)
            if (s[c].node >= node_count || s[c].height < 0 || s[c].height > 31)
                return fail(CQX_ERR_INVALID_DAG, "SubDAG %d invalid (node %u, height %d)", c, s[c].node, s[c].height);
    } else {
        st = cqx::find_subdags(nodes, node_count, root, s);
        if (st != CQX_OK) return st;
    }
    cqx_dvolume* v = new (std::nothrow) cqx_dvolume;
    if (!v) return fail(CQX_ERR_OOM, "out of host memory");
    v->node_count = node_count;
    v->capacity = node_count;
    v->root = root;
    std::memcpy(v->hsub, s, sizeof(s));
    v->max_height = 0;
    for (int c = 0; c < 8; c++)
        if (s[c].node > 0) v->max_height = std::max(v->max_height, int(s[c].height));
    const size_t bytes = size_t(node_count) * 32;
    if (hipMalloc(&v->nodes, bytes) != hipSuccess || hipMalloc(&v->sub, sizeof(s)) != hipSuccess) {
        release(v->nodes);
        delete v;
        return fail(CQX_ERR_OOM, "device allocation of %zu bytes failed", bytes);
    }
    if (hipMemcpy(v->nodes, nodes, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(v->sub, s, sizeof(s), hipMemcpyHostToDevice) != hipSuccess) {
        release(v->nodes);
        release(v->sub);
        delete v;
        return fail(CQX_ERR_HIP, "upload of the node store failed");
    }
    *out = v;
    return CQX_OK;
}

cqx_status cqx_volume_update(cqx_ctx* ctx, cqx_dvolume* vol, const uint32_t* nodes, uint64_t node_count, uint64_t first,
                             uint32_t root, const cqx_subdag sub[8])
{
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    if (!vol || !nodes) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (first > vol->node_count || first < CQX_MATERIAL_COUNT || node_count < first)
        return fail(CQX_ERR_INVALID_ARG, "delta [%llu, %llu) does not extend the uploaded store of %llu nodes",
                    (unsigned long long)first, (unsigned long long)node_count, (unsigned long long)vol->node_count);
    if (node_count >= 0xffffffffull) return fail(CQX_ERR_LIMIT, "node count does not fit 32-bit indices");
    if (root >= node_count) return fail(CQX_ERR_INVALID_DAG, "root %u out of range", root);
    // Appended nodes may only point at material nodes or at nodes with smaller indices: with the
    // uploaded part already validated, the store stays acyclic.
    for (uint64_t i = first; i < node_count; i++)
        for (int c = 0; c < 8; c++) {
            const uint32_t ch = nodes[i * 8 + c];
            if (ch >= CQX_MATERIAL_COUNT && ch >= i)
                return fail(CQX_ERR_INVALID_DAG, "appended node %llu child %d = %u is not an earlier node",
                            (unsigned long long)i, c, ch);
        }
    cqx_subdag s[8];
    if (sub) {
        std::memcpy(s, sub, sizeof(s));
        for (int c = 0; c < 8; c++)
            if (s[c].node >= node_count || s[c].height < 0 || s[c].height > 31)
                return fail(CQX_ERR_INVALID_DAG, "SubDAG %d invalid", c);
    } else {
        st = cqx::find_subdags(nodes, node_count, root, s);
        if (st != CQX_OK) return st;
    }
    int max_height = 0;
    for (int c = 0; c < 8; c++)
        if (s[c].node > 0) max_height = std::max(max_height, int(s[c].height));
    HIP_TRY(hipDeviceSynchronize()); // launches in flight may read the store
    const uint32_t* old_nodes = vol->nodes;
    if (node_count > vol->capacity) {
        const uint64_t cap = std::max<uint64_t>(node_count, vol->capacity + vol->capacity / 2);
        uint32_t* p = nullptr;
        if (hipMalloc(&p, size_t(cap) * 32) != hipSuccess) return fail(CQX_ERR_OOM, "device store growth failed");
        if (hipMemcpy(p, vol->nodes, size_t(first) * 32, hipMemcpyDeviceToDevice) != hipSuccess) {
            release(p);
            return fail(CQX_ERR_HIP, "device store copy failed");
        }
        release(vol->nodes);
        vol->nodes = p;
        vol->capacity = cap;
    }
    if (node_count > first)
        HIP_TRY(hipMemcpy(vol->nodes + first * 8, nodes + first * 8, size_t(node_count - first) * 32,
                          hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(vol->sub, s, sizeof(s), hipMemcpyHostToDevice));
    const uint64_t kOffset32Nodes = uint64_t(1) << 27; // the child-offset width the traversal uses
    if (ctx->pt_plan && (old_nodes != vol->nodes || max_height != vol->max_height ||
                         (vol->node_count <= kOffset32Nodes) != (node_count <= kOffset32Nodes))) {
        // its captured graph holds the old pointer, stack size and child-offset width
        cqx::destroy_pt_plan(ctx->pt_plan);
        ctx->pt_plan = nullptr;
    }
    vol->node_count = node_count;
    vol->root = root;
    std::memcpy(vol->hsub, s, sizeof(s));
    vol->max_height = max_height;
    return CQX_OK;
}

cqx_status cqx_volume_free(cqx_ctx* ctx, cqx_dvolume* vol)
{
    if (!vol) return CQX_OK;
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    HIP_TRY(hipDeviceSynchronize());
    if (ctx->pt_plan) {
        cqx::destroy_pt_plan(ctx->pt_plan); // a captured graph may reference this volume
        ctx->pt_plan = nullptr;
    }
    release(vol->nodes);
    release(vol->sub);
    cqx::destroy_device_edits(vol->edits);
    delete vol;
    return CQX_OK;
}

cqx_status cqx_volume_info(const cqx_dvolume* vol, const uint32_t** dev_nodes, uint64_t* node_count, uint32_t* root,
                           cqx_subdag sub[8])
{
    if (!vol) return fail(CQX_ERR_INVALID_ARG, "null volume");
    if (dev_nodes) *dev_nodes = vol->nodes;
    if (node_count) *node_count = vol->node_count;
    if (root) *root = vol->root;
    if (sub) std::memcpy(sub, vol->hsub, sizeof(vol->hsub));
    return CQX_OK;
}

cqx_status cqx_intersect(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n, uint32_t flags,
                         float max_footprint, cqx_hit* hits, const cqx_launch_cfg* cfg, void* stream)
{
    const cqx::TraceRange range("cqx_intersect");
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    if (!vol) return fail(CQX_ERR_INVALID_ARG, "null volume");
    if (flags & ~CQX_SURFACE_PROPERTIES) return fail(CQX_ERR_INVALID_ARG, "unknown flags 0x%x", flags);
    if (n == 0) return CQX_OK;
    if (!rays || !hits) return fail(CQX_ERR_INVALID_ARG, "null ray or hit buffer");
    if (n > (uint64_t(1) << 38)) return fail(CQX_ERR_LIMIT, "%llu rays exceed the 2^38 limit", (unsigned long long)n);
    cqx_launch_cfg eff;
    if (!cfg || cfg->variant == 0) {
        // "auto": the autotuner's choice for this volume depth and batch size, if there is one
        if (const cqx_launch_cfg* t = cqx::tuned_cfg(ctx, vol->max_height, n)) {
            eff = *t;
            if (cfg && cfg->max_iters) eff.max_iters = cfg->max_iters;
            cfg = &eff;
        }
    }
    return cqx::launch_intersect(ctx, vol, rays, n, nullptr, flags, max_footprint, hits, cfg, hipStream_t(stream));
}

cqx_status cqx_intersect_noalloc(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n, uint32_t flags,
                                 float max_footprint, cqx_hit* hits, const cqx_launch_cfg* cfg, void* stream)
{
    const cqx::TraceRange range("cqx_intersect_noalloc");
    // Allocates nothing: validation, the tuned-launch lookup and kernel launches only; the
    // counters and binning buffers are the device's static scratch (cleared by the launches).
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    if (!vol) return fail(CQX_ERR_INVALID_ARG, "null volume");
    if (flags & ~CQX_SURFACE_PROPERTIES) return fail(CQX_ERR_INVALID_ARG, "unknown flags 0x%x", flags);
    if (n == 0) return CQX_OK;
    if (!rays || !hits) return fail(CQX_ERR_INVALID_ARG, "null ray or hit buffer");
    if (n > (uint64_t(1) << 38)) return fail(CQX_ERR_LIMIT, "%llu rays exceed the 2^38 limit", (unsigned long long)n);
    if (!ctx->noalloc_scratch) return fail(CQX_ERR_HIP, "static scratch of device %d not set up", ctx->device);
    cqx_launch_cfg eff;
    if (!cfg || cfg->variant == 0) {
        if (const cqx_launch_cfg* t = cqx::tuned_cfg(ctx, vol->max_height, n)) {
            eff = *t;
            if (cfg && cfg->max_iters) eff.max_iters = cfg->max_iters;
            cfg = &eff;
        }
    }
    if (cfg && cfg->bin_bits && n > CQX_NOALLOC_MAX_BINNED_RAYS)
        return fail(CQX_ERR_LIMIT, "cqx_intersect_noalloc bins at most %u rays (%llu given); use cqx_intersect",
                    (unsigned)CQX_NOALLOC_MAX_BINNED_RAYS, (unsigned long long)n);
    cqx::ScratchView sv;
    sv.counters = ctx->noalloc_scratch;
    sv.bins = static_cast<unsigned char*>(ctx->noalloc_scratch) + kNoallocCounterBytes;
    sv.bins_bytes = kNoallocBinBytes;
    return cqx::launch_intersect(ctx, vol, rays, n, nullptr, flags, max_footprint, hits, cfg, hipStream_t(stream), &sv);
}

cqx_status cqx_intersect_batched(cqx_ctx* ctx, const cqx_batch* batches, uint32_t n_batches, const cqx_launch_cfg* cfg,
                                 void* stream_ptr)
{
    const cqx::TraceRange range("cqx_intersect_batched");
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    if (n_batches == 0) return CQX_OK;
    if (!batches) return fail(CQX_ERR_INVALID_ARG, "null batch array");
    if (cfg && cfg->bin_bits) return fail(CQX_ERR_INVALID_ARG, "ray binning applies to cqx_intersect only");
    std::vector<BatchDesc> desc(n_batches);
    uint64_t total = 0;
    int max_height = 0;
    for (uint32_t i = 0; i < n_batches; i++) {
        const cqx_batch& b = batches[i];
        if (!b.volume) return fail(CQX_ERR_INVALID_ARG, "batch %u has no volume", i);
        if (b.flags & ~CQX_SURFACE_PROPERTIES) return fail(CQX_ERR_INVALID_ARG, "batch %u: unknown flags", i);
        if (b.n && (!b.rays || !b.hits)) return fail(CQX_ERR_INVALID_ARG, "batch %u: null ray or hit buffer", i);
        desc[i].nodes = b.volume->nodes;
        desc[i].sub = b.volume->sub;
        desc[i].rays = b.rays;
        desc[i].hits = b.hits;
        desc[i].first = total;
        desc[i].flags = b.flags;
        desc[i].max_footprint = b.max_footprint;
        total += b.n;
        if (b.n) max_height = std::max(max_height, b.volume->max_height);
    }
    if (total == 0) return CQX_OK;
    if (total > (uint64_t(1) << 38)) return fail(CQX_ERR_LIMIT, "more than 2^38 rays in one launch");
    const uint32_t block = (cfg && cfg->block) ? cfg->block : 256u;
    if (block % 64 != 0 || block == 0 || block > 256) return fail(CQX_ERR_INVALID_ARG, "bad block size %u", block);
    uint32_t rows = uint32_t(max_height) + 1;
    bool short_stack = false;
    if (cfg && cfg->stack_rows) {
        const uint32_t k = cfg->stack_rows;
        if (k < 2 || (k & (k - 1)) != 0) return fail(CQX_ERR_INVALID_ARG, "stack_rows %u must be a power of two >= 2", k);
        if (cfg->variant != 0 && cfg->variant != 2 && cfg->variant != 3)
            return fail(CQX_ERR_INVALID_ARG, "short stacks (stack_rows) need variant 2 or 3");
        if (k < rows) {
            rows = k;
            short_stack = true;
        }
    }
    const size_t lds = size_t(rows) * block * sizeof(uint32_t);
    if (lds + kSubdagCopyBytes > size_t(ctx->lds_per_block))
        return fail(CQX_ERR_LIMIT, "stack needs %zu B of LDS per block", lds + kSubdagCopyBytes);
    hipStream_t s = hipStream_t(stream_ptr);
    // The descriptor table is per context: wait for an earlier batched launch on this stream.
    HIP_TRY(hipStreamSynchronize(s));
    const size_t bytes = desc.size() * sizeof(BatchDesc);
    if (bytes > ctx->batch_bytes) {
        release(ctx->batch_desc);
        ctx->batch_desc = nullptr;
        ctx->batch_bytes = 0;
        if (hipMalloc(&ctx->batch_desc, bytes) != hipSuccess) return fail(CQX_ERR_OOM, "descriptor allocation failed");
        ctx->batch_bytes = bytes;
    }
    HIP_TRY(hipMemcpy(ctx->batch_desc, desc.data(), bytes, hipMemcpyHostToDevice));
    uint32_t blocks, regroup, refill, chunk;
    st = persistent_shape(ctx, cfg, block, total, blocks, regroup, refill, chunk);
    if (st != CQX_OK) return st;
    BatchSource src{ static_cast<const BatchDesc*>(ctx->batch_desc), n_batches,
                     (cfg && cfg->max_iters) ? cfg->max_iters : cqx::kDefaultMaxIters };
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned long long*>(ctx->d_counter));
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(ctx->d_counter);
    const bool cache = cfg && cfg->variant == 3;
    if (cache && short_stack)
        hipLaunchKernelGGL((k_intersect_persistent<BatchSource, true, false, true>), dim3(blocks), dim3(block), lds, s,
                           src, total, (const uint32_t*)nullptr, counters, regroup, refill, rows, 0u, chunk);
    else if (short_stack)
        hipLaunchKernelGGL((k_intersect_persistent<BatchSource, false, false, true>), dim3(blocks), dim3(block), lds, s,
                           src, total, (const uint32_t*)nullptr, counters, regroup, refill, rows, 0u, chunk);
    else if (cache)
        hipLaunchKernelGGL((k_intersect_persistent<BatchSource, true>), dim3(blocks), dim3(block), lds, s, src, total,
                           (const uint32_t*)nullptr, counters, regroup, refill, rows, 0u, chunk);
    else
        hipLaunchKernelGGL((k_intersect_persistent<BatchSource, false>), dim3(blocks), dim3(block), lds, s, src, total,
                           (const uint32_t*)nullptr, counters, regroup, refill, rows, 0u, chunk);
    HIP_TRY(hipGetLastError());
    return CQX_OK;
}

cqx_status cqx_intersect_host(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n, uint32_t flags,
                              float max_footprint, cqx_hit* hits)
{
    cqx_status st = check_ctx(ctx);
    if (st != CQX_OK) return st;
    if (n == 0) return CQX_OK;
    if (!rays || !hits) return fail(CQX_ERR_INVALID_ARG, "null ray or hit buffer");
    st = cqx::ensure_scratch(ctx, size_t(n) * (sizeof(cqx_ray) + sizeof(cqx_hit)));
    if (st != CQX_OK) return st;
    cqx_ray* d_rays = static_cast<cqx_ray*>(ctx->scratch);
    cqx_hit* d_hits = reinterpret_cast<cqx_hit*>(static_cast<char*>(ctx->scratch) + size_t(n) * sizeof(cqx_ray));
    HIP_TRY(hipMemcpy(d_rays, rays, size_t(n) * sizeof(cqx_ray), hipMemcpyHostToDevice));
    st = cqx_intersect(ctx, vol, d_rays, n, flags, max_footprint, d_hits, nullptr, nullptr);
    if (st != CQX_OK) return st;
    HIP_TRY(hipMemcpy(hits, d_hits, size_t(n) * sizeof(cqx_hit), hipMemcpyDeviceToHost));
    return CQX_OK;
}

} // extern "C"
