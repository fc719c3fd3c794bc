// Device-side ray vs sparse-voxel-DAG traversal for gfx950: a per-lane restatement of
// Cubiquity's intersectVolume / intersectRayNodeESVO (src/library/raytracing.cpp:134-478) with the
// exact FP32 semantics of the reference's Release build:
//   * IEEE binary32 add/sub/mul and correctly rounded division (built with -ffp-contract=off and
//     -fhip-fp32-correctly-rounded-divide-sqrt, no denormal flushing);
//   * std::min / std::max are compare-selects in the reference's operand order (NaN order kept);
//   * int -> float conversions round to nearest; int32 arithmetic wraps.
//
// The traversal is a resumable per-lane state machine so that a persistent wave can interleave
// traversal steps with refilling finished lanes:
//   ray_setup      intersectVolume's ray reflection and starting octant (raytracing.cpp:415-434)
//   octant_enter   SubDAG of the current octant: slab test and ESVO initialisation (:443-458, :220-250)
//   walk_step      one iteration of the ESVO do/while loop (:253-367)
//   octant_advance next octant in order of axis crossings (:469-475)
//   emit_record    the result record of a finished ray (hit epilogue :302-320, findNearestMaterial)
// A step that ends the walk of a SubDAG only records the outcome (phase kLeave, or kHit / kMiss /
// kAnomaly with the hit child in `start`). Moving on to the next SubDAG (enter_octants) and
// building the record (emit_record, which recomputes every value it needs from the state with the
// walk's own operations) happen later, so that a persistent wave runs that code for many lanes at
// once instead of in almost every traversal step.
// The traversal stack lives in LDS, one column per lane: slot h of a lane is stack[h * stride].
#pragma once

#include "cqx.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cqxd {

constexpr uint32_t kMaterialCount = 256u;
constexpr float kFltMax = 3.402823466e+38f;

__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float min3f(float x, float y, float z) { return smin(smin(x, y), z); }
__device__ __forceinline__ float max3f(float x, float y, float z) { return smax(smax(x, y), z); }

__device__ __forceinline__ int find_msb(uint32_t v) { return v ? 31 - __builtin_clz(v) : -1; }

// Bit k (0 <= k <= 31) of x.
__device__ __forceinline__ uint32_t bit_at(int32_t x, int k) { return (uint32_t(x) >> k) & 1u; }
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return int32_t(uint32_t(a) + uint32_t(b)); }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return int32_t(uint32_t(a) - uint32_t(b)); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return int32_t(uint32_t(a) * uint32_t(b)); }

// findFirstChild, raytracing.cpp:178-196; child id bits x=1, y=2, z=4, also returned as flags.
// The `entry <= 0` term (children behind the origin) is combined as a lane mask, not a branch.
__device__ __forceinline__ uint32_t first_child(float entry, float ox, float oy, float oz, float ix, float iy,
                                                float iz, int32_t cx, int32_t cy, int32_t cz, bool& bx, bool& by,
                                                bool& bz)
{
    const float fx = float(cx), fy = float(cy), fz = float(cz);
    const bool behind = entry <= 0.0f;
    bx = ((fx - ox) * ix < entry) | (behind & (ox >= fx));
    by = ((fy - oy) * iy < entry) | (behind & (oy >= fy));
    bz = ((fz - oz) * iz < entry) | (behind & (oz >= fz));
    return (bx ? 1u : 0u) | (by ? 2u : 0u) | (bz ? 4u : 0u);
}

// First voxel of the hit block entered by the ray (reflected shifted coordinates).
__device__ __forceinline__ int32_t entered_voxel(float o, float d, float t, int32_t lo, int32_t size, bool entry_axis)
{
    if (entry_axis) return lo;
    const float q = o + d * t;
    const float f = floorf(q);
    const int32_t hi = wadd(lo, size - 1);
    if (!(f >= float(lo))) return lo;
    if (f > float(hi)) return hi;
    int32_t v = int32_t(f);
    v = v < lo ? lo : v;
    return v > hi ? hi : v;
}

struct TraceParams {
    const uint32_t* __restrict__ nodes; // node_count x 8
    const cqx_subdag* __restrict__ sub; // 8 SubDAGs
    uint32_t surf;                      // computeSurfaceProperties
    float max_footprint;
    uint32_t max_iters;
    uint32_t fp_disabled; // max_footprint is exactly -1 (MAX_FOOTPRINT_DISABLED)
    // node_count <= 2^27: every child's byte offset fits 32 bits, so child reads use a 32-bit
    // offset from the store's base (fewer address instructions than 64-bit indexing)
    uint32_t offset32 = 0;
};

__host__ __device__ inline bool footprint_is_disabled(float max_footprint) { return max_footprint == -1.0f; }

// hasLargeFootprint = (childNodeSize / tChildExit) > maxFootprint (raytracing.cpp:276), IEEE
// division. With the footprint test disabled (maxFootprint == -1) and a size c = 2^k > 0,
// RN(c / t) > -1 holds exactly when t > 0, t == +0 or t < -c, which needs no division.
__device__ __forceinline__ bool footprint_large(const TraceParams& P, int32_t csize, float tcx)
{
    if (P.fp_disabled && csize > 0) // t > 0 or t == +0: sign bit clear and not NaN
        return __float_as_uint(tcx) <= 0x7f800000u || tcx < -float(csize);
    return (float(csize) / tcx) > P.max_footprint;
}

// kLeave: the walk left the current SubDAG without a hit; the next octant is still to be found.
enum : uint32_t { kIdle = 0, kOctant = 1, kWalk = 2, kLeave = 3, kHit = 4, kMiss = 5, kAnomaly = 6 };

// Phases of a lane whose walk is over and whose record is not yet written.
__device__ __forceinline__ bool is_finished(uint32_t phase) { return phase >= kHit; }

struct RayState {
    float ox, oy, oz;       // reflected, shifted origin
    float dx, dy, dz;       // |direction|
    float ix, iy, iz;       // 1 / |direction| (the same in every ESVO call)
    float distx, disty, distz;
    uint32_t sign_bits;
    uint32_t oct;
    uint32_t node;
    int32_t height, start, csize; // start: the hit child once the phase is kHit
    uint32_t cid;
    int32_t cpx, cpy, cpz;
    float last_exit;
    uint32_t entries, iters;
    uint32_t phase;
    uint32_t nc[8]; // the current node's 8 children (node-cache variant only)
    // Short stack only: the SubDAG's root node, and the highest height whose row is valid
    // (rows of heights height .. stack_top hold the walk's ancestors).
    uint32_t root;
    int32_t stack_top;
};

// Node cache: the whole 32-byte node is loaded when the walk enters it (SubDAG root, descent,
// pop), so that sibling advances read the child from registers instead of memory.
__device__ __forceinline__ void load_node(const TraceParams& P, uint32_t node, uint32_t (&nc)[8])
{
    const uint4* p = reinterpret_cast<const uint4*>(P.nodes) + uint64_t(node) * 2;
    const uint4 a = p[0], b = p[1];
    nc[0] = a.x; nc[1] = a.y; nc[2] = a.z; nc[3] = a.w;
    nc[4] = b.x; nc[5] = b.y; nc[6] = b.z; nc[7] = b.w;
}

// nc[s] for a lane-varying s without dynamic register indexing (a select tree).
__device__ __forceinline__ uint32_t pick8(const uint32_t (&nc)[8], uint32_t s)
{
    const bool b0 = s & 1u, b1 = s & 2u, b2 = s & 4u;
    const uint32_t a0 = b0 ? nc[1] : nc[0], a1 = b0 ? nc[3] : nc[2], a2 = b0 ? nc[5] : nc[4], a3 = b0 ? nc[7] : nc[6];
    const uint32_t c0 = b1 ? a1 : a0, c1 = b1 ? a3 : a2;
    return b2 ? c1 : c0;
}

__device__ __forceinline__ void result_miss(cqx_hit& out)
{
    out.t = 0.0f;
    out.voxel[0] = out.voxel[1] = out.voxel[2] = 0;
    out.node = 0;
    out.info = 0;
}

// Ray setup, raytracing.cpp:415-434.
__device__ __forceinline__ void ray_setup(RayState& S, float rox, float roy, float roz, float rdx, float rdy, float rdz)
{
    const uint32_t sbx = rdx < 0.0f, sby = rdy < 0.0f, sbz = rdz < 0.0f;
    S.sign_bits = sbx | (sby << 1) | (sbz << 2);
    const float sx = float(int32_t(sbx) * -2 + 1), sy = float(int32_t(sby) * -2 + 1), sz = float(int32_t(sbz) * -2 + 1);
    S.ox = (rox + 0.5f) * sx;
    S.oy = (roy + 0.5f) * sy;
    S.oz = (roz + 0.5f) * sz;
    S.dx = fabsf(rdx);
    S.dy = fabsf(rdy);
    S.dz = fabsf(rdz);
    S.ix = 1.0f / S.dx;
    S.iy = 1.0f / S.dy;
    S.iz = 1.0f / S.dz;
    S.distx = (-S.ox) / S.dx;
    S.disty = (-S.oy) / S.dy;
    S.distz = (-S.oz) / S.dz;
    uint32_t oct = 0;
    if (S.distx < 0.0f) { oct += 1; S.distx += kFltMax; }
    if (S.disty < 0.0f) { oct += 2; S.disty += kFltMax; }
    if (S.distz < 0.0f) { oct += 4; S.distz += kFltMax; }
    S.oct = oct;
    S.entries = 0;
    S.iters = 0;
    S.phase = kOctant;
}

// Next octant in order of axis crossings, raytracing.cpp:469-475. The walk is over when all
// octants were visited (a miss) or when no axis steps (the reference would loop forever).
__device__ __forceinline__ void octant_advance(RayState& S)
{
    const float m = min3f(S.distx, S.disty, S.distz);
    uint32_t step = 0;
    if (S.distx <= m) { step += 1; S.distx += kFltMax; }
    if (S.disty <= m) { step += 2; S.disty += kFltMax; }
    if (S.distz <= m) { step += 4; S.distz += kFltMax; }
    S.oct += step;
    S.phase = step == 0 ? kAnomaly : (S.oct > 7 ? kMiss : kOctant);
}

// Short stacks: `smask` + 1 rows per lane (a power of two, fewer than the SubDAG's height + 1), the
// ancestor at height h in row h & smask. A push at height h takes the row of height h + smask + 1;
// a pop to a height whose row was taken descends again from the SubDAG root along the current
// position (the ancestor at a height is the one node on that path), filling the rows of the
// heights it passes. The restart's reads are not node entries or iterations of the walk: the
// records are those of the full stack.
template <bool kShort>
__device__ __forceinline__ uint32_t stack_row(int32_t h, uint32_t smask)
{
    return kShort ? uint32_t(h) & smask : uint32_t(h);
}

// Bit i set when SubDAG i is not empty.
__device__ __forceinline__ uint32_t occupied_octants(const cqx_subdag* sub)
{
    uint32_t occ = 0;
    for (uint32_t i = 0; i < 8; i++) occ |= (sub[i].node > 0 ? 1u : 0u) << i;
    return occ;
}

// Enter the current octant's SubDAG (raytracing.cpp:443-458) and initialise the ESVO walk
// (:220-250), or move on to the next octant. `occ`: bit i clear when SubDAG i is known to be empty
// (its octant is passed over without reading it; all bits set reads every SubDAG, the same result).
// `tab`: the SubDAGs read, P.sub or a copy of it (a kernel's LDS copy).
template <bool kCache = false, bool kShort = false>
__device__ __forceinline__ void octant_enter_from(const cqx_subdag* tab, const TraceParams& P, RayState& S,
                                                  uint32_t* stack, uint32_t stride, uint32_t smask = 0,
                                                  uint32_t occ = 0xffu)
{
    const uint32_t o = S.oct ^ S.sign_bits;
    const cqx_subdag sd = ((occ >> o) & 1u) ? tab[o] : cqx_subdag{};
    if (sd.node > 0) {
        const int32_t sbx = S.sign_bits & 1u, sby = (S.sign_bits >> 1) & 1u, sbz = (S.sign_bits >> 2) & 1u;
        const int32_t size = int32_t(1u << sd.height);
        const int32_t px = wsub(wmul(sd.lower[0], 1 - 2 * sbx), wmul(sbx, size));
        const int32_t py = wsub(wmul(sd.lower[1], 1 - 2 * sby), wmul(sby, size));
        const int32_t pz = wsub(wmul(sd.lower[2], 1 - 2 * sbz), wmul(sbz, size));
        const uint32_t node_size = 1u << sd.height;
        const float nsz = float(node_size);
        const float entry = max3f((float(px) - S.ox) * S.ix, (float(py) - S.oy) * S.iy, (float(pz) - S.oz) * S.iz);
        const float exitt = min3f(((float(px) + nsz) - S.ox) * S.ix, ((float(py) + nsz) - S.oy) * S.iy,
                                  ((float(pz) + nsz) - S.oz) * S.iz);
        if (entry < exitt) {
            S.start = sd.height;
            S.height = sd.height;
            if (kShort) {
                S.root = sd.node;
                S.stack_top = sd.height;
                stack[stack_row<true>(sd.height, smask) * stride] = 0;
            } else {
                for (int h = 0; h <= S.start; h++) stack[h * stride] = 0;
            }
            S.node = sd.node;
            if (kCache) load_node(P, S.node, S.nc);
            S.csize = int32_t(node_size / 2);
            bool bx, by, bz;
            S.cid = first_child(entry, S.ox, S.oy, S.oz, S.ix, S.iy, S.iz, wadd(px, S.csize), wadd(py, S.csize),
                                wadd(pz, S.csize), bx, by, bz);
            S.cpx = wadd(px, bx ? S.csize : 0);
            S.cpy = wadd(py, by ? S.csize : 0);
            S.cpz = wadd(pz, bz ? S.csize : 0);
            S.last_exit = exitt;
            S.entries++;
            S.phase = kWalk;
            return;
        }
    }
    octant_advance(S);
}

template <bool kCache = false, bool kShort = false>
__device__ __forceinline__ void octant_enter(const TraceParams& P, RayState& S, uint32_t* stack, uint32_t stride,
                                             uint32_t smask = 0, uint32_t occ = 0xffu)
{
    octant_enter_from<kCache, kShort>(P.sub, P, S, stack, stride, smask, occ);
}

// The child read of a walk step (raytracing.cpp:265): slot cid ^ signBits of the current node.
// With kTop, the children of internal nodes [256, 256 + top_nodes) are read from `top`, their
// copy in LDS. Split from walk_step_on so that a lane walking two rays can issue both reads
// before it waits for either.
template <bool kCache = false, bool kTop = false>
__device__ __forceinline__ uint32_t walk_fetch(const TraceParams& P, const RayState& S, const uint32_t* top = nullptr,
                                               uint32_t top_nodes = 0)
{
    const uint32_t slot = S.cid ^ S.sign_bits;
    if (kCache) return pick8(S.nc, slot);
    if (kTop && S.node - kMaterialCount < top_nodes) return top[(S.node - kMaterialCount) * 8 + slot];
    if (P.offset32)
        return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(P.nodes) + ((S.node << 5) | (slot << 2)));
    return P.nodes[uint64_t(S.node) * 8 + slot];
}

// One iteration of the ESVO loop, raytracing.cpp:253-367, given the child that walk_fetch read.
template <bool kCache = false, bool kShort = false>
__device__ __forceinline__ void walk_step_on(const TraceParams& P, RayState& S, uint32_t* stack, uint32_t stride,
                                             uint32_t child, uint32_t smask = 0)
{
    if (S.iters >= P.max_iters) {
        S.phase = kAnomaly;
        return;
    }
    S.iters++;
    const int32_t csize = S.csize;
    const float t1x = (float(wadd(S.cpx, csize)) - S.ox) * S.ix;
    const float t1y = (float(wadd(S.cpy, csize)) - S.oy) * S.iy;
    const float t1z = (float(wadd(S.cpz, csize)) - S.oz) * S.iz;
    const float tcx = min3f(t1x, t1y, t1z);
    if (child > 0) {
        const bool internal = child >= kMaterialCount;
        if (internal && footprint_large(P, csize, tcx)) {
            if (S.height <= 1) {
                S.phase = kAnomaly; // deeper than a valid DAG
                return;
            }
            const float t0x = (float(S.cpx) - S.ox) * S.ix, t0y = (float(S.cpy) - S.oy) * S.iy,
                        t0z = (float(S.cpz) - S.oz) * S.iz;
            const float tce = max3f(t0x, t0y, t0z);
            if (tcx < S.last_exit) {
                stack[stack_row<kShort>(S.height, smask) * stride] = S.node;
                if (kShort && S.height + int32_t(smask) < S.stack_top) S.stack_top = S.height + int32_t(smask);
            }
            S.last_exit = tcx;
            S.height--;
            S.node = child;
            if (kCache) load_node(P, S.node, S.nc);
            const int32_t half = int32_t(uint32_t(csize) >> 1); // csize = 2^(height - 1) >= 2 here
            S.csize = half;
            S.entries++;
            bool bx, by, bz;
            S.cid = first_child(tce, S.ox, S.oy, S.oz, S.ix, S.iy, S.iz, wadd(S.cpx, half), wadd(S.cpy, half),
                                wadd(S.cpz, half), bx, by, bz);
            S.cpx = wadd(S.cpx, bx ? half : 0);
            S.cpy = wadd(S.cpy, by ? half : 0);
            S.cpz = wadd(S.cpz, bz ? half : 0);
            return;
        }
        // Hit (raytracing.cpp:302-320): the record is built by emit_record.
        S.start = int32_t(child);
        S.phase = kHit;
        return;
    }
    // Advance to the next sibling, raytracing.cpp:328-332.
    const uint32_t flips = (t1x <= tcx ? 1u : 0u) | (t1y <= tcx ? 2u : 0u) | (t1z <= tcx ? 4u : 0u);
    const uint32_t cid = S.cid ^ flips;
    S.cid = cid;
    const int32_t px = S.cpx, py = S.cpy, pz = S.cpz;
    S.cpx = wadd(px, (flips & 1u) ? csize : 0);
    S.cpy = wadd(py, (flips & 2u) ? csize : 0);
    S.cpz = wadd(pz, (flips & 4u) ? csize : 0);
    if ((cid & flips) != flips) {
        // Pop to the highest boundary crossed, raytracing.cpp:339-364.
        const uint32_t diff = uint32_t(px ^ S.cpx) | uint32_t(py ^ S.cpy) | uint32_t(pz ^ S.cpz);
        const int msb = find_msb(diff);
        const int height = msb + 1;
        S.height = height;
        if (height > S.start) { // left the SubDAG without a hit
            S.phase = kLeave;
            return;
        }
        if (!kShort || height <= S.stack_top) {
            S.node = stack[stack_row<kShort>(height, smask) * stride];
        } else {
            // the row was taken: descend from the SubDAG root to `height` along the current
            // position (bit h - 1 of the position picks the child at height h)
            uint32_t node = S.root;
            for (int32_t h = S.start; h > height; h--) {
                stack[stack_row<true>(h, smask) * stride] = node;
                const uint32_t id = bit_at(S.cpx, h - 1) | (bit_at(S.cpy, h - 1) << 1) | (bit_at(S.cpz, h - 1) << 2);
                node = P.nodes[uint64_t(node) * 8 + (id ^ S.sign_bits)];
            }
            stack[stack_row<true>(height, smask) * stride] = node;
            S.node = node;
            S.stack_top = height + int32_t(smask) < S.start ? height + int32_t(smask) : S.start;
        }
        if (kCache) load_node(P, S.node, S.nc);
        const int32_t ns = int32_t(1u << msb);
        S.csize = ns;
        S.cid = bit_at(S.cpx, msb) | (bit_at(S.cpy, msb) << 1) | (bit_at(S.cpz, msb) << 2);
        // (p >> height << height) + bit msb of p * 2^msb is p with the bits below msb cleared
        // (msb <= 30 here: a crossing at bit 31 is above every SubDAG and left the walk above).
        S.cpx &= -ns;
        S.cpy &= -ns;
        S.cpz &= -ns;
        S.last_exit = 0.0f;
        S.entries++;
    }
}

// One iteration of the ESVO loop (read and step).
template <bool kCache = false, bool kTop = false, bool kShort = false>
__device__ __forceinline__ void walk_step(const TraceParams& P, RayState& S, uint32_t* stack, uint32_t stride,
                                          const uint32_t* top = nullptr, uint32_t top_nodes = 0, uint32_t smask = 0)
{
    walk_step_on<kCache, kShort>(P, S, stack, stride, walk_fetch<kCache, kTop>(P, S, top, top_nodes), smask);
}

// Moves a lane that is between SubDAGs (phase kLeave or kOctant) into the walk of the next SubDAG
// the ray enters, or to the end of its walk. At most 8 octants are tried.
template <bool kCache = false, bool kShort = false>
__device__ __forceinline__ void enter_octants_from(const cqx_subdag* tab, const TraceParams& P, RayState& S,
                                                   uint32_t* stack, uint32_t stride, uint32_t smask = 0,
                                                   uint32_t occ = 0xffu)
{
    if (S.phase == kLeave) octant_advance(S);
    while (S.phase == kOctant) octant_enter_from<kCache, kShort>(tab, P, S, stack, stride, smask, occ);
}

template <bool kCache = false, bool kShort = false>
__device__ __forceinline__ void enter_octants(const TraceParams& P, RayState& S, uint32_t* stack, uint32_t stride,
                                              uint32_t smask = 0, uint32_t occ = 0xffu)
{
    enter_octants_from<kCache, kShort>(P.sub, P, S, stack, stride, smask, occ);
}

// One traversal step of a lane whose walk is in progress.
template <bool kCache = false>
__device__ __forceinline__ void step(const TraceParams& P, RayState& S, uint32_t* stack, uint32_t stride,
                                     uint32_t occ = 0xffu)
{
    if (S.phase == kWalk) walk_step<kCache>(P, S, stack, stride);
    else enter_octants<kCache>(P, S, stack, stride, 0, occ);
}

// Record of a finished ray (phase kHit, kMiss or kAnomaly). For a hit, the child-box entry
// times are recomputed from the state with the walk's own operations, so they are the values
// the hit test saw.
__device__ __forceinline__ void emit_record(const TraceParams& P, RayState& S, cqx_hit& out)
{
    result_miss(out);
    out.entries = S.entries;
    out.iters = S.iters;
    if (S.phase == kAnomaly) {
        out.info = CQX_HIT_ANOMALY;
        return;
    }
    if (S.phase != kHit) return;
    const uint32_t child = uint32_t(S.start);
    const uint32_t slot = S.cid ^ S.sign_bits;
    const float t0x = (float(S.cpx) - S.ox) * S.ix, t0y = (float(S.cpy) - S.oy) * S.iy,
                t0z = (float(S.cpz) - S.oz) * S.iz;
    const float tce = max3f(t0x, t0y, t0z);
    // A SubDAG of height 0 (the chain below the root ends in a single voxel) has child size 0:
    // the hit block is then the SubDAG node itself, of side 1.
    const int32_t bsize = S.csize > 0 ? S.csize : 1;
    uint32_t info = 1u | (slot << 1) | (uint32_t(find_msb(uint32_t(bsize))) << 4);
    if (P.surf) {
        // findNearestMaterial, raytracing.cpp:134-162
        const uint32_t order = (0x76534210u | 0x88888888u) ^ (S.sign_bits * 0x11111111u);
        uint32_t m = child;
        uint32_t entries = S.entries;
        while (m >= kMaterialCount) {
            entries++;
            uint32_t next = 0;
            for (uint32_t ids = order; ids != 0; ids >>= 4) {
                const uint32_t c = P.nodes[uint64_t(m) * 8 + (ids & 7u)];
                if (c > 0) { next = c; break; }
            }
            // An internal node without occupied children (never produced by bake) would make
            // the reference loop forever.
            if (next == 0) {
                out.entries = entries;
                out.info = CQX_HIT_ANOMALY;
                return;
            }
            m = next;
        }
        out.entries = entries;
        // normal = vec3(equal(tEntry, childT0)) * -rayDirSign, raytracing.cpp:317-318
        const float sx = (S.sign_bits & 1u) ? -1.0f : 1.0f, sy = (S.sign_bits & 2u) ? -1.0f : 1.0f,
                    sz = (S.sign_bits & 4u) ? -1.0f : 1.0f;
        const float nx = float(tce == t0x) * -sx, ny = float(tce == t0y) * -sy, nz = float(tce == t0z) * -sz;
        info |= ((nx != 0.0f ? 1u : 0u) | ((__float_as_uint(nx) >> 31) << 1)) << 9;
        info |= ((ny != 0.0f ? 1u : 0u) | ((__float_as_uint(ny) >> 31) << 1)) << 11;
        info |= ((nz != 0.0f ? 1u : 0u) | ((__float_as_uint(nz) >> 31) << 1)) << 13;
        info |= (m & 0xffu) << 16;
    }
    const int32_t vx = entered_voxel(S.ox, S.dx, tce, S.cpx, bsize, t0x == tce);
    const int32_t vy = entered_voxel(S.oy, S.dy, tce, S.cpy, bsize, t0y == tce);
    const int32_t vz = entered_voxel(S.oz, S.dz, tce, S.cpz, bsize, t0z == tce);
    out.voxel[0] = (S.sign_bits & 1u) ? ~vx : vx;
    out.voxel[1] = (S.sign_bits & 2u) ? ~vy : vy;
    out.voxel[2] = (S.sign_bits & 4u) ? ~vz : vz;
    out.t = tce;
    out.node = S.node;
    out.info = info;
}

// Runs one ray to completion (state machine driven in a loop).
__device__ __forceinline__ void trace_ray(const TraceParams& P, float rox, float roy, float roz, float rdx, float rdy,
                                          float rdz, uint32_t* stack, uint32_t stride, cqx_hit& out,
                                          uint32_t occ = 0xffu)
{
    RayState S;
    ray_setup(S, rox, roy, roz, rdx, rdy, rdz);
    while (!is_finished(S.phase)) step(P, S, stack, stride, occ);
    emit_record(P, S, out);
}

} // namespace cqxd
