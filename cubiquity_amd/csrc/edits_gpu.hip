// On-device sphere edits: the region-parallel algorithm of cqx_hv_edit_spheres_mt (edits.cpp) run
// on the device copy of a store, giving the same store byte for byte.
//
//   planning (host, plan_edits): the regions of side 2^hs each sphere's bounds touch, with their
//     spheres in order, and the levels above the regions in Morton order;
//   k_edit_regions: one wave per region descends to the region's node and applies its spheres in
//     order with fillBrush's arithmetic (voxelization.cpp:825-915; EditWalk::fill, here with an
//     explicit stack and each visited node's 8 children tested on the wave's 64 lanes at once),
//     interning changed nodes into the region's own list with its own hash table; a node without listed children is first looked up in the device index over the
//     store. A list that outgrows its slice stops its region, which runs again with a larger one;
//   merge (k_merge_*): the lists laid end to end in region order and merged as merge_regions does:
//     level by level, the lowest list position among equal nodes is kept, the kept ones are
//     appended in list order;
//   k_edit_top: one workgroup rebuilds the levels above the regions bottom-up, each level's changed
//     nodes interned in Morton order (rebuild_levels);
//   k_find_subdags: the new root's SubDAGs (findSubDAG, raytracing.cpp:43-87).
// The index over the store (open addressing on the node hash; among equal nodes the lowest index)
// is built on a volume's first edit and extended with every appended node. Appended nodes only
// point at earlier nodes, and no node that existed before a call changes, so device copies and
// traversal launches of earlier versions stay valid (undo / redo switch roots).
#include "device_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

using cqx::fail;
using cqx::release;

namespace {

constexpr uint32_t kMat = CQX_MATERIAL_COUNT;
constexpr uint32_t kEmptySlot = 0xffffffffu;
constexpr uint32_t kOverflow = 0xfffffffeu; // a region's list outgrew its slice
constexpr int kMaxRegionHeight = 20;        // region_height() returns 5..20
constexpr uint32_t kTopBlock = 1024;

struct DNode {
    uint32_t c[8];
};

__device__ __forceinline__ DNode load_node(const uint32_t* __restrict__ base, uint64_t i)
{
    const uint4* q = reinterpret_cast<const uint4*>(base + i * 8);
    const uint4 a = q[0], b = q[1];
    return DNode{ { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w } };
}

// A node read at device scope (past the CU's L1): inside one kernel, what other waves appended or
// inserted with device-scope stores and atomics is seen.
__device__ __forceinline__ DNode load_node_coherent(const uint32_t* base, uint64_t i)
{
    DNode n;
#pragma unroll
    for (int c = 0; c < 8; c++) n.c[c] = __hip_atomic_load(base + i * 8 + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return n;
}

__device__ __forceinline__ void store_node(uint32_t* __restrict__ base, uint64_t i, const DNode& n)
{
    uint4* q = reinterpret_cast<uint4*>(base + i * 8);
    q[0] = make_uint4(n.c[0], n.c[1], n.c[2], n.c[3]);
    q[1] = make_uint4(n.c[4], n.c[5], n.c[6], n.c[7]);
}

__device__ __forceinline__ bool node_eq(const DNode& a, const DNode& b)
{
    bool e = true;
#pragma unroll
    for (int i = 0; i < 8; i++) e = e && a.c[i] == b.c[i];
    return e;
}

// The host interner's node hash (edits.cpp hash_node); any hash would do, this one is shared.
__device__ __forceinline__ uint64_t hash_node(const DNode& n)
{
    uint64_t h = 0x243f6a8885a308d3ULL;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        h ^= n.c[i];
        h *= 0x9fb21c651e98df25ULL;
        h ^= h >> 29;
    }
    return h;
}

// Index of a store node with this content, or kEmptySlot. kCoherent: for a kernel that inserts
// into the index and appends nodes while it looks them up.
template <bool kCoherent = false>
__device__ uint32_t index_find(const uint32_t* __restrict__ nodes, const uint32_t* __restrict__ slots, uint64_t mask,
                               const DNode& n)
{
    for (uint64_t i = hash_node(n) & mask;; i = (i + 1) & mask) {
        const uint32_t s = kCoherent ? __hip_atomic_load(slots + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : slots[i];
        if (s == kEmptySlot) return kEmptySlot;
        if (node_eq(kCoherent ? load_node_coherent(nodes, s) : load_node(nodes, s), n)) return s;
    }
}

// Adds store node j to the index: an empty slot is claimed with a compare-and-swap; where an equal
// node already sits, the lower index stays.
template <bool kCoherent = false>
__device__ void index_insert(const uint32_t* __restrict__ nodes, uint32_t* __restrict__ slots, uint64_t mask, uint32_t j,
                             const DNode& n)
{
    for (uint64_t i = hash_node(n) & mask;; i = (i + 1) & mask) {
        const uint32_t cur = atomicCAS(&slots[i], kEmptySlot, j);
        if (cur == kEmptySlot) return;
        if (node_eq(kCoherent ? load_node_coherent(nodes, cur) : load_node(nodes, cur), n)) {
            atomicMin(&slots[i], j);
            return;
        }
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint64_t count, uint32_t v)
{
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

__global__ void __launch_bounds__(256) k_index_insert(const uint32_t* __restrict__ nodes, uint64_t from, uint64_t to,
                                                      uint32_t* __restrict__ slots, uint64_t mask)
{
    for (uint64_t j = from + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < to;
         j += uint64_t(gridDim.x) * blockDim.x)
        index_insert(nodes, slots, mask, uint32_t(j), load_node(nodes, j));
}

// ---- the region walk ------------------------------------------------------------------------------

struct DRegionKey {
    uint32_t x, y, z;
};

// SphereBrush (voxelization.h:91-126) in FP32, as SphereBrushF in edits.cpp.
struct Brush {
    float cx, cy, cz, r2;
    float lo[3], hi[3];

    __device__ explicit Brush(const cqx_sphere_edit& e)
    {
        cx = e.centre[0];
        cy = e.centre[1];
        cz = e.centre[2];
        r2 = e.radius * e.radius;
        const float c[3] = { cx, cy, cz };
        for (int a = 0; a < 3; a++) {
            lo[a] = c[a] - e.radius;
            hi[a] = c[a] + e.radius;
        }
    }

    // SphereBrush::contains: FP32 squared distance truncated to int64 (out of range and NaN give
    // INT64_MIN, as the x86-64 conversion does), compared as float with r2, strictly.
    __device__ bool contains(float x, float y, float z) const
    {
        const float dx = x - cx, dy = y - cy, dz = z - cz;
        const float sum = dx * dx + dy * dy + dz * dz;
        int64_t dist;
        if (!(sum >= -9.2233720368547758e18f && sum < 9.2233720368547758e18f)) dist = INT64_MIN;
        else dist = int64_t(sum);
        return float(dist) < r2;
    }

    // SphereBrushF::box_outside: contains() is false everywhere in the box, with margins for its
    // FP32 rounding; FP64 arithmetic, the same operations in the same order.
    __device__ bool box_outside(float x0, float y0, float z0, float x1, float y1, float z1) const
    {
        const double c[3] = { cx, cy, cz }, lo3[3] = { x0, y0, z0 }, hi3[3] = { x1, y1, z1 };
        double dmin = 0.0, dmax = 0.0;
        for (int a = 0; a < 3; a++) {
            const double below = lo3[a] - c[a], above = c[a] - hi3[a];
            const double nearest = below > 0.0 ? below : (above > 0.0 ? above : 0.0);
            const double p = fabs(lo3[a] - c[a]), q = fabs(hi3[a] - c[a]);
            const double farthest = (p < q) ? q : p; // std::max
            dmin += nearest * nearest;
            dmax += farthest * farthest;
        }
        return dmax < 1e18 && dmin > double(r2) * (1.0 + 1e-5) + 16.0;
    }
};

struct RegionArgs {
    const uint32_t* nodes;  // the store (read-only during the walk)
    const uint32_t* index;  // index over the store
    uint64_t index_mask;
    uint32_t root;
    int hs;
    uint32_t base; // store size when the call began: list entries are base + position
    const cqx_sphere_edit* edits;
    const DRegionKey* keys;
    const uint64_t* first;    // keys + 1 offsets into edit_ids
    const uint32_t* edit_ids;
    const uint32_t* run; // the regions this launch runs, heaviest first
    uint32_t n_run;
    uint32_t* next_run;  // work counter: the next entry of `run` to take (zero at launch)
    const uint64_t* pool_off; // each region's slice: pool nodes [pool_off, pool_off + cap)
    const uint32_t* cap;      // power of two
    uint32_t* pool;           // listed nodes, 8 words each
    uint32_t* table;          // each region's hash table: 2 x cap slots from 2 x pool_off
    uint32_t* counts;
    uint32_t* results;
    uint32_t* overflow;
};

struct Region {
    const RegionArgs& A;
    uint64_t poff;
    uint32_t cap;
    uint32_t count = 0;

    __device__ uint32_t child(uint32_t node, uint32_t c) const
    {
        if (node < kMat) return node;
        if (node >= A.base) return A.pool[(poff + (node - A.base)) * 8 + c];
        return A.nodes[uint64_t(node) * 8 + c];
    }

    // RegionStore::intern: 8 equal material children collapse; a node without listed children
    // that is in the store is shared; otherwise the region's list keeps one copy of each node.
    __device__ uint32_t intern(const DNode& n)
    {
        bool uniform = n.c[0] < kMat, has_local = false;
        for (int i = 0; i < 8; i++) {
            uniform = uniform && n.c[i] == n.c[0];
            has_local = has_local || n.c[i] >= A.base;
        }
        if (uniform) return n.c[0];
        if (!has_local) {
            const uint32_t g = index_find(A.nodes, A.index, A.index_mask, n);
            if (g != kEmptySlot) return g;
        }
        const uint64_t mask = uint64_t(cap) * 2 - 1;
        uint32_t* const tab = A.table + poff * 2;
        for (uint64_t i = hash_node(n) & mask;; i = (i + 1) & mask) {
            const uint32_t s = tab[i];
            if (s == kEmptySlot) { // every lane of the wave stores the same values
                if (count >= cap) return kOverflow;
                store_node(A.pool, poff + count, n);
                tab[i] = count;
                return A.base + count++;
            }
            if (node_eq(load_node(A.pool + poff * 8, s), n)) return A.base + s;
        }
    }
};

// One wave walks one region: the DFS of EditWalk::fill (the same children in the same order, the
// same nodes interned in the same order) runs in lockstep on all 64 lanes, and at each node entered
// the lanes evaluate its 8 children at once: lane l takes child l / 8 and corner l % 8 of that
// child's box, so the 64 corner tests of SphereBrush::contains, the 8 overlap tests, the 8
// box_outside tests and the 8 child reads are one step. A frame keeps the node's children and
// the per-child results as 8-bit masks, wave-uniform, in LDS.
struct Frame {
    uint32_t out[8];
    uint32_t node;
    int32_t lo[3];
    uint32_t height;
    uint32_t next;    // next child position to consider
    uint32_t changed;
    uint32_t ok;      // children whose float box overlaps the brush bounds
    uint32_t outside; // children wholly outside the sphere (box_outside)
    uint32_t inside;  // children with all 8 corners inside
};

// Bit c (c = 0..7) set when lane 8c's bit of `ballot` is set.
__device__ __forceinline__ uint32_t group_bits(uint64_t ballot)
{
    uint32_t m = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) m |= uint32_t((ballot >> (8 * c)) & 1u) << c;
    return m;
}

// fillBrush's descent into `node` (height >= 1, lower corner lo); false when no child's float box
// overlaps the brush bounds (geometry.h:346-357) and the node comes back unchanged.
__device__ bool enter(const Region& R, const Brush& b, Frame& f, uint32_t node, uint32_t height, int32_t lx, int32_t ly,
                      int32_t lz, uint32_t lane)
{
    const uint32_t side = 1u << (height - 1);
    const uint32_t c = lane >> 3, k = lane & 7u;
    const uint32_t cx = c & 1u, cy = (c >> 1) & 1u, cz = c >> 2;
    const int32_t x0 = int32_t(uint32_t(lx) + side * cx), y0 = int32_t(uint32_t(ly) + side * cy),
                  z0 = int32_t(uint32_t(lz) + side * cz);
    const float fx0 = float(x0), fy0 = float(y0), fz0 = float(z0);
    const float fx1 = float(int32_t(uint32_t(x0) + (side - 1))), fy1 = float(int32_t(uint32_t(y0) + (side - 1))),
                fz1 = float(int32_t(uint32_t(z0) + (side - 1)));
    const bool ok = !(b.hi[0] < fx0 || b.lo[0] > fx1) && !(b.hi[1] < fy0 || b.lo[1] > fy1) &&
                    !(b.hi[2] < fz0 || b.lo[2] > fz1);
    const uint32_t okm = group_bits(__ballot(ok));
    if (okm == 0) return false;
    // a single voxel's 8 corners are one point
    const bool in = b.contains((k & 1u) ? fx1 : fx0, (k & 2u) ? fy1 : fy0, (k & 4u) ? fz1 : fz0);
    const uint64_t inb = __ballot(in);
    uint32_t allm = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) allm |= (((inb >> (8 * j)) & 0xffull) == 0xffull ? 1u : 0u) << j;
    const bool out = height >= 2 && b.box_outside(fx0, fy0, fz0, fx1, fy1, fz1);
    const uint32_t outm = group_bits(__ballot(out));
    const uint32_t child = R.child(node, c);
    if (k == 0) f.out[c] = child;
    if (lane == 0) {
        f.node = node;
        f.height = height;
        f.lo[0] = lx;
        f.lo[1] = ly;
        f.lo[2] = lz;
        f.next = 0;
        f.changed = 0;
        f.ok = okm;
        f.outside = outm;
        f.inside = allm;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return true;
}

// A value that is the same in every active lane, as a scalar.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// EditWalk::fill (edits.cpp) without recursion, on the whole wave; returns the new node or
// kOverflow (the same value in every lane).
__device__ uint32_t fill(Region& R, const Brush& b, uint32_t mat, uint32_t node0, uint32_t height0, int32_t lx,
                         int32_t ly, int32_t lz, Frame* st, uint32_t lane)
{
    int sp = 0;
    if (!enter(R, b, st[0], node0, height0, lx, ly, lz, lane)) return node0;
    bool returned = false;
    uint32_t ret = 0;
    for (;;) {
        // the frame's fields are the same in every lane: read as scalars (uniform branches)
        Frame& f = st[sp];
        uint32_t next = uni(f.next), changed = uni(f.changed);
        if (returned) { // the child frame of position `next` came back with `ret`
            if (ret == kOverflow) return kOverflow;
            if (ret != uni(f.out[next])) {
                if (lane == 0) f.out[next] = ret;
                changed = 1;
            }
            next++;
            returned = false;
        }
        bool pushed = false;
        const uint32_t height = uni(f.height), okm = uni(f.ok), outm = uni(f.outside), inm = uni(f.inside);
        const uint32_t side = 1u << (height - 1);
        for (; next < 8; next++) {
            if (!((okm >> next) & 1u)) continue;
            const uint32_t child = uni(f.out[next]);
            if (child == mat) continue;
            if ((outm >> next) & 1u) continue; // set only at height >= 2
            const bool all = (inm >> next) & 1u;
            if (height >= 2 && !all) {
                const uint32_t cx = next & 1u, cy = (next >> 1) & 1u, cz = next >> 2;
                const int32_t x0 = int32_t(uni(uint32_t(f.lo[0])) + side * cx),
                              y0 = int32_t(uni(uint32_t(f.lo[1])) + side * cy),
                              z0 = int32_t(uni(uint32_t(f.lo[2])) + side * cz);
                if (enter(R, b, st[sp + 1], child, height - 1, x0, y0, z0, lane)) {
                    if (lane == 0) {
                        f.next = next;
                        f.changed = changed;
                    }
                    sp++;
                    pushed = true;
                    break;
                }
            } else if (all) {
                if (lane == 0) f.out[next] = mat; // child != mat here
                changed = 1;
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (pushed) continue;
        uint32_t res = uni(f.node);
        if (changed) {
            DNode n;
#pragma unroll
            for (int c = 0; c < 8; c++) n.c[c] = uni(f.out[c]);
            res = uni(R.intern(n));
        }
        if (sp == 0 || res == kOverflow) return res;
        sp--;
        ret = res;
        returned = true;
    }
}

// The node at height h that holds key k (granularity 2^h), by descent from the root.
__device__ uint32_t node_at(const uint32_t* __restrict__ nodes, uint32_t root, int h, const DRegionKey& k)
{
    if (h >= 32) return root;
    uint32_t node = root;
    const uint32_t ux = k.x << h, uy = k.y << h, uz = k.z << h;
    for (int level = 32; level > h && node >= kMat; level--) {
        const uint32_t c = ((ux >> (level - 1)) & 1u) | (((uy >> (level - 1)) & 1u) << 1) |
                           (((uz >> (level - 1)) & 1u) << 2);
        node = nodes[uint64_t(node) * 8 + c];
    }
    return node;
}

__global__ void __launch_bounds__(64) k_edit_regions(RegionArgs A)
{
    __shared__ Frame st[kMaxRegionHeight + 1];
    const uint32_t lane = threadIdx.x;
    for (;;) { // one region per wave, taken from the work counter: the heaviest start first
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(A.next_run, 1u);
        t = uint32_t(__shfl(int(t), 0));
        if (t >= A.n_run) break;
        const uint32_t r = A.run[t];
        Region R{ A, A.pool_off[r], A.cap[r] };
        uint32_t* const tab = A.table + R.poff * 2;
        for (uint64_t i = lane; i < uint64_t(R.cap) * 2; i += 64) tab[i] = kEmptySlot;
        __threadfence_block();
        __syncthreads();
        const DRegionKey k = A.keys[r];
        uint32_t node = node_at(A.nodes, A.root, A.hs, k);
        const int32_t lx = int32_t((k.x << A.hs) - 0x80000000u), ly = int32_t((k.y << A.hs) - 0x80000000u),
                      lz = int32_t((k.z << A.hs) - 0x80000000u);
        uint32_t over = 0;
        for (uint64_t j = A.first[r]; j < A.first[r + 1]; j++) {
            const cqx_sphere_edit e = A.edits[A.edit_ids[j]];
            const Brush b(e);
            node = fill(R, b, e.material, node, uint32_t(A.hs), lx, ly, lz, st, lane);
            if (node == kOverflow) {
                over = 1;
                break;
            }
        }
        if (lane == 0) {
            A.results[r] = node;
            A.counts[r] = R.count;
            A.overflow[r] = over;
        }
        __syncthreads();
    }
}

// ---- merge of the regions' lists (merge_regions, edits.cpp) ----------------------------------------

struct MergeArgs {
    uint32_t base;
    uint32_t nr;
    const uint64_t* pool_off;
    const uint32_t* pool;
    const uint32_t* counts;
    const uint64_t* off; // list position of each region's first node (exclusive scan of counts)
    uint64_t total;
    uint32_t* src;      // position -> pool node
    uint32_t* region;   // position -> region
    uint8_t* level;     // position -> level (1 + deepest listed child)
    uint32_t* key;      // position -> 8 words: the node with listed children as base + their kept position
    uint64_t* slot_of;  // position -> its slot in the merge table
    uint32_t* rep;      // position -> kept position of its node
    uint32_t* table;    // merge table
    uint64_t table_mask;
    uint32_t* max_level;
};

__global__ void __launch_bounds__(256) k_merge_flatten(MergeArgs M)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < M.nr; r += gridDim.x * blockDim.x) {
        const uint64_t o = M.off[r], poff = M.pool_off[r];
        uint32_t top = 0;
        for (uint32_t l = 0; l < M.counts[r]; l++) { // children before parents
            const DNode n = load_node(M.pool, poff + l);
            uint32_t lv = 0;
            for (int c = 0; c < 8; c++)
                if (n.c[c] >= M.base) {
                    const uint32_t x = uint32_t(M.level[o + (n.c[c] - M.base)]) + 1u;
                    lv = x > lv ? x : lv;
                }
            M.src[o + l] = uint32_t(poff + l);
            M.region[o + l] = r;
            M.level[o + l] = uint8_t(lv);
            top = lv > top ? lv : top;
        }
        if (M.counts[r]) atomicMax(M.max_level, top);
    }
}

// Level `lv`: the keys (listed children replaced by base + their kept positions, which lower
// levels have settled).
__global__ void __launch_bounds__(256) k_merge_keys(MergeArgs M, uint32_t lv)
{
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < M.total; g += uint64_t(gridDim.x) * blockDim.x) {
        if (M.level[g] != lv) continue;
        DNode k = load_node(M.pool, M.src[g]);
        const uint64_t o = M.off[M.region[g]];
        for (int c = 0; c < 8; c++)
            if (k.c[c] >= M.base) k.c[c] = M.base + M.rep[o + (k.c[c] - M.base)];
        store_node(M.key, g, k);
    }
}

// Level `lv`: the lowest position of each set of equal keys ends up in their slot.
__global__ void __launch_bounds__(256) k_merge_insert(MergeArgs M, uint32_t lv)
{
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < M.total; g += uint64_t(gridDim.x) * blockDim.x) {
        if (M.level[g] != lv) continue;
        const DNode k = load_node(M.key, g);
        for (uint64_t i = hash_node(k) & M.table_mask;; i = (i + 1) & M.table_mask) {
            const uint32_t cur = atomicCAS(&M.table[i], kEmptySlot, uint32_t(g));
            if (cur == kEmptySlot || node_eq(load_node(M.key, cur), k)) {
                if (cur != kEmptySlot) atomicMin(&M.table[i], uint32_t(g));
                M.slot_of[g] = i;
                break;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_merge_rep(MergeArgs M, uint32_t lv)
{
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < M.total; g += uint64_t(gridDim.x) * blockDim.x)
        if (M.level[g] == lv) M.rep[g] = M.table[M.slot_of[g]];
}

// Exclusive prefix sums in one workgroup of 1024 threads (each thread a contiguous run).
__device__ void block_exclusive_scan_u64(uint64_t* __restrict__ v, uint64_t n, uint64_t* __restrict__ total)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x, T = blockDim.x;
    const uint64_t per = (n + T - 1) / T;
    const uint64_t b = uint64_t(t) * per < n ? uint64_t(t) * per : n, e = b + per < n ? b + per : n;
    uint64_t s = 0;
    for (uint64_t i = b; i < e; i++) s += v[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < T; d <<= 1) {
        const uint64_t x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint64_t run = part[t] - s;
    for (uint64_t i = b; i < e; i++) {
        const uint64_t x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == T - 1) *total = part[t];
    __syncthreads();
}

__global__ void __launch_bounds__(1024) k_scan_u64(uint64_t* v, uint64_t n, uint64_t* total)
{
    block_exclusive_scan_u64(v, n, total);
}

// Exclusive prefix sums over many workgroups: tiles of kScanTile values, 4 consecutive values per
// thread. k_scan_tile_sums sums each tile, the tile sums are scanned in one workgroup, and
// k_scan_tiles scans each tile from its tile's offset.
constexpr uint32_t kScanBlock = 1024, kScanTile = 4 * kScanBlock;

__device__ uint64_t block_sum_u64(uint64_t s)
{
    __shared__ uint64_t part[kScanBlock / 64];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    uint64_t t = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) t += part[w];
    return t;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_tile_sums(const uint64_t* __restrict__ v, uint64_t n,
                                                               uint64_t* __restrict__ sums)
{
    const uint64_t first = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * 4;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) s += first + k < n ? v[first + k] : 0u;
    const uint64_t t = block_sum_u64(s);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_tiles(uint64_t* __restrict__ v, uint64_t n,
                                                           const uint64_t* __restrict__ offsets)
{
    __shared__ uint64_t part[kScanBlock];
    const uint64_t first = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * 4;
    uint64_t x[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[k] = first + k < n ? v[first + k] : 0u;
        s += x[k];
    }
    const uint32_t t = threadIdx.x;
    part[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
        const uint64_t y = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    uint64_t run = offsets[blockIdx.x] + part[t] - s;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (first + k < n) {
            v[first + k] = run;
            run += x[k];
        }
}

__global__ void __launch_bounds__(256) k_counts_to_u64(const uint32_t* __restrict__ c, uint64_t* __restrict__ o, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) o[i] = c[i];
}

__global__ void __launch_bounds__(256) k_merge_kept(MergeArgs M, uint64_t* kept)
{
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < M.total; g += uint64_t(gridDim.x) * blockDim.x)
        kept[g] = M.rep[g] == uint32_t(g) ? 1u : 0u;
}

// The kept nodes into the store at base + their rank (fin), listed children mapped to the store
// indices of their kept nodes; the regions' results mapped the same way.
__global__ void __launch_bounds__(256) k_merge_write(MergeArgs M, const uint64_t* __restrict__ fin, uint32_t* __restrict__ nodes,
                                                     uint32_t* __restrict__ results)
{
    for (uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < M.total; g += uint64_t(gridDim.x) * blockDim.x) {
        if (M.rep[g] != uint32_t(g)) continue;
        DNode k = load_node(M.pool, M.src[g]);
        const uint64_t o = M.off[M.region[g]];
        for (int c = 0; c < 8; c++)
            if (k.c[c] >= M.base) k.c[c] = M.base + uint32_t(fin[M.rep[o + (k.c[c] - M.base)]]);
        store_node(nodes, M.base + fin[g], k);
    }
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < M.nr; r += gridDim.x * blockDim.x) {
        const uint32_t x = results[r];
        if (x >= M.base) results[r] = M.base + uint32_t(fin[M.rep[M.off[r] + (x - M.base)]]);
    }
}

// ---- the levels above the regions (rebuild_levels, edits.cpp) ---------------------------------------

struct TopArgs {
    uint32_t* nodes;
    uint32_t* index;
    uint64_t index_mask;
    uint32_t root;
    int hs;
    uint32_t levels;          // 32 - hs
    const uint32_t* lvl_off;  // levels + 2 entry offsets: level l is [lvl_off[l], lvl_off[l + 1])
    const DRegionKey* keys;   // per entry
    const uint32_t* cbeg;     // per entry of level >= 1: its children, entries [cbeg, cend) of level l - 1
    const uint32_t* cend;
    const uint32_t* morton_regions; // level 0 entry -> region
    const uint32_t* results;        // region -> new node
    uint32_t* res;        // per entry: the node after the edits
    uint32_t* cand;       // per entry: 1 when it is a new node to intern
    uint32_t* ckey;       // per entry: 8 words
    uint32_t* crep;       // per entry: the first entry of its level with the same node
    uint64_t* cslot;
    uint32_t* table;      // per level: 2 x pow2 slots from tab_off[l]
    const uint64_t* tab_off;
    uint32_t* out_count;  // in: store size after the merge; out: after the top levels
    uint32_t* out_root;
    uint32_t* old;        // per entry of level >= 1: the node it held before the edits
};

__global__ void __launch_bounds__(kTopBlock) k_edit_top(TopArgs A)
{
    __shared__ uint32_t wave_kept[kTopBlock / 64];
    __shared__ uint32_t count_s;
    const uint32_t t = threadIdx.x, T = blockDim.x;
    if (t == 0) count_s = *A.out_count;
    for (uint32_t i = A.lvl_off[0] + t; i < A.lvl_off[1]; i += T) A.res[i] = A.results[A.morton_regions[i]];
    // The nodes the entries held before the edits, top down: the top entry (height 32, one key)
    // holds the root, and an entry's children hold the children of its node (node_at of every
    // entry, with one read per entry instead of one per level above it).
    if (t == 0) A.old[A.lvl_off[A.levels]] = A.root;
    __threadfence();
    __syncthreads();
    for (uint32_t l = A.levels; l >= 2; l--) {
        for (uint32_t i = A.lvl_off[l] + t; i < A.lvl_off[l + 1]; i += T) {
            const uint32_t p = __hip_atomic_load(A.old + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (uint32_t j = A.cbeg[i]; j < A.cend[i]; j++) {
                const DRegionKey k = A.keys[j];
                const uint32_t id = (k.x & 1u) | (k.y & 1u) << 1 | (k.z & 1u) << 2;
                A.old[j] = p < kMat ? p : A.nodes[uint64_t(p) * 8 + id];
            }
        }
        __threadfence();
        __syncthreads();
    }
    for (uint32_t l = 1; l <= A.levels; l++) {
        const uint32_t b = A.lvl_off[l], e = A.lvl_off[l + 1];
        for (uint32_t i = b + t; i < e; i += T) {
            const uint32_t old = __hip_atomic_load(A.old + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            DNode out;
            for (int c = 0; c < 8; c++) out.c[c] = old < kMat ? old : A.nodes[uint64_t(old) * 8 + c];
            bool changed = false;
            for (uint32_t j = A.cbeg[i]; j < A.cend[i]; j++) {
                const DRegionKey k = A.keys[j];
                const uint32_t id = (k.x & 1u) | (k.y & 1u) << 1 | (k.z & 1u) << 2;
                if (A.res[j] != out.c[id]) {
                    out.c[id] = A.res[j];
                    changed = true;
                }
            }
            uint32_t r = old, cand = 0;
            if (changed) { // Interner::intern: collapse, else the store's node, else new
                bool uniform = out.c[0] < kMat;
                for (int c = 1; c < 8; c++) uniform = uniform && out.c[c] == out.c[0];
                if (uniform) r = out.c[0];
                else {
                    r = index_find<true>(A.nodes, A.index, A.index_mask, out);
                    if (r == kEmptySlot) {
                        cand = 1;
                        store_node(A.ckey, i, out);
                    }
                }
            }
            A.res[i] = r;
            A.cand[i] = cand;
        }
        __threadfence();
        __syncthreads();
        // equal new nodes of this level: the first in Morton order is kept
        const uint64_t tmask = (A.tab_off[l + 1] - A.tab_off[l]) - 1;
        uint32_t* const tab = A.table + A.tab_off[l];
        for (uint32_t i = b + t; i < e; i += T) {
            if (!A.cand[i]) continue;
            const DNode k = load_node_coherent(A.ckey, i);
            for (uint64_t s = hash_node(k) & tmask;; s = (s + 1) & tmask) {
                const uint32_t cur = atomicCAS(&tab[s], kEmptySlot, i);
                if (cur == kEmptySlot || node_eq(load_node_coherent(A.ckey, cur), k)) {
                    if (cur != kEmptySlot) atomicMin(&tab[s], i);
                    A.cslot[i] = s;
                    break;
                }
            }
        }
        __threadfence();
        __syncthreads();
        for (uint32_t i = b + t; i < e; i += T)
            if (A.cand[i]) A.crep[i] = __hip_atomic_load(tab + A.cslot[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        __syncthreads();
        // the kept ones appended in order: a scan over the level in chunks of T entries (each wave
        // ranks its lanes by ballot, then adds the counts of the waves before it)
        uint32_t base = count_s;
        for (uint32_t c0 = b; c0 < e; c0 += T) {
            const uint32_t i = c0 + t;
            const bool keep = i < e && A.cand[i] && A.crep[i] == i;
            const uint64_t bal = __ballot(keep);
            const uint32_t lane = t & 63u, w = t >> 6;
            if (lane == 0) wave_kept[w] = uint32_t(__popcll(bal));
            __syncthreads();
            uint32_t before = 0, all = 0;
            for (uint32_t v = 0; v < T / 64; v++) {
                before += v < w ? wave_kept[v] : 0u;
                all += wave_kept[v];
            }
            if (keep) {
                const uint32_t at = base + before + uint32_t(__popcll(bal & ((1ull << lane) - 1ull)));
                store_node(A.nodes, at, load_node_coherent(A.ckey, i));
                A.res[i] = at; // a kept entry's node; the others copy theirs below
            }
            base += all;
            __syncthreads();
        }
        __threadfence();
        __syncthreads();
        for (uint32_t i = b + t; i < e; i += T)
            if (A.cand[i]) {
                if (A.crep[i] != i) A.res[i] = A.res[A.crep[i]];
                else index_insert<true>(A.nodes, A.index, A.index_mask, A.res[i], load_node_coherent(A.ckey, i));
            }
        if (t == 0) count_s = base;
        __threadfence();
        __syncthreads();
    }
    if (t == 0) {
        *A.out_count = count_s;
        *A.out_root = A.res[A.lvl_off[A.levels]];
    }
}

// findSubDAG (raytracing.cpp:43-87) of every root child, as cqx::find_subdags; one thread.
__global__ void k_find_subdags(const uint32_t* __restrict__ nodes, const uint32_t* __restrict__ root_p,
                               cqx_subdag* __restrict__ out, uint32_t* __restrict__ err)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t root = *root_p;
    *err = 0;
    for (uint32_t c = 0; c < 8; c++) {
        int height = 32;
        uint32_t lower[3] = { 0x80000000u, 0x80000000u, 0x80000000u };
        uint32_t only = c, node = 0, count = 1;
        uint32_t next = root < kMat ? root : nodes[uint64_t(root) * 8 + only];
        while (count == 1) {
            if (height == 0) {
                *err = 1;
                return;
            }
            height--;
            node = next;
            for (int a = 0; a < 3; a++) lower[a] ^= ((only >> a) & 1u) << height;
            count = 0;
            for (uint32_t i = 0; i < 8; i++) {
                const uint32_t ch = node < kMat ? node : nodes[uint64_t(node) * 8 + i];
                if (ch > 0) {
                    next = ch;
                    count++;
                    only = i;
                }
            }
        }
        cqx_subdag s;
        s.lower[0] = int32_t(lower[0]);
        s.lower[1] = int32_t(lower[1]);
        s.lower[2] = int32_t(lower[2]);
        s.height = height;
        s.pad0 = 0;
        s.node = node;
        s.pad1 = 0;
        s.pad2 = 0;
        out[c] = s;
    }
}

uint32_t grid_for(uint64_t n, const cqx_ctx* ctx, uint32_t block)
{
    const uint64_t want = (n + block - 1) / block;
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(want, uint64_t(ctx->cu_count) * 16)));
}

uint64_t pow2_at_least(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

} // namespace

namespace cqx {

// Device buffers of a volume's edits, kept between calls (grown as needed).
struct DeviceEdits {
    uint32_t* index = nullptr; // index over the store
    uint64_t index_cap = 0;
    uint64_t indexed = 0; // store nodes [256, indexed) are in the index
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    Buf bufs[40];
    std::vector<uint32_t> roots{ 0 }; // history (checkpoint / undo / redo), as the host volume's
    size_t current = 0;
};

void destroy_device_edits(DeviceEdits* e)
{
    if (!e) return;
    release(e->index);
    for (auto& b : e->bufs) release(b.p);
    delete e;
}

} // namespace cqx

namespace {

// Buffer slots of DeviceEdits::bufs.
enum : int {
    kBufEdits, kBufKeys, kBufFirst, kBufEditIds, kBufRun, kBufPoolOff, kBufCap, kBufCounts, kBufResults,
    kBufOverflow, kBufPool, kBufTable, kBufOff, kBufSrc, kBufRegion, kBufLevel, kBufKey, kBufSlotOf, kBufRep,
    kBufMergeTable, kBufFin, kBufScalars, kBufLvlOff, kBufTopKeys, kBufCbeg, kBufCend, kBufMorton, kBufRes,
    kBufCand, kBufCkey, kBufTopTable, kBufRoot, kBufCslot, kBufTabOff, kBufSubRoot, kBufNext, kBufScanSums, kBufOld
};

// Buffer `slot` of the volume's edit state with at least `bytes` (contents not kept when grown).
template <typename T>
cqx_status buf(cqx::DeviceEdits* E, int slot, size_t count, T** out)
{
    const size_t bytes = std::max<size_t>(256, count * sizeof(T));
    auto& b = E->bufs[slot];
    if (b.bytes < bytes) {
        release(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (hipMalloc(&b.p, bytes) != hipSuccess) return fail(CQX_ERR_OOM, "edit buffer of %zu bytes failed", bytes);
        b.bytes = bytes;
    }
    *out = static_cast<T*>(b.p);
    return CQX_OK;
}

// Exclusive prefix sums of v[0, n) in place and their total in *total, on the null stream.
cqx_status scan_u64(cqx_ctx* ctx, cqx::DeviceEdits* E, uint64_t* v, uint64_t n, uint64_t* total);

template <typename T>
cqx_status upload(cqx::DeviceEdits* E, int slot, const std::vector<T>& v, T** out)
{
    cqx_status st = buf(E, slot, v.size(), out);
    if (st != CQX_OK) return st;
    if (!v.empty()) HIP_TRY(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return CQX_OK;
}

cqx_status scan_u64(cqx_ctx*, cqx::DeviceEdits* E, uint64_t* v, uint64_t n, uint64_t* total)
{
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles <= 1) {
        hipLaunchKernelGGL(k_scan_u64, dim3(1), dim3(1024), 0, 0, v, n, total);
    } else {
        if (tiles > 0x7fffffffull) return fail(CQX_ERR_LIMIT, "scan of %llu values", (unsigned long long)n);
        uint64_t* sums = nullptr;
        cqx_status st = buf(E, kBufScanSums, tiles, &sums);
        if (st != CQX_OK) return st;
        hipLaunchKernelGGL(k_scan_tile_sums, dim3(uint32_t(tiles)), dim3(kScanBlock), 0, 0, v, n, sums);
        hipLaunchKernelGGL(k_scan_u64, dim3(1), dim3(1024), 0, 0, sums, tiles, total);
        hipLaunchKernelGGL(k_scan_tiles, dim3(uint32_t(tiles)), dim3(kScanBlock), 0, 0, v, n, sums);
    }
    HIP_TRY(hipGetLastError());
    return CQX_OK;
}

// The index covers every store node (grown to keep its load at most a half).
cqx_status ensure_index(cqx_ctx* ctx, cqx_dvolume* vol, uint64_t will_hold)
{
    cqx::DeviceEdits* E = vol->edits;
    if (2 * will_hold > E->index_cap) {
        const uint64_t cap = pow2_at_least(std::max<uint64_t>(1024, 2 * will_hold + will_hold / 2));
        release(E->index);
        E->index = nullptr;
        E->index_cap = 0;
        E->indexed = 0;
        if (hipMalloc(&E->index, cap * 4) != hipSuccess)
            return fail(CQX_ERR_OOM, "edit index of %llu slots failed", (unsigned long long)cap);
        E->index_cap = cap;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(cap, ctx, 256)), dim3(256), 0, 0, E->index, cap, kEmptySlot);
        E->indexed = kMat;
    }
    if (vol->node_count > E->indexed) {
        hipLaunchKernelGGL(k_index_insert, dim3(grid_for(vol->node_count - E->indexed, ctx, 256)), dim3(256), 0, 0,
                           vol->nodes, E->indexed, vol->node_count, E->index, E->index_cap - 1);
        E->indexed = vol->node_count;
    }
    HIP_TRY(hipGetLastError());
    return CQX_OK;
}

// The device store holds at least `count` nodes (grown by copying; pointers change).
cqx_status ensure_store(cqx_ctx* ctx, cqx_dvolume* vol, uint64_t count)
{
    if (count <= vol->capacity) return CQX_OK;
    const uint64_t cap = std::max<uint64_t>(count, vol->capacity + vol->capacity / 2);
    uint32_t* p = nullptr;
    if (hipMalloc(&p, size_t(cap) * 32) != hipSuccess) return fail(CQX_ERR_OOM, "device store growth failed");
    if (hipMemcpy(p, vol->nodes, size_t(vol->node_count) * 32, hipMemcpyDeviceToDevice) != hipSuccess) {
        release(p);
        return fail(CQX_ERR_HIP, "device store copy failed");
    }
    HIP_TRY(hipDeviceSynchronize());
    release(vol->nodes);
    vol->nodes = p;
    vol->capacity = cap;
    if (ctx->pt_plan) { // its captured graph holds the old pointer
        cqx::destroy_pt_plan(ctx->pt_plan);
        ctx->pt_plan = nullptr;
    }
    return CQX_OK;
}

// Sets the volume's root and recomputes its SubDAGs on the device.
cqx_status set_root(cqx_ctx* ctx, cqx_dvolume* vol, uint32_t root)
{
    uint32_t* d = nullptr;
    cqx_status st = buf(vol->edits, kBufSubRoot, 2, &d);
    if (st != CQX_OK) return st;
    HIP_TRY(hipMemcpy(d, &root, 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_find_subdags, dim3(1), dim3(64), 0, 0, vol->nodes, d, vol->sub, d + 1);
    HIP_TRY(hipGetLastError());
    uint32_t err = 0;
    HIP_TRY(hipMemcpy(&err, d + 1, 4, hipMemcpyDeviceToHost));
    if (err) return fail(CQX_ERR_INVALID_DAG, "single-child chain deeper than the tree");
    cqx_subdag s[8];
    HIP_TRY(hipMemcpy(s, vol->sub, sizeof(s), hipMemcpyDeviceToHost));
    int max_height = 0;
    for (int c = 0; c < 8; c++)
        if (s[c].node > 0) max_height = std::max(max_height, int(s[c].height));
    if (ctx->pt_plan && max_height != vol->max_height) { // its graph holds the stack size
        cqx::destroy_pt_plan(ctx->pt_plan);
        ctx->pt_plan = nullptr;
    }
    vol->root = root;
    std::memcpy(vol->hsub, s, sizeof(s));
    vol->max_height = max_height;
    return CQX_OK;
}

cqx_status check_args(cqx_ctx* ctx, cqx_dvolume* vol)
{
    if (!ctx || !vol) return fail(CQX_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!vol->edits) {
        vol->edits = new (std::nothrow) cqx::DeviceEdits;
        if (!vol->edits) return fail(CQX_ERR_OOM, "out of host memory");
        vol->edits->roots[0] = vol->root;
    }
    return CQX_OK;
}

} // namespace

extern "C" {

cqx_status cqx_volume_edit_spheres(cqx_ctx* ctx, cqx_dvolume* vol, const cqx_sphere_edit* edits, uint64_t n,
                                   uint64_t* first_new)
{
    const cqx::TraceRange range("cqx_volume_edit_spheres");
    cqx_status st = check_args(ctx, vol);
    if (st != CQX_OK) return st;
    if (n && !edits) return fail(CQX_ERR_INVALID_ARG, "null edits");
    if (first_new) *first_new = vol->node_count;
    if (n == 0) return CQX_OK;
    for (uint64_t i = 0; i < n; i++) {
        if (edits[i].material >= kMat) return fail(CQX_ERR_INVALID_ARG, "edit %llu: material >= 256", (unsigned long long)i);
        if (!(edits[i].radius >= 0.0f)) return fail(CQX_ERR_INVALID_ARG, "edit %llu: bad radius", (unsigned long long)i);
    }
    const int hs = n >= 16 ? cqx::region_height(edits, n) : -1;
    if (hs < 0)
        return fail(CQX_ERR_LIMIT, "on-device edits take batches of at least 16 spheres, each within 2^23 voxels of the "
                                   "origin and with radius < 2^20 / 2; apply this batch with cqx_hv_edit_spheres");
    using clock = std::chrono::steady_clock;
    const auto ms_since = [](clock::time_point t) {
        return std::chrono::duration<double, std::milli>(clock::now() - t).count();
    };
    clock::time_point t_phase = clock::now();
    double ms[6] = { 0, 0, 0, 0, 0, 0 }; // plan, index, regions, merge, levels above, SubDAGs
    int attempts = 0;
    cqx::EditPlan plan;
    try {
        cqx::plan_edits(edits, n, hs, plan);
    } catch (const std::bad_alloc&) {
        return fail(CQX_ERR_OOM, "out of host memory while planning edits");
    }
    cqx::DeviceEdits* E = vol->edits;
    HIP_TRY(hipDeviceSynchronize()); // launches in flight may read the store or the index
    const uint64_t base = vol->node_count;
    const uint32_t nr = uint32_t(plan.keys.size());
    ms[0] = ms_since(t_phase);
    t_phase = clock::now();
    st = ensure_index(ctx, vol, base);
    if (st != CQX_OK) return st;
    HIP_TRY(hipDeviceSynchronize());
    ms[1] = ms_since(t_phase);
    t_phase = clock::now();

    // ---- regions ----
    std::vector<DRegionKey> keys(nr);
    for (uint32_t r = 0; r < nr; r++) keys[r] = { plan.keys[r].x, plan.keys[r].y, plan.keys[r].z };
    std::vector<uint32_t> cap(nr), run(nr);
    std::vector<uint64_t> pool_off(nr);
    std::vector<double> work(nr);
    uint64_t pool_total = 0;
    // First slice of each region's list: per sphere about twice the voxels of its largest cross
    // section the region can hold (a sphere changes nodes along its surface, a few per voxel
    // column), so that few regions have to run again.
    // CQX_EDIT_FIRST_SLICE=<nodes> fixes the first slices instead (the tests use it to make
    // regions outgrow them).
    const char* fixed_slice = std::getenv("CQX_EDIT_FIRST_SLICE");
    const double fixed = fixed_slice ? std::atof(fixed_slice) : 0.0;
    const double side = std::ldexp(1.0, hs);
    for (uint32_t r = 0; r < nr; r++) {
        double est = 64.0;
        for (uint64_t j = plan.first[r]; j < plan.first[r + 1]; j++) {
            const double rad = edits[plan.edit_ids[j]].radius;
            est += 2.0 * std::min(3.14159265 * rad * rad, side * side) + 16.0;
        }
        work[r] = est;
        if (fixed >= 1.0) est = fixed;
        cap[r] = uint32_t(pow2_at_least(uint64_t(std::min(est, 1048576.0))));
        pool_off[r] = pool_total;
        pool_total += cap[r];
        run[r] = r;
    }
    // the wave that takes the heaviest region (most sphere surface) early finishes early: the
    // phase then lasts about as long as its heaviest region or its average load per wave
    std::stable_sort(run.begin(), run.end(), [&](uint32_t a, uint32_t b) { return work[a] > work[b]; });
    cqx_sphere_edit* d_edits;
    DRegionKey* d_keys;
    uint64_t *d_first, *d_pool_off;
    uint32_t *d_edit_ids, *d_run, *d_cap, *d_counts, *d_results, *d_overflow, *d_pool, *d_table;
    std::vector<cqx_sphere_edit> ev(edits, edits + n);
    if ((st = upload(E, kBufEdits, ev, &d_edits)) != CQX_OK || (st = upload(E, kBufKeys, keys, &d_keys)) != CQX_OK ||
        (st = upload(E, kBufFirst, plan.first, &d_first)) != CQX_OK ||
        (st = upload(E, kBufEditIds, plan.edit_ids, &d_edit_ids)) != CQX_OK ||
        (st = buf(E, kBufCounts, nr, &d_counts)) != CQX_OK || (st = buf(E, kBufResults, nr, &d_results)) != CQX_OK ||
        (st = buf(E, kBufOverflow, nr, &d_overflow)) != CQX_OK)
        return st;
    // A region whose list outgrows its slice runs again with a slice 8 times larger, placed after
    // the others; the pool keeps what the finished regions listed.
    for (int attempt = 0; !run.empty(); attempt++) {
        if (attempt > 6) return fail(CQX_ERR_LIMIT, "an edit region listed more nodes than the retries allow");
        if (base + uint64_t(*std::max_element(cap.begin(), cap.end())) >= 0xfffffffeull)
            return fail(CQX_ERR_LIMIT, "store too large for on-device edits");
        // pool and tables: grown keeping their contents
        auto& pb = E->bufs[kBufPool];
        auto& tb = E->bufs[kBufTable];
        const size_t pool_bytes = size_t(pool_total) * 32, table_bytes = size_t(pool_total) * 8;
        for (auto* bb : { &pb, &tb }) {
            const size_t need = bb == &pb ? pool_bytes : table_bytes;
            if (bb->bytes < need) {
                void* q = nullptr;
                const size_t want = std::max(need, bb->bytes + bb->bytes / 2);
                if (hipMalloc(&q, want) != hipSuccess) return fail(CQX_ERR_OOM, "edit pool of %zu bytes failed", want);
                if (attempt > 0 && bb->p) HIP_TRY(hipMemcpy(q, bb->p, bb->bytes, hipMemcpyDeviceToDevice));
                release(bb->p);
                bb->p = q;
                bb->bytes = want;
            }
        }
        d_pool = static_cast<uint32_t*>(pb.p);
        d_table = static_cast<uint32_t*>(tb.p);
        if ((st = upload(E, kBufRun, run, &d_run)) != CQX_OK || (st = upload(E, kBufPoolOff, pool_off, &d_pool_off)) != CQX_OK ||
            (st = upload(E, kBufCap, cap, &d_cap)) != CQX_OK)
            return st;
        uint32_t* d_next = nullptr;
        if ((st = buf(E, kBufNext, 1, &d_next)) != CQX_OK) return st;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, 0, d_next, uint64_t(1), 0u);
        RegionArgs A{ vol->nodes, E->index, E->index_cap - 1, vol->root, hs, uint32_t(base), d_edits, d_keys, d_first,
                      d_edit_ids, d_run, uint32_t(run.size()), d_next, d_pool_off, d_cap, d_pool, d_table, d_counts,
                      d_results, d_overflow };
        const uint32_t waves = uint32_t(std::min<uint64_t>(run.size(), uint64_t(ctx->cu_count) * 32));
        hipLaunchKernelGGL(k_edit_regions, dim3(waves), dim3(64), 0, 0, A);
        HIP_TRY(hipGetLastError());
        std::vector<uint32_t> over(nr);
        HIP_TRY(hipMemcpy(over.data(), d_overflow, nr * 4, hipMemcpyDeviceToHost));
        std::vector<uint32_t> again;
        for (uint32_t r : run)
            if (over[r]) {
                again.push_back(r);
                cap[r] *= 8;
                pool_off[r] = pool_total;
                pool_total += cap[r];
            }
        if (!again.empty())
            cqx::log_debug("device edits: %zu of %zu regions outgrew their node lists; running them again",
                           again.size(), run.size());
        run.swap(again);
        attempts = attempt + 1;
    }
    ms[2] = ms_since(t_phase);
    t_phase = clock::now();

    // ---- merge ----
    uint64_t *d_off, *d_slot_of, *d_fin, *d_scalars;
    uint32_t *d_src, *d_region, *d_key, *d_rep, *d_mtable;
    uint8_t* d_level;
    if ((st = buf(E, kBufOff, nr + 1, &d_off)) != CQX_OK || (st = buf(E, kBufScalars, 8, &d_scalars)) != CQX_OK) return st;
    hipLaunchKernelGGL(k_counts_to_u64, dim3(grid_for(nr, ctx, 256)), dim3(256), 0, 0, d_counts, d_off, nr);
    if ((st = scan_u64(ctx, E, d_off, uint64_t(nr), d_scalars)) != CQX_OK) return st;
    uint64_t total = 0;
    HIP_TRY(hipMemcpy(&total, d_scalars, 8, hipMemcpyDeviceToHost));
    if (base + total >= 0xfffffffeull) return fail(CQX_ERR_LIMIT, "store too large for on-device edits");
    const uint64_t mcap = pow2_at_least(std::max<uint64_t>(1024, 2 * total));
    if ((st = buf(E, kBufSrc, total, &d_src)) != CQX_OK || (st = buf(E, kBufRegion, total, &d_region)) != CQX_OK ||
        (st = buf(E, kBufLevel, total, &d_level)) != CQX_OK || (st = buf(E, kBufKey, total * 8, &d_key)) != CQX_OK ||
        (st = buf(E, kBufSlotOf, total, &d_slot_of)) != CQX_OK || (st = buf(E, kBufRep, total, &d_rep)) != CQX_OK ||
        (st = buf(E, kBufMergeTable, mcap, &d_mtable)) != CQX_OK || (st = buf(E, kBufFin, total + 1, &d_fin)) != CQX_OK)
        return st;
    uint32_t* d_max_level = reinterpret_cast<uint32_t*>(d_scalars + 1);
    HIP_TRY(hipMemset(d_max_level, 0, 4));
    MergeArgs M{ uint32_t(base), nr, d_pool_off, d_pool, d_counts, d_off, total, d_src, d_region, d_level, d_key,
                 d_slot_of, d_rep, d_mtable, mcap - 1, d_max_level };
    uint64_t kept_total = 0;
    if (total) {
        hipLaunchKernelGGL(k_merge_flatten, dim3(grid_for(nr, ctx, 256)), dim3(256), 0, 0, M);
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(mcap, ctx, 256)), dim3(256), 0, 0, d_mtable, mcap, kEmptySlot);
        uint32_t max_level = 0;
        HIP_TRY(hipMemcpy(&max_level, d_max_level, 4, hipMemcpyDeviceToHost));
        const uint32_t g = grid_for(total, ctx, 256);
        for (uint32_t lv = 0; lv <= max_level; lv++) {
            hipLaunchKernelGGL(k_merge_keys, dim3(g), dim3(256), 0, 0, M, lv);
            hipLaunchKernelGGL(k_merge_insert, dim3(g), dim3(256), 0, 0, M, lv);
            hipLaunchKernelGGL(k_merge_rep, dim3(g), dim3(256), 0, 0, M, lv);
        }
        hipLaunchKernelGGL(k_merge_kept, dim3(g), dim3(256), 0, 0, M, d_fin);
        if ((st = scan_u64(ctx, E, d_fin, total, d_scalars + 2)) != CQX_OK) return st;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpy(&kept_total, d_scalars + 2, 8, hipMemcpyDeviceToHost));
    }
    // ---- the levels above the regions: schedule, capacity ----
    const uint32_t levels = uint32_t(32 - hs);
    std::vector<uint32_t> lvl_off(levels + 2, 0), morton(plan.morton_regions), cbeg, cend;
    std::vector<DRegionKey> tkeys;
    std::vector<uint64_t> tab_off(levels + 2, 0);
    for (uint32_t l = 0; l <= levels; l++) {
        lvl_off[l] = uint32_t(tkeys.size());
        const auto& lk = plan.level_keys[l];
        for (size_t i = 0; i < lk.size(); i++) {
            tkeys.push_back({ lk[i].x, lk[i].y, lk[i].z });
            if (l == 0) {
                cbeg.push_back(0);
                cend.push_back(0);
            } else {
                cbeg.push_back(lvl_off[l - 1] + plan.level_children[l][i]);
                cend.push_back(lvl_off[l - 1] + plan.level_children[l][i + 1]);
            }
        }
        tab_off[l + 1] = tab_off[l] + (l == 0 ? 0 : pow2_at_least(std::max<uint64_t>(2, 2 * lk.size())));
    }
    lvl_off[levels + 1] = uint32_t(tkeys.size());
    const uint64_t ntop = tkeys.size();
    st = ensure_store(ctx, vol, base + kept_total + ntop);
    if (st != CQX_OK) return st;
    st = ensure_index(ctx, vol, base + kept_total + ntop);
    if (st != CQX_OK) return st;
    if (total) {
        hipLaunchKernelGGL(k_merge_write, dim3(grid_for(std::max<uint64_t>(total, nr), ctx, 256)), dim3(256), 0, 0, M,
                           d_fin, vol->nodes, d_results);
        hipLaunchKernelGGL(k_index_insert, dim3(grid_for(std::max<uint64_t>(kept_total, 1), ctx, 256)), dim3(256), 0, 0,
                           vol->nodes, base, base + kept_total, E->index, E->index_cap - 1);
        HIP_TRY(hipGetLastError());
    }
    E->indexed = base + kept_total;

    HIP_TRY(hipDeviceSynchronize());
    ms[3] = ms_since(t_phase);
    t_phase = clock::now();

    // ---- the levels above the regions ----
    uint32_t *d_lvl_off, *d_cbeg, *d_cend, *d_morton, *d_res, *d_cand, *d_ckey, *d_ttable, *d_root;
    DRegionKey* d_tkeys;
    uint64_t *d_tab_off, *d_cslot;
    if ((st = upload(E, kBufLvlOff, lvl_off, &d_lvl_off)) != CQX_OK || (st = upload(E, kBufTopKeys, tkeys, &d_tkeys)) != CQX_OK ||
        (st = upload(E, kBufCbeg, cbeg, &d_cbeg)) != CQX_OK || (st = upload(E, kBufCend, cend, &d_cend)) != CQX_OK ||
        (st = upload(E, kBufMorton, morton, &d_morton)) != CQX_OK || (st = buf(E, kBufRes, ntop, &d_res)) != CQX_OK ||
        (st = buf(E, kBufCand, 2 * ntop, &d_cand)) != CQX_OK || (st = buf(E, kBufCkey, ntop * 8, &d_ckey)) != CQX_OK ||
        (st = buf(E, kBufTopTable, tab_off[levels + 1], &d_ttable)) != CQX_OK ||
        (st = upload(E, kBufTabOff, tab_off, &d_tab_off)) != CQX_OK ||
        (st = buf(E, kBufRoot, 2, &d_root)) != CQX_OK)
        return st;
    if ((st = buf(E, kBufCslot, ntop, &d_cslot)) != CQX_OK) return st;
    uint32_t* d_old = nullptr;
    if ((st = buf(E, kBufOld, ntop, &d_old)) != CQX_OK) return st;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(tab_off[levels + 1], ctx, 256)), dim3(256), 0, 0, d_ttable,
                       tab_off[levels + 1], kEmptySlot);
    const uint32_t count_after_merge = uint32_t(base + kept_total);
    HIP_TRY(hipMemcpy(d_root, &count_after_merge, 4, hipMemcpyHostToDevice));
    TopArgs T{ vol->nodes, E->index, E->index_cap - 1, vol->root, hs, levels, d_lvl_off, d_tkeys, d_cbeg, d_cend,
               d_morton, d_results, d_res, d_cand, d_ckey, d_cand + ntop, d_cslot, d_ttable, d_tab_off, d_root,
               d_root + 1, d_old };
    hipLaunchKernelGGL(k_edit_top, dim3(1), dim3(kTopBlock), 0, 0, T);
    HIP_TRY(hipGetLastError());
    uint32_t out[2] = { 0, 0 };
    HIP_TRY(hipMemcpy(out, d_root, 8, hipMemcpyDeviceToHost));
    const uint64_t kOffset32Nodes = uint64_t(1) << 27; // the child-offset width the traversal uses
    if (ctx->pt_plan && (base <= kOffset32Nodes) != (uint64_t(out[0]) <= kOffset32Nodes)) {
        cqx::destroy_pt_plan(ctx->pt_plan); // its captured graph holds the old offset width
        ctx->pt_plan = nullptr;
    }
    vol->node_count = out[0];
    E->indexed = out[0];
    ms[4] = ms_since(t_phase);
    t_phase = clock::now();
    st = set_root(ctx, vol, out[1]);
    if (st != CQX_OK) return st;
    ms[5] = ms_since(t_phase);
    cqx::log_debug("device edits: %llu spheres over %u regions of side 2^%d (%d runs); %llu listed, %llu appended, "
                   "%llu above; plan %.1f ms, index %.1f ms, regions %.1f ms, merge %.1f ms, levels above %.1f ms, "
                   "SubDAGs %.1f ms", (unsigned long long)n, nr, hs, attempts, (unsigned long long)total,
                   (unsigned long long)kept_total, (unsigned long long)(out[0] - base - kept_total), ms[0], ms[1],
                   ms[2], ms[3], ms[4], ms[5]);
    E->roots.resize(E->current + 1);
    E->roots[E->current] = vol->root;
    return CQX_OK;
}

cqx_status cqx_volume_checkpoint(cqx_ctx* ctx, cqx_dvolume* vol)
{
    cqx_status st = check_args(ctx, vol);
    if (st != CQX_OK) return st;
    cqx::DeviceEdits* E = vol->edits;
    E->roots.resize(E->current + 1);
    E->roots[E->current] = vol->root;
    E->roots.push_back(vol->root);
    E->current++;
    return CQX_OK;
}

cqx_status cqx_volume_undo(cqx_ctx* ctx, cqx_dvolume* vol)
{
    cqx_status st = check_args(ctx, vol);
    if (st != CQX_OK) return st;
    cqx::DeviceEdits* E = vol->edits;
    E->roots[E->current] = vol->root;
    if (E->current > 0) E->current--;
    HIP_TRY(hipDeviceSynchronize());
    return set_root(ctx, vol, E->roots[E->current]);
}

cqx_status cqx_volume_redo(cqx_ctx* ctx, cqx_dvolume* vol)
{
    cqx_status st = check_args(ctx, vol);
    if (st != CQX_OK) return st;
    cqx::DeviceEdits* E = vol->edits;
    E->roots[E->current] = vol->root;
    if (E->current + 1 < E->roots.size()) E->current++;
    HIP_TRY(hipDeviceSynchronize());
    return set_root(ctx, vol, E->roots[E->current]);
}

cqx_status cqx_volume_download(cqx_ctx* ctx, const cqx_dvolume* vol, uint64_t first, uint64_t count, uint32_t* nodes)
{
    if (!ctx || !vol || (count && !nodes)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (first > vol->node_count || count > vol->node_count - first)
        return fail(CQX_ERR_INVALID_ARG, "nodes [%llu, %llu) outside the store of %llu", (unsigned long long)first,
                    (unsigned long long)(first + count), (unsigned long long)vol->node_count);
    HIP_TRY(hipSetDevice(ctx->device));
    if (count) HIP_TRY(hipMemcpy(nodes, vol->nodes + first * 8, size_t(count) * 32, hipMemcpyDeviceToHost));
    return CQX_OK;
}

} // extern "C"
