// Volume edits: sphere fill/carve with the reference's semantics, applied copy-on-write with
// hash-consing so that the GPU can take the result as an append-only delta.
//
// The voxel outcome of each edit is that of Cubiquity's fillBrush(volume, SphereBrush, material)
// (src/library/voxelization.cpp:825-915, voxelization.h:91-126), including its arithmetic:
//   * a child region [lower, lower + 2^h - 1] (integer voxel coordinates, corners converted to
//     float) is skipped unless its float box overlaps the brush box [centre - r, centre + r];
//   * SphereBrush::contains(p): the float distances dx*dx + dy*dy + dz*dz are truncated to int64
//     and compared (as float) with r*r, strictly;
//   * a child whose 8 corners are all inside is set to the material; a child already equal to
//     the material is left alone; otherwise the walk descends (down to single voxels).
// The reference rewrites nodes in place after copying shared ones (NodeStore::setNodeChild,
// storage.cpp:152-167) and deduplicates only at bake. Here every changed node is interned in a
// hash table over the whole store: nothing reachable before the call is modified, new nodes are
// appended, identical nodes are shared and 8 equal material children collapse (bake's rules,
// storage.cpp:243-290), so the store stays canonical and the GPU copy can be updated with the
// appended range only.
#include "cqx_internal.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <new>

namespace cqx {
namespace {

inline uint64_t hash_node(const Node& n)
{
    uint64_t h = 0x243f6a8885a308d3ULL;
    for (int i = 0; i < 8; i++) {
        h ^= n.c[i];
        h *= 0x9fb21c651e98df25ULL;
        h ^= h >> 29;
    }
    return h;
}

// Hash-consing index over a node vector. The slot table lives in the host volume between calls:
// on construction only the nodes appended since the last call (or all, the first time) are
// indexed.
struct Interner {
    std::vector<Node>& nodes;
    std::vector<uint32_t>& slots;
    size_t& indexed;
    static constexpr uint32_t kEmpty = 0xffffffffu;

    Interner(std::vector<Node>& n, std::vector<uint32_t>& s, size_t& idx) : nodes(n), slots(s), indexed(idx)
    {
        size_t cap = 1024;
        while (cap < 2 * nodes.size()) cap *= 2;
        if (slots.size() < cap || indexed < kMaterialCount || indexed > nodes.size()) {
            slots.assign(cap, kEmpty);
            indexed = kMaterialCount;
        }
        for (size_t j = indexed; j < nodes.size(); j++) insert_slot(uint32_t(j));
        indexed = nodes.size();
    }

    void insert_slot(uint32_t j)
    {
        const size_t mask = slots.size() - 1;
        for (size_t i = hash_node(nodes[j]) & mask;; i = (i + 1) & mask) {
            if (slots[i] == kEmpty) {
                slots[i] = j;
                return;
            }
            if (std::memcmp(&nodes[slots[i]], &nodes[j], sizeof(Node)) == 0) return; // keep the first
        }
    }

    uint32_t intern(const Node& n)
    {
        bool uniform = n.c[0] < kMaterialCount;
        for (int i = 1; i < 8 && uniform; i++) uniform = n.c[i] == n.c[0];
        if (uniform) return n.c[0];
        if (2 * (nodes.size() + 1) > slots.size()) {
            slots.assign(slots.size() * 2, kEmpty);
            for (size_t j = kMaterialCount; j < nodes.size(); j++) insert_slot(uint32_t(j));
        }
        const size_t mask = slots.size() - 1;
        for (size_t i = hash_node(n) & mask;; i = (i + 1) & mask) {
            const uint32_t s = slots[i];
            if (s == kEmpty) {
                if (nodes.size() >= 0xfffffffeull) throw std::bad_alloc();
                slots[i] = uint32_t(nodes.size());
                nodes.push_back(n);
                indexed = nodes.size();
                return slots[i];
            }
            if (std::memcmp(&nodes[s], &n, sizeof(Node)) == 0) return s;
        }
    }
};

struct SphereBrushF {
    float cx, cy, cz, r2;
    float lo[3], hi[3]; // bounds: centre -+ radius (float)

    SphereBrushF(const cqx_sphere_edit& e)
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

    // SphereBrush::contains, voxelization.h:105-114
    bool contains(float x, float y, float z) const
    {
        const float dx = x - cx, dy = y - cy, dz = z - cz;
        const float sum = dx * dx + dy * dy + dz * dz;
        int64_t dist;
        // float -> int64 conversion as the x86-64 build performs it (cvttss2si): out-of-range and
        // NaN inputs give INT64_MIN
        if (!(sum >= -9.2233720368547758e18f && sum < 9.2233720368547758e18f)) dist = std::numeric_limits<int64_t>::min();
        else dist = int64_t(sum);
        return float(dist) < r2;
    }
};

struct EditWalk {
    const std::vector<Node>& nodes;
    Interner& in;
    const SphereBrushF& b;
    uint32_t mat;

    uint32_t child_of(uint32_t node, uint32_t c) const { return node < kMaterialCount ? node : nodes[node].c[c]; }

    // fillBrush(volume, brush, matId, nodeIndex, nodeHeight, lower), voxelization.cpp:837-915
    uint32_t fill(uint32_t node, int height, int32_t lx, int32_t ly, int32_t lz)
    {
        const uint32_t ch = uint32_t(height - 1);
        Node out;
        for (int c = 0; c < 8; c++) out.c[c] = child_of(node, uint32_t(c));
        bool changed = false;
        for (uint32_t cz = 0; cz <= 1; cz++)
            for (uint32_t cy = 0; cy <= 1; cy++)
                for (uint32_t cx = 0; cx <= 1; cx++) {
                    const uint32_t id = cz << 2 | cy << 1 | cx;
                    const uint32_t side = 1u << ch;
                    const int32_t x0 = int32_t(uint32_t(lx) + side * cx), y0 = int32_t(uint32_t(ly) + side * cy),
                                  z0 = int32_t(uint32_t(lz) + side * cz);
                    const int32_t x1 = int32_t(uint32_t(x0) + (side - 1)), y1 = int32_t(uint32_t(y0) + (side - 1)),
                                  z1 = int32_t(uint32_t(z0) + (side - 1));
                    const float fx0 = float(x0), fy0 = float(y0), fz0 = float(z0);
                    const float fx1 = float(x1), fy1 = float(y1), fz1 = float(z1);
                    // overlaps(brush.bounds(), childBounds), geometry.h:346-357
                    if (b.hi[0] < fx0 || b.lo[0] > fx1 || b.hi[1] < fy0 || b.lo[1] > fy1 || b.hi[2] < fz0 ||
                        b.lo[2] > fz1)
                        continue;
                    const bool all = b.contains(fx0, fy0, fz0) & b.contains(fx0, fy0, fz1) & b.contains(fx0, fy1, fz0) &
                                     b.contains(fx0, fy1, fz1) & b.contains(fx1, fy0, fz0) & b.contains(fx1, fy0, fz1) &
                                     b.contains(fx1, fy1, fz0) & b.contains(fx1, fy1, fz1);
                    const uint32_t child = out.c[id];
                    if (child == mat) continue;
                    uint32_t nc = child;
                    if (height >= 2 && !all) nc = fill(child, height - 1, x0, y0, z0);
                    else if (all) nc = mat;
                    if (nc != child) {
                        out.c[id] = nc;
                        changed = true;
                    }
                }
        if (!changed) return node;
        return in.intern(out);
    }
};

} // namespace
} // namespace cqx

using namespace cqx;

extern "C" {

cqx_status cqx_hv_edit_spheres(cqx_host_volume* hv, const cqx_sphere_edit* edits, uint64_t n, uint64_t* first_new)
{
    if (!hv || (n && !edits)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (first_new) *first_new = hv->v.nodes.size();
    if (n == 0) return CQX_OK;
    for (uint64_t i = 0; i < n; i++) {
        if (edits[i].material >= kMaterialCount) return fail(CQX_ERR_INVALID_ARG, "edit %llu: material >= 256",
                                                              (unsigned long long)i);
        if (!(edits[i].radius >= 0.0f)) return fail(CQX_ERR_INVALID_ARG, "edit %llu: bad radius", (unsigned long long)i);
    }
    try {
        Interner in(hv->v.nodes, hv->v.edit_slots, hv->v.edit_indexed);
        uint32_t root = hv->v.root;
        for (uint64_t i = 0; i < n; i++) {
            const SphereBrushF b(edits[i]);
            EditWalk w{ hv->v.nodes, in, b, edits[i].material };
            root = w.fill(root, kRootHeight, INT32_MIN, INT32_MIN, INT32_MIN);
        }
        hv->v.root = root;
        return CQX_OK;
    } catch (const std::bad_alloc&) {
        return fail(CQX_ERR_OOM, "out of memory while editing");
    }
}

cqx_status cqx_hv_compact(cqx_host_volume* hv)
{
    if (!hv) return fail(CQX_ERR_INVALID_ARG, "null argument");
    try {
        canonical_order(hv->v.nodes, hv->v.root);
        // renumbering invalidates the edit index and the history's other roots
        hv->v.edit_slots.clear();
        hv->v.edit_indexed = 0;
        hv->v.roots.assign(1, hv->v.root);
        hv->v.current = 0;
        return CQX_OK;
    } catch (const std::bad_alloc&) {
        return fail(CQX_ERR_OOM, "out of memory while compacting");
    }
}

// Volume::checkpoint, storage.cpp:358-371: later undo() returns to the current state. Redo
// states beyond the current one are discarded.
cqx_status cqx_hv_checkpoint(cqx_host_volume* hv)
{
    if (!hv) return fail(CQX_ERR_INVALID_ARG, "null argument");
    hv->v.roots.resize(hv->v.current + 1);
    hv->v.roots[hv->v.current] = hv->v.root;
    hv->v.roots.push_back(hv->v.root);
    hv->v.current++;
    return CQX_OK;
}

// Volume::undo / redo, storage.cpp:373-385 (no-ops at either end of the history).
cqx_status cqx_hv_undo(cqx_host_volume* hv)
{
    if (!hv) return fail(CQX_ERR_INVALID_ARG, "null argument");
    hv->v.roots[hv->v.current] = hv->v.root;
    if (hv->v.current > 0) hv->v.current--;
    hv->v.root = hv->v.roots[hv->v.current];
    return CQX_OK;
}

cqx_status cqx_hv_redo(cqx_host_volume* hv)
{
    if (!hv) return fail(CQX_ERR_INVALID_ARG, "null argument");
    hv->v.roots[hv->v.current] = hv->v.root;
    if (hv->v.current + 1 < hv->v.roots.size()) hv->v.current++;
    hv->v.root = hv->v.roots[hv->v.current];
    return CQX_OK;
}

} // extern "C"
