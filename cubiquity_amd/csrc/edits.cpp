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

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <thread>

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
        if (nodes.size() - indexed >= (size_t(1) << 20)) insert_parallel(indexed, nodes.size());
        else
            for (size_t j = indexed; j < nodes.size(); j++) insert_slot(uint32_t(j));
        indexed = nodes.size();
    }

    // insert_slot for a large range on all threads: a slot goes from empty to an index with a
    // compare-and-swap, and where two nodes are equal the lower index wins (an atomic minimum),
    // so lookups return what the serial insertion would.
    void insert_parallel(size_t from, size_t to)
    {
        const size_t mask = slots.size() - 1;
        uint32_t* const sl = slots.data();
        std::atomic<size_t> next{ from };
        auto work = [&]() {
            constexpr size_t kBatch = 4096;
            for (size_t b; (b = next.fetch_add(kBatch)) < to;)
                for (size_t j = b; j < std::min(to, b + kBatch); j++) {
                    for (size_t i = hash_node(nodes[j]) & mask;; i = (i + 1) & mask) {
                        uint32_t cur = __atomic_load_n(&sl[i], __ATOMIC_ACQUIRE);
                        if (cur == kEmpty) {
                            if (__atomic_compare_exchange_n(&sl[i], &cur, uint32_t(j), false, __ATOMIC_ACQ_REL,
                                                            __ATOMIC_ACQUIRE))
                                break;
                            // another thread took the slot: look at what it holds
                        }
                        if (std::memcmp(&nodes[cur], &nodes[j], sizeof(Node)) == 0) {
                            while (uint32_t(j) < cur &&
                                   !__atomic_compare_exchange_n(&sl[i], &cur, uint32_t(j), false, __ATOMIC_ACQ_REL,
                                                                __ATOMIC_ACQUIRE)) {
                            }
                            break;
                        }
                    }
                }
        };
        const int nt = default_threads();
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; t++) pool.emplace_back(work);
        work();
        for (std::thread& t : pool) t.join();
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

    // Indexes the nodes appended to the vector since the last call (each one distinct from every
    // indexed node), growing the table first if it would pass half full.
    void index_appended()
    {
        if (2 * nodes.size() > slots.size()) {
            size_t cap = std::max<size_t>(1024, slots.size());
            while (cap < 2 * nodes.size()) cap *= 2;
            slots.assign(cap, kEmpty);
            indexed = kMaterialCount;
        }
        if (nodes.size() - indexed >= (size_t(1) << 16)) insert_parallel(indexed, nodes.size());
        else
            for (size_t j = indexed; j < nodes.size(); j++) insert_slot(uint32_t(j));
        indexed = nodes.size();
    }

    // Index of a node with this content, or kEmpty (read-only: safe while no thread interns).
    uint32_t find(const Node& n) const
    {
        const size_t mask = slots.size() - 1;
        for (size_t i = hash_node(n) & mask;; i = (i + 1) & mask) {
            const uint32_t s = slots[i];
            if (s == kEmpty || std::memcmp(&nodes[s], &n, sizeof(Node)) == 0) return s;
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

    // True when contains() is false at every point of the box [x0, x1] x [y0, y1] x [z0, z1]: the
    // exact squared distance from the centre to the box exceeds r2 by far more than contains()'s
    // FP32 rounding (< 1e-6 relative) and int64 truncation (1) can make up, and no point of the box
    // is far enough away for its squared distance to leave the int64 range (contains() counts
    // that as inside). NaN centres or radii never qualify.
    bool box_outside(float x0, float y0, float z0, float x1, float y1, float z1) const
    {
        const double c[3] = { cx, cy, cz }, lo3[3] = { x0, y0, z0 }, hi3[3] = { x1, y1, z1 };
        double dmin = 0.0, dmax = 0.0;
        for (int a = 0; a < 3; a++) {
            const double below = lo3[a] - c[a], above = c[a] - hi3[a];
            const double nearest = below > 0.0 ? below : (above > 0.0 ? above : 0.0);
            const double farthest = std::max(std::fabs(lo3[a] - c[a]), std::fabs(hi3[a] - c[a]));
            dmin += nearest * nearest;
            dmax += farthest * farthest;
        }
        return dmax < 1e18 && dmin > double(r2) * (1.0 + 1e-5) + 16.0;
    }
};

// Node access and interning for an edit walk over the whole store (sequential edits).
struct GlobalStore {
    const std::vector<Node>& nodes;
    Interner& in;
    uint32_t child(uint32_t node, uint32_t c) const { return node < kMaterialCount ? node : nodes[node].c[c]; }
    uint32_t intern(const Node& n) { return in.intern(n); }
};

// The nodes one region creates during parallel edits: a private list whose entries have the
// temporary indices base + i (base = store size when the edits began), over a read-only view of
// the store. A node equal to one already in the store is not created; merge_regions() moves the
// lists into the store afterwards.
struct RegionStore {
    const std::vector<Node>& nodes;
    const Interner& in;
    uint32_t base;
    std::vector<Node> local;
    std::vector<uint32_t> slots; // open addressing over `local`
    static constexpr uint32_t kEmpty = 0xffffffffu;

    uint32_t child(uint32_t node, uint32_t c) const
    {
        if (node < kMaterialCount) return node;
        return node >= base ? local[node - base].c[c] : nodes[node].c[c];
    }

    uint32_t intern(const Node& n)
    {
        bool uniform = n.c[0] < kMaterialCount; // the same collapse rule as Interner::intern
        bool has_local = false;
        for (int i = 0; i < 8; i++) {
            uniform = uniform && n.c[i] == n.c[0];
            has_local = has_local || n.c[i] >= base;
        }
        if (uniform) return n.c[0];
        if (!has_local) {
            const uint32_t g = in.find(n);
            if (g != Interner::kEmpty) return g;
        }
        if (2 * (local.size() + 1) > slots.size()) {
            slots.assign(std::max<size_t>(64, slots.size() * 2), kEmpty);
            for (size_t j = 0; j < local.size(); j++) insert_slot(uint32_t(j));
        }
        const size_t mask = slots.size() - 1;
        for (size_t i = hash_node(n) & mask;; i = (i + 1) & mask) {
            const uint32_t s = slots[i];
            if (s == kEmpty) {
                if (uint64_t(base) + local.size() >= 0xfffffffeull) throw std::bad_alloc();
                slots[i] = uint32_t(local.size());
                local.push_back(n);
                return base + slots[i];
            }
            if (std::memcmp(&local[s], &n, sizeof(Node)) == 0) return base + s;
        }
    }

    void insert_slot(uint32_t j)
    {
        const size_t mask = slots.size() - 1;
        for (size_t i = hash_node(local[j]) & mask;; i = (i + 1) & mask)
            if (slots[i] == kEmpty) {
                slots[i] = j;
                return;
            }
    }

};

template <typename Store>
struct EditWalk {
    Store& st;
    const SphereBrushF& b;
    uint32_t mat;

    // fillBrush(volume, brush, matId, nodeIndex, nodeHeight, lower), voxelization.cpp:837-915
    uint32_t fill(uint32_t node, int height, int32_t lx, int32_t ly, int32_t lz)
    {
        const uint32_t ch = uint32_t(height - 1);
        const uint32_t side = 1u << ch;
        // overlaps(brush.bounds(), childBounds), geometry.h:346-357, decided per axis: a child's
        // extent along an axis depends only on its position (0 or 1) along that axis, so two
        // tests per axis give the overlapping children (bit p of ok[a]: position p on axis a).
        const int32_t lo[3] = { lx, ly, lz };
        float f0[3][2], f1[3][2];
        uint32_t ok[3] = { 0, 0, 0 };
        for (int a = 0; a < 3; a++)
            for (uint32_t p = 0; p <= 1; p++) {
                const int32_t v0 = int32_t(uint32_t(lo[a]) + side * p);
                const int32_t v1 = int32_t(uint32_t(v0) + (side - 1));
                f0[a][p] = float(v0);
                f1[a][p] = float(v1);
                if (!(b.hi[a] < f0[a][p] || b.lo[a] > f1[a][p])) ok[a] |= 1u << p;
            }
        if (!(ok[0] && ok[1] && ok[2])) return node; // no child overlaps: nothing changes
        Node out;
        for (int c = 0; c < 8; c++) out.c[c] = st.child(node, uint32_t(c));
        bool changed = false;
        for (uint32_t cz = 0; cz <= 1; cz++)
            for (uint32_t cy = 0; cy <= 1; cy++)
                for (uint32_t cx = 0; cx <= 1; cx++) {
                    if (!((ok[0] >> cx) & (ok[1] >> cy) & (ok[2] >> cz) & 1u)) continue;
                    const uint32_t id = cz << 2 | cy << 1 | cx;
                    const int32_t x0 = int32_t(uint32_t(lx) + side * cx), y0 = int32_t(uint32_t(ly) + side * cy),
                                  z0 = int32_t(uint32_t(lz) + side * cz);
                    const float fx0 = f0[0][cx], fy0 = f0[1][cy], fz0 = f0[2][cz];
                    const float fx1 = f1[0][cx], fy1 = f1[1][cy], fz1 = f1[2][cz];
                    const uint32_t child = out.c[id];
                    if (child == mat) continue;
                    // a child box wholly outside the sphere would come back unchanged from fill()
                    if (height >= 2 && b.box_outside(fx0, fy0, fz0, fx1, fy1, fz1)) continue;
                    // all 8 corners inside (contains is a pure function: evaluation stops at the first
                    // outside, and a single voxel's 8 corners are one point)
                    const bool all = ch == 0 ? b.contains(fx0, fy0, fz0)
                                             : b.contains(fx0, fy0, fz0) && b.contains(fx1, fy1, fz1) &&
                                                   b.contains(fx0, fy0, fz1) && b.contains(fx0, fy1, fz0) &&
                                                   b.contains(fx0, fy1, fz1) && b.contains(fx1, fy0, fz0) &&
                                                   b.contains(fx1, fy0, fz1) && b.contains(fx1, fy1, fz0);
                    uint32_t nc = child;
                    if (height >= 2 && !all) nc = fill(child, height - 1, x0, y0, z0);
                    else if (all) nc = mat;
                    if (nc != child) {
                        out.c[id] = nc;
                        changed = true;
                    }
                }
        if (!changed) return node;
        return st.intern(out);
    }
};

// ---- parallel edits -----------------------------------------------------------------------------
// The volume is cut into aligned cubes ("regions") of side 2^hs, with hs chosen so that no node
// of height >= hs fits inside any of the spheres: (2^hs - 1) * sqrt(3) > 2 r + 4. Above the
// regions fillBrush then only descends (no node there can have its 8 corners inside a sphere),
// and the float box overlap test is monotone in the integer corners, so applying each sphere, in
// order, to every region its bounds touch and rebuilding the levels above gives the voxels of the
// sequential walk. Regions are independent: each applies its spheres in order into a private node
// list (RegionStore), in parallel; the lists are merged into the store as if interned in region
// order (merge_regions), so the result does not depend on the number of threads.

} // namespace

// Smallest region height for these edits (>= 5), or -1 for sequential edits: spheres too large,
// or bounds not within 2^23 voxels of the origin (there every voxel coordinate the walk converts
// to float is exact, so a region's float box overlaps a sphere's bounds exactly when its integer
// box meets them).
int region_height(const cqx_sphere_edit* edits, uint64_t n)
{
    double rmax = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            const double c = edits[i].centre[a];
            if (!(std::fabs(c) + edits[i].radius < 8388608.0)) return -1;
        }
        rmax = std::max(rmax, double(edits[i].radius));
    }
    for (int hs = 5; hs <= 20; hs++)
        if ((std::ldexp(1.0, hs) - 1.0) * std::sqrt(3.0) > 2.0 * rmax + 4.0) return hs;
    return -1;
}

bool morton_less(const RegionKey& a, const RegionKey& b)
{
    // the axis whose differing bit is the most significant; at equal bits z ranks above y above x
    const uint32_t d[3] = { a.x ^ b.x, a.y ^ b.y, a.z ^ b.z };
    const auto below = [](uint32_t p, uint32_t q) { return p < q && p < (p ^ q); }; // msb(p) < msb(q)
    int axis = 2;
    if (below(d[axis], d[1])) axis = 1;
    if (below(d[axis], d[0])) axis = 0;
    const uint32_t av = axis == 0 ? a.x : axis == 1 ? a.y : a.z, bv = axis == 0 ? b.x : axis == 1 ? b.y : b.z;
    return av < bv;
}

void plan_edits(const cqx_sphere_edit* edits, uint64_t n, int hs, EditPlan& plan)
{
    plan = EditPlan{};
    plan.hs = hs;
    // Regions each edit's bounds touch (one voxel of margin), edits kept in order per region.
    std::vector<std::pair<RegionKey, uint32_t>> touch;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            const float c = edits[i].centre[a], r = edits[i].radius;
            const float blo = c - r, bhi = c + r; // SphereBrush bounds, FP32
            const int64_t vlo = int64_t(std::floor(double(blo))) - 1, vhi = int64_t(std::ceil(double(bhi))) + 1;
            lo[a] = uint32_t(std::max<int64_t>(vlo, INT32_MIN) + (int64_t(1) << 31)) >> hs;
            hi[a] = uint32_t(std::min<int64_t>(vhi, INT32_MAX) + (int64_t(1) << 31)) >> hs;
        }
        for (uint32_t z = lo[2]; z <= hi[2]; z++)
            for (uint32_t y = lo[1]; y <= hi[1]; y++)
                for (uint32_t x = lo[0]; x <= hi[0]; x++) touch.push_back({ RegionKey{ x, y, z }, uint32_t(i) });
    }
    // Stable sort by region key (the edits of a region stay in order): an LSD radix sort on the key
    // as one integer relative to the smallest key, when it spans less than 2^40; otherwise a
    // comparison sort.
    uint32_t kmin[3] = { UINT32_MAX, UINT32_MAX, UINT32_MAX }, kmax[3] = { 0, 0, 0 };
    for (const auto& t : touch) {
        const uint32_t k[3] = { t.first.x, t.first.y, t.first.z };
        for (int a = 0; a < 3; a++) {
            kmin[a] = std::min(kmin[a], k[a]);
            kmax[a] = std::max(kmax[a], k[a]);
        }
    }
    uint64_t span[3];
    for (int a = 0; a < 3; a++) span[a] = touch.empty() ? 1 : uint64_t(kmax[a] - kmin[a]) + 1;
    if (!touch.empty() && span[0] * span[1] < (uint64_t(1) << 40) / span[2]) {
        std::vector<uint64_t> key(touch.size());
        for (size_t j = 0; j < touch.size(); j++)
            key[j] = (uint64_t(touch[j].first.x - kmin[0]) * span[1] + (touch[j].first.y - kmin[1])) * span[2] +
                     (touch[j].first.z - kmin[2]);
        std::vector<uint32_t> idx(touch.size()), tmp(touch.size());
        for (size_t j = 0; j < idx.size(); j++) idx[j] = uint32_t(j);
        const uint64_t top = span[0] * span[1] * span[2];
        for (int shift = 0; shift < 40 && (top >> shift) > 0; shift += 11) {
            std::vector<uint32_t> count(2049, 0);
            for (uint32_t j : idx) count[((key[j] >> shift) & 2047u) + 1]++;
            for (int d = 0; d < 2048; d++) count[d + 1] += count[d];
            for (uint32_t j : idx) tmp[count[(key[j] >> shift) & 2047u]++] = j;
            idx.swap(tmp);
        }
        std::vector<std::pair<RegionKey, uint32_t>> sorted(touch.size());
        for (size_t j = 0; j < idx.size(); j++) sorted[j] = touch[idx[j]];
        touch.swap(sorted);
    } else {
        std::stable_sort(touch.begin(), touch.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    }
    plan.edit_ids.resize(touch.size());
    for (size_t j = 0; j < touch.size(); j++) {
        plan.edit_ids[j] = touch[j].second;
        if (j == 0 || !(touch[j].first == touch[j - 1].first)) {
            plan.keys.push_back(touch[j].first);
            plan.first.push_back(j);
        }
    }
    plan.first.push_back(touch.size());
    const size_t nr = plan.keys.size();
    plan.morton_regions.resize(nr);
    for (size_t r = 0; r < nr; r++) plan.morton_regions[r] = uint32_t(r);
    // Morton order: the interleaved code of the keys' 32 bits per axis, z above y above x at each
    // bit (morton_less's rule: the axis of the most significant differing bit decides), sorted as
    // 96-bit integers.
    {
        const auto spread = [](uint32_t v) { // bits 0..20 of v to bits 0, 3, ..., 60
            uint64_t x = v & 0x1fffffu;
            x = (x | (x << 32)) & 0x1f00000000ffffull;
            x = (x | (x << 16)) & 0x1f0000ff0000ffull;
            x = (x | (x << 8)) & 0x100f00f00f00f00full;
            x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
            x = (x | (x << 2)) & 0x1249249249249249ull;
            return x;
        };
        struct Coded {
            uint64_t hi, lo; // bits 21..31 and bits 0..20 of each axis, interleaved
            uint32_t region;
            bool operator<(const Coded& o) const { return hi != o.hi ? hi < o.hi : lo < o.lo; }
        };
        std::vector<Coded> code(nr);
        for (size_t r = 0; r < nr; r++) {
            const RegionKey& k = plan.keys[r];
            code[r] = { spread(k.x >> 21) | (spread(k.y >> 21) << 1) | (spread(k.z >> 21) << 2),
                        spread(k.x) | (spread(k.y) << 1) | (spread(k.z) << 2), uint32_t(r) };
        }
        std::sort(code.begin(), code.end());
        for (size_t r = 0; r < nr; r++) plan.morton_regions[r] = code[r].region;
    }
    const int levels = kRootHeight - hs;
    plan.level_keys.assign(size_t(levels) + 1, {});
    plan.level_children.assign(size_t(levels) + 1, {});
    for (size_t r = 0; r < nr; r++) plan.level_keys[0].push_back(plan.keys[plan.morton_regions[r]]);
    for (int l = 1; l <= levels; l++) {
        const std::vector<RegionKey>& below = plan.level_keys[size_t(l) - 1];
        std::vector<RegionKey>& up = plan.level_keys[size_t(l)];
        std::vector<uint32_t>& ch = plan.level_children[size_t(l)];
        for (size_t j = 0; j < below.size(); j++) {
            const RegionKey k{ below[j].x >> 1, below[j].y >> 1, below[j].z >> 1 };
            if (up.empty() || !(up.back() == k)) {
                up.push_back(k);
                ch.push_back(uint32_t(j));
            }
        }
        ch.push_back(uint32_t(below.size()));
    }
}

namespace {

// The node at height hs whose region is `k`, found by descending from the root (Volume::voxel's
// walk, storage.cpp:481-499).
uint32_t region_node(const std::vector<Node>& nodes, uint32_t root, int hs, const RegionKey& k)
{
    if (hs >= kRootHeight) return root;
    uint32_t node = root;
    const uint32_t ux = k.x << hs, uy = k.y << hs, uz = k.z << hs;
    for (int h = kRootHeight; h > hs && node >= kMaterialCount; h--) {
        const uint32_t c = ((ux >> (h - 1)) & 1u) | (((uy >> (h - 1)) & 1u) << 1) | (((uz >> (h - 1)) & 1u) << 2);
        node = nodes[node].c[c];
    }
    return node;
}

// Runs f(i) for i in [0, count) on up to `threads` threads, `chunk` indices at a time.
template <typename F>
void parallel_for(size_t count, int threads, size_t chunk, F&& f)
{
    std::atomic<size_t> next{ 0 };
    auto work = [&]() {
        for (size_t b; (b = next.fetch_add(chunk)) < count;)
            for (size_t i = b; i < std::min(count, b + chunk); i++) f(i);
    };
    const int nt = int(std::max<size_t>(1, std::min<size_t>(size_t(threads), (count + chunk - 1) / chunk)));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
}

// Moves the regions' node lists into the store and maps each region's result to a store index.
// The store is the one that interning every listed node, region by region and in list order
// (children before parents), would give, built on all threads:
//   * a listed node never equals a node that was in the store before the edits (one without
//     listed children was looked up when it was created; one with a listed child points at a new
//     node), so interning in order keeps the first of each set of equal listed nodes, their
//     listed children replaced by their kept nodes, and appends the kept ones in order;
//   * equal nodes are found level by level (a node's level is one more than its deepest listed
//     child's), in a table of list positions where the lowest position wins.
void merge_regions(std::vector<Node>& nodes, Interner& in, uint32_t base, std::vector<RegionStore>& stores,
                   std::vector<uint32_t>& results, int threads)
{
    const size_t nr = stores.size();
    std::vector<uint64_t> off(nr + 1, 0); // list position of each region's first node
    for (size_t r = 0; r < nr; r++) off[r + 1] = off[r] + stores[r].local.size();
    const uint64_t total = off[nr];
    if (uint64_t(base) + total >= 0xfffffffeull) throw std::bad_alloc();
    std::vector<uint32_t> region(total);
    std::vector<uint8_t> level(total);
    parallel_for(nr, threads, 64, [&](size_t r) {
        const std::vector<Node>& loc = stores[r].local;
        for (size_t i = 0; i < loc.size(); i++) {
            uint32_t l = 0;
            for (int c = 0; c < 8; c++) {
                const uint32_t ch = loc[i].c[c];
                if (ch >= base) l = std::max<uint32_t>(l, uint32_t(level[off[r] + (ch - base)]) + 1u);
            }
            region[off[r] + i] = uint32_t(r);
            level[off[r] + i] = uint8_t(l); // at most the region height (< 32)
        }
    });
    uint32_t levels = 0;
    for (uint8_t l : level) levels = std::max<uint32_t>(levels, uint32_t(l) + 1u);
    std::vector<uint64_t> lstart(size_t(levels) + 1, 0);
    for (uint8_t l : level) lstart[size_t(l) + 1]++;
    for (uint32_t l = 0; l < levels; l++) lstart[l + 1] += lstart[l];
    std::vector<uint32_t> by_level(total); // positions grouped by level, ascending within a level
    {
        std::vector<uint64_t> fill(lstart.begin(), lstart.end() - 1);
        for (uint64_t g = 0; g < total; g++) by_level[fill[level[g]]++] = uint32_t(g);
    }
    // Written before they are read, so left uninitialised: the kept position of each position's
    // node, each listed node with its listed children replaced by base + their kept positions,
    // and the slot holding each position's key.
    std::unique_ptr<uint32_t[]> rep(new uint32_t[total]);
    std::unique_ptr<Node[]> key(new Node[total]);
    std::unique_ptr<uint64_t[]> slot_of(new uint64_t[total]);
    constexpr uint32_t kFree = 0xffffffffu;
    size_t cap = 1024;
    while (cap < 2 * total) cap *= 2;
    std::unique_ptr<uint32_t[]> slots(new uint32_t[cap]);
    uint32_t* const sl = slots.get();
    parallel_for(cap, threads, size_t(1) << 16, [&](size_t i) { sl[i] = kFree; });
    const size_t mask = cap - 1;
    for (uint32_t l = 0; l < levels; l++) {
        const uint64_t b = lstart[l], e = lstart[l + 1];
        // the key is stored before the position is published in a slot (release), and read after
        // the slot is read (acquire)
        parallel_for(e - b, threads, 4096, [&](size_t j) {
            const uint32_t g = by_level[b + j];
            const uint32_t r = region[g];
            Node k = stores[r].local[g - off[r]];
            for (int c = 0; c < 8; c++)
                if (k.c[c] >= base) k.c[c] = base + rep[off[r] + (k.c[c] - base)];
            key[g] = k;
            for (size_t i = hash_node(k) & mask;; i = (i + 1) & mask) {
                uint32_t cur = __atomic_load_n(&sl[i], __ATOMIC_ACQUIRE);
                if (cur == kFree &&
                    __atomic_compare_exchange_n(&sl[i], &cur, g, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
                    slot_of[g] = i;
                    return;
                }
                // a failed exchange left the slot's position in cur
                if (std::memcmp(&key[cur], &k, sizeof(Node)) == 0) {
                    while (g < cur &&
                           !__atomic_compare_exchange_n(&sl[i], &cur, g, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
                    }
                    slot_of[g] = i;
                    return;
                }
            }
        });
        // the lowest position of each set of equal keys is now in their slot
        parallel_for(e - b, threads, 4096, [&](size_t j) {
            const uint32_t g = by_level[b + j];
            rep[g] = sl[slot_of[g]];
        });
    }
    std::vector<uint32_t> fin(total); // store index of each kept position
    uint32_t next = base;
    for (uint64_t g = 0; g < total; g++)
        if (rep[g] == g) fin[g] = next++;
    nodes.resize(next);
    parallel_for(total, threads, 4096, [&](size_t g) {
        if (rep[g] != g) return;
        const uint32_t r = region[g];
        Node k = stores[r].local[g - off[r]];
        for (int c = 0; c < 8; c++)
            if (k.c[c] >= base) k.c[c] = fin[rep[off[r] + (k.c[c] - base)]];
        nodes[fin[g]] = k;
    });
    for (size_t r = 0; r < nr; r++)
        if (results[r] >= base) results[r] = fin[rep[off[r] + (results[r] - base)]];
    in.index_appended();
}

// Rebuilds the levels above the regions, bottom-up: at each height hs + l the nodes that contain
// edited regions, in Morton order (plan.level_keys), take their changed children and the changed
// ones are interned in that order. `results` holds each region's new node (by region index).
uint32_t rebuild_levels(GlobalStore& g, uint32_t root, const EditPlan& plan, const std::vector<uint32_t>& results)
{
    std::vector<uint32_t> cur(plan.morton_regions.size()), next;
    for (size_t i = 0; i < cur.size(); i++) cur[i] = results[plan.morton_regions[i]];
    const int levels = kRootHeight - plan.hs;
    for (int l = 1; l <= levels; l++) {
        const int h = plan.hs + l;
        const std::vector<RegionKey>& keys = plan.level_keys[size_t(l)];
        const std::vector<RegionKey>& below = plan.level_keys[size_t(l) - 1];
        const std::vector<uint32_t>& ch = plan.level_children[size_t(l)];
        next.assign(keys.size(), 0);
        for (size_t i = 0; i < keys.size(); i++) {
            const uint32_t old = region_node(g.nodes, root, h, keys[i]);
            Node out;
            for (int c = 0; c < 8; c++) out.c[c] = g.child(old, uint32_t(c));
            bool changed = false;
            for (uint32_t j = ch[i]; j < ch[i + 1]; j++) {
                const uint32_t id = (below[j].x & 1u) | (below[j].y & 1u) << 1 | (below[j].z & 1u) << 2;
                if (cur[j] != out.c[id]) {
                    out.c[id] = cur[j];
                    changed = true;
                }
            }
            next[i] = changed ? g.intern(out) : old;
        }
        cur.swap(next);
    }
    return cur[0];
}

} // namespace
} // namespace cqx

using namespace cqx;

extern "C" {

cqx_status cqx_hv_edit_spheres_mt(cqx_host_volume* hv, const cqx_sphere_edit* edits, uint64_t n, int32_t threads,
                                  uint64_t* first_new)
{
    if (!hv || (n && !edits)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (first_new) *first_new = hv->v.nodes.size();
    if (n == 0) return CQX_OK;
    for (uint64_t i = 0; i < n; i++) {
        if (edits[i].material >= kMaterialCount) return fail(CQX_ERR_INVALID_ARG, "edit %llu: material >= 256",
                                                              (unsigned long long)i);
        if (!(edits[i].radius >= 0.0f)) return fail(CQX_ERR_INVALID_ARG, "edit %llu: bad radius", (unsigned long long)i);
    }
    if (threads <= 0) threads = default_threads();
    std::vector<Node>& nodes = hv->v.nodes;
    try {
        Interner in(nodes, hv->v.edit_slots, hv->v.edit_indexed);
        GlobalStore g{ nodes, in };
        const int hs = (threads > 1 && n >= 16) ? region_height(edits, n) : -1;
        if (hs < 0) {
            uint32_t root = hv->v.root;
            for (uint64_t i = 0; i < n; i++) {
                const SphereBrushF b(edits[i]);
                EditWalk<GlobalStore> w{ g, b, edits[i].material };
                root = w.fill(root, kRootHeight, INT32_MIN, INT32_MIN, INT32_MIN);
            }
            hv->v.root = root;
            return CQX_OK;
        }
        using clock = std::chrono::steady_clock;
        const auto ms_since = [](clock::time_point t) {
            return std::chrono::duration<double, std::milli>(clock::now() - t).count();
        };
        const clock::time_point t_start = clock::now();
        EditPlan plan;
        plan_edits(edits, n, hs, plan);
        const std::vector<RegionKey>& keys = plan.keys;
        const std::vector<uint64_t>& first = plan.first;
        const size_t nr = keys.size();
        const uint32_t base = uint32_t(nodes.size());
        std::vector<RegionStore> stores;
        stores.reserve(nr);
        for (size_t r = 0; r < nr; r++) stores.push_back(RegionStore{ nodes, in, base, {}, {} });
        std::vector<uint32_t> results(nr);
        std::atomic<size_t> next{ 0 };
        std::atomic<bool> oom{ false };
        auto work = [&]() {
            try {
                for (size_t r; (r = next.fetch_add(1)) < nr;) {
                    const RegionKey& k = keys[r];
                    uint32_t node = region_node(nodes, hv->v.root, hs, k);
                    const int32_t lx = int32_t((k.x << hs) - 0x80000000u), ly = int32_t((k.y << hs) - 0x80000000u),
                                  lz = int32_t((k.z << hs) - 0x80000000u);
                    for (uint64_t j = first[r]; j < first[r + 1]; j++) {
                        const cqx_sphere_edit& e = edits[plan.edit_ids[j]];
                        const SphereBrushF b(e);
                        EditWalk<RegionStore> w{ stores[r], b, e.material };
                        node = w.fill(node, hs, lx, ly, lz);
                    }
                    results[r] = node;
                }
            } catch (const std::bad_alloc&) {
                oom = true;
                next = nr;
            }
        };
        const double plan_ms = ms_since(t_start);
        const clock::time_point t_apply = clock::now();
        const int nt = int(std::min<size_t>(size_t(threads), nr));
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; t++) pool.emplace_back(work);
        work();
        for (std::thread& t : pool) t.join();
        if (oom) throw std::bad_alloc();
        const double apply_ms = ms_since(t_apply);
        const clock::time_point t_merge = clock::now();
        merge_regions(nodes, in, base, stores, results, threads);
        stores.clear();
        const double merge_ms = ms_since(t_merge);
        const clock::time_point t_top = clock::now();
        hv->v.root = rebuild_levels(g, hv->v.root, plan, results);
        log_debug("edits: %llu spheres over %zu regions of side 2^%d on %d threads; plan %.1f ms, apply %.1f ms, "
                  "merge %.1f ms, top levels %.1f ms; %zu nodes appended",
                  (unsigned long long)n, nr, hs, nt, plan_ms, apply_ms, merge_ms, ms_since(t_top),
                  nodes.size() - size_t(base));
        return CQX_OK;
    } catch (const std::bad_alloc&) {
        return fail(CQX_ERR_OOM, "out of memory while editing");
    }
}

cqx_status cqx_hv_edit_spheres(cqx_host_volume* hv, const cqx_sphere_edit* edits, uint64_t n, uint64_t* first_new)
{
    return cqx_hv_edit_spheres_mt(hv, edits, n, 0, first_new);
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
