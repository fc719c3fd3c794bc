// Internal declarations shared by the host-side sources of libcqx.so.
#pragma once

#include "cqx.h"

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

namespace cqx {

// Records `code` and a printf-style message as the calling thread's last error; returns code.
cqx_status fail(cqx_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
// A debug note (cqx_set_debug_handler; without one, stderr when CQX_PROGRESS=1).
void log_debug(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
// True when a debug or progress handler is set or CQX_PROGRESS=1 (progress worth computing).
bool progress_wanted();
void report_progress(const char* task, int first_step, int current_step, int last_step);

constexpr uint32_t kMaterialCount = CQX_MATERIAL_COUNT;
constexpr int kRootHeight = 32;

struct Node {
    uint32_t c[8];
};
static_assert(sizeof(Node) == 32, "node is 32 bytes");

// Host node store + root: the library's view of a reference Volume.
struct HostVolume {
    std::vector<Node> nodes; // includes the 256 material nodes
    uint32_t root = 0;
    // Edit history (Volume::checkpoint/undo/redo, storage.cpp:358-385): edits never modify an
    // existing node, so a checkpoint only needs the root index. roots[current] is brought up to
    // date from `root` whenever the history is used.
    std::vector<uint32_t> roots{ 0 };
    size_t current = 0;
    // Hash-consing index over `nodes` kept between edit calls (empty until the first edit,
    // dropped by operations that renumber nodes).
    std::vector<uint32_t> edit_slots;
    size_t edit_indexed = 0; // nodes [256, edit_indexed) are in edit_slots
};

// A NodeStore holding only the 256 material nodes (storage.cpp:110-122).
std::vector<Node> material_nodes();

cqx_status validate_dag(const uint32_t* nodes, uint64_t node_count, uint32_t root);
// Breadth-first renumbering from the root; drops unreachable nodes (generator.cpp).
void canonical_order(std::vector<Node>& nodes, uint32_t& root);
cqx_status find_subdags(const uint32_t* nodes, uint64_t node_count, uint32_t root, cqx_subdag out[8]);

int default_threads();

// ---- sphere edits: the region plan shared by the host and device edit paths (edits.cpp) ----

// A region of side 2^hs, by its index along each axis of the offset coordinates u = v + 2^31.
struct RegionKey {
    uint32_t x, y, z;
    bool operator<(const RegionKey& o) const { return x != o.x ? x < o.x : y != o.y ? y < o.y : z < o.z; }
    bool operator==(const RegionKey& o) const { return x == o.x && y == o.y && z == o.z; }
};

// Morton order of keys (bit b of z above bit b of y above bit b of x, higher bits first): a node's
// children (key >> 1 shared) stay together and in child-id order.
bool morton_less(const RegionKey& a, const RegionKey& b);

struct EditPlan {
    int hs = -1;                    // region height; -1: the edits are applied in order, no regions
    std::vector<RegionKey> keys;    // edited regions in ascending (x, y, z) order: the merge order
    std::vector<uint64_t> first;    // keys.size() + 1 offsets into edit_ids
    std::vector<uint32_t> edit_ids; // per region, the edits whose bounds touch it, in edit order
    // The levels above the regions, bottom-up. Level 0 holds the regions in Morton order
    // (morton_regions: the region behind each entry); level l >= 1 (height hs + l, up to the root
    // at 32) the nodes containing edited regions in Morton order, each covering the level l - 1
    // entries [level_children[l][i], level_children[l][i + 1]). A level l - 1 entry with key k is
    // child (k.x & 1) | (k.y & 1) << 1 | (k.z & 1) << 2 of its node, whose key is k >> 1.
    std::vector<uint32_t> morton_regions;
    std::vector<std::vector<RegionKey>> level_keys;
    std::vector<std::vector<uint32_t>> level_children;
};

// Smallest region height for these edits (>= 5), or -1 when they must be applied in order.
int region_height(const cqx_sphere_edit* edits, uint64_t n);
// The regions each edit touches and the levels above them (hs >= 0).
void plan_edits(const cqx_sphere_edit* edits, uint64_t n, int hs, EditPlan& plan);

} // namespace cqx

struct cqx_host_volume {
    cqx::HostVolume v;
};
