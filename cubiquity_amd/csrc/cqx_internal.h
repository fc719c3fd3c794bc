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

} // namespace cqx

struct cqx_host_volume {
    cqx::HostVolume v;
};
