// Host side of the node store: .dag I/O, validation, SubDAG precompute, point queries and the
// expansion of compact GPU records into the reference's RayVolumeIntersection fields.
#include "cqx_internal.h"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>

namespace cqx {

static thread_local std::string g_last_error;
static std::atomic<cqx_message_fn> g_debug_handler{ nullptr }, g_warning_handler{ nullptr };
static std::atomic<cqx_progress_fn> g_progress_handler{ nullptr };
static std::mutex g_handler_lock; // one handler call at a time

static bool progress_env()
{
    const char* env = std::getenv("CQX_PROGRESS");
    return env && env[0] == '1';
}

cqx_status fail(cqx_status code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (cqx_message_fn f = g_warning_handler.load()) {
        std::lock_guard<std::mutex> g(g_handler_lock);
        f(buf);
    }
    return code;
}

void log_debug(const char* fmt, ...)
{
    cqx_message_fn f = g_debug_handler.load();
    if (!f && !progress_env()) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> g(g_handler_lock);
    if (f) f(buf);
    else std::fprintf(stderr, "[cqx] %s\n", buf);
}

bool progress_wanted() { return g_debug_handler.load() || g_progress_handler.load() || progress_env(); }

void report_progress(const char* task, int first_step, int current_step, int last_step)
{
    if (cqx_progress_fn f = g_progress_handler.load()) {
        std::lock_guard<std::mutex> g(g_handler_lock);
        f(task, first_step, current_step, last_step);
    }
}

void clear_error() { g_last_error.clear(); }

// All cores, or OMP_NUM_THREADS when the environment sets a CPU share.
int default_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (const char* env = std::getenv("OMP_NUM_THREADS")) {
        const int k = std::atoi(env);
        if (k > 0 && (n == 0 || unsigned(k) < n)) n = unsigned(k);
    }
    return n ? int(n) : 1;
}

std::vector<Node> material_nodes()
{
    std::vector<Node> v(kMaterialCount);
    for (uint32_t i = 0; i < kMaterialCount; i++)
        for (int c = 0; c < 8; c++) v[i].c[c] = i;
    return v;
}

// Iterative DFS from the root: every followed index in range, no cycle, and at most 32 internal
// levels (root at height 32, an internal node at height 1 has only material children).
cqx_status validate_dag(const uint32_t* nodes, uint64_t node_count, uint32_t root)
{
    if (!nodes) return fail(CQX_ERR_INVALID_ARG, "nodes is null");
    if (node_count < kMaterialCount)
        return fail(CQX_ERR_INVALID_DAG, "store has %llu nodes, fewer than the %u material nodes",
                    (unsigned long long)node_count, kMaterialCount);
    if (node_count >= 0xffffffffull)
        return fail(CQX_ERR_LIMIT, "node count %llu does not fit 32-bit indices", (unsigned long long)node_count);
    if (root >= node_count)
        return fail(CQX_ERR_INVALID_DAG, "root %u out of range (%llu nodes)", root, (unsigned long long)node_count);
    for (uint32_t i = 0; i < kMaterialCount; i++)
        for (int c = 0; c < 8; c++)
            if (nodes[uint64_t(i) * 8 + c] != i)
                return fail(CQX_ERR_INVALID_DAG, "material node %u is not self-referencing", i);
    if (root < kMaterialCount) return CQX_OK;

    // depth[n]: 0 = unvisited, 0xff = on the DFS path, else 1 + longest internal chain below.
    std::vector<uint8_t> depth(node_count, 0);
    struct Frame { uint32_t node; uint8_t next; };
    std::vector<Frame> stack;
    stack.reserve(64);
    stack.push_back({ root, 0 });
    depth[root] = 0xff;
    while (!stack.empty()) {
        Frame& f = stack.back();
        if (f.next == 8) {
            uint8_t d = 1;
            for (int c = 0; c < 8; c++) {
                uint32_t ch = nodes[uint64_t(f.node) * 8 + c];
                if (ch >= kMaterialCount && depth[ch] + 1 > d) d = uint8_t(depth[ch] + 1);
            }
            depth[f.node] = d;
            stack.pop_back();
            continue;
        }
        uint32_t ch = nodes[uint64_t(f.node) * 8 + f.next++];
        if (ch >= node_count)
            return fail(CQX_ERR_INVALID_DAG, "node %u child %d = %u out of range (%llu nodes)", f.node,
                        f.next - 1, ch, (unsigned long long)node_count);
        if (ch < kMaterialCount) continue;
        if (depth[ch] == 0xff) return fail(CQX_ERR_INVALID_DAG, "cycle through node %u", ch);
        if (depth[ch] != 0) continue;
        if (stack.size() >= 32)
            return fail(CQX_ERR_INVALID_DAG, "DAG deeper than 32 levels below node %u", f.node);
        depth[ch] = 0xff;
        stack.push_back({ ch, 0 });
    }
    if (depth[root] > kRootHeight)
        return fail(CQX_ERR_INVALID_DAG, "DAG has %u internal levels, more than 32", depth[root]);
    return CQX_OK;
}

// findSubDAG, src/library/raytracing.cpp:43-87: follow the chain of single-child nodes below
// each root child to the first split.
cqx_status find_subdags(const uint32_t* nodes, uint64_t node_count, uint32_t root, cqx_subdag out[8])
{
    if (!nodes || !out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (root >= node_count) return fail(CQX_ERR_INVALID_DAG, "root %u out of range", root);
    for (uint32_t c = 0; c < 8; c++) {
        int height = kRootHeight;
        uint32_t lower[3] = { 0x80000000u, 0x80000000u, 0x80000000u };
        uint32_t only = c;
        uint32_t next = nodes[uint64_t(root) * 8 + only];
        uint32_t node = 0;
        uint32_t count = 1;
        while (count == 1) {
            if (height == 0) return fail(CQX_ERR_INVALID_DAG, "single-child chain deeper than the tree");
            height--;
            node = next;
            if (node >= node_count) return fail(CQX_ERR_INVALID_DAG, "child index %u out of range", node);
            for (int a = 0; a < 3; a++) lower[a] ^= ((only >> a) & 1u) << height;
            count = 0;
            const uint32_t* n = nodes + uint64_t(node) * 8;
            for (uint32_t i = 0; i < 8; i++)
                if (n[i] > 0) { next = n[i]; count++; only = i; }
        }
        cqx_subdag s;
        std::memset(&s, 0, sizeof(s));
        for (int a = 0; a < 3; a++) s.lower[a] = int32_t(lower[a]);
        s.height = height;
        s.node = node;
        out[c] = s;
    }
    return CQX_OK;
}

} // namespace cqx

using namespace cqx;

extern "C" {

const char* cqx_last_error(void) { return g_last_error.c_str(); }

void cqx_set_debug_handler(cqx_message_fn fn) { g_debug_handler.store(fn); }
void cqx_set_warning_handler(cqx_message_fn fn) { g_warning_handler.store(fn); }
void cqx_set_progress_handler(cqx_progress_fn fn) { g_progress_handler.store(fn); }

const char* cqx_version(void) { return "cqx 0.1 (gfx950)"; }

cqx_status cqx_validate_dag(const uint32_t* nodes, uint64_t node_count, uint32_t root)
{
    return validate_dag(nodes, node_count, root);
}

cqx_status cqx_find_subdags(const uint32_t* nodes, uint64_t node_count, uint32_t root, cqx_subdag out[8])
{
    return find_subdags(nodes, node_count, root, out);
}

cqx_status cqx_hv_load_dag(const char* path, cqx_host_volume** out)
{
    if (!path || !out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CQX_ERR_IO, "cannot open %s", path);
    uint32_t hdr[2];
    if (std::fread(hdr, 4, 2, f) != 2) { std::fclose(f); return fail(CQX_ERR_IO, "%s: short header", path); }
    cqx_host_volume* hv = new (std::nothrow) cqx_host_volume;
    if (!hv) { std::fclose(f); return fail(CQX_ERR_OOM, "out of host memory"); }
    hv->v.nodes = material_nodes();
    const uint64_t n = hdr[1];
    try {
        hv->v.nodes.resize(kMaterialCount + n);
    } catch (...) {
        delete hv; std::fclose(f);
        return fail(CQX_ERR_OOM, "cannot hold %llu nodes", (unsigned long long)n);
    }
    if (n && std::fread(hv->v.nodes.data() + kMaterialCount, sizeof(Node), n, f) != n) {
        delete hv; std::fclose(f);
        return fail(CQX_ERR_IO, "%s: truncated node array", path);
    }
    std::fclose(f);
    // The file holds full-array indices; Volume::load adds sharedNodesBegin() - MaterialCount = 0.
    hv->v.root = hdr[0];
    cqx_status st = validate_dag(&hv->v.nodes[0].c[0], hv->v.nodes.size(), hv->v.root);
    if (st != CQX_OK) { delete hv; return st; }
    *out = hv;
    return CQX_OK;
}

cqx_status cqx_hv_save_dag(const cqx_host_volume* hv, const char* path)
{
    if (!hv || !path) return fail(CQX_ERR_INVALID_ARG, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CQX_ERR_IO, "cannot create %s", path);
    const uint64_t n = hv->v.nodes.size() - kMaterialCount;
    uint32_t hdr[2] = { hv->v.root, uint32_t(n) };
    bool ok = std::fwrite(hdr, 4, 2, f) == 2;
    ok = ok && std::fwrite(hv->v.nodes.data() + kMaterialCount, sizeof(Node), n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? CQX_OK : fail(CQX_ERR_IO, "write to %s failed", path);
}

cqx_status cqx_hv_from_nodes(const uint32_t* nodes, uint64_t node_count, uint32_t root, cqx_host_volume** out)
{
    if (!out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    cqx_status st = validate_dag(nodes, node_count, root);
    if (st != CQX_OK) return st;
    cqx_host_volume* hv = new (std::nothrow) cqx_host_volume;
    if (!hv) return fail(CQX_ERR_OOM, "out of host memory");
    try {
        hv->v.nodes.resize(node_count);
    } catch (...) {
        delete hv;
        return fail(CQX_ERR_OOM, "cannot hold %llu nodes", (unsigned long long)node_count);
    }
    std::memcpy(hv->v.nodes.data(), nodes, node_count * sizeof(Node));
    hv->v.root = root;
    *out = hv;
    return CQX_OK;
}

void cqx_hv_free(cqx_host_volume* hv) { delete hv; }

const uint32_t* cqx_hv_nodes(const cqx_host_volume* hv) { return hv ? &hv->v.nodes[0].c[0] : nullptr; }

uint64_t cqx_hv_node_count(const cqx_host_volume* hv) { return hv ? hv->v.nodes.size() : 0; }

uint32_t cqx_hv_root(const cqx_host_volume* hv) { return hv ? hv->v.root : 0; }

// Volume::voxel, src/library/storage.cpp:481-499, with the signed-to-unsigned mapping of
// make_position_unsigned (storage.cpp:45-50) and extract_next_child_id (storage.cpp:57-67).
cqx_status cqx_hv_voxel(const cqx_host_volume* hv, int32_t x, int32_t y, int32_t z, uint32_t* material)
{
    if (!hv || !material) return fail(CQX_ERR_INVALID_ARG, "null argument");
    uint32_t u[3] = { uint32_t(x) + 0x80000000u, uint32_t(y) + 0x80000000u, uint32_t(z) + 0x80000000u };
    uint32_t node = hv->v.root;
    for (int level = 0; node >= kMaterialCount; level++) {
        if (level > kRootHeight) return fail(CQX_ERR_INVALID_DAG, "descent deeper than the tree");
        uint32_t c = (u[0] >> 31) | ((u[1] >> 31) << 1) | ((u[2] >> 31) << 2);
        u[0] <<= 1; u[1] <<= 1; u[2] <<= 1;
        node = hv->v.nodes[node].c[c];
    }
    *material = node;
    return CQX_OK;
}

cqx_status cqx_hits_to_intersections(const cqx_ray* rays, const cqx_hit* hits, uint64_t n, uint32_t flags,
                                     cqx_intersection* out)
{
    if (n && (!rays || !hits || !out)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    const bool surf = (flags & CQX_SURFACE_PROPERTIES) != 0;
    for (uint64_t i = 0; i < n; i++) {
        const cqx_hit& h = hits[i];
        cqx_intersection& r = out[i];
        std::memset(&r, 0, sizeof(r));
        if (!(h.info & CQX_HIT_FLAG)) continue;
        r.hit = 1;
        r.distance = double(h.t);
        for (int a = 0; a < 3; a++) {
            volatile float p = rays[i].d[a] * h.t; // keep multiply and add separate
            r.position[a] = rays[i].o[a] + p;
        }
        if (surf) {
            r.material = (h.info >> 16) & 0xffu;
            for (int a = 0; a < 3; a++) {
                uint32_t code = (h.info >> (9 + 2 * a)) & 3u;
                float mag = (code & 1u) ? 1.0f : 0.0f;
                r.normal[a] = (code & 2u) ? -mag : mag;
            }
        }
    }
    return CQX_OK;
}

} // extern "C"
