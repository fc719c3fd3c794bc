// Seeded synthetic workloads: the heightfield + spheres volume generator with a parallel,
// top-down-pruned, bottom-up hash-consed DAG builder, and the camera / incoherent ray sets.
//
// The reference builds volumes voxel by voxel (Volume::setVoxel, storage.cpp:396-438) and then
// deduplicates with Volume::bake (NodeStore::merge, storage.cpp:208-290); that is O(N^3 * 32)
// and takes hours at 1024^3. This builder produces the same canonical DAG directly: identical
// nodes share one index and a node whose 8 children are the same material collapses into that
// material (the isMaterialNode(Node) rule of storage.cpp:69-75 used by merge_node, :261-263).
#include "cqx_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <mutex>
#include <thread>

namespace cqx {
namespace {

// PCG32 (M.E. O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically Good
// Algorithms for Random Number Generation", 2014; pcg32_srandom_r / pcg32_random_r).
struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t initstate, uint64_t initseq)
    {
        state = 0;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
    }
    double uniform() { return next() * (1.0 / 4294967296.0); }
};

// splitmix64 finaliser (Steele, Lea, Flood 2014)
inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Value-noise lattice value in [-1, 1) for octave o at integer lattice point (ix, iz).
inline double lattice(uint64_t seed, int o, int64_t ix, int64_t iz)
{
    uint64_t k = mix64(seed + 0x9e3779b97f4a7c15ULL * uint64_t(o + 1));
    k = mix64(k ^ uint64_t(uint32_t(int32_t(ix))));
    k = mix64(k ^ (uint64_t(uint32_t(int32_t(iz))) << 32));
    return double(k >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// Bilinear value noise with smoothstep weights.
inline double value_noise(uint64_t seed, int o, double x, double z)
{
    const double fx = std::floor(x), fz = std::floor(z);
    const int64_t ix = int64_t(fx), iz = int64_t(fz);
    const double tx = x - fx, tz = z - fz;
    const double sx = tx * tx * (3.0 - 2.0 * tx);
    const double sz = tz * tz * (3.0 - 2.0 * tz);
    const double v00 = lattice(seed, o, ix, iz), v10 = lattice(seed, o, ix + 1, iz);
    const double v01 = lattice(seed, o, ix, iz + 1), v11 = lattice(seed, o, ix + 1, iz + 1);
    const double a = v00 + (v10 - v00) * sx;
    const double b = v01 + (v11 - v01) * sx;
    return a + (b - a) * sz;
}

double height_at(const cqx_gen_params& p, int64_t x, int64_t z)
{
    const double u = double(x) / p.period, v = double(z) / p.period;
    double sum = 0.0, norm = 0.0, amp = 1.0, freq = 1.0;
    for (int o = 0; o < p.octaves; o++) {
        sum += amp * value_noise(p.seed, o, u * freq, v * freq);
        norm += amp;
        amp *= p.gain;
        freq *= p.lacunarity;
    }
    return p.base_height + p.amplitude * (sum / norm);
}

// Integer solid threshold: voxel y is below height h iff y < ceil(h).
inline int32_t ceil_clamped(double h)
{
    double c = std::ceil(h);
    if (!(c > -1073741824.0)) return -1073741824;
    if (c > 1073741824.0) return 1073741824;
    return int32_t(c);
}

struct Sphere {
    double cx, cy, cz, r2;
    uint32_t mat;
};

inline bool inside(const Sphere& s, int64_t x, int64_t y, int64_t z)
{
    const double dx = double(x) - s.cx, dy = double(y) - s.cy, dz = double(z) - s.cz;
    return dx * dx + dy * dy + dz * dz <= s.r2;
}

struct NodeEq {
    bool operator()(const Node& a, const Node& b) const { return std::memcmp(&a, &b, sizeof(Node)) == 0; }
};

inline uint64_t node_hash(const Node& n)
{
    uint64_t h = 0x243f6a8885a308d3ULL;
    for (int i = 0; i < 8; i++) {
        h ^= n.c[i];
        h *= 0x9fb21c651e98df25ULL;
        h ^= h >> 29;
    }
    return h;
}

// Open-addressing hash-consing table over a node vector whose first 256 entries are the
// material nodes. intern() returns the canonical index of a node given its children.
class NodeTable {
public:
    std::vector<Node> nodes;

    NodeTable() : nodes(material_nodes()), slots_(1024, kEmpty) {}

    uint32_t intern(const Node& n)
    {
        bool uniform = n.c[0] < kMaterialCount;
        for (int i = 1; i < 8 && uniform; i++) uniform = n.c[i] == n.c[0];
        if (uniform) return n.c[0];
        if (2 * (nodes.size() + 1) > slots_.size()) grow();
        const size_t mask = slots_.size() - 1;
        for (size_t i = node_hash(n) & mask;; i = (i + 1) & mask) {
            const uint32_t s = slots_[i];
            if (s == kEmpty) {
                slots_[i] = uint32_t(nodes.size());
                nodes.push_back(n);
                return slots_[i];
            }
            if (NodeEq()(nodes[s], n)) return s;
        }
    }

    // Frees the hash index (the nodes stay): for when no more nodes will be interned.
    void drop_index() { std::vector<uint32_t>().swap(slots_); }

private:
    static constexpr uint32_t kEmpty = 0xffffffffu;
    std::vector<uint32_t> slots_;

    void grow()
    {
        std::vector<uint32_t> s(slots_.size() * 2, kEmpty);
        const size_t mask = s.size() - 1;
        for (size_t j = kMaterialCount; j < nodes.size(); j++) {
            size_t i = node_hash(nodes[j]) & mask;
            while (s[i] != kEmpty) i = (i + 1) & mask;
            s[i] = uint32_t(j);
        }
        slots_.swap(s);
    }
};

class Builder {
public:
    explicit Builder(const cqx_gen_params& p, int threads) : p_(p), threads_(threads)
    {
        d_ = p.log2_size;
        n_ = int64_t(1) << d_;
        rockc_ = ceil_clamped(p.rock_height);
        make_spheres();
        make_heights();
    }

    // Returns the canonical reference of the volume node (height d, lower corner 0).
    uint32_t build(NodeTable& global)
    {
        const int top = std::max(kBrick, d_ - 3);
        const int64_t tside = int64_t(1) << top;
        const int64_t per_axis = n_ / tside;
        const int64_t ntasks = per_axis * per_axis * per_axis;
        std::vector<uint32_t> all(spheres_.size());
        for (size_t i = 0; i < all.size(); i++) all[i] = uint32_t(i);

        const int workers = int(std::max<int64_t>(1, std::min<int64_t>(threads_, ntasks)));
        // Each task is built into a worker-local table, then merged into the global table under a
        // lock (children before parents, as created) and the local table starts afresh, so that
        // memory stays near the size of the result. The merge order does not matter: the store is
        // renumbered breadth-first from the root afterwards.
        std::vector<uint32_t> task_ref(ntasks);
        std::atomic<int64_t> next{ 0 }, done{ 0 };
        std::mutex merge_lock;
        // progress (cqx_set_progress_handler / debug notes) each time another tenth of the regions
        // is built
        const bool progress = progress_wanted();
        if (progress) report_progress("Building volume", 0, 0, int(ntasks));
        auto work = [&]() {
            NodeTable local;
            std::vector<uint32_t> remap;
            for (;;) {
                const int64_t t = next.fetch_add(1);
                if (t >= ntasks) break;
                const int64_t x0 = (t % per_axis) * tside, y0 = ((t / per_axis) % per_axis) * tside,
                              z0 = (t / (per_axis * per_axis)) * tside;
                const uint32_t r = build_region(local, x0, y0, z0, top, all);
                const std::vector<Node>& ln = local.nodes;
                remap.assign(ln.size(), 0);
                {
                    std::lock_guard<std::mutex> g(merge_lock);
                    for (size_t j = kMaterialCount; j < ln.size(); j++) {
                        Node n = ln[j];
                        for (int c = 0; c < 8; c++)
                            if (n.c[c] >= kMaterialCount) n.c[c] = remap[n.c[c]];
                        remap[j] = global.intern(n);
                    }
                }
                task_ref[t] = r < kMaterialCount ? r : remap[r];
                local = NodeTable();
                const int64_t k = ++done;
                if (progress && (k * 10) / ntasks != ((k - 1) * 10) / ntasks) {
                    std::lock_guard<std::mutex> g(merge_lock);
                    log_debug("built %lld of %lld regions, %zu nodes", (long long)k, (long long)ntasks,
                              global.nodes.size());
                    report_progress("Building volume", 0, int(k), int(ntasks));
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < workers; w++) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        auto task_global = [&](int64_t t) { return task_ref[t]; };
        return build_top(global, 0, 0, 0, d_, top, per_axis, task_global);
    }

private:
    static constexpr int kBrick = 3; // dense evaluation below 8^3 regions

    const cqx_gen_params& p_;
    int threads_;
    int d_;
    int64_t n_;
    int32_t rockc_;
    std::vector<Sphere> spheres_;
    std::vector<std::vector<int32_t>> hmin_, hmax_; // [level][(z >> l) * (n >> l) + (x >> l)]

    void make_spheres()
    {
        Pcg32 rng(p_.seed, 0x5fe7e5u);
        const double n = double(n_);
        spheres_.resize(size_t(std::max(0, p_.n_spheres)));
        for (int k = 0; k < p_.n_spheres; k++) {
            Sphere& s = spheres_[k];
            s.cx = rng.uniform() * n;
            s.cy = rng.uniform() * n;
            s.cz = rng.uniform() * n;
            const double r = p_.r_min + (p_.r_max - p_.r_min) * rng.uniform();
            s.r2 = r * r;
            s.mat = 3u + uint32_t(k % (p_.n_materials - 2));
        }
    }

    void make_heights()
    {
        hmin_.resize(d_ + 1);
        hmax_.resize(d_ + 1);
        hmin_[0].resize(size_t(n_ * n_));
        std::atomic<int64_t> next{ 0 };
        auto rows = [&]() {
            for (;;) {
                const int64_t z = next.fetch_add(1);
                if (z >= n_) break;
                for (int64_t x = 0; x < n_; x++) hmin_[0][size_t(z * n_ + x)] = ceil_clamped(height_at(p_, x, z));
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < threads_; w++) pool.emplace_back(rows);
        rows();
        for (auto& th : pool) th.join();
        hmax_[0] = hmin_[0];
        for (int l = 1; l <= d_; l++) {
            const int64_t s = n_ >> l, ps = n_ >> (l - 1);
            hmin_[l].resize(size_t(s * s));
            hmax_[l].resize(size_t(s * s));
            for (int64_t z = 0; z < s; z++)
                for (int64_t x = 0; x < s; x++) {
                    int32_t mn = INT32_MAX, mx = INT32_MIN;
                    for (int dz = 0; dz < 2; dz++)
                        for (int dx = 0; dx < 2; dx++) {
                            const size_t i = size_t((2 * z + dz) * ps + 2 * x + dx);
                            mn = std::min(mn, hmin_[l - 1][i]);
                            mx = std::max(mx, hmax_[l - 1][i]);
                        }
                    hmin_[l][size_t(z * s + x)] = mn;
                    hmax_[l][size_t(z * s + x)] = mx;
                }
        }
    }

    int32_t column_height(int64_t x, int64_t z) const { return hmin_[0][size_t(z * n_ + x)]; }

    uint32_t terrain(int64_t x, int64_t y, int64_t z) const
    {
        if (y >= column_height(x, z)) return 0;
        return y < rockc_ ? 2u : 1u;
    }

    // Some voxel centre of [x0, x0+s)^3 may lie inside sphere k (conservative; the per-voxel test
    // is monotone in each |coordinate - centre|, so the nearest point bounds every voxel).
    bool overlaps(const Sphere& sp, int64_t x0, int64_t y0, int64_t z0, int64_t s) const
    {
        auto axis = [](double c, int64_t lo, int64_t hi) {
            if (c < double(lo)) return double(lo) - c;
            if (c > double(hi)) return double(hi) - c;
            return 0.0;
        };
        const double dx = axis(sp.cx, x0, x0 + s - 1), dy = axis(sp.cy, y0, y0 + s - 1),
                     dz = axis(sp.cz, z0, z0 + s - 1);
        return dx * dx + dy * dy + dz * dz <= sp.r2;
    }

    // Every voxel centre of the region is inside the sphere iff the farthest corner is (same
    // monotonicity argument).
    bool contains(const Sphere& sp, int64_t x0, int64_t y0, int64_t z0, int64_t s) const
    {
        auto far = [](double c, int64_t lo, int64_t hi) {
            double a = double(lo) - c, b = double(hi) - c;
            return std::fabs(a) > std::fabs(b) ? a : b;
        };
        const double dx = far(sp.cx, x0, x0 + s - 1), dy = far(sp.cy, y0, y0 + s - 1),
                     dz = far(sp.cz, z0, z0 + s - 1);
        return dx * dx + dy * dy + dz * dz <= sp.r2;
    }

    uint32_t build_region(NodeTable& t, int64_t x0, int64_t y0, int64_t z0, int h,
                          const std::vector<uint32_t>& cand)
    {
        const int64_t s = int64_t(1) << h;
        std::vector<uint32_t> sph;
        for (uint32_t k : cand)
            if (overlaps(spheres_[k], x0, y0, z0, s)) sph.push_back(k);
        if (!sph.empty()) {
            const Sphere& last = spheres_[sph.back()];
            if (contains(last, x0, y0, z0, s)) return last.mat;
        } else {
            const size_t i = size_t((z0 >> h) * (n_ >> h) + (x0 >> h));
            const int32_t hmn = hmin_[h][i], hmx = hmax_[h][i];
            if (y0 >= hmx) return 0;
            if (y0 + s - 1 < hmn) {
                if (y0 + s - 1 < rockc_) return 2;
                if (y0 >= rockc_) return 1;
            }
        }
        if (h == kBrick) return build_brick(t, x0, y0, z0, sph);
        Node n;
        const int64_t half = s / 2;
        for (uint32_t c = 0; c < 8; c++)
            n.c[c] = build_region(t, x0 + ((c & 1) ? half : 0), y0 + ((c & 2) ? half : 0),
                                  z0 + ((c & 4) ? half : 0), h - 1, sph);
        return t.intern(n);
    }

    uint32_t build_brick(NodeTable& t, int64_t x0, int64_t y0, int64_t z0, const std::vector<uint32_t>& sph)
    {
        constexpr int B = 1 << kBrick;
        uint32_t cell[B * B * B]; // current level, x fastest
        for (int z = 0; z < B; z++)
            for (int y = 0; y < B; y++)
                for (int x = 0; x < B; x++) {
                    const int64_t gx = x0 + x, gy = y0 + y, gz = z0 + z;
                    uint32_t m = 0;
                    bool in_sphere = false;
                    for (size_t k = sph.size(); k-- > 0;)
                        if (inside(spheres_[sph[k]], gx, gy, gz)) { m = spheres_[sph[k]].mat; in_sphere = true; break; }
                    if (!in_sphere) m = terrain(gx, gy, gz);
                    cell[(z * B + y) * B + x] = m;
                }
        for (int side = B; side > 1; side /= 2) {
            const int hs = side / 2;
            for (int z = 0; z < hs; z++)
                for (int y = 0; y < hs; y++)
                    for (int x = 0; x < hs; x++) {
                        Node n;
                        for (uint32_t c = 0; c < 8; c++) {
                            const int cx = 2 * x + (c & 1), cy = 2 * y + ((c >> 1) & 1), cz = 2 * z + ((c >> 2) & 1);
                            n.c[c] = cell[(cz * side + cy) * side + cx];
                        }
                        cell[(z * hs + y) * hs + x] = t.intern(n);
                    }
        }
        return cell[0];
    }

    template <typename F>
    uint32_t build_top(NodeTable& g, int64_t x0, int64_t y0, int64_t z0, int h, int top, int64_t per_axis, F& task)
    {
        if (h == top) {
            const int64_t tside = int64_t(1) << top;
            const int64_t t = (x0 / tside) + (y0 / tside) * per_axis + (z0 / tside) * per_axis * per_axis;
            return task(t);
        }
        const int64_t half = (int64_t(1) << h) / 2;
        Node n;
        for (uint32_t c = 0; c < 8; c++)
            n.c[c] = build_top(g, x0 + ((c & 1) ? half : 0), y0 + ((c & 2) ? half : 0), z0 + ((c & 4) ? half : 0),
                               h - 1, top, per_axis, task);
        return g.intern(n);
    }
};

// Dense grid placed at an arbitrary lower corner: top-down over the full 2^32 space, visiting only
// regions that intersect the grid's box.
class DenseBuilder {
public:
    DenseBuilder(const uint8_t* grid, const int64_t n[3], const int64_t lo[3]) : g_(grid)
    {
        for (int a = 0; a < 3; a++) {
            n_[a] = n[a];
            lo_[a] = lo[a];
        }
    }

    uint32_t build(NodeTable& t, int64_t x0, int64_t y0, int64_t z0, int h)
    {
        const int64_t s = int64_t(1) << h;
        const int64_t p[3] = { x0, y0, z0 };
        for (int a = 0; a < 3; a++)
            if (p[a] + s <= lo_[a] || p[a] >= lo_[a] + n_[a]) return 0;
        if (h == 0) return g_[((z0 - lo_[2]) * n_[1] + (y0 - lo_[1])) * n_[0] + (x0 - lo_[0])];
        const int64_t half = s / 2;
        Node nd;
        for (uint32_t c = 0; c < 8; c++)
            nd.c[c] = build(t, x0 + ((c & 1) ? half : 0), y0 + ((c & 2) ? half : 0), z0 + ((c & 4) ? half : 0), h - 1);
        return t.intern(nd);
    }

private:
    const uint8_t* g_;
    int64_t n_[3], lo_[3];
};

} // namespace

// Renumber the internal nodes breadth-first from the root (children visited in slot order, each
// node numbered at its first visit). The result is independent of construction order (so of
// thread scheduling), keeps the top levels together at the front, and gives the children of a
// node neighbouring indices. Unreachable nodes are dropped.
void canonical_order(std::vector<Node>& nodes, uint32_t& root)
{
    if (root < kMaterialCount) {
        nodes.resize(kMaterialCount);
        return;
    }
    // The output doubles as the breadth-first queue: out[i] holds the old children of the i-th
    // node reached until its turn comes, then its children get their new indices.
    std::vector<uint32_t> remap(nodes.size(), 0xffffffffu);
    std::vector<Node> out;
    out.reserve(nodes.size());
    out.insert(out.end(), nodes.begin(), nodes.begin() + kMaterialCount);
    remap[root] = kMaterialCount;
    out.push_back(nodes[root]);
    for (size_t i = kMaterialCount; i < out.size(); i++) {
        for (int c = 0; c < 8; c++) {
            const uint32_t ch = out[i].c[c];
            if (ch < kMaterialCount) continue;
            if (remap[ch] == 0xffffffffu) {
                remap[ch] = uint32_t(out.size());
                out.push_back(nodes[ch]);
            }
            out[i].c[c] = remap[ch];
        }
    }
    std::vector<uint32_t>().swap(remap);
    out.shrink_to_fit();
    nodes.swap(out);
    root = kMaterialCount;
}

} // namespace cqx

using namespace cqx;

extern "C" {

cqx_status cqx_gen_default_params(int32_t log2_size, uint64_t seed, cqx_gen_params* out)
{
    if (!out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (log2_size < 1 || log2_size > 20) return fail(CQX_ERR_LIMIT, "log2_size %d outside [1, 20]", log2_size);
    const double s = std::ldexp(1.0, log2_size - 7);
    const double l2s = double(log2_size - 7);
    cqx_gen_params p;
    std::memset(&p, 0, sizeof(p));
    p.log2_size = log2_size;
    p.octaves = std::max(1, 4 + int(std::lround(2.0 / 3.0 * l2s)));
    p.n_spheres = int32_t(std::lround(32.0 * s * s));
    p.n_materials = 3;
    p.seed = seed;
    p.gain = 0.5;
    p.lacunarity = 2.0;
    p.period = 64.0 * s;
    p.base_height = 64.0 * s;
    p.amplitude = 40.0 * s;
    p.rock_height = 40.0 * s;
    p.r_min = 4.0 * std::cbrt(s);
    p.r_max = 16.0 * std::cbrt(s);
    *out = p;
    return CQX_OK;
}

cqx_status cqx_hv_generate(const cqx_gen_params* p, int threads, cqx_host_volume** out)
{
    if (!p || !out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (p->log2_size < 3 || p->log2_size > 20) return fail(CQX_ERR_LIMIT, "log2_size %d outside [3, 20]", p->log2_size);
    if (p->n_materials < 3 || p->n_materials > 255) return fail(CQX_ERR_INVALID_ARG, "n_materials must be in [3, 255]");
    if (p->octaves < 1 || p->n_spheres < 0 || !(p->period > 0)) return fail(CQX_ERR_INVALID_ARG, "bad generator parameters");
    if (threads <= 0) threads = default_threads();
    try {
        NodeTable g;
        Builder b(*p, threads);
        uint32_t v = b.build(g);
        // Chain from the root (height 32) to the volume node at height d placed at [0, N)^3:
        // root child 7 is the all-positive octant, below it child 0 (the lower corner) each level.
        for (int h = p->log2_size + 1; h <= 31; h++) {
            Node n{};
            n.c[0] = v;
            v = g.intern(n);
        }
        Node r{};
        r.c[7] = v;
        uint32_t root = g.intern(r);
        if (g.nodes.size() >= 0xffffffffull) return fail(CQX_ERR_LIMIT, "volume needs more than 2^32 nodes");
        g.drop_index();
        log_debug("renumbering %zu nodes", g.nodes.size());
        canonical_order(g.nodes, root);
        cqx_host_volume* hv = new cqx_host_volume;
        hv->v.nodes = std::move(g.nodes);
        hv->v.root = root;
        *out = hv;
        return CQX_OK;
    } catch (const std::bad_alloc&) {
        return fail(CQX_ERR_OOM, "out of host memory while building the volume");
    }
}

cqx_status cqx_hv_from_dense(const uint8_t* grid, const int32_t size[3], const int32_t lower[3], cqx_host_volume** out)
{
    if (!grid || !size || !lower || !out) return fail(CQX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int64_t n[3], lo[3];
    for (int a = 0; a < 3; a++) {
        n[a] = size[a];
        lo[a] = lower[a];
        if (n[a] < 1) return fail(CQX_ERR_INVALID_ARG, "grid size must be positive");
        if (lo[a] + n[a] - 1 > INT32_MAX) return fail(CQX_ERR_LIMIT, "grid extends past the volume's last voxel");
    }
    if (n[0] * n[1] * n[2] > (int64_t(1) << 31)) return fail(CQX_ERR_LIMIT, "grid larger than 2^31 voxels");
    try {
        NodeTable t;
        DenseBuilder b(grid, n, lo);
        const int64_t m = -(int64_t(1) << 31);
        uint32_t root = b.build(t, m, m, m, kRootHeight);
        canonical_order(t.nodes, root);
        cqx_host_volume* hv = new cqx_host_volume;
        hv->v.nodes = std::move(t.nodes);
        hv->v.root = root;
        *out = hv;
        return CQX_OK;
    } catch (const std::bad_alloc&) {
        return fail(CQX_ERR_OOM, "out of host memory");
    }
}

cqx_status cqx_camera_rays(const double eye[3], const double target[3], double fov_y_degrees, uint32_t width,
                           uint32_t height, cqx_ray* out)
{
    if (!eye || !target || !out || !width || !height) return fail(CQX_ERR_INVALID_ARG, "bad argument");
    double f[3] = { target[0] - eye[0], target[1] - eye[1], target[2] - eye[2] };
    const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0)) return fail(CQX_ERR_INVALID_ARG, "eye equals target");
    for (double& c : f) c /= fl;
    // right = normalize(cross(forward, +y)), up = cross(right, forward)
    double r[3] = { -f[2], 0.0, f[0] };
    const double rl = std::sqrt(r[0] * r[0] + r[2] * r[2]);
    if (!(rl > 0)) return fail(CQX_ERR_INVALID_ARG, "view direction parallel to +y");
    for (double& c : r) c /= rl;
    const double u[3] = { r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0] };
    const double scale = 2.0 * std::tan(fov_y_degrees * 3.14159265358979323846 / 360.0);
    const double aspect = double(width) / double(height);
    for (uint32_t y = 0; y < height; y++)
        for (uint32_t x = 0; x < width; x++) {
            const double xo = (x + 0.5 - width / 2.0) / width * aspect * scale;
            const double yo = (y + 0.5 - height / 2.0) / height * scale;
            double d[3];
            for (int a = 0; a < 3; a++) d[a] = f[a] + r[a] * xo - u[a] * yo;
            const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            cqx_ray& ray = out[uint64_t(y) * width + x];
            for (int a = 0; a < 3; a++) {
                ray.o[a] = float(eye[a]);
                ray.d[a] = float(d[a] / dl);
            }
        }
    return CQX_OK;
}

cqx_status cqx_incoherent_rays(const cqx_host_volume* hv, int32_t log2_size, uint64_t seed, uint64_t n, int threads,
                               cqx_ray* out)
{
    if (!hv || (n && !out)) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (log2_size < 1 || log2_size > 30) return fail(CQX_ERR_LIMIT, "log2_size out of range");
    if (threads <= 0) threads = default_threads();
    const double side = std::ldexp(1.0, log2_size);
    std::atomic<uint64_t> next{ 0 };
    auto work = [&]() {
        constexpr uint64_t kChunk = 4096;
        for (;;) {
            const uint64_t c0 = next.fetch_add(kChunk);
            if (c0 >= n) break;
            const uint64_t c1 = std::min(n, c0 + kChunk);
            for (uint64_t i = c0; i < c1; i++) {
                Pcg32 rng(seed, i);
                float o[3];
                for (int attempt = 0; attempt < 4096; attempt++) {
                    for (int a = 0; a < 3; a++) o[a] = float(rng.uniform() * side - 0.5);
                    int32_t v[3];
                    for (int a = 0; a < 3; a++) v[a] = int32_t(std::floor(double(o[a]) + 0.5));
                    uint32_t m = 0;
                    cqx_hv_voxel(hv, v[0], v[1], v[2], &m);
                    if (m == 0) break;
                }
                const double z = 2.0 * rng.uniform() - 1.0;
                const double phi = 2.0 * 3.14159265358979323846 * rng.uniform();
                const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
                cqx_ray& ray = out[i];
                for (int a = 0; a < 3; a++) ray.o[a] = o[a];
                ray.d[0] = float(rr * std::cos(phi));
                ray.d[1] = float(rr * std::sin(phi));
                ray.d[2] = float(z);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < threads; w++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return CQX_OK;
}

} // extern "C"
