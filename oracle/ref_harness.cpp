// Test infrastructure only: a thin extern "C" shim over the UNMODIFIED Cubiquity library sources,
// which oracle/Makefile compiles in place from the reference checkout into oracle/_ref/.
// Nothing in the product links this; tests and the bench's CPU-baseline leg load it with ctypes.
//
// The shim only calls the reference API:
//   Volume::load                 src/library/storage.cpp:505-528
//   findSubDAGs                  src/library/raytracing.cpp:89-97
//   intersectVolume              src/library/raytracing.cpp:397-478
//   traceRayRef                  src/library/raytracing.cpp:521-535
//   Volume::voxel                src/library/storage.cpp:481-499
//   Volume::setVoxel / bake      src/library/storage.cpp:396-438, 387-394
//   NodeStore::countNodes        src/library/storage.cpp:169-190
//   Internals::fnv1a / bit_mix   src/library/base.cpp:62-78
// and records each call's result as plain numbers.
#include "raytracing.h"
#include "storage.h"
#include "utility.h"
#include "voxelization.h"

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <chrono>

using namespace Cubiquity;

extern "C" {

// Result of one reference raycast, widened to plain fields (RayVolumeIntersection,
// src/library/raytracing.h:48-55).
struct ref_result {
    int32_t hit;
    uint32_t material;
    double distance;
    float position[3];
    float normal[3];
};

void* ref_volume_load(const char* path)
{
    Volume* v = new Volume();
    if (!v->load(path)) { delete v; return nullptr; }
    return v;
}

void* ref_volume_new(void) { return new Volume(); }

void ref_volume_free(void* h) { delete static_cast<Volume*>(h); }

uint32_t ref_volume_root(void* h) { return static_cast<Volume*>(h)->rootNodeIndex(); }

void ref_volume_set_voxel(void* h, int32_t x, int32_t y, int32_t z, uint32_t m)
{
    static_cast<Volume*>(h)->setVoxel(x, y, z, static_cast<MaterialId>(m));
}

void ref_volume_save(void* h, const char* path) { static_cast<Volume*>(h)->save(path); }

// fillBrush(volume, SphereBrush(centre, radius), material), voxelization.cpp:825-835
void ref_fill_sphere(void* h, float cx, float cy, float cz, float radius, uint32_t m)
{
    fillBrush(*static_cast<Volume*>(h), SphereBrush(cx, cy, cz, radius), static_cast<MaterialId>(m));
}

// Dense read-out of a box of voxels (x fastest), via Volume::voxel.
void ref_volume_dense(void* h, int32_t x0, int32_t y0, int32_t z0, int32_t nx, int32_t ny, int32_t nz, uint8_t* out)
{
    Volume* v = static_cast<Volume*>(h);
    for (int32_t z = 0; z < nz; z++)
        for (int32_t y = 0; y < ny; y++)
            for (int32_t x = 0; x < nx; x++)
                out[(size_t(z) * ny + y) * nx + x] = v->voxel(x0 + x, y0 + y, z0 + z);
}

void ref_volume_bake(void* h) { static_cast<Volume*>(h)->bake(); }

// Volume::checkpoint / undo / redo, storage.cpp:358-385
void ref_volume_checkpoint(void* h) { static_cast<Volume*>(h)->checkpoint(); }
void ref_volume_undo(void* h) { static_cast<Volume*>(h)->undo(); }
void ref_volume_redo(void* h) { static_cast<Volume*>(h)->redo(); }

uint32_t ref_volume_count_nodes(void* h) { return static_cast<Volume*>(h)->countNodes(); }

uint32_t ref_volume_voxel(void* h, int32_t x, int32_t y, int32_t z)
{
    return static_cast<Volume*>(h)->voxel(x, y, z);
}

// Copies the 8 SubDAG records (32 B each) into out[64] as int32.
void ref_find_subdags(void* h, int32_t* out)
{
    Volume* v = static_cast<Volume*>(h);
    SubDAGArray s = findSubDAGs(Internals::getNodes(*v), Internals::getRootNodeIndex(*v));
    static_assert(sizeof(SubDAG) == 32, "SubDAG layout");
    std::memcpy(out, s.data(), sizeof(SubDAG) * 8);
}

static void fill(ref_result& r, const RayVolumeIntersection& x)
{
    r.hit = x.hit ? 1 : 0;
    r.material = x.material;
    r.distance = x.distance;
    r.position[0] = x.position.x; r.position[1] = x.position.y; r.position[2] = x.position.z;
    r.normal[0] = x.normal.x; r.normal[1] = x.normal.y; r.normal[2] = x.normal.z;
}

// rays: n x 6 floats (ox, oy, oz, dx, dy, dz)
void ref_intersect(void* h, const float* rays, uint64_t n, int surf, float max_footprint,
                   ref_result* out)
{
    Volume* v = static_cast<Volume*>(h);
    SubDAGArray s = findSubDAGs(Internals::getNodes(*v), Internals::getRootNodeIndex(*v));
    for (uint64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        fill(out[i], intersectVolume(*v, s, r[0], r[1], r[2], r[3], r[4], r[5], surf != 0, max_footprint));
    }
}

void ref_trace_ref(void* h, const float* rays, uint64_t n, ref_result* out)
{
    Volume* v = static_cast<Volume*>(h);
    for (uint64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        fill(out[i], traceRayRef(*v, r[0], r[1], r[2], r[3], r[4], r[5]));
    }
}

// All-core CPU timing of the reference raycast: `threads` std::threads over static interleaved
// chunks of `chunk` rays, each calling the single-ray intersectVolume. Returns seconds of the
// best of `reps` runs (after one warm-up) and the hit count of the last run. SubDAGs are
// computed outside the timed region.
double ref_time_intersect(void* h, const float* rays, uint64_t n, int surf, float max_footprint,
                          int threads, uint64_t chunk, int reps, uint64_t* hit_count)
{
    Volume* v = static_cast<Volume*>(h);
    const SubDAGArray s = findSubDAGs(Internals::getNodes(*v), Internals::getRootNodeIndex(*v));
    if (threads < 1) threads = 1;
    if (chunk < 1) chunk = 4096;
    std::vector<uint64_t> hits(threads);
    double best = 1e30;
    for (int rep = -1; rep < reps; rep++) {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++) {
            pool.emplace_back([&, t]() {
                uint64_t count = 0;
                for (uint64_t c0 = uint64_t(t) * chunk; c0 < n; c0 += uint64_t(threads) * chunk) {
                    const uint64_t c1 = (c0 + chunk < n) ? c0 + chunk : n;
                    for (uint64_t i = c0; i < c1; i++) {
                        const float* r = rays + 6 * i;
                        RayVolumeIntersection x = intersectVolume(*v, s, r[0], r[1], r[2], r[3], r[4], r[5],
                                                                  surf != 0, max_footprint);
                        count += x.hit ? 1 : 0;
                    }
                }
                hits[t] = count;
            });
        }
        for (auto& th : pool) th.join();
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rep >= 0 && dt < best) best = dt;
    }
    uint64_t total = 0;
    for (uint64_t c : hits) total += c;
    if (hit_count) *hit_count = total;
    return best;
}

uint64_t ref_bit_mix(uint64_t bits) { return Internals::bit_mix(bits); }

uint64_t ref_fnv1a(const void* data, int64_t length, uint64_t seed) { return Internals::fnv1a(data, length, seed); }

} // extern "C"
