// cubiquity_hip.hpp: header-only C++ host API over cqx.h, shaped like Cubiquity's raycast API
// so that a reference caller can switch its batched raycasts to the MI355X with few changes.
//
//   Cubiquity (reference)                              this header
//   findSubDAGs(nodes, root)        raytracing.h:69    cqx_hip::findSubDAGs(nodes, count, root)
//   intersectVolume(volume, subDAGs, ox..dz, surf, fp) cqx_hip::intersectVolume(gpuVolume, rays,
//                                   raytracing.h:72-76     surf, fp) -> one record per ray
//   traceRayRef(volume, ox..dz)     raytracing.h:78-81 cqx_hip::traceRayRef(gpuVolume, rays)
//   RayVolumeIntersection           raytracing.h:48-55 cqx_intersection (same fields)
//   NodeStore::rawBytesPtr/Count    storage.h:101-102  GpuVolume(ctx, rawBytes, byteCount, root)
//   Volume (load, save, voxel,      storage.h:135-178  cqx_hip::HostVolume
//     checkpoint, undo, redo, bake)
//   fillBrush(volume, SphereBrush(c, r), material)     HostVolume::fillSphere / fillSpheres
//                                   voxelization.h:91-126, voxelization.cpp:825-915
//   Metadata::load (materials)      metadata.cpp:34-65 cqx_hip::loadMaterials
//   PathtracingDemo::traceSingleRay pathtracing_demo.cpp:150-194  cqx_hip::demoTrace
//
// Errors become cqx_hip::Error exceptions on this side of the ABI (the C ABI itself never throws).
#pragma once

#include "cqx.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cqx_hip {

class Error : public std::runtime_error {
public:
    Error(cqx_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    cqx_status status;
};

inline void check(cqx_status s)
{
    if (s != CQX_OK) throw Error(s, std::string("cqx error ") + std::to_string(s) + ": " + cqx_last_error());
}

using SubDAGArray = std::vector<cqx_subdag>; // 8 entries, == the reference's std::array<SubDAG, 8>

// findSubDAGs (raytracing.cpp:89-97) over a raw store of node_count x 8 uint32.
inline SubDAGArray findSubDAGs(const uint32_t* nodes, uint64_t node_count, uint32_t root)
{
    SubDAGArray s(8);
    check(cqx_find_subdags(nodes, node_count, root, s.data()));
    return s;
}

// One GPU (device index) and its work counters.
class Context {
public:
    explicit Context(int device = 0) { check(cqx_create(device, &ctx_)); }
    ~Context() { cqx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cqx_ctx* get() const { return ctx_; }

private:
    cqx_ctx* ctx_ = nullptr;
};

// The host-side volume: a node store and root with the reference Volume's operations.
class HostVolume {
public:
    // Volume(filename) / Volume::load, storage.cpp:505-528
    explicit HostVolume(const std::string& dag_path) { check(cqx_hv_load_dag(dag_path.c_str(), &hv_)); }
    ~HostVolume() { cqx_hv_free(hv_); }
    HostVolume(const HostVolume&) = delete;
    HostVolume& operator=(const HostVolume&) = delete;

    void save(const std::string& dag_path) const { check(cqx_hv_save_dag(hv_, dag_path.c_str())); }
    uint32_t voxel(int32_t x, int32_t y, int32_t z) const
    {
        uint32_t m = 0;
        check(cqx_hv_voxel(hv_, x, y, z, &m));
        return m;
    }
    // fillBrush(volume, SphereBrush(centre, radius), material); returns the first appended node
    uint64_t fillSphere(float cx, float cy, float cz, float radius, uint32_t material)
    {
        const cqx_sphere_edit e{ { cx, cy, cz }, radius, material };
        return fillSpheres(&e, 1);
    }
    // Many spheres, applied in order (region-parallel on the host's cores)
    uint64_t fillSpheres(const cqx_sphere_edit* edits, uint64_t n, int32_t threads = 0)
    {
        uint64_t first = 0;
        check(cqx_hv_edit_spheres_mt(hv_, edits, n, threads, &first));
        return first;
    }
    void checkpoint() { check(cqx_hv_checkpoint(hv_)); }
    void undo() { check(cqx_hv_undo(hv_)); }
    void redo() { check(cqx_hv_redo(hv_)); }
    // Volume::bake's storage effect: unreachable nodes dropped, nodes renumbered
    void bake() { check(cqx_hv_compact(hv_)); }

    const uint32_t* nodes() const { return cqx_hv_nodes(hv_); }
    uint64_t nodeCount() const { return cqx_hv_node_count(hv_); }
    uint32_t root() const { return cqx_hv_root(hv_); }
    cqx_host_volume* get() const { return hv_; }

private:
    cqx_host_volume* hv_ = nullptr;
};

// A node store resident in HBM. `raw_bytes` is the whole store including the 256 material nodes,
// exactly what NodeStore::rawBytesPtr()/rawBytesCount() return (storage.h:101-102).
class GpuVolume {
public:
    GpuVolume(Context& ctx, const void* raw_bytes, uint64_t byte_count, uint32_t root, const cqx_subdag* sub = nullptr)
        : ctx_(&ctx)
    {
        if (byte_count % 32 != 0) throw Error(CQX_ERR_INVALID_ARG, "node store size is not a multiple of 32 bytes");
        check(cqx_volume_upload(ctx.get(), static_cast<const uint32_t*>(raw_bytes), byte_count / 32, root, sub, &vol_));
    }
    GpuVolume(Context& ctx, const HostVolume& hv) : GpuVolume(ctx, hv.nodes(), hv.nodeCount() * 32, hv.root()) {}
    // After edits (first = the first appended node) or undo/redo (first = nodeCount()): copies the
    // appended nodes and takes the new root and SubDAGs.
    void update(const HostVolume& hv, uint64_t first)
    {
        check(cqx_volume_update(ctx_->get(), vol_, hv.nodes(), hv.nodeCount(), first, hv.root(), nullptr));
    }
    ~GpuVolume() { cqx_volume_free(ctx_->get(), vol_); }
    GpuVolume(const GpuVolume&) = delete;
    GpuVolume& operator=(const GpuVolume&) = delete;
    cqx_dvolume* get() const { return vol_; }
    Context& context() const { return *ctx_; }

private:
    Context* ctx_;
    cqx_dvolume* vol_ = nullptr;
};

// Batched intersectVolume: the reference's per-ray semantics (raytracing.cpp:397-478) for every
// ray, results in the reference's RayVolumeIntersection layout. Host buffers in and out.
inline std::vector<cqx_intersection> intersectVolume(const GpuVolume& vol, const std::vector<cqx_ray>& rays,
                                                     bool computeSurfaceProperties,
                                                     float maxFootprint = CQX_MAX_FOOTPRINT_DISABLED)
{
    const uint32_t flags = computeSurfaceProperties ? CQX_SURFACE_PROPERTIES : 0u;
    std::vector<cqx_hit> hits(rays.size());
    check(cqx_intersect_host(vol.context().get(), vol.get(), rays.data(), rays.size(), flags, maxFootprint,
                             hits.data()));
    std::vector<cqx_intersection> out(rays.size());
    check(cqx_hits_to_intersections(rays.data(), hits.data(), rays.size(), flags, out.data()));
    return out;
}

// The same on device buffers (rays and hits already in HBM), asynchronous on `stream`.
inline void intersectVolumeDevice(const GpuVolume& vol, const cqx_ray* d_rays, uint64_t n, bool computeSurfaceProperties,
                                  float maxFootprint, cqx_hit* d_hits, void* stream = nullptr)
{
    check(cqx_intersect(vol.context().get(), vol.get(), d_rays, n, computeSurfaceProperties ? CQX_SURFACE_PROPERTIES : 0u,
                        maxFootprint, d_hits, nullptr, stream));
}

// traceRayRef (raytracing.cpp:480-535) for a batch, on the GPU: the reference's brute-force raycast.
inline std::vector<cqx_intersection> traceRayRef(const GpuVolume& vol, const std::vector<cqx_ray>& rays)
{
    std::vector<cqx_intersection> out(rays.size());
    check(cqx_trace_ref_host(vol.context().get(), vol.get(), rays.data(), rays.size(), out.data()));
    return out;
}

// PathtracingDemo::traceSingleRay (pathtracing_demo.cpp:150-194) for one primary ray per pixel,
// on device buffers: the reference CPU tracer's shading with per-pixel generator states.
inline void demoTrace(const GpuVolume& vol, const cqx_ray* d_rays, uint64_t n, const cqx_demo_params& params,
                      float* d_rgb, void* stream = nullptr)
{
    check(cqx_demo_trace(vol.context().get(), vol.get(), d_rays, n, &params, nullptr, d_rgb, stream));
}

// Materials of a volume's .toml metadata (Metadata::load, metadata.cpp:34-65), indexed by id.
inline std::vector<cqx_material> loadMaterials(const std::string& toml_path)
{
    std::vector<cqx_material> m(256);
    uint32_t count = 0;
    check(cqx_materials_load(toml_path.c_str(), m.data(), &count, nullptr, nullptr));
    m.resize(count);
    return m;
}

// Reads a .dag file (Volume::load, storage.cpp:505-528) into a raw store + root.
inline std::pair<std::vector<uint32_t>, uint32_t> loadDag(const std::string& path)
{
    cqx_host_volume* hv = nullptr;
    check(cqx_hv_load_dag(path.c_str(), &hv));
    const uint32_t* n = cqx_hv_nodes(hv);
    std::vector<uint32_t> nodes(n, n + cqx_hv_node_count(hv) * 8);
    const uint32_t root = cqx_hv_root(hv);
    cqx_hv_free(hv);
    return { std::move(nodes), root };
}

} // namespace cqx_hip
