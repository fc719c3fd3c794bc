// cubiquity_hip_reference.hpp: the batched intersectVolume overload in Cubiquity's own types, for
// code compiled inside the Cubiquity tree (it includes the library's raytracing.h and storage.h,
// which are not part of this repository). It adds, next to the reference's per-ray functions
// (src/library/raytracing.h:72-76),
//
//   void Cubiquity::intersectVolume(const Volume&, const SubDAGArray&, std::span<const Ray3f> rays,
//                                   bool computeSurfaceProperties, float maxFootprint,
//                                   std::span<RayVolumeIntersection> out);
//
// with out[i] bit-identical to intersectVolume(volume, subDAGs, rays[i], computeSurfaceProperties,
// maxFootprint) for every i, computed on the GPU through the C ABI (cqx.h). This form uploads the
// volume's node store (NodeStore's raw bytes, storage.h:101-102) on every call, because the
// reference store changes its unshared nodes in place and a device copy cannot follow them; for
// many calls on one version of a volume, upload once (cqx_hip::GpuVolume) and use the second
// overload. Needs C++20 (std::span), as the reference does.
#pragma once

#include "cubiquity_hip.hpp"

#include "raytracing.h" // Cubiquity: RayVolumeIntersection, SubDAG(Array), intersectVolume
#include "storage.h"    // Cubiquity: Volume, Internals::getNodes / getRootNodeIndex, NodeStore

#include <cstddef>
#include <span>
#include <stdexcept>
#include <vector>

namespace Cubiquity {

// The SubDAG array is handed to the C ABI as is: the layouts must agree (raytracing.h:57-65).
static_assert(sizeof(SubDAG) == sizeof(cqx_subdag), "SubDAG is 32 bytes");
static_assert(offsetof(SubDAG, lowerBound) == offsetof(cqx_subdag, lower), "SubDAG::lowerBound");
static_assert(offsetof(SubDAG, nodeHeight) == offsetof(cqx_subdag, height), "SubDAG::nodeHeight");
static_assert(offsetof(SubDAG, nodeIndex) == offsetof(cqx_subdag, node), "SubDAG::nodeIndex");
static_assert(sizeof(SubDAGArray) == 8 * sizeof(cqx_subdag), "SubDAGArray is 8 SubDAGs");
static_assert(sizeof(Internals::Node) == 32, "a node is 8 x u32");

namespace hip_detail {

inline cqx_hip::Context& context()
{
    static cqx_hip::Context ctx(0); // GPU 0, for the life of the process
    return ctx;
}

inline void to_reference(const cqx_intersection& h, RayVolumeIntersection& r)
{
    r.hit = h.hit != 0;
    r.distance = h.distance;
    r.material = h.material;
    r.position = vec3(h.position[0], h.position[1], h.position[2]);
    r.normal = vec3(h.normal[0], h.normal[1], h.normal[2]);
}

} // namespace hip_detail

// Batched raytracing.h:76 on a store already resident in HBM.
inline void intersectVolume(const cqx_hip::GpuVolume& gpu, std::span<const Ray3f> rays, bool computeSurfaceProperties,
                            float maxFootprint, std::span<RayVolumeIntersection> out)
{
    if (out.size() < rays.size()) throw std::invalid_argument("intersectVolume: output span shorter than the rays");
    std::vector<cqx_ray> r(rays.size());
    for (std::size_t i = 0; i < rays.size(); i++)
        r[i] = { { rays[i].mOrigin.x, rays[i].mOrigin.y, rays[i].mOrigin.z },
                 { rays[i].mDir.x, rays[i].mDir.y, rays[i].mDir.z } };
    const std::vector<cqx_intersection> h = cqx_hip::intersectVolume(gpu, r, computeSurfaceProperties, maxFootprint);
    for (std::size_t i = 0; i < rays.size(); i++) hip_detail::to_reference(h[i], out[i]);
}

// Batched raytracing.h:76: the volume's store and these SubDAGs uploaded, then every ray traced.
inline void intersectVolume(const Volume& volume, const SubDAGArray& subDAGs, std::span<const Ray3f> rays,
                            bool computeSurfaceProperties, float maxFootprint, std::span<RayVolumeIntersection> out)
{
    const Internals::NodeStore& store = Internals::getNodes(volume);
    // rawBytesCount() is a u32 and wraps past 2^27 nodes; the node count does not
    const uint64_t bytes = uint64_t(store.unsharedNodesEnd()) * sizeof(Internals::Node);
    cqx_hip::GpuVolume gpu(hip_detail::context(), store.rawBytesPtr(), bytes, Internals::getRootNodeIndex(volume),
                           reinterpret_cast<const cqx_subdag*>(subDAGs.data()));
    intersectVolume(gpu, rays, computeSurfaceProperties, maxFootprint, out);
}

} // namespace Cubiquity
