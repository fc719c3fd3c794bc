// Test program (not part of the library): the batched intersectVolume overload in Cubiquity's own
// types (include/cubiquity_hip_reference.hpp), compiled against the reference library's headers
// and linked with the compiled, unmodified reference library (oracle/_ref) and libcqx.so.
//
// For a .dag and a ray file (n x 6 float32) it loads the volume with the reference's Volume,
// computes its SubDAGs with the reference's findSubDAGs, traces every ray with the reference's
// per-ray intersectVolume (src/library/raytracing.cpp:407-478) and with the batched GPU overload,
// and compares every field bit for bit. Exit status 0 when all rays agree.
//
// Usage: reference_dropin <volume.dag> <rays.f32> <computeSurfaceProperties 0|1> <maxFootprint>
#include "cubiquity_hip_reference.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <span>
#include <vector>

// The reference result layout a caller holds (raytracing.h:48-55): the overload fills it by field.
static_assert(sizeof(Cubiquity::RayVolumeIntersection) == 48, "RayVolumeIntersection is 48 bytes");
static_assert(offsetof(Cubiquity::RayVolumeIntersection, hit) == 0, "hit at 0");
static_assert(offsetof(Cubiquity::RayVolumeIntersection, distance) == 8, "distance at 8");
static_assert(offsetof(Cubiquity::RayVolumeIntersection, material) == 16, "material at 16");
static_assert(offsetof(Cubiquity::RayVolumeIntersection, position) == 20, "position at 20");
static_assert(offsetof(Cubiquity::RayVolumeIntersection, normal) == 32, "normal at 32");

static bool same_bits(const void* a, const void* b, std::size_t n) { return std::memcmp(a, b, n) == 0; }

static bool same(const Cubiquity::RayVolumeIntersection& a, const Cubiquity::RayVolumeIntersection& b)
{
    const float pa[6] = { a.position.x, a.position.y, a.position.z, a.normal.x, a.normal.y, a.normal.z };
    const float pb[6] = { b.position.x, b.position.y, b.position.z, b.normal.x, b.normal.y, b.normal.z };
    return a.hit == b.hit && same_bits(&a.distance, &b.distance, sizeof(double)) && a.material == b.material &&
           same_bits(pa, pb, sizeof(pa));
}

int main(int argc, char** argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <volume.dag> <rays.f32> <surf 0|1> <maxFootprint>\n", argv[0]);
        return 2;
    }
    Cubiquity::Volume volume(argv[1]);
    std::ifstream f(argv[2], std::ios::binary | std::ios::ate);
    if (!f) {
        std::fprintf(stderr, "cannot read %s\n", argv[2]);
        return 2;
    }
    const std::size_t bytes = std::size_t(f.tellg());
    f.seekg(0);
    std::vector<float> raw(bytes / sizeof(float));
    f.read(reinterpret_cast<char*>(raw.data()), std::streamsize(raw.size() * sizeof(float)));
    const std::size_t n = raw.size() / 6;
    std::vector<Cubiquity::Ray3f> rays(n);
    for (std::size_t i = 0; i < n; i++)
        rays[i] = Cubiquity::Ray3f(Cubiquity::vec3(raw[6 * i], raw[6 * i + 1], raw[6 * i + 2]),
                                   Cubiquity::vec3(raw[6 * i + 3], raw[6 * i + 4], raw[6 * i + 5]));
    const bool surf = std::atoi(argv[3]) != 0;
    const float fp = float(std::atof(argv[4]));

    const Cubiquity::SubDAGArray sub =
        Cubiquity::findSubDAGs(Cubiquity::Internals::getNodes(volume), Cubiquity::Internals::getRootNodeIndex(volume));
    std::vector<Cubiquity::RayVolumeIntersection> gpu(n);
    Cubiquity::intersectVolume(volume, sub, std::span<const Cubiquity::Ray3f>(rays), surf, fp,
                               std::span<Cubiquity::RayVolumeIntersection>(gpu));
    std::size_t differ = 0, hits = 0;
    for (std::size_t i = 0; i < n; i++) {
        const Cubiquity::RayVolumeIntersection ref = Cubiquity::intersectVolume(volume, sub, rays[i], surf, fp);
        hits += ref.hit ? 1 : 0;
        if (!same(ref, gpu[i])) {
            if (differ < 5)
                std::fprintf(stderr, "ray %zu: reference hit %d t %.9g m %u, GPU hit %d t %.9g m %u\n", i, int(ref.hit),
                             ref.distance, ref.material, int(gpu[i].hit), gpu[i].distance, gpu[i].material);
            differ++;
        }
    }
    std::printf("reference_dropin: %zu rays, %zu hits, %zu differ\n", n, hits, differ);
    return differ == 0 ? 0 : 1;
}
