// Uses the C++ drop-in API (include/cubiquity_hip.hpp) the way a Cubiquity caller would:
//   dropin_check <volume.dag> <rays.f32> <out.bin> <surface 0|1> <max_footprint>
//     load a .dag, compute the SubDAGs, upload the raw node bytes, raycast a batch
//   dropin_check edit <volume.dag> <edits.f32> <rays.f32> <hits.bin> <ref.bin> <out.dag>
//     an edit session: fillSphere, two checkpointed batches of spheres (each uploaded as a
//     delta), undo of the second (a root-only update); then intersectVolume and traceRayRef on
//     the GPU copy, and the edited store saved for comparison
// rays.f32: n x 6 floats; edits.f32: n x 5 floats (centre, radius, material); results are
// cqx_intersection records (40 bytes each).
#include "cubiquity_hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static std::vector<float> read_floats(const char* path)
{
    std::vector<float> v;
    FILE* f = std::fopen(path, "rb");
    if (!f) return v;
    float x;
    while (std::fread(&x, sizeof(x), 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}

static bool write_records(const char* path, const std::vector<cqx_intersection>& r)
{
    FILE* o = std::fopen(path, "wb");
    if (!o) return false;
    const bool ok = std::fwrite(r.data(), sizeof(cqx_intersection), r.size(), o) == r.size();
    std::fclose(o);
    return ok;
}

static int edit_session(char** argv)
{
    cqx_hip::HostVolume hv(argv[2]);
    const std::vector<float> e = read_floats(argv[3]);
    const std::vector<float> rf = read_floats(argv[4]);
    std::vector<cqx_sphere_edit> edits(e.size() / 5);
    for (size_t i = 0; i < edits.size(); i++)
        edits[i] = { { e[5 * i], e[5 * i + 1], e[5 * i + 2] }, e[5 * i + 3], uint32_t(e[5 * i + 4]) };
    std::vector<cqx_ray> rays(rf.size() / 6);
    for (size_t i = 0; i < rays.size(); i++)
        rays[i] = { { rf[6 * i], rf[6 * i + 1], rf[6 * i + 2] }, { rf[6 * i + 3], rf[6 * i + 4], rf[6 * i + 5] } };
    if (edits.size() < 3) return 2;

    cqx_hip::Context ctx(0);
    cqx_hip::GpuVolume gpu(ctx, hv);
    const cqx_sphere_edit& e0 = edits[0];
    gpu.update(hv, hv.fillSphere(e0.centre[0], e0.centre[1], e0.centre[2], e0.radius, e0.material));
    const size_t half = 1 + (edits.size() - 1) / 2;
    hv.checkpoint();
    gpu.update(hv, hv.fillSpheres(edits.data() + 1, half - 1));
    hv.checkpoint();
    gpu.update(hv, hv.fillSpheres(edits.data() + half, edits.size() - half));
    hv.undo();
    gpu.update(hv, hv.nodeCount()); // root and SubDAGs only

    if (!write_records(argv[5], cqx_hip::intersectVolume(gpu, rays, true))) return 4;
    if (!write_records(argv[6], cqx_hip::traceRayRef(gpu, rays))) return 4;
    hv.save(argv[7]);
    std::printf("%zu edits, %zu rays, %llu nodes\n", edits.size(), rays.size(), (unsigned long long)hv.nodeCount());
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 8 && std::string(argv[1]) == "edit") {
        try {
            return edit_session(argv);
        } catch (const cqx_hip::Error& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    }
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s volume.dag rays.f32 out.bin surface max_footprint\n", argv[0]);
        return 2;
    }
    try {
        auto [nodes, root] = cqx_hip::loadDag(argv[1]);
        const cqx_hip::SubDAGArray sub = cqx_hip::findSubDAGs(nodes.data(), nodes.size() / 8, root);

        FILE* f = std::fopen(argv[2], "rb");
        if (!f) return 3;
        std::vector<cqx_ray> rays;
        cqx_ray r;
        while (std::fread(&r, sizeof(r), 1, f) == 1) rays.push_back(r);
        std::fclose(f);

        cqx_hip::Context ctx(0);
        cqx_hip::GpuVolume vol(ctx, nodes.data(), nodes.size() * sizeof(uint32_t), root, sub.data());
        const std::vector<cqx_intersection> res =
            cqx_hip::intersectVolume(vol, rays, std::atoi(argv[4]) != 0, float(std::atof(argv[5])));

        FILE* o = std::fopen(argv[3], "wb");
        if (!o) return 3;
        const bool ok = std::fwrite(res.data(), sizeof(cqx_intersection), res.size(), o) == res.size();
        std::fclose(o);
        std::printf("%zu rays\n", rays.size());
        return ok ? 0 : 4;
    } catch (const cqx_hip::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
