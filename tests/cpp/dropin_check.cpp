// Uses the C++ drop-in API (include/cubiquity_hip.hpp) the way a Cubiquity caller would:
// load a .dag, compute the SubDAGs, upload the raw node bytes, raycast a batch.
//   dropin_check <volume.dag> <rays.f32> <out.bin> <surface 0|1> <max_footprint>
// rays.f32: n x 6 floats; out.bin: n cqx_intersection records (40 bytes each).
#include "cubiquity_hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv)
{
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
