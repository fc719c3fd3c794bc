// Test program (not part of the library): the host library's multi-threaded paths built with a
// sanitizer (ThreadSanitizer, or AddressSanitizer + UndefinedBehaviorSanitizer) from the host
// sources alone (cubiquity_amd/csrc/*.cpp; no HIP code is involved):
//   * the parallel generator / DAG builder (cqx_hv_generate, 8 threads),
//   * region-parallel sphere edits (cqx_hv_edit_spheres_mt, 8 threads): the regions' private
//     node lists, the lock-free compare-and-swap interner over the existing store and the
//     parallel level-by-level merge, over two frames (the second reuses the interner's index),
//   * checkpoint / undo / redo and compaction.
// The stores of the 8-thread run and of the in-order run (1 thread) must be byte-identical.
// Exit status 0 on success; the sanitizer makes the process fail on a report.
#include "cqx.h"

#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

static int check(cqx_status s, const char* what)
{
    if (s != CQX_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, s, cqx_last_error());
        return 1;
    }
    return 0;
}

static std::vector<cqx_sphere_edit> frame(int f, int n, float side)
{
    std::mt19937 rng(7000 + f);
    std::uniform_real_distribution<float> pos(0.0f, side), rad(2.0f, 9.0f);
    std::uniform_int_distribution<int> mat(0, 8);
    std::vector<cqx_sphere_edit> e(n);
    for (auto& s : e) {
        s.centre[0] = pos(rng);
        s.centre[1] = pos(rng);
        s.centre[2] = pos(rng);
        s.radius = rad(rng);
        s.material = uint32_t(mat(rng));
    }
    return e;
}

static int run(int threads, std::vector<uint32_t>& store, uint32_t& root)
{
    cqx_gen_params p;
    if (check(cqx_gen_default_params(9, 3, &p), "cqx_gen_default_params")) return 1;
    p.n_materials = 8;
    cqx_host_volume* hv = nullptr;
    if (check(cqx_hv_generate(&p, threads, &hv), "cqx_hv_generate")) return 1;
    uint64_t first = 0;
    for (int f = 0; f < 2; f++) {
        const std::vector<cqx_sphere_edit> e = frame(f, 1500, 512.0f);
        if (check(cqx_hv_checkpoint(hv), "cqx_hv_checkpoint")) return 1;
        if (check(cqx_hv_edit_spheres_mt(hv, e.data(), e.size(), threads, &first), "cqx_hv_edit_spheres_mt")) return 1;
    }
    if (check(cqx_hv_undo(hv), "cqx_hv_undo") || check(cqx_hv_redo(hv), "cqx_hv_redo")) return 1;
    if (check(cqx_hv_compact(hv), "cqx_hv_compact")) return 1;
    const uint32_t* n = cqx_hv_nodes(hv);
    store.assign(n, n + cqx_hv_node_count(hv) * 8);
    root = cqx_hv_root(hv);
    cqx_hv_free(hv);
    return 0;
}

int main()
{
    std::vector<uint32_t> a, b;
    uint32_t ra = 0, rb = 0;
    if (run(8, a, ra) || run(1, b, rb)) return 1;
    const bool same = ra == rb && a == b;
    std::printf("host_race_check: %zu nodes, root %u, 8 threads %s 1 thread\n", a.size() / 8, ra,
                same ? "==" : "!=");
    return same ? 0 : 1;
}
