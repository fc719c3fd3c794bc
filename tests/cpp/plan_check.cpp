// Test infrastructure: the edit plan (cqx::plan_edits, cubiquity_amd/csrc/edits.cpp) on random
// sphere batches, against its definition: every region a sphere's bounds touch (one voxel of
// margin) is listed with that sphere, in edit order; the regions are in ascending key order; the
// level-0 entries are the regions in Morton order by pairwise comparison (cqx::morton_less, the
// rule the interleaved codes must reproduce, also for keys either side of a power-of-two boundary
// of the region grid); each level's entries are the distinct parents of the level below, in order.
// Prints "plan ok" and exits 0, or prints the first mismatch and exits 1.
#include "cqx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

static int fail(const char* what, int batch)
{
    std::printf("batch %d: %s\n", batch, what);
    return 1;
}

int main()
{
    std::mt19937_64 rng(12345);
    for (int batch = 0; batch < 40; batch++) {
        // centres around the origin (the region grid's keys cross 2^(31 - hs) there), in a far
        // corner, or spread over a large cube; radii up to 32 (region height 6) or 4 (height 5)
        const double span = batch % 3 == 0 ? 200.0 : batch % 3 == 1 ? 5000.0 : 2.0e6;
        const double offset = batch % 4 == 3 ? 3.0e6 : 0.0;
        const float rmax = batch % 2 ? 32.0f : 4.0f;
        std::uniform_real_distribution<double> u(-span, span);
        std::uniform_real_distribution<float> ur(0.5f, rmax);
        std::vector<cqx_sphere_edit> e(200 + batch * 50);
        for (auto& s : e) {
            for (int a = 0; a < 3; a++) s.centre[a] = float(offset + u(rng));
            s.radius = ur(rng);
            s.material = uint32_t(rng() % 256);
        }
        const int hs = cqx::region_height(e.data(), e.size());
        if (hs < 5) return fail("no region height", batch);
        cqx::EditPlan plan;
        cqx::plan_edits(e.data(), e.size(), hs, plan);
        const size_t nr = plan.keys.size();
        if (plan.first.size() != nr + 1 || plan.first.back() != plan.edit_ids.size()) return fail("offsets", batch);
        for (size_t r = 1; r < nr; r++)
            if (!(plan.keys[r - 1] < plan.keys[r])) return fail("keys not ascending", batch);
        // the touched regions of every edit, from the definition
        std::vector<std::pair<cqx::RegionKey, uint32_t>> want;
        for (uint32_t i = 0; i < e.size(); i++) {
            uint32_t lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                const float blo = e[i].centre[a] - e[i].radius, bhi = e[i].centre[a] + e[i].radius;
                lo[a] = uint32_t(int64_t(std::floor(double(blo))) - 1 + (int64_t(1) << 31)) >> hs;
                hi[a] = uint32_t(int64_t(std::ceil(double(bhi))) + 1 + (int64_t(1) << 31)) >> hs;
            }
            for (uint32_t z = lo[2]; z <= hi[2]; z++)
                for (uint32_t y = lo[1]; y <= hi[1]; y++)
                    for (uint32_t x = lo[0]; x <= hi[0]; x++) want.push_back({ cqx::RegionKey{ x, y, z }, i });
        }
        std::stable_sort(want.begin(), want.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        if (want.size() != plan.edit_ids.size()) return fail("touch count", batch);
        for (size_t r = 0, j = 0; r < nr; r++)
            for (uint64_t k = plan.first[r]; k < plan.first[r + 1]; k++, j++)
                if (!(want[j].first == plan.keys[r]) || want[j].second != plan.edit_ids[k]) return fail("touch list", batch);
        // Morton order of the regions
        if (plan.morton_regions.size() != nr) return fail("morton size", batch);
        std::vector<uint32_t> ref(nr);
        for (uint32_t r = 0; r < nr; r++) ref[r] = r;
        std::sort(ref.begin(), ref.end(), [&](uint32_t a, uint32_t b) { return cqx::morton_less(plan.keys[a], plan.keys[b]); });
        if (ref != plan.morton_regions) return fail("morton order", batch);
        // levels
        const int levels = 32 - hs;
        if (int(plan.level_keys.size()) != levels + 1) return fail("level count", batch);
        for (int l = 1; l <= levels; l++) {
            const auto& below = plan.level_keys[size_t(l) - 1];
            const auto& up = plan.level_keys[size_t(l)];
            const auto& ch = plan.level_children[size_t(l)];
            if (ch.size() != up.size() + 1 || ch.back() != below.size()) return fail("children offsets", batch);
            for (size_t i = 0; i < up.size(); i++)
                for (uint32_t j = ch[i]; j < ch[i + 1]; j++) {
                    const cqx::RegionKey p{ below[j].x >> 1, below[j].y >> 1, below[j].z >> 1 };
                    if (!(p == up[i])) return fail("parent", batch);
                }
            for (size_t i = 1; i < up.size(); i++)
                if (up[i - 1] == up[i]) return fail("duplicate parent", batch);
        }
        if (plan.level_keys[size_t(levels)].size() != 1) return fail("root level", batch);
    }
    std::printf("plan ok\n");
    return 0;
}
