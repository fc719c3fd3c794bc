// Test infrastructure: the device traversal core (cubiquity_amd/csrc/traverse_core.h) compiled
// for the host, driven the way the kernels drive it, so that its per-lane logic can be checked
// against the oracle on the CPU before any GPU run:
//   lane        trace_ray per ray (k_intersect_lane, the path-tracer kernels)
//   persistent  64-lane waves with the persistent kernel's service-pass policy: walk steps for
//               walking lanes only; every `regroup` steps, once `refill` lanes wait (or none
//               walks), waiting lanes move on to their next SubDAG, finished lanes emit their
//               record, idle lanes take new rays (k_intersect_persistent)
//   top         persistent, with the children of internal nodes [256, 256 + 512) read from a copy
//               (the LDS node cache of variant 4)
//   pair        persistent with two rays per lane (variant 5): both slots' child reads are taken
//               before either slot steps on its read
//   shortK      persistent with a short stack of K rows per lane (K a power of two): rows reused
//               modulo K, pops past them restart from the SubDAG root
//   traceref    traceRayRef per ray (trace_ref_core.h, the cqx_trace_ref kernel's walk); writes
//               40-byte RayVolumeIntersection records
//   census      persistent, and also prints a census of the walk loop on stdout: lane slots of
//               walking and waiting lanes, service passes, and which path each walk step took
//               (descend, advance, advance + pop, leave the SubDAG, hit or stop), per lane and
//               per wave step (a wave step runs every path one of its lanes takes)
// Usage: core_emulation lane|persistent|top|pair|traceref nodes.bin sub.bin rays.bin out.bin surf
//                       max_footprint max_iters(traceref: root) refill regroup
#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__
#include "traverse_core.h"
#include "trace_ref_core.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string_view>
#include <vector>

using namespace cqxd;

static std::vector<char> slurp(const char* path)
{
    std::vector<char> v;
    FILE* f = std::fopen(path, "rb");
    if (!f) return v;
    std::fseek(f, 0, SEEK_END);
    v.resize(size_t(std::ftell(f)));
    std::fseek(f, 0, SEEK_SET);
    if (std::fread(v.data(), 1, v.size(), f) != v.size()) v.clear();
    std::fclose(f);
    return v;
}

int main(int argc, char** argv)
{
    if (argc != 11) {
        std::fprintf(stderr, "usage: %s lane|persistent nodes sub rays out surf fp max_iters refill regroup\n", argv[0]);
        return 2;
    }
    const bool use_top = std::string_view(argv[1]) == "top";
    const bool pair = std::string_view(argv[1]) == "pair";
    const bool short_stack = std::string_view(argv[1]).substr(0, 5) == "short";
    const uint32_t smask = short_stack ? uint32_t(std::atoi(argv[1] + 5)) - 1u : 0u;
    const bool census = std::string_view(argv[1]) == "census";
    const bool persistent = use_top || short_stack || census || std::string_view(argv[1]) == "persistent";
    const std::vector<char> nb = slurp(argv[2]), sb = slurp(argv[3]), rb = slurp(argv[4]);
    if (nb.empty() || sb.size() != 8 * sizeof(cqx_subdag)) return 3;
    TraceParams P;
    P.nodes = reinterpret_cast<const uint32_t*>(nb.data());
    P.sub = reinterpret_cast<const cqx_subdag*>(sb.data());
    P.surf = uint32_t(std::atoi(argv[6]));
    P.max_footprint = float(std::atof(argv[7]));
    P.max_iters = uint32_t(std::strtoul(argv[8], nullptr, 10));
    P.fp_disabled = footprint_is_disabled(P.max_footprint) ? 1u : 0u;
    P.offset32 = nb.size() / 32 <= (uint64_t(1) << 27) ? 1u : 0u;
    const int refill = std::atoi(argv[9]), regroup = std::atoi(argv[10]);
    const float* rays = reinterpret_cast<const float*>(rb.data());
    const uint64_t n = rb.size() / 24;
    std::vector<cqx_hit> out(n);
    std::vector<uint32_t> stack(64 * 33);
    const uint64_t internal = nb.size() / 32 - kMaterialCount;
    const uint32_t top_nodes = uint32_t(internal < 512 ? internal : 512);
    const std::vector<uint32_t> top(P.nodes + kMaterialCount * 8, P.nodes + (kMaterialCount + top_nodes) * 8);

    if (std::string_view(argv[1]) == "traceref") {
        std::vector<cqx_intersection> res(n);
        std::vector<uint32_t> sn(33);
        std::vector<uint8_t> sx(33);
        const uint32_t root = uint32_t(std::strtoul(argv[8], nullptr, 10)); // the max_iters slot carries the root
        for (uint64_t i = 0; i < n; i++) {
            cqx_ray r;
            std::memcpy(&r, rays + 6 * i, sizeof(r));
            res[i] = trace_ref_ray(P.nodes, root, r, sn.data(), sx.data(), 1);
        }
        FILE* f = std::fopen(argv[5], "wb");
        if (!f) return 4;
        const size_t w = std::fwrite(res.data(), sizeof(cqx_intersection), n, f);
        std::fclose(f);
        return w == n ? 0 : 5;
    }
    if (pair) {
        // 64 lanes x 2 slots; slot j of lane l uses stack column 2 * l + j
        std::vector<uint32_t> stack2(128 * 33);
        RayState S[128];
        uint64_t idx[128] = {};
        for (RayState& s : S) s.phase = kIdle;
        uint64_t next = 0;
        for (;;) {
            int walking = 0;
            for (const RayState& s : S) walking += s.phase == kWalk;
            if (128 - walking >= refill || walking == 0) {
                for (int q = 0; q < 128; q++) {
                    enter_octants(P, S[q], &stack2[q], 128);
                    if (is_finished(S[q].phase)) {
                        emit_record(P, S[q], out[idx[q]]);
                        S[q].phase = kIdle;
                    }
                }
                // slot 0 of every lane first, then slot 1 (the kernel's rank order)
                for (int j = 0; j < 2; j++)
                    for (int l = 0; l < 64; l++) {
                        RayState& s = S[2 * l + j];
                        if (s.phase == kIdle && next < n) {
                            const float* r = rays + 6 * next;
                            idx[2 * l + j] = next++;
                            ray_setup(s, r[0], r[1], r[2], r[3], r[4], r[5]);
                            enter_octants(P, s, &stack2[2 * l + j], 128);
                        }
                    }
                int live = 0;
                for (const RayState& s : S) live += s.phase != kIdle;
                if (next >= n && live == 0) break;
            }
            for (int st = 0; st < regroup; st++) {
                uint32_t child[128];
                bool w[128];
                for (int q = 0; q < 128; q++) {
                    w[q] = S[q].phase == kWalk;
                    child[q] = w[q] ? walk_fetch(P, S[q]) : 0u;
                }
                for (int q = 0; q < 128; q++)
                    if (w[q]) walk_step_on(P, S[q], &stack2[q], 128, child[q]);
            }
        }
    } else if (!persistent) {
        for (uint64_t i = 0; i < n; i++) {
            const float* r = rays + 6 * i;
            trace_ray(P, r[0], r[1], r[2], r[3], r[4], r[5], stack.data(), 1, out[i]);
        }
    } else {
        RayState S[64];
        uint64_t idx[64] = {};
        for (RayState& s : S) s.phase = kIdle;
        uint64_t next = 0;
        const uint32_t occ = occupied_octants(P.sub); // as the persistent kernel passes empty octants over
        // census: lane slots, service passes, steps per path (per lane; per wave step with the path)
        enum { kD, kA, kP, kL, kH, kPaths };
        uint64_t slots_walk = 0, slots_wait = 0, services = 0, wave_steps = 0, lane_paths[kPaths] = {},
                 wave_paths[kPaths] = {};
        for (;;) {
            int walking = 0;
            for (const RayState& s : S) walking += s.phase == kWalk;
            if (64 - walking >= refill || walking == 0) {
                services++;
                for (int l = 0; l < 64; l++) {
                    if (short_stack) enter_octants<false, true>(P, S[l], &stack[l], 64, smask, occ);
                    else enter_octants(P, S[l], &stack[l], 64, 0, occ);
                }
                for (int l = 0; l < 64; l++)
                    if (is_finished(S[l].phase)) {
                        emit_record(P, S[l], out[idx[l]]);
                        S[l].phase = kIdle;
                    }
                for (int l = 0; l < 64; l++)
                    if (S[l].phase == kIdle && next < n) {
                        const float* r = rays + 6 * next;
                        idx[l] = next++;
                        ray_setup(S[l], r[0], r[1], r[2], r[3], r[4], r[5]);
                        if (short_stack) enter_octants<false, true>(P, S[l], &stack[l], 64, smask, occ);
                        else enter_octants(P, S[l], &stack[l], 64, 0, occ);
                    }
                int live = 0;
                for (const RayState& s : S) live += s.phase != kIdle;
                if (next >= n && live == 0) break;
            }
            for (int s = 0; s < regroup; s++) {
                bool any[kPaths] = {};
                int walked = 0;
                for (int l = 0; l < 64; l++)
                    if (S[l].phase == kWalk) {
                        const int32_t h0 = S[l].height;
                        const uint32_t e0 = S[l].entries;
                        if (use_top) walk_step<false, true>(P, S[l], &stack[l], 64, top.data(), top_nodes);
                        else if (short_stack) walk_step<false, false, true>(P, S[l], &stack[l], 64, nullptr, 0, smask);
                        else walk_step(P, S[l], &stack[l], 64);
                        const int path = S[l].phase == kLeave ? kL
                                         : S[l].phase != kWalk ? kH
                                         : S[l].height < h0    ? kD
                                         : S[l].entries != e0  ? kP
                                                               : kA;
                        lane_paths[path]++;
                        any[path] = true;
                        walked++;
                    }
                slots_walk += uint64_t(walked);
                slots_wait += uint64_t(64 - walked);
                if (walked) {
                    wave_steps++;
                    for (int k = 0; k < kPaths; k++) wave_paths[k] += any[k] ? 1u : 0u;
                }
            }
        }
        if (census)
            std::printf("census rays %llu services %llu wave_steps %llu slots_walk %llu slots_wait %llu "
                        "lane_D %llu lane_A %llu lane_P %llu lane_L %llu lane_H %llu "
                        "wave_D %llu wave_A %llu wave_P %llu wave_L %llu wave_H %llu\n",
                        (unsigned long long)n, (unsigned long long)services, (unsigned long long)wave_steps,
                        (unsigned long long)slots_walk, (unsigned long long)slots_wait,
                        (unsigned long long)lane_paths[kD], (unsigned long long)lane_paths[kA],
                        (unsigned long long)lane_paths[kP], (unsigned long long)lane_paths[kL],
                        (unsigned long long)lane_paths[kH], (unsigned long long)wave_paths[kD],
                        (unsigned long long)wave_paths[kA], (unsigned long long)wave_paths[kP],
                        (unsigned long long)wave_paths[kL], (unsigned long long)wave_paths[kH]);
    }
    FILE* f = std::fopen(argv[5], "wb");
    if (!f) return 4;
    const size_t w = std::fwrite(out.data(), sizeof(cqx_hit), n, f);
    std::fclose(f);
    return w == n ? 0 : 5;
}
