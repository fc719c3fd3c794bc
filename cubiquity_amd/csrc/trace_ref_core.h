// Device core of traceRayRef (src/library/raytracing.cpp:480-535 IntersectionFinder /
// traceRayRef, utility.h:60-116 visitVolumeNodes / visitChildNodes, geometry.h:460-486 ray-box
// intersection) for one ray.
//
// Every node whose integer box, dilated by 0.5, the ray pierces (FP64 slab test with the entry
// and exit stored as float, kept when exit >= entry) is visited depth first, children in id order
// 0..7; among the non-empty material nodes with exit > 0 the first one with the smallest entry is
// the hit. The walk keeps a stack of (node, next child) per height (slot h at [h * stride]) and
// the current node's bounds in registers: descending into child c flips bit (h - 1) of the lower
// (c's axis bit set) or upper bound of each axis, and ascending flips it back.
#pragma once

#include "cqx.h"

#include <hip/hip_runtime.h>
#include <cfloat>
#include <climits>
#include <stdint.h>

namespace cqxd {

constexpr uint64_t kTraceRefMaxVisits = 1ull << 28; // a valid (acyclic, <= 32-level) DAG ends far earlier

struct Finder {
    double o[3], inv[3];
    bool hit = false;
    double best = DBL_MAX;
    uint32_t material = 0;

    // IntersectionFinder::operator(): returns whether the node's children are to be visited.
    __device__ __forceinline__ bool visit(uint32_t node, const int32_t (&lo)[3], const int32_t (&hi)[3])
    {
        double mn[3], mx[3];
        for (int a = 0; a < 3; a++) {
            const double l = ((double(lo[a]) - 0.5) - o[a]) * inv[a];
            const double u = ((double(hi[a]) + 0.5) - o[a]) * inv[a];
            mn[a] = (u < l) ? u : l; // std::min(lower, upper)
            mx[a] = (l < u) ? u : l; // std::max(lower, upper)
        }
        double e = mn[0];
        e = (e < mn[1]) ? mn[1] : e;
        e = (e < mn[2]) ? mn[2] : e;
        double x = mx[0];
        x = (mx[1] < x) ? mx[1] : x;
        x = (mx[2] < x) ? mx[2] : x;
        const float entry = float(e), exit = float(x);
        if (!(exit >= entry)) return false;
        if (node < 256u && node > 0 && exit > 0.0f && double(entry) < best) {
            hit = true;
            best = double(entry);
            material = node;
        }
        return true;
    }
};

// Child c of a node at height h: collapse the faces selected by ~c (utility.h:98-104). The same
// call undoes it.
__device__ __forceinline__ void flip_child(int32_t (&lo)[3], int32_t (&hi)[3], uint32_t c, uint32_t flip)
{
    for (int a = 0; a < 3; a++) {
        if ((c >> a) & 1u) lo[a] = int32_t(uint32_t(lo[a]) ^ flip);
        else hi[a] = int32_t(uint32_t(hi[a]) ^ flip);
    }
}

// traceRayRef for one ray; stk_node / stk_next hold 33 slots each (heights 0..32) at [h * stride].
__device__ __forceinline__ cqx_intersection trace_ref_ray(const uint32_t* __restrict__ nodes, uint32_t root,
                                                          const cqx_ray& r, uint32_t* stk_node, uint8_t* stk_next,
                                                          uint32_t stride)
{
    Finder f;
    for (int a = 0; a < 3; a++) {
        f.o[a] = double(r.o[a]);
        f.inv[a] = 1.0 / double(r.d[a]);
    }
    int32_t lo[3] = { INT_MIN, INT_MIN, INT_MIN }, hi[3] = { INT_MAX, INT_MAX, INT_MAX };
    bool anomaly = false;
    if (f.visit(root, lo, hi) && root >= 256u) {
        uint32_t h = 32, node = root, c = 0;
        uint64_t visits = 0;
        for (;;) {
            if (c == 8) {
                if (h == 32) break; // the root's children are done
                h++;
                node = stk_node[h * stride];
                c = stk_next[h * stride];
                flip_child(lo, hi, c - 1, 1u << (h - 1)); // back to the parent's bounds
                continue;
            }
            if (++visits > kTraceRefMaxVisits) {
                anomaly = true;
                break;
            }
            const uint32_t child = nodes[uint64_t(node) * 8 + c];
            const uint32_t flip = 1u << (h - 1);
            flip_child(lo, hi, c, flip);
            if (f.visit(child, lo, hi) && child >= 256u && h > 1) {
                stk_node[h * stride] = node;
                stk_next[h * stride] = uint8_t(c + 1);
                node = child;
                h--;
                c = 0;
            } else {
                flip_child(lo, hi, c, flip);
                c++;
            }
        }
    }
    cqx_intersection res;
    res.hit = (f.hit && !anomaly) ? 1 : 0;
    res.material = anomaly ? 0u : f.material;
    res.distance = (f.hit && !anomaly) ? f.best : DBL_MAX;
    for (int a = 0; a < 3; a++) res.position[a] = res.normal[a] = 0.0f;
    return res;
}

} // namespace cqxd
