// Reference-shading mode: the reference CPU path tracer's per-pixel shading,
// PathtracingDemo::traceSingleRay (src/application/commands/view/pathtracing_demo.cpp:150-194),
// for a batch of primary rays on the gfx950 traversal.
//
// Per pixel p with primary ray r (all FP32 unless noted, operations in the reference's order):
//   i0 = intersectVolume(r, surface properties, maxFootprint)                         :153-156
//   miss: colour (0.8, 0.8, 1.0)                                                      :151
//   hit:  surf0 = colours[m0] x noise(pos0)                         surfaceColour, :44-59
//         direct0 = gather(pos0, n0)                                gatherLighting, :82-117
//         reflected ray from pos0 + n0 x 0.01 along normalize(n0 + randomPointInUnitSphere())
//         i1 = intersectVolume(reflected, surface properties, maxFootprint)
//         indirect0 = i1 hit ? surf1 x gather(pos1, n1) : 0
//         colour = pow(surf0 x (direct0 + indirect0), 1/2.2)                           :187-190
//   gather(pos, n): shadow rays (no surface properties) from pos + n x 0.001:
//         sun: along normalize(1, -2, 10); if unoccluded += 0.1 x max(dot(sunDir, n), 0)
//         sky: along normalize(n + randomPointInUnitSphere()); if unoccluded += 1.5
//   noise(pos) = float(0.9 + 0.1 x ((fnv1a(int(pos + 0.499)) & 0xff) / 255)) (the mapping in
//         double: the reference's literals are doubles)               positionBasedNoise, :34-42
//   randomPointInUnitSphere: s = uint32(bit_mix(s)); the three 10-bit fields of s mapped by
//         (v - 511.5) / 511.5 (tabulated per block), repeated while the squared length is >= 1
//                                                                                      :61-80
// The reference draws from ONE generator state shared by all pixels in scan order, so a pixel's
// draws depend on how many draws every earlier pixel made (which depends on their hits). Here each
// pixel has its own state, uint32(bit_mix(seed x 2^32 + p + 1)) (17, the reference's initial
// state, if that is 0); the draw function, and the draw order within a pixel (sky at i0,
// reflection, sky at i1), are the reference's. A draw gives up after 1024 generator steps with
// (0, 0, 0): the reference loops forever on a state cycle that never lands inside the sphere.
// Rays the traversal flags as anomalies (the reference would not return) count as misses.
//
// Wavefront: trace primaries -> k_demo_primary (the sun and sky shadow rays and the reflected ray
// of each hit, in one list compacted with one atomic per list per 1024-thread block) -> trace
// that list -> k_demo_bounce (first visibilities; second hits and their shadow rays) -> trace ->
// k_demo_visibility -> k_demo_resolve. The first shadow rays and the reflected rays are
// independent, so one traversal launch serves both. Ray counts stay on the device; the traversal
// kernels read them.
// Visibility is kept per (pixel, shadow ray) and summed in the reference's order at the end.
#include "device_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

using namespace cqxd;
using cqx::fail;
using cqx::release;

namespace {

constexpr int kMaxDraws = 1024;
// shadow-ray kinds: a shadow ray's tag is pixel x 4 + kind, also its visibility byte. In the
// first-hit list (where second shadow rays never occur) kind kRefl marks a reflected ray.
constexpr uint32_t kSun0 = 0, kSky0 = 1, kSun1 = 2, kSky1 = 3;
constexpr uint32_t kRefl = 2;

struct DemoDev {
    const float* colours; // 256 x 3, device
    float sun[3];
    float gamma;
    uint32_t add_noise, include_sun, include_sky, accumulate;
    uint64_t seed;
};

// Internals::bit_mix, src/library/base.cpp:72-77 (the 'xmxmx' mixer)
__device__ __forceinline__ uint64_t bit_mix(uint64_t b)
{
    b = ((b >> 32) ^ b) * 0x0e9846af9b1a615dull;
    b = ((b >> 32) ^ b) * 0x0e9846af9b1a615dull;
    return (b >> 28) ^ b;
}

// Internals::fnv1a (src/library/base.cpp:62-69) over the 12 bytes of an ivec3
__device__ __forceinline__ uint64_t fnv1a_ivec3(const int32_t (&q)[3])
{
    uint64_t h = 0xcbf29ce484222325ull;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const uint32_t w = uint32_t(q[a]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            h ^= (w >> (8 * k)) & 0xffu;
            h *= 0x00000100000001B3ull;
        }
    }
    return h;
}

__device__ __forceinline__ float position_noise(const float (&pos)[3])
{
    int32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; a++) q[a] = int32_t(pos[a] + 0.499f);
    const uint32_t h = uint32_t(fnv1a_ivec3(q));
    const float noise = float(h & 0xffu) / 255.0f;
    return float((double(noise) * 0.1) + 0.9);
}

// A coordinate of a random point is (f - 511.5) / 511.5 for a 10-bit field f: 1024 possible FP32
// values. Each block tabulates them and their FP32 squares in LDS once, so a draw's rejection test
// is three lookups and two adds, with the same values the arithmetic gives.
struct PointTable {
    float v[1024];
    float sq[1024];
};

// Called by every thread of the block (it synchronises the block).
__device__ __forceinline__ void fill_point_table(PointTable& t)
{
    for (uint32_t i = threadIdx.x; i < 1024; i += blockDim.x) {
        const float v = (float(i) - 511.5f) / 511.5f;
        t.v[i] = v;
        t.sq[i] = v * v;
    }
    __syncthreads();
}

__device__ __forceinline__ void random_point(uint32_t& s, float (&r)[3], const PointTable& t)
{
    for (int k = 0; k < kMaxDraws; k++) {
        s = uint32_t(bit_mix(s));
        const uint32_t f0 = s & 0x3FFu, f1 = (s >> 10) & 0x3FFu, f2 = (s >> 20) & 0x3FFu;
        float d = 0.0f;
        d = d + t.sq[f0];
        d = d + t.sq[f1];
        d = d + t.sq[f2];
        if (!(d >= 1.0f)) {
            r[0] = t.v[f0];
            r[1] = t.v[f1];
            r[2] = t.v[f2];
            return;
        }
    }
    r[0] = r[1] = r[2] = 0.0f;
}

// linalg normalize: a / sqrt(dot(a, a)), dot a left fold from 0
__device__ __forceinline__ void normalize3(const float (&v)[3], float (&out)[3])
{
    float d = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) d = d + v[a] * v[a];
    const float len = sqrtf(d);
#pragma unroll
    for (int a = 0; a < 3; a++) out[a] = v[a] / len;
}

__device__ __forceinline__ uint32_t pixel_state(uint64_t seed, uint64_t p)
{
    const uint32_t s = uint32_t(bit_mix((seed << 32) + p + 1u));
    return s ? s : 17u;
}

// Face normal (components -1, -0, +0, +1) from a hit record's normal code.
__device__ __forceinline__ void face_normal(uint32_t info, float (&n)[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const uint32_t code = (info >> (9 + 2 * a)) & 3u;
        const float mag = (code & 1u) ? 1.0f : 0.0f;
        n[a] = (code & 2u) ? -mag : mag;
    }
}

// Hit position o + d t (intersectVolume, raytracing.cpp:464) and surfaceColour.
__device__ __forceinline__ void hit_surface(const DemoDev& d, const cqx_ray& r, const cqx_hit& h, float (&pos)[3],
                                            float (&nrm)[3], float (&col)[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) pos[a] = r.o[a] + r.d[a] * h.t;
    face_normal(h.info, nrm);
    const uint32_t m = (h.info >> 16) & 0xffu;
#pragma unroll
    for (int a = 0; a < 3; a++) col[a] = d.colours[m * 3 + a];
    if (d.add_noise) {
        const float noise = position_noise(pos);
#pragma unroll
        for (int a = 0; a < 3; a++) col[a] = col[a] * noise;
    }
}

// The shadow rays of gatherLighting at (pos, n); the sky direction takes one draw from s.
__device__ __forceinline__ void shadow_rays(const DemoDev& d, const float (&pos)[3], const float (&n)[3], uint32_t& s,
                                            const PointTable& pts, cqx_ray& sun, cqx_ray& sky)
{
    float so[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float offset = n[a] * 0.001f;
        so[a] = pos[a] + offset;
        sun.o[a] = so[a];
        sun.d[a] = d.sun[a];
        sky.o[a] = so[a];
    }
    if (d.include_sky) {
        float pt[3], v[3], dir[3];
        random_point(s, pt, pts);
#pragma unroll
        for (int a = 0; a < 3; a++) v[a] = n[a] + pt[a];
        normalize3(v, dir);
#pragma unroll
        for (int a = 0; a < 3; a++) sky.d[a] = dir[a];
    }
}

__global__ void k_demo_zero(uint32_t* __restrict__ counts)
{
    if (threadIdx.x < 4) counts[threadIdx.x] = 0u;
}

// First hits: surface colour, the two shadow rays and the reflected ray; clears the pixel's
// visibility bytes.
__global__ void __launch_bounds__(cqx::kCompactBlock) k_demo_primary(DemoDev d, uint64_t n, const cqx_ray* __restrict__ rays,
                                                      const cqx_hit* __restrict__ hits, uint32_t* __restrict__ state,
                                                      uint32_t* __restrict__ info0, float* __restrict__ surf0,
                                                      uint32_t* __restrict__ vis, cqx_ray* __restrict__ sec_rays,
                                                      uint32_t* __restrict__ sec_tags, uint32_t* __restrict__ sec_count)
{
    __shared__ PointTable pts;
    fill_point_table(pts);
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool hit = false;
    cqx_ray sun, sky, refl;
    if (p < n) {
        const cqx_hit h = hits[p];
        info0[p] = h.info;
        vis[p] = 0u;
        hit = (h.info & CQX_HIT_FLAG) != 0;
        if (hit) {
            float pos[3], nrm[3], col[3];
            hit_surface(d, rays[p], h, pos, nrm, col);
#pragma unroll
            for (int a = 0; a < 3; a++) surf0[p * 3 + a] = col[a];
            uint32_t s = pixel_state(d.seed, p);
            shadow_rays(d, pos, nrm, s, pts, sun, sky);
            float pt[3], v[3], dir[3];
            random_point(s, pt, pts);
#pragma unroll
            for (int a = 0; a < 3; a++) v[a] = nrm[a] + pt[a];
            normalize3(v, dir);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                refl.o[a] = pos[a] + nrm[a] * 0.01f;
                refl.d[a] = dir[a];
            }
            state[p] = s;
        }
    }
    const bool take[3] = { hit && d.include_sun, hit && d.include_sky, hit };
    uint32_t* const counts[3] = { sec_count, sec_count, sec_count };
    uint32_t k[3];
    cqx::block_slots<cqx::kCompactBlock>(take, counts, k);
    const cqx_ray r[3] = { sun, sky, refl };
    const uint32_t kind[3] = { kSun0, kSky0, kRefl };
#pragma unroll
    for (int l = 0; l < 3; l++)
        if (take[l]) {
            sec_rays[k[l]] = r[l];
            sec_tags[k[l]] = uint32_t(p) * 4u + kind[l];
        }
}

// A shadow ray that hits nothing lets its light through.
__global__ void __launch_bounds__(256) k_demo_visibility(const uint32_t* __restrict__ count, uint64_t cap,
                                                         const uint32_t* __restrict__ tags,
                                                         const cqx_hit* __restrict__ hits, uint8_t* __restrict__ vis)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t c = *count;
    if (i >= c || i >= cap) return;
    vis[tags[i]] = (hits[i].info & CQX_HIT_FLAG) ? 0u : 1u;
}

// The traced first-hit list: a shadow ray's visibility; for a reflected ray, the second hit's
// surface colour and the second gatherLighting's shadow rays.
__global__ void __launch_bounds__(cqx::kCompactBlock) k_demo_bounce(DemoDev d, uint64_t cap, const uint32_t* __restrict__ sec_count,
                                                     const cqx_ray* __restrict__ sec_rays,
                                                     const uint32_t* __restrict__ sec_tags,
                                                     const cqx_hit* __restrict__ sec_hits, const uint32_t* __restrict__ state,
                                                     uint8_t* __restrict__ vis, uint32_t* __restrict__ info1,
                                                     float* __restrict__ surf1, cqx_ray* __restrict__ sh_rays,
                                                     uint32_t* __restrict__ sh_tags, uint32_t* __restrict__ sh_count)
{
    __shared__ PointTable pts;
    fill_point_table(pts);
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t c = *sec_count;
    bool hit = false;
    uint32_t p = 0;
    cqx_ray sun, sky;
    if (i < c && i < cap) {
        const uint32_t tag = sec_tags[i];
        const cqx_hit h = sec_hits[i];
        p = tag >> 2;
        if ((tag & 3u) != kRefl) {
            vis[tag] = (h.info & CQX_HIT_FLAG) ? 0u : 1u; // a shadow ray that hits nothing lets its light through
        } else {
            info1[p] = h.info;
            hit = (h.info & CQX_HIT_FLAG) != 0;
            if (hit) {
                float pos[3], nrm[3], col[3];
                hit_surface(d, sec_rays[i], h, pos, nrm, col);
#pragma unroll
                for (int a = 0; a < 3; a++) surf1[uint64_t(p) * 3 + a] = col[a];
                uint32_t s = state[p];
                shadow_rays(d, pos, nrm, s, pts, sun, sky);
            }
        }
    }
    const bool take[2] = { hit && d.include_sun, hit && d.include_sky };
    uint32_t* const counts[2] = { sh_count, sh_count };
    uint32_t k[2];
    cqx::block_slots<cqx::kCompactBlock>(take, counts, k);
    if (take[0]) {
        sh_rays[k[0]] = sun;
        sh_tags[k[0]] = p * 4u + kSun1;
    }
    if (take[1]) {
        sh_rays[k[1]] = sky;
        sh_tags[k[1]] = p * 4u + kSky1;
    }
}

// gatherLighting's sum from the two visibilities, in the reference's order
__device__ __forceinline__ void gather(const DemoDev& d, const float (&n)[3], bool sun_vis, bool sky_vis,
                                       float (&out)[3])
{
    out[0] = out[1] = out[2] = 0.0f;
    if (d.include_sun && sun_vis) {
        float dt = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; a++) dt = dt + d.sun[a] * n[a];
        const float m = (dt < 0.0f) ? 0.0f : dt;
#pragma unroll
        for (int a = 0; a < 3; a++) out[a] = out[a] + 0.1f * m;
    }
    if (d.include_sky && sky_vis) {
#pragma unroll
        for (int a = 0; a < 3; a++) out[a] = out[a] + 1.5f;
    }
}

__global__ void __launch_bounds__(256) k_demo_resolve(DemoDev d, uint64_t n, const uint32_t* __restrict__ info0,
                                                      const uint32_t* __restrict__ info1,
                                                      const float* __restrict__ surf0, const float* __restrict__ surf1,
                                                      const uint8_t* __restrict__ vis, float* __restrict__ out)
{
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float px[3] = { 0.8f, 0.8f, 1.0f };
    const uint32_t i0 = info0[p];
    if (i0 & CQX_HIT_FLAG) {
        float n0[3], direct0[3], indirect0[3] = { 0.0f, 0.0f, 0.0f };
        face_normal(i0, n0);
        gather(d, n0, vis[p * 4 + kSun0] != 0, vis[p * 4 + kSky0] != 0, direct0);
        const uint32_t i1 = info1[p];
        if (i1 & CQX_HIT_FLAG) {
            float n1[3], direct1[3];
            face_normal(i1, n1);
            gather(d, n1, vis[p * 4 + kSun1] != 0, vis[p * 4 + kSky1] != 0, direct1);
#pragma unroll
            for (int a = 0; a < 3; a++) indirect0[a] = surf1[p * 3 + a] * direct1[a];
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float v = surf0[p * 3 + a] * (direct0[a] + indirect0[a]);
            px[a] = d.gamma == 1.0f ? v : powf(v, d.gamma); // pow(v, 1) = v
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) out[p * 3 + a] = d.accumulate ? out[p * 3 + a] + px[a] : px[a];
}

} // namespace

namespace cqx {

// Per-context buffers of the mode, carved out of one allocation, for up to `cap` pixels.
struct DemoPlan {
    uint64_t cap = 0;
    void* mem = nullptr;
    float* colours = nullptr;
    uint32_t* counts = nullptr;
    cqx_hit* hits0 = nullptr;
    uint32_t* state = nullptr;
    uint32_t* info0 = nullptr;
    uint32_t* info1 = nullptr;
    float* surf0 = nullptr;
    float* surf1 = nullptr;
    uint32_t* vis = nullptr; // 4 bytes per pixel
    cqx_ray* sec_rays = nullptr; // first-hit list: sun and sky shadow rays, reflected rays (3 per pixel)
    uint32_t* sec_tags = nullptr;
    cqx_hit* sec_hits = nullptr;
    cqx_ray* sh_rays = nullptr; // second-hit shadow rays (2 per pixel)
    uint32_t* sh_tags = nullptr;
    cqx_hit* sh_hits = nullptr;
};

void destroy_demo_plan(DemoPlan* p)
{
    if (!p) return;
    release(p->mem);
    delete p;
}

} // namespace cqx

namespace {

using cqx::DemoPlan;

cqx_status demo_plan(cqx_ctx* ctx, uint64_t n, DemoPlan** out)
{
    DemoPlan* p = ctx->demo_plan;
    if (p && p->cap >= n) {
        *out = p;
        return CQX_OK;
    }
    HIP_TRY(hipDeviceSynchronize()); // launches in flight may use the old buffers
    cqx::destroy_demo_plan(p);
    ctx->demo_plan = nullptr;
    p = new (std::nothrow) DemoPlan;
    if (!p) return fail(CQX_ERR_OOM, "out of host memory");
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t o_col = take(256 * 3 * sizeof(float)), o_cnt = take(4 * sizeof(uint32_t));
    const size_t o_h0 = take(n * sizeof(cqx_hit)), o_st = take(n * 4), o_i0 = take(n * 4), o_i1 = take(n * 4);
    const size_t o_s0 = take(n * 12), o_s1 = take(n * 12), o_vis = take(n * 4);
    const size_t o_ser = take(3 * n * sizeof(cqx_ray)), o_set = take(3 * n * 4), o_seh = take(3 * n * sizeof(cqx_hit));
    const size_t o_shr = take(2 * n * sizeof(cqx_ray)), o_sht = take(2 * n * 4), o_shh = take(2 * n * sizeof(cqx_hit));
    if (hipMalloc(&p->mem, off) != hipSuccess) {
        delete p;
        return fail(CQX_ERR_OOM, "device allocation of %zu bytes for %llu pixels failed", off, (unsigned long long)n);
    }
    char* b = static_cast<char*>(p->mem);
    p->cap = n;
    p->colours = reinterpret_cast<float*>(b + o_col);
    p->counts = reinterpret_cast<uint32_t*>(b + o_cnt);
    p->hits0 = reinterpret_cast<cqx_hit*>(b + o_h0);
    p->state = reinterpret_cast<uint32_t*>(b + o_st);
    p->info0 = reinterpret_cast<uint32_t*>(b + o_i0);
    p->info1 = reinterpret_cast<uint32_t*>(b + o_i1);
    p->surf0 = reinterpret_cast<float*>(b + o_s0);
    p->surf1 = reinterpret_cast<float*>(b + o_s1);
    p->vis = reinterpret_cast<uint32_t*>(b + o_vis);
    p->sec_rays = reinterpret_cast<cqx_ray*>(b + o_ser);
    p->sec_tags = reinterpret_cast<uint32_t*>(b + o_set);
    p->sec_hits = reinterpret_cast<cqx_hit*>(b + o_seh);
    p->sh_rays = reinterpret_cast<cqx_ray*>(b + o_shr);
    p->sh_tags = reinterpret_cast<uint32_t*>(b + o_sht);
    p->sh_hits = reinterpret_cast<cqx_hit*>(b + o_shh);
    ctx->demo_plan = p;
    *out = p;
    return CQX_OK;
}

} // namespace

extern "C" {

cqx_status cqx_demo_trace(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                          const cqx_demo_params* params, const cqx_launch_cfg* cfg, float* out_rgb, void* stream)
{
    const cqx::TraceRange range("cqx_demo_trace");
    if (!ctx || !vol || !params) return fail(CQX_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (n == 0) return CQX_OK;
    if (!rays || !out_rgb || !params->colours) return fail(CQX_ERR_INVALID_ARG, "null ray, colour or output buffer");
    if (n > (uint64_t(1) << 30)) return fail(CQX_ERR_LIMIT, "%llu pixels exceed the 2^30 limit", (unsigned long long)n);
    hipStream_t s = hipStream_t(stream);
    DemoPlan* p = nullptr;
    cqx_status st = demo_plan(ctx, n, &p);
    if (st != CQX_OK) return st;
    HIP_TRY(hipMemcpyAsync(p->colours, params->colours, 256 * 3 * sizeof(float), hipMemcpyHostToDevice, s));

    DemoDev d;
    d.colours = p->colours;
    {
        // normalize(vec3(1, -2, 10)) as the reference computes it in FP32
        const float v[3] = { 1.0f, -2.0f, 10.0f };
        float sq = 0.0f;
        for (int a = 0; a < 3; a++) sq = sq + v[a] * v[a];
        const float len = std::sqrt(sq);
        for (int a = 0; a < 3; a++) d.sun[a] = v[a] / len;
    }
    d.gamma = params->gamma != 0.0f ? params->gamma : float(1.0 / 2.2);
    d.add_noise = params->add_noise ? 1u : 0u;
    d.include_sun = params->include_sun ? 1u : 0u;
    d.include_sky = params->include_sky ? 1u : 0u;
    d.accumulate = params->accumulate ? 1u : 0u;
    d.seed = params->seed;
    const float fp = params->max_footprint;

    // "auto" (cfg null or variant 0): the primaries (coherent camera rays) one per lane, 64 per
    // workgroup; the secondary lists (incoherent) on a light persistent grid, 16 waves per CU.
    // The frame's batches are too small for the persistent default (32 waves per CU) to pay.
    // Only the fields the caller left at zero take these values.
    cqx_launch_cfg c1{};
    if (cfg) c1 = *cfg;
    cqx_launch_cfg c2 = c1;
    if (c1.variant == 0) {
        c1.variant = 1;
        if (!c1.block) c1.block = 64;
        c2.variant = 2;
        if (!c2.waves_per_cu) c2.waves_per_cu = 16;
    }
    // secondary launches read their ray count on the device: no binning prepass there
    c2.bin_bits = 0;
    const uint32_t g1 = uint32_t((n + 255) / 256), g2 = uint32_t((2 * n + 255) / 256);
    const uint32_t gc = uint32_t((n + cqx::kCompactBlock - 1) / cqx::kCompactBlock);
    const uint32_t gc3 = uint32_t((3 * n + cqx::kCompactBlock - 1) / cqx::kCompactBlock);
    uint32_t* cnt = p->counts; // [0] first-hit list, [1] second shadow rays
    uint8_t* vis = reinterpret_cast<uint8_t*>(p->vis);
    hipLaunchKernelGGL(k_demo_zero, dim3(1), dim3(64), 0, s, cnt);
    st = cqx::launch_intersect(ctx, vol, rays, n, nullptr, CQX_SURFACE_PROPERTIES, fp, p->hits0, &c1, s);
    if (st != CQX_OK) return st;
    hipLaunchKernelGGL(k_demo_primary, dim3(gc), dim3(cqx::kCompactBlock), 0, s, d, n, rays, p->hits0, p->state,
                       p->info0, p->surf0, p->vis, p->sec_rays, p->sec_tags, cnt + 0);
    // Surface properties are for the reflected rays; a shadow ray reads only the hit flag, which
    // does not depend on them.
    st = cqx::launch_intersect(ctx, vol, p->sec_rays, 3 * n, cnt + 0, CQX_SURFACE_PROPERTIES, fp, p->sec_hits, &c2, s);
    if (st != CQX_OK) return st;
    hipLaunchKernelGGL(k_demo_bounce, dim3(gc3), dim3(cqx::kCompactBlock), 0, s, d, 3 * n, cnt + 0, p->sec_rays,
                       p->sec_tags, p->sec_hits, p->state, vis, p->info1, p->surf1, p->sh_rays, p->sh_tags, cnt + 1);
    st = cqx::launch_intersect(ctx, vol, p->sh_rays, 2 * n, cnt + 1, 0u, fp, p->sh_hits, &c2, s);
    if (st != CQX_OK) return st;
    hipLaunchKernelGGL(k_demo_visibility, dim3(g2), dim3(256), 0, s, cnt + 1, 2 * n, p->sh_tags, p->sh_hits, vis);
    hipLaunchKernelGGL(k_demo_resolve, dim3(g1), dim3(256), 0, s, d, n, p->info0, p->info1, p->surf0, p->surf1,
                       reinterpret_cast<const uint8_t*>(p->vis), out_rgb);
    HIP_TRY(hipGetLastError());
    return CQX_OK;
}

} // extern "C"
