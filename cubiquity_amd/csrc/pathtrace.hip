// Path tracer and its path-replay adjoint on the gfx950 traversal.
//
// Shading model (no reference counterpart: the reference's CPU and GLSL tracers use sun + sky
// shadow rays and one bounce, src/application/commands/view/pathtracing_demo.cpp:81-190; the
// conventions kept from them are the hit position o + d t, the bounce-ray origin offset along the
// face normal, pathtracing_demo.cpp:139, and the light-blue background (0.8, 0.8, 1.0), :124):
//   * per path (pixel p, sample s): PCG32 with initstate (seed << 32) + p and stream s;
//   * camera ray through (x + u, y + v) (two draws), pinhole, y down;
//   * at each hit: L += beta * emission[m]; if fewer than `bounces` scatterings so far, sample a
//     cosine-weighted direction about the face normal (two draws), beta *= albedo[m];
//   * a ray that escapes adds beta * sky; a path ends after `bounces` scatterings.
// Lambertian BRDF albedo/pi with cosine-weighted sampling makes the throughput weight albedo.
// sin/cos of 2 pi u use the quadrant reduction + Taylor polynomials below and sqrt is correctly
// rounded, so the same arithmetic on the CPU reproduces every path bit for bit.
//
// Forward, two forms with identical per-path arithmetic:
//   wavefront  generate -> (intersect, shade) x (bounces + 1), captured once in a hipGraph per
//              plan and replayed; ray counts live on the device, the persistent traversal kernel
//              reads them; surviving rays are compacted by one atomic per
//              1024-thread block
//   megakernel one lane per path, grid-stride
// Both write per-path radiance; k_pt_resolve sums the samples of a pixel in sample order.
// Adjoint (path replay): runs the wavefront pass again (each path's PCG stream makes it trace the
// same paths), which records one byte per path vertex (its material) and how the path ended; then
// a sweep kernel rebuilds each path's throughputs in the forward's FP32 order and runs backwards,
//   G_K = sky if escaped else 0;  G_j = E_j + A_j G_{j+1};
//   dV/dE_j = beta_j,  dV/dA_j = beta_j G_{j+1},
// scaled by dL/dpixel / spp; per-block partial sums in LDS, reduced over blocks in FP64.
#include "device_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

using namespace cqxd;
using cqx::fail;
using cqx::release;

namespace {

constexpr int kMaxVertices = 65; // bounces + 1 <= 65 (at most 64 bounces)
// Path record of the wavefront pass: vmat[j * P + p] is
// the material of path p's vertex j, vstat[p] its vertex count K, plus kEscaped when the path
// ended in the sky. Paths that end by an anomaly or a failed scatter have kEscaped clear.
constexpr uint32_t kEscaped = 0x80u;

struct PtDevParams {
    float eye[3];
    float fwd[3], right[3], up[3];
    float kx, ky, half_w, half_h;
    float sky[3];
    float offset;
    float max_footprint;
    uint32_t width, height, spp, bounces;
    uint32_t pad;
    uint64_t seed;
    float albedo[256 * 3];   // by material id
    float emission[256 * 3]; // by material id
};

// ---- PCG32 (O'Neill 2014) ---------------------------------------------------------------------
struct Pcg {
    uint64_t state, inc;
};

__device__ __forceinline__ uint32_t pcg_next(Pcg& r)
{
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}

__device__ __forceinline__ Pcg pcg_seed(uint64_t initstate, uint64_t initseq)
{
    Pcg r;
    r.state = 0;
    r.inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r.state += initstate;
    pcg_next(r);
    return r;
}

// uniform in [0, 1) with 24 random bits (exact conversion and scaling)
__device__ __forceinline__ float pcg_uniform(Pcg& r) { return float(pcg_next(r) >> 8) * (1.0f / 16777216.0f); }

// cos and sin of 2 pi u for u in [0, 1): quadrant q = floor(4u), theta = (4u - q) pi/2 in [0, pi/2),
// Taylor polynomials of degree 11 (sin) and 12 (cos), then the quadrant rotation.
__device__ __forceinline__ void sincos_2pi(float u, float& c, float& s)
{
    const float u4 = u * 4.0f;
    const float q = floorf(u4);
    const float th = (u4 - q) * 1.57079632679489662f;
    const float t2 = th * th;
    const float sp = th * (1.0f + t2 * (-1.66666667e-1f + t2 * (8.33333333e-3f + t2 * (-1.98412698e-4f +
                     t2 * (2.75573192e-6f + t2 * -2.50521084e-8f)))));
    const float cp = 1.0f + t2 * (-0.5f + t2 * (4.16666667e-2f + t2 * (-1.38888889e-3f + t2 * (2.48015873e-5f +
                     t2 * (-2.75573192e-7f + t2 * 2.08767570e-9f)))));
    const int qi = int(q);
    if (qi == 0) { c = cp; s = sp; }
    else if (qi == 1) { c = -sp; s = cp; }
    else if (qi == 2) { c = -cp; s = -sp; }
    else { c = sp; s = -cp; }
}

__device__ __forceinline__ void camera_ray(const PtDevParams& d, uint32_t x, uint32_t y, float jx, float jy,
                                           float* o, float* dir)
{
    const float cx = ((float(x) + jx) - d.half_w) * d.kx;
    const float cy = ((float(y) + jy) - d.half_h) * d.ky;
    float v[3];
    for (int a = 0; a < 3; a++) v[a] = (d.fwd[a] + d.right[a] * cx) - d.up[a] * cy;
    const float len = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int a = 0; a < 3; a++) {
        o[a] = d.eye[a];
        dir[a] = v[a] / len;
    }
}

// Face normal (components -1, -0, +0, +1) from a hit record's normal code.
__device__ __forceinline__ void decode_normal(uint32_t info, float* n)
{
    for (int a = 0; a < 3; a++) {
        const uint32_t code = (info >> (9 + 2 * a)) & 3u;
        const float mag = (code & 1u) ? 1.0f : 0.0f;
        n[a] = (code & 2u) ? -mag : mag;
    }
}

// Cosine-weighted direction about the face normal n and the next ray's origin
// (pos + n * offset). Orthonormal basis after Duff et al., "Building an Orthonormal Basis,
// Revisited", JCGT 6(1), 2017. Returns false for a degenerate (zero) normal.
__device__ __forceinline__ bool scatter(const PtDevParams& d, Pcg& rng, const float* pos, const float* n, float* o,
                                        float* dir)
{
    const float nl = sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    if (!(nl > 0.0f)) return false;
    const float nx = n[0] / nl, ny = n[1] / nl, nz = n[2] / nl;
    const float u1 = pcg_uniform(rng);
    const float u2 = pcg_uniform(rng);
    const float r = sqrtf(u1);
    float c, s;
    sincos_2pi(u2, c, s);
    const float lx = r * c, ly = r * s, lz = sqrtf(1.0f - u1);
    const float sign = copysignf(1.0f, nz);
    const float a = -1.0f / (sign + nz);
    const float b = nx * ny * a;
    const float t0 = 1.0f + sign * nx * nx * a, t1 = sign * b, t2 = -sign * nx;
    const float u0 = b, uu1 = sign + ny * ny * a, u2b = -ny;
    dir[0] = (t0 * lx + u0 * ly) + nx * lz;
    dir[1] = (t1 * lx + uu1 * ly) + ny * lz;
    dir[2] = (t2 * lx + u2b * ly) + nz * lz;
    for (int k = 0; k < 3; k++) o[k] = pos[k] + n[k] * d.offset;
    return true;
}

__device__ __forceinline__ void path_start(const PtDevParams& d, uint64_t p, Pcg& rng, float* o, float* dir,
                                           uint32_t& pixel)
{
    pixel = uint32_t(p / d.spp);
    const uint32_t s = uint32_t(p % d.spp);
    rng = pcg_seed((d.seed << 32) + uint64_t(pixel), uint64_t(s));
    const float jx = pcg_uniform(rng);
    const float jy = pcg_uniform(rng);
    camera_ray(d, pixel % d.width, pixel / d.width, jx, jy, o, dir);
}

// ---- wavefront kernels -----------------------------------------------------------------------

// Camera rays and generator states of all paths. The paths' throughput (1) and radiance (0) are
// not stored: the shading pass at depth 0, which sees every path, starts from those values.
__global__ void k_pt_generate(const PtDevParams* __restrict__ dp, uint64_t P, cqx_ray* __restrict__ rays,
                              uint64_t* __restrict__ rng_state, uint32_t* __restrict__ count)
{
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p == 0) *count = uint32_t(P);
    if (p >= P) return;
    const PtDevParams& d = *dp;
    Pcg rng;
    float o[3], dir[3];
    uint32_t pixel;
    path_start(d, p, rng, o, dir, pixel);
    cqx_ray r;
    for (int a = 0; a < 3; a++) {
        r.o[a] = o[a];
        r.d[a] = dir[a];
    }
    rays[p] = r;
    rng_state[p] = rng.state;
}

__global__ void k_set_u32(uint32_t* __restrict__ x, uint32_t v)
{
    if (threadIdx.x == 0) *x = v;
}

// kRadiance: the render's shading pass (throughput and radiance kept per path). Without it, the
// adjoint's replay: the same paths, recording each vertex's material (vmat, vstat) only.
template <bool kRadiance>
__global__ void __launch_bounds__(cqx::kCompactBlock) k_pt_shade(const PtDevParams* __restrict__ dp, uint64_t P, uint32_t depth,
                                                  const uint32_t* __restrict__ count_cur,
                                                  const cqx_ray* __restrict__ rays_cur,
                                                  const uint32_t* __restrict__ paths_cur,
                                                  const cqx_hit* __restrict__ hits, float* __restrict__ beta,
                                                  float* __restrict__ L, uint64_t* __restrict__ rng_state,
                                                  cqx_ray* __restrict__ rays_next, uint32_t* __restrict__ paths_next,
                                                  uint32_t* __restrict__ count_next, uint8_t* __restrict__ vmat,
                                                  uint8_t* __restrict__ vstat)
{
    const PtDevParams& d = *dp;
    const uint32_t n = *count_cur;
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool cont = false;
    uint32_t p = 0;
    cqx_ray next;
    if (i < n) {
        // depth 0 holds every path in order, with throughput 1 and radiance 0
        p = depth == 0 ? uint32_t(i) : paths_cur[i];
        const cqx_hit h = hits[i];
        float b[3] = { 1.0f, 1.0f, 1.0f }, l[3] = { 0.0f, 0.0f, 0.0f };
        if (kRadiance && depth > 0)
            for (int a = 0; a < 3; a++) {
                b[a] = beta[a * P + p];
                l[a] = L[a * P + p];
            }
        if (h.info & CQX_HIT_ANOMALY) {
            // traversal gave up on this ray: the path ends without a contribution
            if (vstat) vstat[p] = uint8_t(depth);
        } else if (!(h.info & CQX_HIT_FLAG)) {
            for (int a = 0; a < 3; a++) l[a] += b[a] * d.sky[a];
            if (vstat) vstat[p] = uint8_t(depth | kEscaped);
        } else {
            const uint32_t m = (h.info >> 16) & 0xffu;
            if (vstat) {
                vmat[depth * P + p] = uint8_t(m);
                vstat[p] = uint8_t(depth + 1);
            }
            for (int a = 0; a < 3; a++) l[a] += b[a] * d.emission[m * 3 + a];
            if (depth < d.bounces) {
                const cqx_ray r = rays_cur[i];
                float pos[3], nrm[3];
                for (int a = 0; a < 3; a++) pos[a] = r.o[a] + r.d[a] * h.t;
                decode_normal(h.info, nrm);
                Pcg rng;
                rng.state = rng_state[p];
                rng.inc = ((uint64_t(p % d.spp)) << 1u) | 1u;
                if (scatter(d, rng, pos, nrm, next.o, next.d)) {
                    rng_state[p] = rng.state;
                    if (kRadiance)
                        for (int a = 0; a < 3; a++) beta[a * P + p] = b[a] * d.albedo[m * 3 + a];
                    cont = true;
                }
            }
        }
        if (kRadiance)
            for (int a = 0; a < 3; a++) L[a * P + p] = l[a];
    }
    // compact the surviving rays
    const bool take[1] = { cont };
    uint32_t* const counts[1] = { count_next };
    uint32_t slot[1];
    cqx::block_slots<cqx::kCompactBlock>(take, counts, slot);
    if (cont) {
        rays_next[slot[0]] = next;
        paths_next[slot[0]] = p;
    }
}

// Per-pixel mean of the samples, summed in sample order. With `staged`, each block first copies
// its pixels' samples of one channel (contiguous in L) to LDS with coalesced reads, then each
// thread sums its pixel's row there; otherwise threads read L directly (spp too large for LDS).
__global__ void __launch_bounds__(256) k_pt_resolve(const PtDevParams* __restrict__ dp, uint64_t P,
                                                    const float* __restrict__ L, float* __restrict__ image,
                                                    uint32_t staged)
{
    extern __shared__ float tile[]; // blockDim.x x spp floats when staged
    const PtDevParams& d = *dp;
    const uint32_t pixels = d.width * d.height, spp = d.spp;
    const uint32_t q0 = blockIdx.x * blockDim.x; // < pixels: the grid covers the pixels once
    const uint32_t q = q0 + threadIdx.x;
    const uint32_t npix = pixels - q0 < blockDim.x ? pixels - q0 : blockDim.x;
    for (int a = 0; a < 3; a++) {
        float sum = 0.0f;
        if (staged) {
            const float* src = L + a * P + uint64_t(q0) * spp;
            for (uint32_t k = threadIdx.x; k < npix * spp; k += blockDim.x) tile[k] = src[k];
            __syncthreads();
            if (q < pixels)
                for (uint32_t s = 0; s < spp; s++) sum += tile[threadIdx.x * spp + s];
            __syncthreads();
        } else if (q < pixels) {
            for (uint32_t s = 0; s < spp; s++) sum += L[a * P + uint64_t(q) * spp + s];
        }
        if (q < pixels) image[uint64_t(q) * 3 + a] = sum / float(spp);
    }
}

// ---- megakernel forward ------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_pt_megakernel(const PtDevParams* __restrict__ dp, TraceParams T, uint64_t P,
                                                       float* __restrict__ L)
{
    extern __shared__ uint32_t lds_stack[];
    const PtDevParams& d = *dp;
    const uint32_t occ = occupied_octants(T.sub);
    const uint64_t stride_all = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < P; p += stride_all) {
        Pcg rng;
        float o[3], dir[3];
        uint32_t pixel;
        path_start(d, p, rng, o, dir, pixel);
        float b[3] = { 1.0f, 1.0f, 1.0f }, l[3] = { 0.0f, 0.0f, 0.0f };
        for (uint32_t depth = 0;; depth++) {
            cqx_hit h;
            trace_ray(T, o[0], o[1], o[2], dir[0], dir[1], dir[2], lds_stack + threadIdx.x, blockDim.x, h, occ);
            if (h.info & CQX_HIT_ANOMALY) break;
            if (!(h.info & CQX_HIT_FLAG)) {
                for (int a = 0; a < 3; a++) l[a] += b[a] * d.sky[a];
                break;
            }
            const uint32_t m = (h.info >> 16) & 0xffu;
            for (int a = 0; a < 3; a++) l[a] += b[a] * d.emission[m * 3 + a];
            if (depth >= d.bounces) break;
            float pos[3], nrm[3];
            for (int a = 0; a < 3; a++) pos[a] = o[a] + dir[a] * h.t;
            decode_normal(h.info, nrm);
            if (!scatter(d, rng, pos, nrm, o, dir)) break;
            for (int a = 0; a < 3; a++) b[a] = b[a] * d.albedo[m * 3 + a];
        }
        for (int a = 0; a < 3; a++) L[a * P + p] = l[a];
    }
}

// ---- adjoint (path replay) -----------------------------------------------------------------------

// The backward sweep over each path's recorded vertices. The throughputs beta_j are rebuilt as the
// forward pass formed them (beta_0 = 1, beta_{j+1} = beta_j * albedo[m_j], FP32 in that order).
// The sweep takes the vertices in segments of kSegment, last segment first: a segment's betas are
// rebuilt from beta_0 by the same products (so a path of up to 64 bounces needs registers for one
// segment only), then its vertices are swept from last to first.
//
// The gradient sums have a fixed order, so the result is the same bit for bit on every run: each
// wave takes 64 consecutive paths at a time. With at most kAccMaterials materials (kAcc), each
// lane adds its paths' contributions, vertex by vertex in order, into its own registers, one row per
// material, and after its last path the wave sums each row by a fixed butterfly into its row of
// per-material totals in LDS. With more materials, vertex slot by vertex slot, the wave groups its
// lanes by material (the first remaining lane's material, then the next); a group's six
// contributions are summed over the wave by a fixed butterfly (lanes outside the group add 0) and
// lane 0 adds the sums to the wave's row. No atomics are involved; the block adds its waves' rows
// in wave order.
constexpr uint32_t kAdjointBlock = 256;
constexpr int kSegment = 4; // vertices per segment: a path of up to 3 bounces is one
constexpr uint32_t kAdjointWaves = kAdjointBlock / 64;
constexpr int kAccMaterials = 16;

template <bool kAcc>
__global__ void __launch_bounds__(kAdjointBlock) k_pt_adjoint(const PtDevParams* __restrict__ dp, uint64_t P,
                                                              const uint8_t* __restrict__ vmat,
                                                              const uint8_t* __restrict__ vstat,
                                                              const float* __restrict__ dimage, uint32_t M,
                                                              float* __restrict__ partials)
{
    float acc[kAcc ? kAccMaterials : 1][6]; // kAcc: this lane's sums, row m - 1 for material m
#pragma unroll
    for (int mm = 0; mm < (kAcc ? kAccMaterials : 1); mm++)
#pragma unroll
        for (int k = 0; k < 6; k++) acc[mm][k] = 0.0f;
    __shared__ float g[kAdjointWaves][256 * 6]; // per wave: [material][dA rgb, dE rgb]
    for (uint32_t k = threadIdx.x; k < kAdjointWaves * 256 * 6; k += blockDim.x) (&g[0][0])[k] = 0.0f;
    __syncthreads();
    const PtDevParams& d = *dp;
    const float inv_spp = 1.0f / float(d.spp);
    const uint32_t lane = threadIdx.x & 63u;
    float* const gw = g[threadIdx.x >> 6];
    const uint64_t stride_all = uint64_t(gridDim.x) * blockDim.x;
    // `base` is the same for the 64 lanes of a wave, so every lane runs every iteration
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u); base < P; base += stride_all) {
        const uint64_t p = base + lane;
        const bool valid = p < P;
        const uint32_t stat = valid ? vstat[p] : 0u;
        const int K = int(stat & ~kEscaped);
        const bool escaped = (stat & kEscaped) != 0;
        const uint32_t pixel = valid ? uint32_t(p) / d.spp : 0u; // P < 2^32
        float delta[3];
        for (int a = 0; a < 3; a++) delta[a] = valid ? dimage[uint64_t(pixel) * 3 + a] * inv_spp : 0.0f;
        float G[3];
        for (int a = 0; a < 3; a++) G[a] = escaped ? d.sky[a] : 0.0f;
        // the wave's most vertices (a wave-uniform segment count)
        int kw = K;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(kw, off);
            kw = o > kw ? o : kw;
        }
        for (int seg = (kw + kSegment - 1) / kSegment - 1; seg >= 0; seg--) {
            const int j0 = seg * kSegment;
            float b[3] = { 1.0f, 1.0f, 1.0f };
            for (int j = 0; j < j0 && j < K; j++) {
                const uint32_t m = vmat[uint64_t(j) * P + p];
                for (int a = 0; a < 3; a++) b[a] = b[a] * d.albedo[m * 3 + a];
            }
            uint32_t vm[kSegment];
            float vb[kSegment][3];
#pragma unroll
            for (int i = 0; i < kSegment; i++) {
                vm[i] = 0u;
                if (j0 + i < K) {
                    const uint32_t m = vmat[uint64_t(j0 + i) * P + p];
                    vm[i] = m;
                    for (int a = 0; a < 3; a++) {
                        vb[i][a] = b[a];
                        b[a] = b[a] * d.albedo[m * 3 + a];
                    }
                }
            }
#pragma unroll
            for (int i = kSegment - 1; i >= 0; i--) {
                const bool live = j0 + i < K;
                const uint32_t m = vm[i];
                float c[6] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
                if (live) {
                    for (int a = 0; a < 3; a++) {
                        const float w = delta[a] * vb[i][a];
                        c[a] = w * G[a];
                        c[3 + a] = w;
                        G[a] = d.emission[m * 3 + a] + d.albedo[m * 3 + a] * G[a];
                    }
                }
                if constexpr (kAcc) {
                    if (live) {
#pragma unroll
                        for (int mm = 0; mm < kAccMaterials; mm++) {
                            const bool hit = m == uint32_t(mm + 1);
#pragma unroll
                            for (int k = 0; k < 6; k++) acc[mm][k] += hit ? c[k] : 0.0f;
                        }
                    }
                    continue;
                }
                uint64_t todo = __ballot(live);
                while (todo) {
                    const uint32_t m0 = uint32_t(__shfl(int(m), __ffsll((unsigned long long)todo) - 1));
                    const bool in = live && m == m0;
                    todo &= ~__ballot(in);
#pragma unroll
                    for (int k = 0; k < 6; k++) {
                        float v = in ? c[k] : 0.0f;
#pragma unroll
                        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                        if (lane == 0) gw[m0 * 6 + k] += v;
                    }
                }
            }
        }
    }
    if constexpr (kAcc) {
#pragma unroll
        for (int mm = 0; mm < kAccMaterials; mm++)
#pragma unroll
            for (int k = 0; k < 6; k++) {
                float v = acc[mm][k];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == 0) gw[(mm + 1) * 6 + k] = v;
            }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < M * 6; k += blockDim.x) { // material ids 1..M
        float sum = 0.0f;
        for (uint32_t w = 0; w < kAdjointWaves; w++) sum += g[w][6 + k];
        partials[uint64_t(blockIdx.x) * M * 6 + k] = sum;
    }
}

// One workgroup per output value ((material - 1) * 6 + component): strided FP64 sums over the
// block partials, then a fixed-order tree in LDS (deterministic for a given grid).
__global__ void __launch_bounds__(256) k_pt_reduce(const float* __restrict__ partials, uint32_t blocks, uint32_t M,
                                                   float* __restrict__ d_albedo, float* __restrict__ d_emission)
{
    __shared__ double acc[256];
    const uint32_t k = blockIdx.x;
    double sum = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) sum += double(partials[uint64_t(b) * M * 6 + k]);
    acc[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w /= 2) {
        if (threadIdx.x < w) acc[threadIdx.x] += acc[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t m = k / 6, c = k % 6;
        if (c < 3) d_albedo[m * 3 + c] = float(acc[0]);
        else d_emission[m * 3 + (c - 3)] = float(acc[0]);
    }
}

} // namespace

namespace cqx {

struct PtPlan {
    const cqx_dvolume* vol = nullptr;
    uint32_t width = 0, height = 0, spp = 0, bounces = 0;
    float max_footprint = -1.0f;
    cqx_launch_cfg cfg{};  // traversal launch of the bounce rays
    cqx_launch_cfg cfg0{}; // traversal launch of the camera rays (depth 0)
    uint64_t P = 0;
    PtDevParams* dparams = nullptr;
    cqx_ray* rays[2] = { nullptr, nullptr };
    uint32_t* paths[2] = { nullptr, nullptr };
    cqx_hit* hits = nullptr;
    float* beta = nullptr;
    float* L = nullptr;
    uint64_t* rng = nullptr;
    uint32_t* counts = nullptr;
    uint8_t* vmat = nullptr;  // (bounces + 1) x P vertex materials
    uint8_t* vstat = nullptr; // P path records (vertex count, escaped)
    float* partials = nullptr;
    size_t partial_floats = 0;
    hipGraph_t graph = nullptr;        // the render: radiance, no path record
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_replay = nullptr; // the adjoint's replay: path record, no radiance
    hipGraphExec_t exec_replay = nullptr;
};

void destroy_pt_plan(PtPlan* p)
{
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->exec_replay) (void)hipGraphExecDestroy(p->exec_replay);
    if (p->graph_replay) (void)hipGraphDestroy(p->graph_replay);
    release(p->dparams);
    for (int k = 0; k < 2; k++) {
        release(p->rays[k]);
        release(p->paths[k]);
    }
    release(p->hits);
    release(p->beta);
    release(p->L);
    release(p->rng);
    release(p->counts);
    release(p->vmat);
    release(p->vstat);
    release(p->partials);
    delete p;
}

} // namespace cqx

namespace {

using cqx::PtPlan;

cqx_status fill_params(const cqx_pt_params* pp, PtDevParams& d)
{
    std::memset(&d, 0, sizeof(d));
    if (!pp->width || !pp->height || !pp->spp) return fail(CQX_ERR_INVALID_ARG, "image size and spp must be positive");
    if (pp->bounces > 64) return fail(CQX_ERR_LIMIT, "at most 64 bounces");
    if (pp->n_materials > 255) return fail(CQX_ERR_INVALID_ARG, "at most 255 materials");
    if (pp->n_materials && (!pp->albedo || !pp->emission)) return fail(CQX_ERR_INVALID_ARG, "null material table");
    double f[3] = { pp->target[0] - pp->eye[0], pp->target[1] - pp->eye[1], pp->target[2] - pp->eye[2] };
    const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0)) return fail(CQX_ERR_INVALID_ARG, "eye equals target");
    for (double& c : f) c /= fl;
    double r[3] = { -f[2], 0.0, f[0] };
    const double rl = std::sqrt(r[0] * r[0] + r[2] * r[2]);
    if (!(rl > 0)) return fail(CQX_ERR_INVALID_ARG, "view direction parallel to +y");
    for (double& c : r) c /= rl;
    const double u[3] = { r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0] };
    const double scale = 2.0 * std::tan(pp->fov_y_degrees * 3.14159265358979323846 / 360.0);
    const double aspect = double(pp->width) / double(pp->height);
    for (int a = 0; a < 3; a++) {
        d.eye[a] = float(pp->eye[a]);
        d.fwd[a] = float(f[a]);
        d.right[a] = float(r[a]);
        d.up[a] = float(u[a]);
        d.sky[a] = pp->sky[a];
    }
    d.kx = float(aspect * scale / pp->width);
    d.ky = float(scale / pp->height);
    d.half_w = 0.5f * float(pp->width);
    d.half_h = 0.5f * float(pp->height);
    d.offset = pp->bounce_offset;
    d.max_footprint = pp->max_footprint;
    d.width = pp->width;
    d.height = pp->height;
    d.spp = pp->spp;
    d.bounces = pp->bounces;
    d.seed = pp->seed;
    for (uint32_t m = 1; m <= pp->n_materials; m++)
        for (int a = 0; a < 3; a++) {
            d.albedo[m * 3 + a] = pp->albedo[(m - 1) * 3 + a];
            d.emission[m * 3 + a] = pp->emission[(m - 1) * 3 + a];
        }
    return CQX_OK;
}

template <typename T>
cqx_status alloc(T** p, size_t count)
{
    if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess)
        return fail(CQX_ERR_OOM, "device allocation of %zu bytes failed", sizeof(T) * count);
    return CQX_OK;
}

// Returns the context's plan for this volume and image shape, (re)building it when anything
// differs. The wavefront buffers and graph are added on first use.
cqx_status get_plan(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_pt_params* pp, const cqx_launch_cfg* cfg,
                    PtPlan** out)
{
    cqx_launch_cfg c{};
    if (cfg) c = *cfg;
    c.bin_bits = 0; // the bounce launches read their ray count on the device: no binning prepass
    // "auto" (variant 0): the camera rays (depth 0, coherent: a pixel's samples are adjacent) one
    // per lane, 64 per workgroup; the bounce rays on persistent waves at 24 per CU. On the 1024^3
    // 1280x720x16 frame's camera rays one ray per lane took 0.77 ms against 0.90 ms persistent;
    // over the whole frame persistent waves at 24 per CU gave 4.13 ms against 4.38 at 32 per CU,
    // 4.41 at 16 and 5.1 with one ray per lane.
    // Only the fields the caller left at zero take these values.
    cqx_launch_cfg c0 = c;
    if (c.variant == 0) {
        c0.variant = 1;
        if (!c0.block) c0.block = 64;
        if (!c.waves_per_cu) c.waves_per_cu = 24;
    }
    PtPlan* p = ctx->pt_plan;
    if (p && p->vol == vol && p->width == pp->width && p->height == pp->height && p->spp == pp->spp &&
        p->bounces == pp->bounces && p->max_footprint == pp->max_footprint &&
        std::memcmp(&p->cfg, &c, sizeof(c)) == 0) {
        *out = p;
        return CQX_OK;
    }
    HIP_TRY(hipDeviceSynchronize());
    cqx::destroy_pt_plan(p);
    ctx->pt_plan = nullptr;
    p = new (std::nothrow) PtPlan;
    if (!p) return fail(CQX_ERR_OOM, "out of host memory");
    p->vol = vol;
    p->width = pp->width;
    p->height = pp->height;
    p->spp = pp->spp;
    p->bounces = pp->bounces;
    p->max_footprint = pp->max_footprint;
    p->cfg = c;
    p->cfg0 = c0;
    p->P = uint64_t(pp->width) * pp->height * pp->spp;
    cqx_status st = alloc(&p->dparams, 1);
    if (st == CQX_OK) st = alloc(&p->L, 3 * p->P);
    if (st != CQX_OK) {
        cqx::destroy_pt_plan(p);
        return st;
    }
    ctx->pt_plan = p;
    *out = p;
    return CQX_OK;
}

// Allocates the wavefront buffers and captures generate -> (reset next count, intersect, shade) x
// (bounces + 1) once into a graph.
// The wavefront graph of the render (replay false) or of the adjoint's replay (replay true),
// captured once per plan; the buffers they share are allocated with the first.
cqx_status ensure_wavefront(cqx_ctx* ctx, PtPlan* p, bool replay)
{
    if (replay ? p->exec_replay : p->exec) return CQX_OK;
    const uint64_t P = p->P;
    cqx_status st = CQX_OK;
    if (!p->hits) {
        for (int k = 0; k < 2 && st == CQX_OK; k++) {
            st = alloc(&p->rays[k], P);
            if (st == CQX_OK) st = alloc(&p->paths[k], P);
        }
        if (st == CQX_OK) st = alloc(&p->hits, P);
        if (st == CQX_OK) st = alloc(&p->beta, 3 * P);
        if (st == CQX_OK) st = alloc(&p->rng, P);
        if (st == CQX_OK) st = alloc(&p->counts, 64);
    }
    if (st == CQX_OK && replay && !p->vstat) { // the adjoint's path record
        st = alloc(&p->vmat, (p->bounces + 1) * P);
        if (st == CQX_OK) st = alloc(&p->vstat, P);
    }
    if (st != CQX_OK) return st;
    hipStream_t s = ctx->capture_stream;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const uint32_t gb = uint32_t((P + 255) / 256);
    const uint32_t gc = uint32_t((P + cqx::kCompactBlock - 1) / cqx::kCompactBlock);
    hipLaunchKernelGGL(k_pt_generate, dim3(gb), dim3(256), 0, s, p->dparams, P, p->rays[0], p->rng, p->counts + 0);
    cqx_status cst = CQX_OK;
    for (uint32_t depth = 0; depth <= p->bounces && cst == CQX_OK; depth++) {
        const int cur = depth & 1, nxt = cur ^ 1;
        uint32_t* cnt_cur = p->counts + cur * 16;
        uint32_t* cnt_next = p->counts + nxt * 16;
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, cnt_next, 0u);
        cst = cqx::launch_intersect(ctx, p->vol, p->rays[cur], P, cnt_cur, CQX_SURFACE_PROPERTIES, p->max_footprint,
                                    p->hits, depth == 0 ? &p->cfg0 : &p->cfg, s);
        if (replay)
            hipLaunchKernelGGL(k_pt_shade<false>, dim3(gc), dim3(cqx::kCompactBlock), 0, s, p->dparams, P, depth, cnt_cur,
                               p->rays[cur], p->paths[cur], p->hits, p->beta, p->L, p->rng, p->rays[nxt], p->paths[nxt],
                               cnt_next, p->vmat, p->vstat);
        else
            hipLaunchKernelGGL(k_pt_shade<true>, dim3(gc), dim3(cqx::kCompactBlock), 0, s, p->dparams, P, depth, cnt_cur,
                               p->rays[cur], p->paths[cur], p->hits, p->beta, p->L, p->rng, p->rays[nxt], p->paths[nxt],
                               cnt_next, nullptr, nullptr);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    (replay ? p->graph_replay : p->graph) = g;
    if (cst != CQX_OK) return cst;
    if (e != hipSuccess || !g) return fail(CQX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    HIP_TRY(hipGraphInstantiate(replay ? &p->exec_replay : &p->exec, g, nullptr, nullptr, 0));
    return CQX_OK;
}

cqx_status check(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_pt_params* pp)
{
    if (!ctx || !vol || !pp) return fail(CQX_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    const uint64_t P = uint64_t(pp->width) * pp->height * pp->spp;
    if (P == 0) return fail(CQX_ERR_INVALID_ARG, "empty image");
    if (P >= (uint64_t(1) << 32)) return fail(CQX_ERR_LIMIT, "at most 2^32 - 1 paths per call");
    return CQX_OK;
}

TraceParams trace_params(const cqx_dvolume* vol, const cqx_pt_params* pp, const cqx_launch_cfg* cfg)
{
    TraceParams T;
    T.nodes = vol->nodes;
    T.sub = vol->sub;
    T.surf = 1;
    T.max_footprint = pp->max_footprint;
    T.max_iters = (cfg && cfg->max_iters) ? cfg->max_iters : cqx::kDefaultMaxIters;
    T.fp_disabled = footprint_is_disabled(pp->max_footprint) ? 1u : 0u;
    T.offset32 = vol->node_count <= (uint64_t(1) << 27) ? 1u : 0u;
    return T;
}

} // namespace

extern "C" {

cqx_status cqx_pathtrace(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_pt_params* pp, uint32_t mode,
                         const cqx_launch_cfg* cfg, float* image, void* stream_ptr)
{
    const cqx::TraceRange range("cqx_pathtrace");
    cqx_status st = check(ctx, vol, pp);
    if (st != CQX_OK) return st;
    if (!image) return fail(CQX_ERR_INVALID_ARG, "null image");
    if (mode > 1) return fail(CQX_ERR_INVALID_ARG, "mode must be 0 (wavefront) or 1 (megakernel)");
    PtDevParams host;
    st = fill_params(pp, host);
    if (st != CQX_OK) return st;
    PtPlan* p = nullptr;
    st = get_plan(ctx, vol, pp, cfg, &p);
    if (st != CQX_OK) return st;
    if (mode == 0) {
        st = ensure_wavefront(ctx, p, false);
        if (st != CQX_OK) return st;
    }
    hipStream_t s = hipStream_t(stream_ptr);
    HIP_TRY(hipMemcpyAsync(p->dparams, &host, sizeof(host), hipMemcpyHostToDevice, s));
    if (mode == 0) {
        const cqx::TraceRange bounces("wavefront graph: camera rays, then per bounce traversal and shading");
        HIP_TRY(hipGraphLaunch(p->exec, s));
    } else {
        const cqx::TraceRange mega("megakernel: one lane per path");
        const uint32_t block = 256;
        const size_t lds = size_t(vol->max_height + 1) * block * sizeof(uint32_t);
        if (lds > size_t(ctx->lds_per_block)) return fail(CQX_ERR_LIMIT, "stack needs %zu B of LDS", lds);
        const uint32_t waves = (cfg && cfg->waves_per_cu) ? cfg->waves_per_cu : 16u;
        const uint64_t want = (uint64_t(ctx->cu_count) * waves * 64 + block - 1) / block;
        const uint32_t blocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(want, (p->P + block - 1) / block)));
        hipLaunchKernelGGL(k_pt_megakernel, dim3(blocks), dim3(block), lds, s, p->dparams, trace_params(vol, pp, cfg),
                           p->P, p->L);
    }
    const uint32_t pixels = pp->width * pp->height;
    const size_t tile = size_t(256) * pp->spp * sizeof(float);
    const bool staged = tile <= 65536;
    hipLaunchKernelGGL(k_pt_resolve, dim3((pixels + 255) / 256), dim3(256), staged ? tile : 0, s, p->dparams, p->P,
                       p->L, image, staged ? 1u : 0u);
    HIP_TRY(hipGetLastError());
    // the plan's parameter block is read by the kernels just queued; keep `host` alive until then
    HIP_TRY(hipStreamSynchronize(s));
    return CQX_OK;
}

cqx_status cqx_pathtrace_adjoint(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_pt_params* pp,
                                 const cqx_launch_cfg* cfg, const float* d_image, float* d_albedo, float* d_emission,
                                 void* stream_ptr)
{
    const cqx::TraceRange range("cqx_pathtrace_adjoint");
    cqx_status st = check(ctx, vol, pp);
    if (st != CQX_OK) return st;
    if (!d_image || !d_albedo || !d_emission) return fail(CQX_ERR_INVALID_ARG, "null buffer");
    if (pp->bounces + 1 > uint32_t(kMaxVertices))
        return fail(CQX_ERR_LIMIT, "the adjoint supports at most %d bounces", kMaxVertices - 1);
    if (pp->n_materials == 0) return fail(CQX_ERR_INVALID_ARG, "no materials to differentiate");
    PtDevParams host;
    st = fill_params(pp, host);
    if (st != CQX_OK) return st;
    PtPlan* p = nullptr;
    st = get_plan(ctx, vol, pp, cfg, &p);
    if (st == CQX_OK) st = ensure_wavefront(ctx, p, true);
    if (st != CQX_OK) return st;
    const uint32_t block = kAdjointBlock;
    const uint32_t blocks =
        uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(ctx->cu_count) * 8, (p->P + block - 1) / block)));
    const size_t need = size_t(blocks) * pp->n_materials * 6;
    if (need > p->partial_floats) {
        HIP_TRY(hipDeviceSynchronize());
        release(p->partials);
        p->partials = nullptr;
        p->partial_floats = 0;
        st = alloc(&p->partials, need);
        if (st != CQX_OK) return st;
        p->partial_floats = need;
    }
    hipStream_t s = hipStream_t(stream_ptr);
    HIP_TRY(hipMemcpyAsync(p->dparams, &host, sizeof(host), hipMemcpyHostToDevice, s));
    // replay the paths (the replay graph records each path's vertices), then sweep them backwards
    {
        const cqx::TraceRange replay("adjoint: path replay (wavefront graph)");
        HIP_TRY(hipGraphLaunch(p->exec_replay, s));
    }
    {
        const cqx::TraceRange sweep("adjoint: backward sweep and reduction");
        if (pp->n_materials <= uint32_t(kAccMaterials))
            hipLaunchKernelGGL(k_pt_adjoint<true>, dim3(blocks), dim3(block), 0, s, p->dparams, p->P, p->vmat, p->vstat,
                               d_image, pp->n_materials, p->partials);
        else
            hipLaunchKernelGGL(k_pt_adjoint<false>, dim3(blocks), dim3(block), 0, s, p->dparams, p->P, p->vmat, p->vstat,
                               d_image, pp->n_materials, p->partials);
        hipLaunchKernelGGL(k_pt_reduce, dim3(pp->n_materials * 6), dim3(256), 0, s, p->partials, blocks,
                           pp->n_materials, d_albedo, d_emission);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return CQX_OK;
}

} // extern "C"
