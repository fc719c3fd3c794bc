/* Test infrastructure only (the oracle): a C restatement of the reference's CPU path tracer's
 * shading, PathtracingDemo::traceSingleRay and its helpers
 * (src/application/commands/view/pathtracing_demo.cpp:34-194), over the oracle traversal.
 *
 *   orc_bit_mix, orc_fnv1a      Internals::bit_mix / fnv1a (src/library/base.cpp:62-78); checked
 *                               against the compiled reference library on random inputs
 *   orc_position_noise          positionBasedNoise, pathtracing_demo.cpp:34-42
 *   orc_random_point            randomPointInUnitSphere, :61-80 (one draw from a given state)
 *   orc_demo_trace              traceSingleRay (:150-194) with surfaceColour (:44-59) and
 *                               gatherLighting (:82-117), for a batch of primary rays
 *
 * rng_mode 1 runs the reference's generator: ONE state shared by all pixels in order (the
 * reference starts it at 17, pathtracing_demo.cpp:32). rng_mode 0 gives every pixel its own state,
 * uint32(bit_mix(seed * 2^32 + pixel + 1)) (17 if that is 0), the scheme of the GPU entry point;
 * the draw function and the draw order within a pixel are the reference's either way.
 *
 * Every float operation is written out in the reference's order (linalg's dot is a left fold from
 * 0, normalize is a / sqrt(dot(a, a)), the noise mapping is computed in double because of its
 * double literals); the file is compiled with -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

uint64_t orc_bit_mix(uint64_t b)
{
    b = ((b >> 32) ^ b) * UINT64_C(0x0e9846af9b1a615d);
    b = ((b >> 32) ^ b) * UINT64_C(0x0e9846af9b1a615d);
    return (b >> 28) ^ b;
}

uint64_t orc_fnv1a(const void* data, int64_t len, uint64_t seed)
{
    const uint8_t* p = (const uint8_t*)data;
    for (int64_t i = 0; i < len; i++) {
        seed ^= p[i];
        seed *= UINT64_C(0x00000100000001B3);
    }
    return seed;
}

float orc_position_noise(const float pos[3])
{
    int32_t q[3];
    for (int a = 0; a < 3; a++) q[a] = (int32_t)(pos[a] + 0.499f); /* vec3(0.499) is 0.499f */
    const uint32_t h = (uint32_t)orc_fnv1a(q, sizeof(q), UINT64_C(0xcbf29ce484222325));
    float noise = (float)(h & 0xffu) / 255.0f;
    noise = (float)(((double)noise * 0.1) + 0.9);
    return noise;
}

/* The reference loops until a point lands inside the unit sphere; a state cycle that never does
 * would loop forever there. Here the draw gives up after ORC_DEMO_MAX_DRAWS steps with (0, 0, 0). */
#define ORC_DEMO_MAX_DRAWS 1024

int orc_random_point(uint32_t* state, float out[3])
{
    for (int k = 1; k <= ORC_DEMO_MAX_DRAWS; k++) {
        const uint32_t s = (uint32_t)orc_bit_mix(*state);
        *state = s;
        float r[3] = { (float)(s & 0x3FFu), (float)((s >> 10) & 0x3FFu), (float)((s >> 20) & 0x3FFu) };
        for (int a = 0; a < 3; a++) r[a] = (r[a] - 511.5f) / 511.5f;
        float d = 0.0f;
        for (int a = 0; a < 3; a++) d = d + r[a] * r[a];
        if (!(d >= 1.0f)) {
            memcpy(out, r, sizeof(r));
            return k;
        }
    }
    out[0] = out[1] = out[2] = 0.0f;
    return ORC_DEMO_MAX_DRAWS;
}

static void normalize3(const float v[3], float out[3])
{
    float d = 0.0f;
    for (int a = 0; a < 3; a++) d = d + v[a] * v[a];
    const float len = sqrtf(d);
    for (int a = 0; a < 3; a++) out[a] = v[a] / len;
}

uint32_t orc_demo_pixel_state(uint64_t seed, uint64_t pixel)
{
    const uint32_t s = (uint32_t)orc_bit_mix((seed << 32) + pixel + 1u);
    return s ? s : 17u;
}

typedef struct {
    const uint32_t* nodes;
    const orc_subdag* sub;
    const orc_demo_params* p;
    float sun[3];
} demo_ctx;

/* RayVolumeIntersection of one ray (position o + d t from the FP32 distance, face normal) */
static void trace(const demo_ctx* c, const float o[3], const float d[3], int surf, orc_result* r)
{
    const float ray[6] = { o[0], o[1], o[2], d[0], d[1], d[2] };
    orc_hit h;
    orc_intersect_one(c->nodes, c->sub, ray, surf, c->p->max_footprint, 1u << 22, &h);
    orc_hits_to_results(ray, &h, 1, surf, r);
}

static void surface_colour(const demo_ctx* c, const orc_result* r, float out[3])
{
    for (int a = 0; a < 3; a++) out[a] = c->p->colours[r->material * 3 + a];
    if (c->p->add_noise) {
        const float noise = orc_position_noise(r->position);
        for (int a = 0; a < 3; a++) out[a] = out[a] * noise;
    }
}

static void gather_lighting(const demo_ctx* c, const float pos[3], const float n[3], uint32_t* state, float out[3])
{
    float offset[3], so[3];
    for (int a = 0; a < 3; a++) {
        offset[a] = n[a] * 0.001f;
        so[a] = pos[a] + offset[a];
        out[a] = 0.0f;
    }
    if (c->p->include_sun) {
        orc_result r;
        trace(c, so, c->sun, 0, &r);
        if (!r.hit) {
            float d = 0.0f;
            for (int a = 0; a < 3; a++) d = d + c->sun[a] * n[a];
            const float m = (d < 0.0f) ? 0.0f : d; /* std::max(d, 0.0f) */
            for (int a = 0; a < 3; a++) out[a] = out[a] + 0.1f * m;
        }
    }
    if (c->p->include_sky) {
        float pt[3], v[3], dir[3];
        orc_random_point(state, pt);
        for (int a = 0; a < 3; a++) v[a] = n[a] + pt[a];
        normalize3(v, dir);
        orc_result r;
        trace(c, so, dir, 0, &r);
        if (!r.hit)
            for (int a = 0; a < 3; a++) out[a] = out[a] + 1.5f;
    }
}

static void trace_single_ray(const demo_ctx* c, const float* ray, uint32_t* state, float out[3])
{
    out[0] = 0.8f;
    out[1] = 0.8f;
    out[2] = 1.0f;
    orc_result i0;
    trace(c, ray, ray + 3, 1, &i0);
    if (!i0.hit) return;
    float surf0[3], direct0[3], pt[3], v[3], rdir[3], ro[3], indirect0[3] = { 0.0f, 0.0f, 0.0f };
    surface_colour(c, &i0, surf0);
    gather_lighting(c, i0.position, i0.normal, state, direct0);
    orc_random_point(state, pt);
    for (int a = 0; a < 3; a++) v[a] = i0.normal[a] + pt[a];
    normalize3(v, rdir);
    for (int a = 0; a < 3; a++) ro[a] = i0.position[a] + i0.normal[a] * 0.01f;
    orc_result i1;
    trace(c, ro, rdir, 1, &i1);
    if (i1.hit) {
        float surf1[3], direct1[3];
        surface_colour(c, &i1, surf1);
        gather_lighting(c, i1.position, i1.normal, state, direct1);
        for (int a = 0; a < 3; a++) indirect0[a] = surf1[a] * direct1[a];
    }
    const float gamma = c->p->gamma != 0.0f ? c->p->gamma : (float)(1.0 / 2.2);
    for (int a = 0; a < 3; a++) {
        const float v = surf0[a] * (direct0[a] + indirect0[a]);
        out[a] = gamma == 1.0f ? v : powf(v, gamma); /* pow(v, 1) = v */
    }
}

typedef struct {
    const demo_ctx* c;
    const float* rays;
    float* out;
    uint64_t begin, end;
} demo_job;

static void store(const orc_demo_params* p, float* dst, const float px[3])
{
    for (int a = 0; a < 3; a++) dst[a] = p->accumulate ? dst[a] + px[a] : px[a];
}

static void* demo_worker(void* arg)
{
    const demo_job* j = (const demo_job*)arg;
    for (uint64_t i = j->begin; i < j->end; i++) {
        uint32_t s = orc_demo_pixel_state(j->c->p->seed, i);
        float px[3];
        trace_single_ray(j->c, j->rays + 6 * i, &s, px);
        store(j->c->p, j->out + 3 * i, px);
    }
    return NULL;
}

void orc_demo_trace(const uint32_t* nodes, const orc_subdag* sub, const float* rays, uint64_t n,
                    const orc_demo_params* p, int rng_mode, uint32_t* serial_state, float* out_rgb, int threads)
{
    demo_ctx c;
    c.nodes = nodes;
    c.sub = sub;
    c.p = p;
    const float sun_v[3] = { 1.0f, -2.0f, 10.0f };
    normalize3(sun_v, c.sun);
    if (rng_mode == 1) {
        uint32_t s = *serial_state;
        for (uint64_t i = 0; i < n; i++) {
            float px[3];
            trace_single_ray(&c, rays + 6 * i, &s, px);
            store(p, out_rgb + 3 * i, px);
        }
        *serial_state = s;
        return;
    }
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    demo_job jobs[256];
    const uint64_t per = (n + (uint64_t)threads - 1) / (uint64_t)threads;
    for (int t = 0; t < threads; t++) {
        jobs[t].c = &c;
        jobs[t].rays = rays;
        jobs[t].out = out_rgb;
        jobs[t].begin = (uint64_t)t * per < n ? (uint64_t)t * per : n;
        jobs[t].end = jobs[t].begin + per < n ? jobs[t].begin + per : n;
        pthread_create(&tid[t], NULL, demo_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}
