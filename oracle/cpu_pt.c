/* Test infrastructure only (the oracle): CPU restatement of the path tracer and of its material
 * gradient, written independently of the GPU renderer from the same specification (the comment
 * on cqx_pt_params in include/cqx.h). The reference has no counterpart: its CPU tracer
 * (src/application/commands/view/pathtracing_demo.cpp:81-229) uses sun and sky shadow rays, one
 * bounce and a global serial RNG. Conventions kept from it: hit position o + d t
 * (raytracing.cpp:464), bounce origin offset along the face normal (pathtracing_demo.cpp:139),
 * background (0.8, 0.8, 1.0) (:124).
 *
 * orc_pathtrace reproduces the GPU image bit for bit (FP32, same operation order, traversal by
 * orc_intersect_one). orc_pathtrace_f64 follows the same paths (their geometry does not depend on
 * the material tables) but accumulates throughput, radiance and gradients in FP64; it is the
 * reference for the GPU adjoint, and the image it returns bounds the GPU's FP32 rounding.
 *
 * Pinning (each piece by something outside this file):
 *   PCG32        O'Neill's published pcg32 demo stream (initstate 42, initseq 54);
 *   sincos2pi    every FP32 input k / 2^24 of the uniform draws against FP64 sin/cos;
 *   bounce       moments of the cosine-weighted law (unit length, n.d >= 0, E[n.d] = 2/3,
 *                E[(n.d)^2] = 1/2, the CDF of n.d, uniform azimuth) on axis and edge normals;
 *   camera ray   the FP64 pinhole through the jittered pixel point (camera.cpp:12-38 geometry);
 *   estimator    closed forms: a furnace (camera inside a closed shell of one material gives
 *                e (1 + a + ... + a^B) per pixel), a lit floor (e + a sky for B >= 1, e for
 *                B = 0), an empty volume (sky), emission linearity;
 *   gradient     central finite differences of the FP64 image (a polynomial in the tables).
 * A planted slip in any of these (tests/test_oracle_mutations.py) fails one of the checks.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MAXV 65 /* vertices of a path: bounces + 1, at most 64 bounces */

typedef struct { uint64_t state, inc; } pcg_t;

static uint32_t pcg_step(pcg_t* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}

static pcg_t pcg_init(uint64_t initstate, uint64_t initseq)
{
    pcg_t r = { 0, (initseq << 1u) | 1u };
    pcg_step(&r);
    r.state += initstate;
    pcg_step(&r);
    return r;
}

static float pcg_u(pcg_t* r) { return (float)(pcg_step(r) >> 8) * (1.0f / 16777216.0f); }

/* cos/sin of 2 pi u: quadrant reduction, Taylor polynomials (degree 11 / 12) on [0, pi/2). */
static void sincos2pi(float u, float* c, float* s)
{
    float u4 = u * 4.0f;
    float q = floorf(u4);
    float th = (u4 - q) * 1.57079632679489662f;
    float t2 = th * th;
    float sp = th * (1.0f + t2 * (-1.66666667e-1f + t2 * (8.33333333e-3f + t2 * (-1.98412698e-4f +
               t2 * (2.75573192e-6f + t2 * -2.50521084e-8f)))));
    float cp = 1.0f + t2 * (-0.5f + t2 * (4.16666667e-2f + t2 * (-1.38888889e-3f + t2 * (2.48015873e-5f +
               t2 * (-2.75573192e-7f + t2 * 2.08767570e-9f)))));
    switch ((int)q) {
    case 0: *c = cp; *s = sp; break;
    case 1: *c = -sp; *s = cp; break;
    case 2: *c = -cp; *s = -sp; break;
    default: *c = sp; *s = -cp; break;
    }
}

typedef struct {
    float eye[3], fwd[3], right[3], up[3];
    float kx, ky, half_w, half_h;
    float albedo[256 * 3], emission[256 * 3];
} cam_t;

static int setup(const orc_pt_params* p, cam_t* c)
{
    memset(c, 0, sizeof(*c));
    double f[3] = { p->target[0] - p->eye[0], p->target[1] - p->eye[1], p->target[2] - p->eye[2] };
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0)) return -1;
    for (int a = 0; a < 3; a++) f[a] /= fl;
    double r[3] = { -f[2], 0.0, f[0] };
    double rl = sqrt(r[0] * r[0] + r[2] * r[2]);
    if (!(rl > 0)) return -1;
    for (int a = 0; a < 3; a++) r[a] /= rl;
    double u[3] = { r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0] };
    double scale = 2.0 * tan(p->fov_y_degrees * 3.14159265358979323846 / 360.0);
    double aspect = (double)p->width / (double)p->height;
    for (int a = 0; a < 3; a++) {
        c->eye[a] = (float)p->eye[a];
        c->fwd[a] = (float)f[a];
        c->right[a] = (float)r[a];
        c->up[a] = (float)u[a];
    }
    c->kx = (float)(aspect * scale / p->width);
    c->ky = (float)(scale / p->height);
    c->half_w = 0.5f * (float)p->width;
    c->half_h = 0.5f * (float)p->height;
    for (uint32_t m = 1; m <= p->n_materials && m < 256; m++)
        for (int a = 0; a < 3; a++) {
            c->albedo[m * 3 + a] = p->albedo[(m - 1) * 3 + a];
            c->emission[m * 3 + a] = p->emission[(m - 1) * 3 + a];
        }
    return 0;
}

/* Start of path number `path` (pixel = path / spp, sample = path % spp): RNG and camera ray. */
static void start(const orc_pt_params* p, const cam_t* c, uint64_t path, pcg_t* rng, float ray[6], uint32_t* pixel)
{
    uint32_t px = (uint32_t)(path / p->spp), s = (uint32_t)(path % p->spp);
    *rng = pcg_init((p->seed << 32) + (uint64_t)px, (uint64_t)s);
    float jx = pcg_u(rng), jy = pcg_u(rng);
    uint32_t x = px % p->width, y = px / p->width;
    float cx = (((float)x + jx) - c->half_w) * c->kx;
    float cy = (((float)y + jy) - c->half_h) * c->ky;
    float v[3];
    for (int a = 0; a < 3; a++) v[a] = (c->fwd[a] + c->right[a] * cx) - c->up[a] * cy;
    float len = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int a = 0; a < 3; a++) {
        ray[a] = c->eye[a];
        ray[3 + a] = v[a] / len;
    }
    *pixel = px;
}

/* Cosine-weighted bounce about the face normal (Duff et al. 2017 orthonormal basis); the next
 * origin is pos + n * offset. Returns 0 for a zero normal. */
static int bounce(const orc_pt_params* p, pcg_t* rng, const float pos[3], const float n[3], float ray[6])
{
    float nl = sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    if (!(nl > 0.0f)) return 0;
    float nx = n[0] / nl, ny = n[1] / nl, nz = n[2] / nl;
    float u1 = pcg_u(rng), u2 = pcg_u(rng);
    float r = sqrtf(u1), c, s;
    sincos2pi(u2, &c, &s);
    float lx = r * c, ly = r * s, lz = sqrtf(1.0f - u1);
    float sg = copysignf(1.0f, nz);
    float a = -1.0f / (sg + nz);
    float b = nx * ny * a;
    float t0 = 1.0f + sg * nx * nx * a, t1 = sg * b, t2 = -sg * nx;
    float b0 = b, b1 = sg + ny * ny * a, b2 = -ny;
    ray[3] = (t0 * lx + b0 * ly) + nx * lz;
    ray[4] = (t1 * lx + b1 * ly) + ny * lz;
    ray[5] = (t2 * lx + b2 * ly) + nz * lz;
    for (int k = 0; k < 3; k++) ray[k] = pos[k] + n[k] * p->bounce_offset;
    return 1;
}

static void normal_of(uint32_t info, float n[3])
{
    for (int a = 0; a < 3; a++) {
        uint32_t code = (info >> (9 + 2 * a)) & 3u;
        float mag = (code & 1u) ? 1.0f : 0.0f;
        n[a] = (code & 2u) ? -mag : mag;
    }
}

/* One path: the materials of its hit vertices (<= bounces + 1), whether it escaped to the sky,
 * and the FP32 radiance with the GPU's operation order. */
typedef struct {
    int K, escaped;
    uint32_t m[MAXV];
    float L[3];
} path_t;

static void trace_path(const uint32_t* nodes, const orc_subdag* sub, const orc_pt_params* p, const cam_t* c,
                       uint64_t path, path_t* out, uint32_t* pixel)
{
    pcg_t rng;
    float ray[6];
    start(p, c, path, &rng, ray, pixel);
    float b[3] = { 1.0f, 1.0f, 1.0f };
    out->K = 0;
    out->escaped = 0;
    out->L[0] = out->L[1] = out->L[2] = 0.0f;
    for (uint32_t depth = 0;; depth++) {
        orc_hit h;
        orc_intersect_one(nodes, sub, ray, 1, p->max_footprint, 1u << 22, &h);
        if (h.info & 0x8000u) break;
        if (!(h.info & 1u)) {
            for (int a = 0; a < 3; a++) out->L[a] += b[a] * p->sky[a];
            out->escaped = 1;
            break;
        }
        uint32_t m = (h.info >> 16) & 0xffu;
        if (out->K < MAXV) out->m[out->K] = m;
        out->K++;
        for (int a = 0; a < 3; a++) out->L[a] += b[a] * c->emission[m * 3 + a];
        if (depth >= p->bounces) break;
        float pos[3], n[3];
        for (int a = 0; a < 3; a++) pos[a] = ray[a] + ray[3 + a] * h.t;
        normal_of(h.info, n);
        if (!bounce(p, &rng, pos, n, ray)) break;
        for (int a = 0; a < 3; a++) b[a] = b[a] * c->albedo[m * 3 + a];
    }
}

typedef struct {
    const uint32_t* nodes;
    const orc_subdag* sub;
    const orc_pt_params* p;
    const cam_t* c;
    float* L;           /* per path radiance, 3 x P (FP32 mode) */
    const double* w;    /* d_image (FP64 mode) */
    double* L64;        /* per path radiance, 3 x P (FP64 mode) */
    double grad[256 * 6];
    int f64, threads, tid;
} job_t;

static void* worker(void* arg)
{
    job_t* j = (job_t*)arg;
    const orc_pt_params* p = j->p;
    const cam_t* c = j->c;
    uint64_t P = (uint64_t)p->width * p->height * p->spp;
    memset(j->grad, 0, sizeof(j->grad));
    for (uint64_t path = (uint64_t)j->tid; path < P; path += (uint64_t)j->threads) {
        path_t t;
        uint32_t pixel;
        trace_path(j->nodes, j->sub, p, c, path, &t, &pixel);
        if (!j->f64) {
            for (int a = 0; a < 3; a++) j->L[a * P + path] = t.L[a];
            continue;
        }
        /* FP64 replay of the same vertices */
        double beta[MAXV + 1][3];
        for (int a = 0; a < 3; a++) beta[0][a] = 1.0;
        double L[3] = { 0, 0, 0 };
        int K = t.K < MAXV ? t.K : MAXV;
        for (int k = 0; k < K; k++)
            for (int a = 0; a < 3; a++) {
                L[a] += beta[k][a] * (double)c->emission[t.m[k] * 3 + a];
                beta[k + 1][a] = beta[k][a] * (double)c->albedo[t.m[k] * 3 + a];
            }
        if (t.escaped)
            for (int a = 0; a < 3; a++) L[a] += beta[K][a] * (double)p->sky[a];
        if (j->L64)
            for (int a = 0; a < 3; a++) j->L64[a * P + path] = L[a];
        if (j->w) {
            double G[3];
            for (int a = 0; a < 3; a++) G[a] = t.escaped ? (double)p->sky[a] : 0.0;
            for (int k = K - 1; k >= 0; k--) {
                uint32_t m = t.m[k];
                for (int a = 0; a < 3; a++) {
                    double wk = j->w[(uint64_t)pixel * 3 + a] / (double)p->spp * beta[k][a];
                    j->grad[m * 6 + a] += wk * G[a];
                    j->grad[m * 6 + 3 + a] += wk;
                    G[a] = (double)c->emission[m * 3 + a] + (double)c->albedo[m * 3 + a] * G[a];
                }
            }
        }
    }
    return NULL;
}

static int run(const uint32_t* nodes, const orc_subdag* sub, const orc_pt_params* p, int f64, const double* w,
               float* L, double* L64, double* grad, int threads)
{
    cam_t* c = (cam_t*)malloc(sizeof(cam_t));
    if (!c || setup(p, c) != 0) { free(c); return -1; }
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    job_t* jobs = (job_t*)calloc((size_t)threads, sizeof(job_t));
    pthread_t th[64];
    for (int t = 0; t < threads; t++) {
        job_t jb;
        memset(&jb, 0, sizeof(jb));
        jb.nodes = nodes; jb.sub = sub; jb.p = p; jb.c = c; jb.L = L; jb.w = w; jb.L64 = L64;
        jb.f64 = f64; jb.threads = threads; jb.tid = t;
        jobs[t] = jb;
        pthread_create(&th[t], NULL, worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    if (grad) {
        memset(grad, 0, sizeof(double) * 256 * 6);
        for (int t = 0; t < threads; t++)
            for (int k = 0; k < 256 * 6; k++) grad[k] += jobs[t].grad[k];
    }
    free(jobs);
    free(c);
    return 0;
}

int orc_pathtrace(const uint32_t* nodes, const orc_subdag* subdags, const orc_pt_params* p, float* image, int threads)
{
    uint64_t P = (uint64_t)p->width * p->height * p->spp;
    float* L = (float*)malloc(sizeof(float) * 3 * P);
    if (!L) return -1;
    if (run(nodes, subdags, p, 0, NULL, L, NULL, NULL, threads) != 0) { free(L); return -1; }
    for (uint64_t q = 0; q < (uint64_t)p->width * p->height; q++)
        for (int a = 0; a < 3; a++) {
            float sum = 0.0f;
            for (uint32_t s = 0; s < p->spp; s++) sum += L[a * P + q * p->spp + s];
            image[q * 3 + a] = sum / (float)p->spp;
        }
    free(L);
    return 0;
}

int orc_pathtrace_f64(const uint32_t* nodes, const orc_subdag* subdags, const orc_pt_params* p,
                      const double* d_image, double* image, double* d_albedo, double* d_emission, int threads)
{
    uint64_t P = (uint64_t)p->width * p->height * p->spp;
    double* L64 = image ? (double*)malloc(sizeof(double) * 3 * P) : NULL;
    double grad[256 * 6];
    if (run(nodes, subdags, p, 1, d_image, NULL, L64, grad, threads) != 0) { free(L64); return -1; }
    if (image) {
        for (uint64_t q = 0; q < (uint64_t)p->width * p->height; q++)
            for (int a = 0; a < 3; a++) {
                double sum = 0.0;
                for (uint32_t s = 0; s < p->spp; s++) sum += L64[a * P + q * p->spp + s];
                image[q * 3 + a] = sum / (double)p->spp;
            }
        free(L64);
    }
    for (uint32_t m = 1; m <= p->n_materials && m < 256; m++)
        for (int a = 0; a < 3; a++) {
            if (d_albedo) d_albedo[(m - 1) * 3 + a] = grad[m * 6 + a];
            if (d_emission) d_emission[(m - 1) * 3 + a] = grad[m * 6 + 3 + a];
        }
    return 0;
}

/* ---- entry points for the pinning tests (each runs one piece of the path tracer alone) ---- */

void orc_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out)
{
    pcg_t r = pcg_init(initstate, initseq);
    for (uint32_t i = 0; i < n; i++) out[i] = pcg_step(&r);
}

void orc_pcg_uniform(uint64_t initstate, uint64_t initseq, uint32_t n, float* out)
{
    pcg_t r = pcg_init(initstate, initseq);
    for (uint32_t i = 0; i < n; i++) out[i] = pcg_u(&r);
}

void orc_sincos2pi(const float* u, uint64_t n, float* c, float* s)
{
    for (uint64_t i = 0; i < n; i++) sincos2pi(u[i], &c[i], &s[i]);
}

/* `count` bounces about the face normal n from the origin `pos`, all drawing from one PCG stream
 * (two draws each, as along a path); rays: count x 6. Returns the number of rays written (0 for a
 * zero normal). */
uint32_t orc_pt_bounce_rays(const float n[3], const float pos[3], float bounce_offset, uint64_t initstate,
                            uint64_t initseq, uint32_t count, float* rays)
{
    orc_pt_params p;
    memset(&p, 0, sizeof(p));
    p.bounce_offset = bounce_offset;
    pcg_t r = pcg_init(initstate, initseq);
    for (uint32_t i = 0; i < count; i++)
        if (!bounce(&p, &r, pos, n, rays + 6 * (uint64_t)i)) return i;
    return count;
}

/* The camera ray of path number `path` and the pixel jitter (jx, jy) it was drawn with. */
int orc_pt_camera_ray(const orc_pt_params* p, uint64_t path, float ray[6], float jitter[2])
{
    cam_t* c = (cam_t*)malloc(sizeof(cam_t));
    if (!c || setup(p, c) != 0) { free(c); return -1; }
    pcg_t rng;
    uint32_t pixel;
    start(p, c, path, &rng, ray, &pixel);
    pcg_t again = pcg_init((p->seed << 32) + (uint64_t)pixel, (uint64_t)(path % p->spp));
    jitter[0] = pcg_u(&again);
    jitter[1] = pcg_u(&again);
    free(c);
    return 0;
}
