/* Test infrastructure only (the oracle): a direct per-voxel restatement of the synthetic volume
 * generator, written independently of the product's builder, so that tests can check the
 * builder's DAG voxel by voxel. The generator has no reference counterpart (the reference
 * builds volumes with Volume::setVoxel, src/library/storage.cpp:396-438, from app-side noise
 * functions); its specification is the comment on cqx_gen_params in include/cqx.h.
 *
 * Pinning: the builder's DAG must agree with orc_gen_dense on every voxel, and the compiled
 * reference's Volume::voxel on the same .dag file must agree with both.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* PCG32, M.E. O'Neill 2014 (pcg32_srandom_r / pcg32_random_r). */
typedef struct { uint64_t state, inc; } pcg32;

static uint32_t pcg_next(pcg32* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}

static void pcg_seed(pcg32* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0;
    r->inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r->state += initstate;
    pcg_next(r);
}

static double pcg_uniform(pcg32* r) { return pcg_next(r) * (1.0 / 4294967296.0); }

/* splitmix64 finaliser */
static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static double lattice_value(uint64_t seed, int octave, int64_t ix, int64_t iz)
{
    uint64_t k = mix64(seed + 0x9e3779b97f4a7c15ULL * (uint64_t)(octave + 1));
    k = mix64(k ^ (uint64_t)(uint32_t)(int32_t)ix);
    k = mix64(k ^ ((uint64_t)(uint32_t)(int32_t)iz << 32));
    return (double)(k >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

static double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

static double noise2(uint64_t seed, int octave, double x, double z)
{
    double fx = floor(x), fz = floor(z);
    int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double sx = smooth(x - fx), sz = smooth(z - fz);
    double v00 = lattice_value(seed, octave, ix, iz);
    double v10 = lattice_value(seed, octave, ix + 1, iz);
    double v01 = lattice_value(seed, octave, ix, iz + 1);
    double v11 = lattice_value(seed, octave, ix + 1, iz + 1);
    double a = v00 + (v10 - v00) * sx;
    double b = v01 + (v11 - v01) * sx;
    return a + (b - a) * sz;
}

double orc_gen_height(const orc_gen_params* p, int64_t x, int64_t z)
{
    double u = (double)x / p->period, v = (double)z / p->period;
    double sum = 0.0, norm = 0.0, amp = 1.0, freq = 1.0;
    for (int o = 0; o < p->octaves; o++) {
        sum += amp * noise2(p->seed, o, u * freq, v * freq);
        norm += amp;
        amp *= p->gain;
        freq *= p->lacunarity;
    }
    return p->base_height + p->amplitude * (sum / norm);
}

void orc_gen_spheres(const orc_gen_params* p, double* s)
{
    pcg32 r;
    pcg_seed(&r, p->seed, 0x5fe7e5u);
    double n = ldexp(1.0, p->log2_size);
    for (int k = 0; k < p->n_spheres; k++) {
        double cx = pcg_uniform(&r) * n;
        double cy = pcg_uniform(&r) * n;
        double cz = pcg_uniform(&r) * n;
        double rad = p->r_min + (p->r_max - p->r_min) * pcg_uniform(&r);
        s[5 * k + 0] = cx;
        s[5 * k + 1] = cy;
        s[5 * k + 2] = cz;
        s[5 * k + 3] = rad * rad;
        s[5 * k + 4] = (double)(3 + k % (p->n_materials - 2));
    }
}

static uint32_t material_with_height(const orc_gen_params* p, const double* s, double h, int64_t x, int64_t y,
                                     int64_t z)
{
    for (int k = p->n_spheres - 1; k >= 0; k--) {
        const double* sp = s + 5 * k;
        double dx = (double)x - sp[0], dy = (double)y - sp[1], dz = (double)z - sp[2];
        if (dx * dx + dy * dy + dz * dz <= sp[3]) return (uint32_t)sp[4];
    }
    if ((double)y < h) return (double)y < p->rock_height ? 2u : 1u;
    return 0u;
}

uint32_t orc_gen_material(const orc_gen_params* p, const double* spheres, int64_t x, int64_t y, int64_t z)
{
    int64_t n = (int64_t)1 << p->log2_size;
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return 0;
    return material_with_height(p, spheres, orc_gen_height(p, x, z), x, y, z);
}

typedef struct {
    const orc_gen_params* p;
    const double* spheres;
    uint8_t* grid;
    int threads, tid;
} dense_job;

static void* dense_worker(void* arg)
{
    dense_job* j = (dense_job*)arg;
    int64_t n = (int64_t)1 << j->p->log2_size;
    for (int64_t z = j->tid; z < n; z += j->threads)
        for (int64_t x = 0; x < n; x++) {
            double h = orc_gen_height(j->p, x, z);
            for (int64_t y = 0; y < n; y++)
                j->grid[(z * n + y) * n + x] = (uint8_t)material_with_height(j->p, j->spheres, h, x, y, z);
        }
    return NULL;
}

void orc_gen_dense(const orc_gen_params* p, uint8_t* grid, int threads)
{
    double* s = (double*)malloc(sizeof(double) * 5 * (size_t)(p->n_spheres > 0 ? p->n_spheres : 1));
    orc_gen_spheres(p, s);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    dense_job jobs[256];
    for (int t = 0; t < threads; t++) {
        dense_job j = { p, s, grid, threads, t };
        jobs[t] = j;
        pthread_create(&th[t], NULL, dense_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(s);
}

typedef struct {
    const uint8_t* grid;
    int log2_size;
    const uint32_t* nodes;
    uint32_t root;
    int threads, tid;
    uint64_t bad;
} cmp_job;

static void* cmp_worker(void* arg)
{
    cmp_job* j = (cmp_job*)arg;
    int64_t n = (int64_t)1 << j->log2_size;
    uint64_t bad = 0;
    for (int64_t z = j->tid; z < n; z += j->threads)
        for (int64_t y = 0; y < n; y++)
            for (int64_t x = 0; x < n; x++)
                if (orc_voxel(j->nodes, j->root, (int32_t)x, (int32_t)y, (int32_t)z) != j->grid[(z * n + y) * n + x])
                    bad++;
    j->bad = bad;
    return NULL;
}

uint64_t orc_compare_dense(const uint8_t* grid, int log2_size, const uint32_t* nodes, uint32_t root, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    cmp_job jobs[256];
    for (int t = 0; t < threads; t++) {
        cmp_job j = { grid, log2_size, nodes, root, threads, t, 0 };
        jobs[t] = j;
        pthread_create(&th[t], NULL, cmp_worker, &jobs[t]);
    }
    uint64_t bad = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        bad += jobs[t].bad;
    }
    return bad;
}
