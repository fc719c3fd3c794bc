/* Test infrastructure only (the oracle): a scalar, FP-faithful C restatement of Cubiquity's ray vs
 * sparse-voxel-DAG traversal and of its brute-force checker. tests/ and bench.py's CPU-baseline
 * leg use it as the checker; the product never links or calls it.
 *
 * Every function cites the reference code it restates (paths relative to the reference
 * checkout). Floating-point semantics follow the reference's Release build: IEEE binary32,
 * no contraction (built with -ffp-contract=off), std::min/std::max compare-select order,
 * round-to-nearest int->float conversions, two's-complement wrapping integer arithmetic.
 *
 * Pinning (tests/test_oracle_*.py): bit-exact against the compiled reference (oracle/_ref) on
 * seeded camera, incoherent and special-case rays; against an independent FP64 dense brute
 * force (orc_dense_brute) on small volumes; against closed forms (axis-aligned rays into
 * solid boxes) and against Volume::voxel / root descent for the derived voxel and node
 * outputs.
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MATERIAL_COUNT 256u /* src/library/storage.h:55 */
#define ROOT_HEIGHT 32      /* src/library/raytracing.cpp:12 */

/* ---- scalar helpers ------------------------------------------------------------------------ */

/* std::min(a,b) == (b < a) ? b : a and std::max(a,b) == (a < b) ? b : a, nested as in
 * min3/max3 (src/library/raytracing.cpp:168-176). NaN and signed-zero results depend on the
 * operand order, so the order is kept. */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
static inline float min3f(const float v[3]) { return smin(smin(v[0], v[1]), v[2]); }
static inline float max3f(const float v[3]) { return smax(smax(v[0], v[1]), v[2]); }

/* GLSL findMSB() for an unsigned argument: -1 for zero (src/library/raytracing.cpp:16-24). */
static inline int find_msb(uint32_t v)
{
    int r = -1;
    while (v) { r++; v >>= 1; }
    return r;
}

/* Wrapping int32 arithmetic (the reference's ivec3 ops compile to wrapping machine ops). */
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
/* (x >> h) << h with an arithmetic right shift; x86 masks shift counts to 5 bits, which only
 * matters for h == 32, a state in which the traversal loop exits without using the value. */
static inline int32_t floor_to(int32_t x, int h)
{
    h &= 31;
    return (int32_t)((uint32_t)(x >> h) << h);
}

/* ---- SubDAGs: src/library/raytracing.cpp:43-97 ---------------------------------------------- */

/* Descend from root child `c` along the chain of single-child nodes to the first node with more
 * than one occupied child (or a material node). Returns 0, or -1 when the chain is longer than
 * the tree can be (a malformed store). */
static int find_subdag(const uint32_t* nodes, uint64_t node_count, uint32_t root, uint32_t c,
                       orc_subdag* out)
{
    int child_height = ROOT_HEIGHT;
    int32_t lower[3] = { INT_MIN, INT_MIN, INT_MIN };
    uint32_t only = c;
    uint32_t next = nodes[(uint64_t)root * 8 + only];
    uint32_t node = 0;
    uint32_t count = 1;
    while (count == 1) {
        if (child_height == 0) return -1;
        child_height--;
        node = next;
        if ((uint64_t)node >= node_count) return -1;
        for (int a = 0; a < 3; a++)
            lower[a] = (int32_t)((uint32_t)lower[a] ^ (((only >> a) & 1u) << child_height));
        count = 0;
        for (uint32_t i = 0; i < 8; i++) {
            uint32_t ch = nodes[(uint64_t)node * 8 + i];
            if (ch > 0) { next = ch; count++; only = i; }
        }
    }
    memset(out, 0, sizeof(*out));
    out->lower[0] = lower[0]; out->lower[1] = lower[1]; out->lower[2] = lower[2];
    out->height = child_height;
    out->node = node;
    return 0;
}

int orc_find_subdags(const uint32_t* nodes, uint64_t node_count, uint32_t root, orc_subdag out[8])
{
    for (uint32_t c = 0; c < 8; c++)
        if (find_subdag(nodes, node_count, root, c, &out[c]) != 0) return -1;
    return 0;
}

/* ---- traversal: src/library/raytracing.cpp:134-478 ----------------------------------------- */

typedef struct {
    const uint32_t* nodes;
    uint32_t max_iters; /* 0 = unbounded */
    uint32_t iters;
    uint32_t entries;
    int anomaly;
} walk_ctx;

/* findNearestMaterial, src/library/raytracing.cpp:134-162: near-to-far child order
 * 0,1,2,4,3,5,6,7 (packed nibbles 0x76534210 with a marker bit per nibble), reflected by the
 * ray's sign bits, first occupied child taken, repeated until a material node is reached. */
static uint32_t nearest_material(walk_ctx* w, uint32_t node, uint32_t sign_bits)
{
    const uint32_t order = (0x76534210u | 0x88888888u) ^ (sign_bits * 0x11111111u);
    while (node >= MATERIAL_COUNT) {
        w->entries++;
        uint32_t next = 0;
        for (uint32_t ids = order; ids != 0; ids >>= 4) {
            uint32_t ch = w->nodes[(uint64_t)node * 8 + (ids & 7u)];
            if (ch > 0) { next = ch; break; }
        }
        /* an internal node without occupied children: the reference would loop forever */
        if (next == 0) { w->anomaly = 1; return 0; }
        node = next;
    }
    return node;
}

/* findFirstChild, src/library/raytracing.cpp:178-196. Returns child id bits (x=1, y=2, z=4). */
static inline uint32_t first_child(float entry, const float o[3], const float inv[3],
                                   const int32_t centre[3])
{
    uint32_t id = 0;
    for (int a = 0; a < 3; a++) {
        float tm = ((float)centre[a] - o[a]) * inv[a];
        if (tm < entry) id |= 1u << a;
    }
    if (entry <= 0) {
        for (int a = 0; a < 3; a++)
            if (o[a] >= (float)centre[a]) id |= 1u << a;
    }
    return id;
}

/* Voxel of the hit block that the ray enters first, in reflected shifted coordinates: on entry
 * axes the block's near face, elsewhere floor(o' + d' t) clamped into the block. (Not exposed by
 * the reference; defined identically in the product kernel.) */
static inline int32_t entered_voxel(float o, float d, float t, int32_t lo, int32_t size, int entry_axis)
{
    if (entry_axis) return lo;
    float q = o + d * t;
    float f = floorf(q);
    int32_t hi = wadd(lo, size - 1);
    if (!(f >= (float)lo)) return lo; /* also catches NaN */
    if (f > (float)hi) return hi;
    int32_t v = (int32_t)f;
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

/* intersectRayNodeESVO, src/library/raytracing.cpp:213-371. o and d are the reflected, shifted
 * ray (d >= 0 per axis), sign = +-1 per axis, sign_bits the reflection mask. Returns 1 on hit. */
static int esvo(walk_ctx* w, uint32_t node, const int32_t pos[3], int height, const float o[3],
                const float d[3], const float sign[3], uint32_t sign_bits, int surf,
                float max_footprint, orc_hit* hit)
{
    const uint32_t node_size = 1u << height;
    float inv[3], t0[3], t1[3];
    for (int a = 0; a < 3; a++) {
        inv[a] = 1.0f / d[a];
        t0[a] = ((float)pos[a] - o[a]) * inv[a];
        t1[a] = (((float)pos[a] + (float)node_size) - o[a]) * inv[a];
    }
    const float node_entry = max3f(t0);
    const float node_exit = min3f(t1);
    if (!(node_entry < node_exit)) return 0;

    const int start = height;
    int32_t csize = (int32_t)(node_size / 2);
    int32_t centre[3] = { wadd(pos[0], csize), wadd(pos[1], csize), wadd(pos[2], csize) };
    uint32_t cid = first_child(node_entry, o, inv, centre);
    int32_t cpos[3];
    for (int a = 0; a < 3; a++) cpos[a] = wadd(pos[a], ((cid >> a) & 1u) ? csize : 0);

    float last_exit = node_exit;
    uint32_t stack[ROOT_HEIGHT + 1];
    memset(stack, 0, sizeof(stack));
    w->entries++;

    for (;;) {
        if (w->max_iters && w->iters >= w->max_iters) { w->anomaly = 1; return 0; }
        w->iters++;
        float ct1[3];
        for (int a = 0; a < 3; a++) ct1[a] = ((float)wadd(cpos[a], csize) - o[a]) * inv[a];
        const float tcx = min3f(ct1);
        const uint32_t slot = cid ^ sign_bits;
        const uint32_t child = w->nodes[(uint64_t)node * 8 + slot];

        if (child > 0) {
            float ct0[3];
            for (int a = 0; a < 3; a++) ct0[a] = ((float)cpos[a] - o[a]) * inv[a];
            const float tce = max3f(ct0);
            const int internal = child >= MATERIAL_COUNT;
            const int large = ((float)csize / tcx) > max_footprint;
            if (internal && large) {
                if (height <= 1) { w->anomaly = 1; return 0; } /* deeper than a valid DAG */
                if (tcx < last_exit) stack[height] = node;
                last_exit = tcx;
                height--;
                node = child;
                csize /= 2;
                w->entries++;
                int32_t c2[3] = { wadd(cpos[0], csize), wadd(cpos[1], csize), wadd(cpos[2], csize) };
                cid = first_child(tce, o, inv, c2);
                for (int a = 0; a < 3; a++) cpos[a] = wadd(cpos[a], ((cid >> a) & 1u) ? csize : 0);
            } else {
                /* A SubDAG of height 0 (the chain below the root ends in a single voxel) has
                 * child size 0: the hit block is then the SubDAG node itself, of side 1. */
                const int32_t bsize = csize > 0 ? csize : 1;
                uint32_t info = 1u | (slot << 1);
                info |= (uint32_t)find_msb((uint32_t)bsize) << 4;
                uint32_t material = 0;
                if (surf) {
                    material = nearest_material(w, child, sign_bits);
                    if (w->anomaly) return 0;
                    for (int a = 0; a < 3; a++) {
                        /* vec3(equal(tEntry, childT0)) * -rayDirSign, raytracing.cpp:317-318 */
                        float n = (float)(tce == ct0[a]) * -sign[a];
                        uint32_t code = (n != 0.0f ? 1u : 0u) | (signbit(n) ? 2u : 0u);
                        info |= code << (9 + 2 * a);
                    }
                }
                info |= (material & 0xffu) << 16;
                hit->t = tce;
                for (int a = 0; a < 3; a++) {
                    int32_t v = entered_voxel(o[a], d[a], tce, cpos[a], bsize, ct0[a] == tce);
                    /* un-reflect: reflected voxel v' spans [v', v'+1); real voxel = -v'-1 */
                    hit->voxel[a] = ((sign_bits >> a) & 1u) ? ~v : v;
                }
                hit->node = node;
                hit->info = info;
                return 1;
            }
        } else {
            /* advance to the next sibling, raytracing.cpp:328-332 */
            uint32_t flips = 0;
            for (int a = 0; a < 3; a++)
                if (ct1[a] <= tcx) flips |= 1u << a;
            cid ^= flips;
            int32_t old[3] = { cpos[0], cpos[1], cpos[2] };
            for (int a = 0; a < 3; a++)
                if ((flips >> a) & 1u) cpos[a] = wadd(cpos[a], csize);
            /* pop when a flipped axis wrapped, raytracing.cpp:339-364 */
            if ((cid & flips) != flips) {
                uint32_t diff = (uint32_t)(old[0] ^ cpos[0]) | (uint32_t)(old[1] ^ cpos[1]) |
                                (uint32_t)(old[2] ^ cpos[2]);
                int msb = find_msb(diff);
                height = msb + 1;
                node = (height <= ROOT_HEIGHT) ? stack[height] : 0;
                csize = (int32_t)(1u << (msb & 31));
                cid = 0;
                for (int a = 0; a < 3; a++) {
                    uint32_t bit = (uint32_t)(cpos[a] >> (msb & 31)) & 1u;
                    cid |= bit << a;
                    cpos[a] = wadd(floor_to(cpos[a], height), bit ? csize : 0);
                }
                last_exit = 0.0f;
                if (height <= start) w->entries++;
            }
        }
        if (height > start) return 0;
    }
}

/* intersectVolume, src/library/raytracing.cpp:407-478. */
void orc_intersect_one(const uint32_t* nodes, const orc_subdag* subdags, const float* r,
                          int surf, float max_footprint, uint32_t max_iters, orc_hit* hit)
{
    memset(hit, 0, sizeof(*hit));
    walk_ctx w = { nodes, max_iters, 0, 0, 0 };

    int32_t sb[3];
    float sign[3], o[3], d[3], dist[3];
    uint32_t sign_bits = 0;
    for (int a = 0; a < 3; a++) {
        sb[a] = r[3 + a] < 0.0f;       /* lessThan(dir, 0): -0 counts as positive */
        sign_bits |= (uint32_t)sb[a] << a;
        sign[a] = (float)(sb[a] * -2 + 1);
        o[a] = (r[a] + 0.5f) * sign[a]; /* two ops: add, then multiply */
        d[a] = fabsf(r[3 + a]);
    }
    uint32_t oct = 0;
    for (int a = 0; a < 3; a++) {
        dist[a] = (-o[a]) / d[a];
        if (dist[a] < 0.0) { oct += 1u << a; dist[a] += FLT_MAX; }
    }

    do {
        const orc_subdag* s = &subdags[oct ^ sign_bits];
        if (s->node > 0) {
            const int32_t size = (int32_t)(1u << s->height);
            int32_t lb[3];
            for (int a = 0; a < 3; a++)
                lb[a] = wsub(wmul(s->lower[a], (int32_t)sign[a]), wmul(sb[a], size));
            orc_hit h;
            memset(&h, 0, sizeof(h));
            int got = esvo(&w, s->node, lb, s->height, o, d, sign, sign_bits, surf, max_footprint, &h);
            if (got) { *hit = h; break; }
            if (w.anomaly) break;
        }
        const float m = min3f(dist);
        uint32_t step = 0;
        if (dist[0] <= m) { step += 1; dist[0] += FLT_MAX; }
        if (dist[1] <= m) { step += 2; dist[1] += FLT_MAX; }
        if (dist[2] <= m) { step += 4; dist[2] += FLT_MAX; }
        if (step == 0) { w.anomaly = 1; break; } /* the reference would loop forever here */
        oct += step;
    } while (oct <= 7);

    if (w.anomaly) {
        memset(hit, 0, sizeof(*hit));
        hit->info = 1u << 15; /* reported as a miss */
    }
    hit->entries = w.entries;
    hit->iters = w.iters;
}

void orc_intersect(const uint32_t* nodes, const orc_subdag* subdags, const float* rays, uint64_t n,
                   int surf, float max_footprint, uint32_t max_iters, orc_hit* out)
{
    for (uint64_t i = 0; i < n; i++)
        orc_intersect_one(nodes, subdags, rays + 6 * i, surf, max_footprint, max_iters, &out[i]);
}

typedef struct {
    const uint32_t* nodes;
    const orc_subdag* subdags;
    const float* rays;
    uint64_t n, chunk;
    int surf, threads, tid;
    float max_footprint;
    uint32_t max_iters;
    orc_hit* out;
    uint64_t hits;
} mt_job;

static void* mt_worker(void* arg)
{
    mt_job* j = (mt_job*)arg;
    uint64_t hits = 0;
    orc_hit scratch;
    for (uint64_t c0 = (uint64_t)j->tid * j->chunk; c0 < j->n; c0 += (uint64_t)j->threads * j->chunk) {
        uint64_t c1 = c0 + j->chunk < j->n ? c0 + j->chunk : j->n;
        for (uint64_t i = c0; i < c1; i++) {
            orc_hit* h = j->out ? &j->out[i] : &scratch;
            orc_intersect_one(j->nodes, j->subdags, j->rays + 6 * i, j->surf, j->max_footprint, j->max_iters, h);
            hits += h->info & 1u;
        }
    }
    j->hits = hits;
    return NULL;
}

static uint64_t run_mt(const uint32_t* nodes, const orc_subdag* subdags, const float* rays, uint64_t n,
                       int surf, float max_footprint, uint32_t max_iters, orc_hit* out, int threads,
                       uint64_t chunk)
{
    if (threads < 1) threads = 1;
    if (threads > 512) threads = 512;
    pthread_t th[512];
    mt_job jobs[512];
    for (int t = 0; t < threads; t++) {
        mt_job j = { nodes, subdags, rays, n, chunk, surf, threads, t, max_footprint, max_iters, out, 0 };
        jobs[t] = j;
        pthread_create(&th[t], NULL, mt_worker, &jobs[t]);
    }
    uint64_t hits = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        hits += jobs[t].hits;
    }
    return hits;
}

void orc_intersect_mt(const uint32_t* nodes, const orc_subdag* subdags, const float* rays, uint64_t n,
                      int surf, float max_footprint, uint32_t max_iters, orc_hit* out, int threads)
{
    run_mt(nodes, subdags, rays, n, surf, max_footprint, max_iters, out, threads, 4096);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* All-core timing protocol: one warm-up, best of `reps`; results are not stored. */
double orc_time_intersect(const uint32_t* nodes, const orc_subdag* subdags, const float* rays,
                          uint64_t n, int surf, float max_footprint, int threads, uint64_t chunk,
                          int reps, uint64_t* hit_count)
{
    double best = 1e30;
    uint64_t hits = 0;
    for (int rep = -1; rep < reps; rep++) {
        double t0 = now_s();
        hits = run_mt(nodes, subdags, rays, n, surf, max_footprint, 0, NULL, threads, chunk ? chunk : 4096);
        double dt = now_s() - t0;
        if (rep >= 0 && dt < best) best = dt;
    }
    if (hit_count) *hit_count = hits;
    return best;
}

/* Expand records into the reference's RayVolumeIntersection fields:
 * position = o + d * (float)distance on the original ray (raytracing.cpp:464), normal from the
 * packed code, material and normal zero unless surface properties were requested. */
void orc_hits_to_results(const float* rays, const orc_hit* hits, uint64_t n, int surf, orc_result* out)
{
    for (uint64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        const orc_hit* h = &hits[i];
        orc_result* o = &out[i];
        memset(o, 0, sizeof(*o));
        if (!(h->info & 1u)) continue;
        o->hit = 1;
        o->distance = (double)h->t;
        for (int a = 0; a < 3; a++) o->position[a] = r[a] + r[3 + a] * h->t;
        if (surf) {
            o->material = (h->info >> 16) & 0xffu;
            for (int a = 0; a < 3; a++) {
                uint32_t code = (h->info >> (9 + 2 * a)) & 3u;
                float mag = (code & 1u) ? 1.0f : 0.0f;
                o->normal[a] = (code & 2u) ? -mag : mag;
            }
        }
    }
}

/* ---- point queries: src/library/storage.cpp:45-67, 481-499 --------------------------------- */

uint32_t orc_voxel(const uint32_t* nodes, uint32_t root, int32_t x, int32_t y, int32_t z)
{
    uint32_t u[3] = { (uint32_t)x + 0x80000000u, (uint32_t)y + 0x80000000u, (uint32_t)z + 0x80000000u };
    uint32_t node = root;
    while (node >= MATERIAL_COUNT) {
        uint32_t c = (u[0] >> 31) | ((u[1] >> 31) << 1) | ((u[2] >> 31) << 2);
        u[0] <<= 1; u[1] <<= 1; u[2] <<= 1;
        node = nodes[(uint64_t)node * 8 + c];
    }
    return node;
}

/* Voxels of the box [lower, lower + size) as [z][y][x] bytes, by root descent. */
void orc_dense_readout(const uint32_t* nodes, uint32_t root, const int32_t lower[3], const int32_t size[3], uint8_t* out)
{
    for (int32_t z = 0; z < size[2]; z++)
        for (int32_t y = 0; y < size[1]; y++)
            for (int32_t x = 0; x < size[0]; x++)
                out[((size_t)z * size[1] + y) * size[0] + x] =
                    (uint8_t)orc_voxel(nodes, root, lower[0] + x, lower[1] + y, lower[2] + z);
}

/* path[h] = index of the node at height h on the descent to voxel (x,y,z); heights below the
 * first material node repeat that material. */
void orc_descend(const uint32_t* nodes, uint32_t root, int32_t x, int32_t y, int32_t z, uint32_t path[33])
{
    uint32_t u[3] = { (uint32_t)x + 0x80000000u, (uint32_t)y + 0x80000000u, (uint32_t)z + 0x80000000u };
    uint32_t node = root;
    for (int h = ROOT_HEIGHT; h >= 0; h--) {
        path[h] = node;
        if (h == 0) break;
        uint32_t c = (u[0] >> 31) | ((u[1] >> 31) << 1) | ((u[2] >> 31) << 2);
        u[0] <<= 1; u[1] <<= 1; u[2] <<= 1;
        if (node >= MATERIAL_COUNT) node = nodes[(uint64_t)node * 8 + c];
    }
}

/* ---- traceRayRef: src/library/raytracing.cpp:480-535 + utility.h:60-116 + geometry.h:460-486 */

typedef struct {
    const uint32_t* nodes;
    double o[3], d[3];
    int hit;
    double best;
    uint32_t material;
} ref_finder;

/* IntersectionFinder::operator(): dilate the integer box by 0.5, slab-test in double, store the
 * entry/exit as float; record the nearest non-empty material node with exit > 0. */
static int finder_visit(ref_finder* f, uint32_t node, const int32_t lo[3], const int32_t hi[3])
{
    double mn[3], mx[3];
    for (int a = 0; a < 3; a++) {
        double inv = 1.0 / f->d[a];
        double l = (((double)lo[a] - 0.5) - f->o[a]) * inv;
        double u = (((double)hi[a] + 0.5) - f->o[a]) * inv;
        mn[a] = (u < l) ? u : l; /* std::min(lower, upper) */
        mx[a] = (l < u) ? u : l; /* std::max(lower, upper) */
    }
    double e = mn[0]; e = (e < mn[1]) ? mn[1] : e; e = (e < mn[2]) ? mn[2] : e;
    double x = mx[0]; x = (mx[1] < x) ? mx[1] : x; x = (mx[2] < x) ? mx[2] : x;
    float entry = (float)e, exit = (float)x;
    if (!(exit >= entry)) return 0;
    if (node < MATERIAL_COUNT && node > 0) {
        if (exit > 0.0 && entry < f->best) {
            f->hit = 1;
            f->best = entry;
            f->material = node;
        }
    }
    return 1;
}

static void visit_children(ref_finder* f, uint32_t node, const int32_t lo[3], const int32_t hi[3], uint32_t height)
{
    const uint32_t ch_height = height - 1;
    const uint32_t flip = 1u << ch_height;
    for (uint32_t c = 0; c < 8; c++) {
        const uint32_t child = f->nodes[(uint64_t)node * 8 + c];
        int32_t clo[3], chi[3];
        for (int a = 0; a < 3; a++) {
            clo[a] = lo[a];
            chi[a] = hi[a];
            /* collapse the faces selected by ~childId (utility.h:102-104) */
            if ((c >> a) & 1u) clo[a] = (int32_t)((uint32_t)clo[a] ^ flip);
            else chi[a] = (int32_t)((uint32_t)chi[a] ^ flip);
        }
        if (finder_visit(f, child, clo, chi) && child >= MATERIAL_COUNT)
            visit_children(f, child, clo, chi, ch_height);
    }
}

void orc_trace_ref(const uint32_t* nodes, uint32_t root, const float* rays, uint64_t n, orc_result* out)
{
    for (uint64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        ref_finder f;
        f.nodes = nodes;
        for (int a = 0; a < 3; a++) { f.o[a] = (double)r[a]; f.d[a] = (double)r[3 + a]; }
        f.hit = 0;
        f.best = DBL_MAX;
        f.material = 0;
        const int32_t lo[3] = { INT_MIN, INT_MIN, INT_MIN }, hi[3] = { INT_MAX, INT_MAX, INT_MAX };
        if (finder_visit(&f, root, lo, hi) && root >= MATERIAL_COUNT)
            visit_children(&f, root, lo, hi, ROOT_HEIGHT);
        memset(&out[i], 0, sizeof(out[i]));
        out[i].hit = f.hit;
        out[i].material = f.material;
        out[i].distance = f.hit ? f.best : DBL_MAX;
    }
}

/* ---- independent FP64 dense brute force (no reference code path) --------------------------- */

/* For a dense n^3 grid of materials (x fastest, then y, then z) whose voxel v spans
 * [v-0.5, v+0.5]^3, return for each ray the nearest entry t over all occupied voxels whose slab
 * interval [entry, exit] is non-empty with exit > 0, the second-nearest distinct entry (for
 * ambiguity screening), the material and the voxel. Axis-parallel components are handled
 * explicitly rather than through IEEE infinities. */
void orc_dense_brute(const uint8_t* grid, int n, const float* rays, uint64_t count,
                     int32_t* hit, double* t_near, double* t_second, uint32_t* material,
                     int32_t* voxel)
{
    for (uint64_t i = 0; i < count; i++) {
        const float* r = rays + 6 * i;
        double o[3] = { r[0], r[1], r[2] }, d[3] = { r[3], r[4], r[5] };
        double best = INFINITY, second = INFINITY;
        uint32_t mat = 0;
        int32_t bv[3] = { 0, 0, 0 };
        for (int z = 0; z < n; z++)
            for (int y = 0; y < n; y++)
                for (int x = 0; x < n; x++) {
                    uint8_t m = grid[((size_t)z * n + y) * n + x];
                    if (!m) continue;
                    int v[3] = { x, y, z };
                    double lo = -INFINITY, hi = INFINITY;
                    int miss = 0;
                    for (int a = 0; a < 3 && !miss; a++) {
                        double b0 = v[a] - 0.5, b1 = v[a] + 0.5;
                        if (d[a] == 0.0) {
                            if (o[a] < b0 || o[a] > b1) miss = 1;
                        } else {
                            double t0 = (b0 - o[a]) / d[a], t1 = (b1 - o[a]) / d[a];
                            if (t0 > t1) { double s = t0; t0 = t1; t1 = s; }
                            if (t0 > lo) lo = t0;
                            if (t1 < hi) hi = t1;
                        }
                    }
                    if (miss || !(lo <= hi) || !(hi > 0.0)) continue;
                    if (lo < best) {
                        second = best;
                        best = lo; mat = m; bv[0] = x; bv[1] = y; bv[2] = z;
                    } else if (lo < second) {
                        second = lo; /* includes exact ties with best */
                    }
                }
        hit[i] = best < INFINITY;
        t_near[i] = best;
        t_second[i] = second;
        material[i] = mat;
        voxel[3 * i] = bv[0]; voxel[3 * i + 1] = bv[1]; voxel[3 * i + 2] = bv[2];
    }
}
