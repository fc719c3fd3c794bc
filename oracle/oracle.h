/* Test infrastructure only (the oracle). A plain-C restatement of Cubiquity's ray traversal and of
 * the synthetic workloads, used by tests/ and by bench.py's CPU-baseline leg as the checker.
 * The product never links, loads or calls anything declared here.
 *
 * Record layouts are byte-identical to the product's (include/cqx.h) so that tests compare
 * arrays directly; they are restated here so that the oracle stays self-contained.
 */
#ifndef CUBIQUITY_ORACLE_H
#define CUBIQUITY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* == reference SubDAG (src/library/raytracing.h:57-65), 32 B */
typedef struct {
    int32_t lower[3];
    int32_t height;
    uint32_t pad0, node, pad1, pad2;
} orc_subdag;

/* One traversal result, 32 B.
 *   t        reference RayVolumeIntersection::distance (an FP32 value), 0 on miss
 *   voxel    the first voxel of the hit block that the ray enters (volume coordinates)
 *   node     index of the node whose child was hit (the traversal's current node)
 *   info     bit 0 hit | bits 1-3 child slot in `node` | bits 4-8 log2 of the hit block's side
 *            | bits 9-14 normal code (axis i: bit 9+2i = |n_i|==1, bit 10+2i = signbit(n_i))
 *            | bit 15 anomaly (iteration cap or stalled octant walk) | bits 16-23 material
 *   entries  node entries (SubDAG roots entered + descents + pops that stay inside the SubDAG +
 *            findNearestMaterial steps); each entry reads one 32-B node
 *   iters    traversal loop iterations (child reads) */
typedef struct {
    float t;
    int32_t voxel[3];
    uint32_t node;
    uint32_t info;
    uint32_t entries;
    uint32_t iters;
} orc_hit;

/* reference RayVolumeIntersection as plain fields */
typedef struct {
    int32_t hit;
    uint32_t material;
    double distance;
    float position[3];
    float normal[3];
} orc_result;

/* ---- traversal (cpu_ref.c) ---- */
int orc_find_subdags(const uint32_t* nodes, uint64_t node_count, uint32_t root, orc_subdag out[8]);
/* one ray: r = {ox, oy, oz, dx, dy, dz} */
void orc_intersect_one(const uint32_t* nodes, const orc_subdag* subdags, const float* r, int surf,
                       float max_footprint, uint32_t max_iters, orc_hit* hit);
void orc_intersect(const uint32_t* nodes, const orc_subdag* subdags, const float* rays, uint64_t n,
                   int surf, float max_footprint, uint32_t max_iters, orc_hit* out);
void orc_intersect_mt(const uint32_t* nodes, const orc_subdag* subdags, const float* rays, uint64_t n,
                      int surf, float max_footprint, uint32_t max_iters, orc_hit* out, int threads);
void orc_hits_to_results(const float* rays, const orc_hit* hits, uint64_t n, int surf, orc_result* out);
void orc_trace_ref(const uint32_t* nodes, uint32_t root, const float* rays, uint64_t n, orc_result* out);
uint32_t orc_voxel(const uint32_t* nodes, uint32_t root, int32_t x, int32_t y, int32_t z);
void orc_dense_readout(const uint32_t* nodes, uint32_t root, const int32_t lower[3], const int32_t size[3],
                       uint8_t* out);
void orc_descend(const uint32_t* nodes, uint32_t root, int32_t x, int32_t y, int32_t z,
                 uint32_t path[33]);
double orc_time_intersect(const uint32_t* nodes, const orc_subdag* subdags, const float* rays,
                          uint64_t n, int surf, float max_footprint, int threads, uint64_t chunk,
                          int reps, uint64_t* hit_count);

/* ---- dense FP64 brute force over a small dense grid (cpu_ref.c) ---- */
void orc_dense_brute(const uint8_t* grid, int n, const float* rays, uint64_t count,
                     int32_t* hit, double* t_near, double* t_second, uint32_t* material,
                     int32_t* voxel);

/* ---- generator restatement (cpu_gen.c) ---- */
/* same field order as the product's cqx_gen_params */
typedef struct {
    int32_t log2_size;
    int32_t octaves;
    int32_t n_spheres;
    int32_t n_materials;
    uint64_t seed;
    double gain, lacunarity, period;
    double base_height, amplitude, rock_height;
    double r_min, r_max;
} orc_gen_params;

double orc_gen_height(const orc_gen_params* p, int64_t x, int64_t z);
/* spheres: n_spheres x {cx, cy, cz, r*r, material} */
void orc_gen_spheres(const orc_gen_params* p, double* spheres);
uint32_t orc_gen_material(const orc_gen_params* p, const double* spheres, int64_t x, int64_t y, int64_t z);
/* dense N^3 grid, x fastest then y then z */
void orc_gen_dense(const orc_gen_params* p, uint8_t* grid, int threads);
/* number of voxels whose material differs between the dense grid and the DAG */
uint64_t orc_compare_dense(const uint8_t* grid, int log2_size, const uint32_t* nodes, uint32_t root, int threads);

/* ---- path tracer restatement (cpu_pt.c) ---- */
/* same field order as the product's cqx_pt_params */
typedef struct {
    uint32_t width, height, spp, bounces;
    double eye[3], target[3], fov_y_degrees;
    uint32_t n_materials;
    const float* albedo;
    const float* emission;
    float sky[3];
    float bounce_offset;
    float max_footprint;
    uint64_t seed;
} orc_pt_params;

/* FP32 forward image (height x width x 3), arithmetic identical to the GPU renderer */
int orc_pathtrace(const uint32_t* nodes, const orc_subdag* subdags, const orc_pt_params* p, float* image,
                  int threads);
/* FP64 reference: image and the gradient of sum_pixels w . image with respect to the material
 * tables (w = d_image, height x width x 3), over the same paths (geometry in FP32, exactly as the
 * GPU traces it; throughput, radiance and gradients in FP64). Any output may be NULL. */
int orc_pathtrace_f64(const uint32_t* nodes, const orc_subdag* subdags, const orc_pt_params* p,
                      const double* d_image, double* image, double* d_albedo, double* d_emission, int threads);

/* pieces of the path tracer, for the pinning tests */
void orc_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out);
void orc_pcg_uniform(uint64_t initstate, uint64_t initseq, uint32_t n, float* out);
void orc_sincos2pi(const float* u, uint64_t n, float* c, float* s);
uint32_t orc_pt_bounce_rays(const float n[3], const float pos[3], float bounce_offset, uint64_t initstate,
                            uint64_t initseq, uint32_t count, float* rays);
int orc_pt_camera_ray(const orc_pt_params* p, uint64_t path, float ray[6], float jitter[2]);

/* ---- the reference CPU tracer's shading (cpu_demo.c) ---- */
/* same field order as the product's cqx_demo_params */
typedef struct {
    const float* colours; /* 256 x 3 material colours */
    float max_footprint;
    float gamma; /* 0 = the reference's 1/2.2 */
    uint32_t add_noise, include_sun, include_sky, accumulate;
    uint64_t seed;
} orc_demo_params;

uint64_t orc_bit_mix(uint64_t bits);
uint64_t orc_fnv1a(const void* data, int64_t len, uint64_t seed);
float orc_position_noise(const float pos[3]);
/* one randomPointInUnitSphere draw from *state; returns the number of generator steps taken */
int orc_random_point(uint32_t* state, float out[3]);
uint32_t orc_demo_pixel_state(uint64_t seed, uint64_t pixel);
/* traceSingleRay for n primary rays; out_rgb: n x 3. rng_mode 0: per-pixel states (threads may
 * be > 1); rng_mode 1: the reference's single generator in pixel order from *serial_state. */
void orc_demo_trace(const uint32_t* nodes, const orc_subdag* sub, const float* rays, uint64_t n,
                    const orc_demo_params* p, int rng_mode, uint32_t* serial_state, float* out_rgb, int threads);

#ifdef __cplusplus
}
#endif
#endif
