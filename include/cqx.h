/* cqx.h: C ABI of the MI355X (gfx950) sparse-voxel-DAG traversal library (libcqx.so).
 *
 * Drop-in boundary for Cubiquity's raycast path. Plain pointers and sizes, no C++ types,
 * every function returns a cqx_status (0 = OK, < 0 = error) and never throws; the message of
 * the last error on the calling thread is available from cqx_last_error().
 *
 * Reference interfaces replaced (paths relative to the Cubiquity checkout):
 *   SubDAG / findSubDAGs              src/library/raytracing.h:57-69, raytracing.cpp:43-97
 *   intersectVolume (per ray)         src/library/raytracing.h:71-76, raytracing.cpp:397-478
 *   RayVolumeIntersection             src/library/raytracing.h:48-55
 *   NodeStore raw bytes + SubDAGArray upload to the GPU
 *                                     src/application/commands/view/gpu_pathtracing_viewer.cpp:43-57
 *   Volume::load / save (.dag)        src/library/storage.cpp:505-542, 192-206
 *   Volume::voxel                     src/library/storage.cpp:481-499
 *
 * Threading: one cqx_ctx per host thread, or calls on a context serialized by the caller.
 * Host buffers are caller-owned; device volumes are owned by the context that uploaded them.
 */
#ifndef CQX_H
#define CQX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t cqx_status;
#define CQX_OK 0
#define CQX_ERR_INVALID_ARG (-1) /* null pointer, bad size or flag */
#define CQX_ERR_INVALID_DAG (-2) /* child index out of range, cycle, deeper than 32 levels */
#define CQX_ERR_HIP (-3)         /* a HIP runtime call failed */
#define CQX_ERR_OOM (-4)         /* host or device allocation failed */
#define CQX_ERR_IO (-5)          /* file could not be read or written */
#define CQX_ERR_NO_DEVICE (-6)   /* no usable GPU */
#define CQX_ERR_LIMIT (-7)       /* input exceeds a documented limit */
#define CQX_ERR_PARSE (-8)       /* a metadata file is not valid (see cqx_materials_load) */

/* Number of material ids; node indices below this are material (leaf) nodes whose children
 * all point at themselves. Matches Internals::MaterialCount, src/library/storage.h:55. */
#define CQX_MATERIAL_COUNT 256u

/* == SubDAG, src/library/raytracing.h:57-65 (32 B, same field order and padding). */
typedef struct {
    int32_t lower[3]; /* lower corner in signed voxel coordinates */
    int32_t height;   /* node height (side = 2^height voxels) */
    uint32_t pad0;
    uint32_t node;    /* node index; 0 = empty octant */
    uint32_t pad1, pad2;
} cqx_subdag;

/* One ray, FP32, 24 B: the six floats intersectVolume takes (raytracing.h:72-75). */
typedef struct {
    float o[3];
    float d[3];
} cqx_ray;

/* One traversal result, 32 B (the GPU's compact record).
 *   t        RayVolumeIntersection::distance (an FP32 value); 0 on miss
 *   voxel    first voxel of the hit block entered by the ray, signed volume coordinates
 *   node     index of the node whose child was hit
 *   info     bit 0 hit | bits 1-3 child slot in `node` | bits 4-8 log2 of the hit block's side
 *            | bits 9-14 normal code (axis i: bit 9+2i set if |n_i| == 1, bit 10+2i = signbit)
 *            | bit 15 anomaly (iteration cap reached or octant walk stalled; reported as miss)
 *            | bits 16-23 material (0 unless CQX_SURFACE_PROPERTIES)
 *   entries  node entries (SubDAG roots entered, descents, pops inside the SubDAG, nearest-
 *            material steps); each reads one 32-B node
 *   iters    traversal loop iterations (child reads)
 * The reference's position is o + d * t (FP32 multiply then add) and its normal is decoded from
 * the normal code; cqx_hits_to_intersections() produces both. */
typedef struct {
    float t;
    int32_t voxel[3];
    uint32_t node;
    uint32_t info;
    uint32_t entries;
    uint32_t iters;
} cqx_hit;

#define CQX_HIT_FLAG 0x1u
#define CQX_HIT_ANOMALY 0x8000u

/* == RayVolumeIntersection, src/library/raytracing.h:48-55, as plain fields. */
typedef struct {
    int32_t hit;
    uint32_t material;
    double distance;
    float position[3];
    float normal[3];
} cqx_intersection;

/* intersect flags */
#define CQX_SURFACE_PROPERTIES 0x1u /* computeSurfaceProperties (material + normal) */

/* reference MAX_FOOTPRINT_DISABLED, src/library/raytracing.h:71 */
#define CQX_MAX_FOOTPRINT_DISABLED (-1.0f)

/* ---- errors ---------------------------------------------------------------------------------- */
const char* cqx_last_error(void);
const char* cqx_version(void);

/* ---- messages --------------------------------------------------------------------------------
 * Cubiquity's set_debug_handler / set_warning_handler / setProgressHandler
 * (src/library/base.h:102-107, base.cpp:86-131): process-wide hooks, NULL by default.
 *   warning   every error as it is raised (the text cqx_last_error() then returns on that thread)
 *   debug     notes from long operations (volume generation progress with node counts,
 *             renumbering); without a debug handler, CQX_PROGRESS=1 in the environment prints
 *             them on stderr
 *   progress  (task, first_step, current_step, last_step) from long operations, the first call
 *             at first_step and the last at last_step (volume generation: regions built)
 * Handlers may be called from the library's worker threads; calls are serialised. */
typedef void (*cqx_message_fn)(const char* message);
typedef void (*cqx_progress_fn)(const char* task, int first_step, int current_step, int last_step);
void cqx_set_debug_handler(cqx_message_fn fn);
void cqx_set_warning_handler(cqx_message_fn fn);
void cqx_set_progress_handler(cqx_progress_fn fn);

/* ---- host node store (replaces Volume/NodeStore access on the path) --------------------------- */

/* Nodes are stored as in the reference's NodeStore: node_count x uint32[8], indices 0..255 are
 * the material nodes {i,i,i,i,i,i,i,i}, internal nodes start at 256 (storage.cpp:110-122). */
typedef struct cqx_host_volume cqx_host_volume;

/* Volume::load, storage.cpp:505-528 + NodeStore::read, storage.cpp:192-199 */
cqx_status cqx_hv_load_dag(const char* path, cqx_host_volume** out);
/* Volume::save's file layout, storage.cpp:530-542 (the store is written as it is, not re-baked) */
cqx_status cqx_hv_save_dag(const cqx_host_volume* hv, const char* path);
/* Copy a raw store (incl. the 256 material nodes) and validate it. */
cqx_status cqx_hv_from_nodes(const uint32_t* nodes, uint64_t node_count, uint32_t root,
                             cqx_host_volume** out);
void cqx_hv_free(cqx_host_volume* hv);
const uint32_t* cqx_hv_nodes(const cqx_host_volume* hv);
uint64_t cqx_hv_node_count(const cqx_host_volume* hv);
uint32_t cqx_hv_root(const cqx_host_volume* hv);
/* Volume::voxel, storage.cpp:481-499 */
cqx_status cqx_hv_voxel(const cqx_host_volume* hv, int32_t x, int32_t y, int32_t z, uint32_t* material);

/* findSubDAGs, raytracing.cpp:89-97, over a raw store. Validates every index it follows. */
cqx_status cqx_find_subdags(const uint32_t* nodes, uint64_t node_count, uint32_t root,
                            cqx_subdag out[8]);

/* Checks that every child index is < node_count, that the DAG reachable from `root` is acyclic
 * and at most 32 levels deep (root at height 32) and that node_count fits the 32-bit indices. */
cqx_status cqx_validate_dag(const uint32_t* nodes, uint64_t node_count, uint32_t root);

/* Expand compact records into reference RayVolumeIntersection fields (host). */
cqx_status cqx_hits_to_intersections(const cqx_ray* rays, const cqx_hit* hits, uint64_t n,
                                     uint32_t flags, cqx_intersection* out);

/* ---- seeded synthetic workloads (the benchmark's volumes and rays) --------------------------- */

/* Heightfield + spheres generator. Voxel (x,y,z) of [0, 2^log2_size)^3 is
 *   sphere material  if inside any sphere (last sphere wins): (x-cx)^2+(y-cy)^2+(z-cz)^2 <= r^2
 *   2 (rock)         else if y < height(x,z) and y < rock_height
 *   1 (ground)       else if y < height(x,z)
 *   0 (empty)        otherwise
 * height(x,z) = base_height + amplitude * fbm(x/period, z/period), fbm = normalised sum of
 * `octaves` octaves of 2D value noise (gain, lacunarity). Sphere k: centre U[0,N)^3,
 * radius U[r_min, r_max], material 3 + k mod (n_materials - 2), drawn from PCG32(seed, stream
 * 0x5fe7e5). Everything is evaluated in FP64. */
typedef struct {
    int32_t log2_size;
    int32_t octaves;
    int32_t n_spheres;
    int32_t n_materials; /* material ids used: 1..n_materials (>= 3) */
    uint64_t seed;
    double gain, lacunarity, period;
    double base_height, amplitude, rock_height;
    double r_min, r_max;
} cqx_gen_params;

/* The scaled defaults for a side of 2^log2_size (s = N/128): period 64 s, base 64 s,
 * amplitude 40 s, rock below 40 s, octaves 4 + round(2/3 log2 s), 32 s^2 spheres with radius
 * [4, 16] s^(1/3), gain 0.5, lacunarity 2, 3 materials. */
cqx_status cqx_gen_default_params(int32_t log2_size, uint64_t seed, cqx_gen_params* out);
/* Build the volume as a deduplicated DAG (the shape Volume::bake produces, storage.cpp:208-290),
 * placed at [0, N)^3. `threads` <= 0 uses all cores. */
cqx_status cqx_hv_generate(const cqx_gen_params* p, int threads, cqx_host_volume** out);

/* Sphere fill / carve (material 0) edits with the voxel outcome of the reference's
 * fillBrush(volume, SphereBrush(centre, radius), material) applied in order
 * (src/library/voxelization.cpp:825-915, voxelization.h:91-126), including its FP32 containment
 * test (squared distance truncated to int64, compared with radius^2, strictly). Copy-on-write
 * with hash-consing: no node reachable before the call changes; new nodes are appended from
 * index *first_new on, shared when identical, and 8 equal material children collapse. */
typedef struct {
    float centre[3];
    float radius;
    uint32_t material;
} cqx_sphere_edit;

cqx_status cqx_hv_edit_spheres(cqx_host_volume* hv, const cqx_sphere_edit* edits, uint64_t n,
                               uint64_t* first_new);
/* The same with an explicit number of host threads (0 = all cores or OMP_NUM_THREADS, 1 = in
 * order on the calling thread). With more than one thread and at least 16 edits, the volume is
 * cut into aligned regions too large to lie inside any of the spheres; each region applies its
 * spheres in order, regions in parallel, and the levels above are rebuilt once. The voxels are
 * those of the in-order walk for any thread count (spheres farther than 2^23 voxels from the
 * origin are applied in order). */
cqx_status cqx_hv_edit_spheres_mt(cqx_host_volume* hv, const cqx_sphere_edit* edits, uint64_t n,
                                  int32_t threads, uint64_t* first_new);
/* Drops unreachable nodes and renumbers breadth-first (what Volume::bake does for storage,
 * storage.cpp:387-394). Invalidates device copies (upload again). */
cqx_status cqx_hv_compact(cqx_host_volume* hv);

/* Edit history, replacing Volume::checkpoint / undo / redo (storage.cpp:358-385; declared in
 * storage.h). checkpoint() records the current root; undo() and redo() move along the recorded
 * roots and are no-ops at either end; a checkpoint after an undo discards the redo states.
 * Edits append nodes and never modify existing ones, so every recorded state stays valid in the
 * store and on the device: after undo/redo, cqx_volume_update with first == node_count and the
 * new root brings a device copy up to date without copying nodes. cqx_hv_compact clears the
 * history (it renumbers nodes). */
cqx_status cqx_hv_checkpoint(cqx_host_volume* hv);
cqx_status cqx_hv_undo(cqx_host_volume* hv);
cqx_status cqx_hv_redo(cqx_host_volume* hv);

/* ---- material metadata ------------------------------------------------------------------------
 * The .toml file next to a .dag (Metadata::load / Metadata::save,
 * src/application/base/metadata.cpp:34-109; Material, metadata.h:11-35): an optional
 * `dimensions = [x, y, z]` and an array of [[materials]] tables, one per material id from 0, each
 * with an optional `name` and `base_color` (missing colour: Material::DefaultColor 0.8 grey) and,
 * as an extension for the path tracer, an optional `emission` (missing: 0). Other keys and tables
 * are skipped. */
#define CQX_MATERIAL_HAS_NAME 1u
#define CQX_MATERIAL_HAS_BASE_COLOR 2u
#define CQX_MATERIAL_HAS_EMISSION 4u
typedef struct {
    char name[64]; /* NUL-terminated */
    float base_color[3];
    float emission[3];
    uint32_t flags; /* CQX_MATERIAL_HAS_* : which keys the file holds */
} cqx_material;

/* Reads up to 256 materials into `materials` (capacity 256) and their number into `count`.
 * `dimensions` / `has_dimensions` (may be NULL) receive the dimensions key if present (a DAG's
 * metadata should not hold one; serialize.cpp:64-69 ignores it). CQX_ERR_PARSE names the line. */
cqx_status cqx_materials_load(const char* path, cqx_material* materials, uint32_t* count, int32_t dimensions[3],
                              uint32_t* has_dimensions);
/* Writes the materials in Metadata::save's layout (colours with four decimals). */
cqx_status cqx_materials_save(const char* path, const cqx_material* materials, uint32_t count);
/* The viewer's 256-entry colour table (Viewer::onVolumeModified, viewer.cpp:49-54): material i's
 * base_color for i < count, purple (1, 0, 1) for the rest. */
cqx_status cqx_materials_colour_table(const cqx_material* materials, uint32_t count, float colours[256 * 3]);

/* Build the canonical (baked) DAG of a dense grid of material ids, x fastest then y then z,
 * whose voxel (0,0,0) sits at signed voxel coordinates `lower` (any placement, including across
 * the origin). The counterpart of filling a Volume with setVoxel and calling bake
 * (storage.cpp:396-438, 387-394). */
cqx_status cqx_hv_from_dense(const uint8_t* grid, const int32_t size[3], const int32_t lower[3],
                             cqx_host_volume** out);

/* Pinhole camera rays through pixel centres, y down, built in FP64 and rounded to FP32
 * (the convention of src/application/commands/view/camera.cpp:12-38). Row-major, width*height. */
cqx_status cqx_camera_rays(const double eye[3], const double target[3], double fov_y_degrees,
                           uint32_t width, uint32_t height, cqx_ray* out);
/* Incoherent rays: ray i draws from PCG32(seed, i): origin U[0,N)^3 - 0.5 (the volume's extent,
 * voxel v spanning [v-0.5, v+0.5)), re-drawn until the voxel at floor(origin + 0.5) is empty;
 * direction uniform on the sphere (z = 2u-1, phi = 2 pi v), computed in FP64, rounded to FP32. */
cqx_status cqx_incoherent_rays(const cqx_host_volume* hv, int32_t log2_size, uint64_t seed,
                               uint64_t n, int threads, cqx_ray* out);

/* ---- device ------------------------------------------------------------------------------------ */

typedef struct cqx_ctx cqx_ctx;
typedef struct cqx_dvolume cqx_dvolume;

cqx_status cqx_create(int device, cqx_ctx** out);
cqx_status cqx_destroy(cqx_ctx* ctx);

/* Upload a store (node_count x 32 B incl. the material nodes, as the reference's SSBO upload
 * does) and its SubDAGs (NULL: computed here). Validates the store first. */
cqx_status cqx_volume_upload(cqx_ctx* ctx, const uint32_t* nodes, uint64_t node_count, uint32_t root,
                             const cqx_subdag sub[8], cqx_dvolume** out);
cqx_status cqx_volume_free(cqx_ctx* ctx, cqx_dvolume* vol);
/* Append-only delta after host edits (cqx_hv_edit_spheres): nodes [first, node_count) of the
 * host store are copied to the device (the first `first` nodes must be unchanged since the
 * upload), and the root and SubDAGs (NULL: recomputed) are replaced. Appended nodes may only
 * reference material nodes or earlier nodes. Waits for the device to be idle. (The reference
 * re-uploads nothing after edits: its GPU viewer has no onVolumeModified,
 * gpu_pathtracing_viewer.h:14-26.) */
cqx_status cqx_volume_update(cqx_ctx* ctx, cqx_dvolume* vol, const uint32_t* nodes, uint64_t node_count,
                             uint64_t first, uint32_t root, const cqx_subdag sub[8]);
/* Sphere fill / carve edits applied on the device to the device copy of the store
 * (fillBrush(volume, SphereBrush(centre, radius), material) in order, voxelization.cpp:825-915,
 * NodeStore::setNodeChild's copy-on-write, storage.cpp:152-167; the reference has no device
 * edits). The result is cqx_hv_edit_spheres_mt's with more than one thread, node for node: the
 * same region-parallel algorithm, hash-consing against an index over the device store (built on
 * the volume's first edit, 4 bytes x 2-4 x node_count of device memory, kept up to date), new
 * nodes appended from *first_new on; the root and SubDAGs are replaced. Batches of fewer than 16
 * spheres, spheres farther than 2^23 voxels from the origin and radii too large for a region of
 * side 2^20 are refused with CQX_ERR_LIMIT (apply those with cqx_hv_edit_spheres). Synchronizes
 * the device. */
cqx_status cqx_volume_edit_spheres(cqx_ctx* ctx, cqx_dvolume* vol, const cqx_sphere_edit* edits, uint64_t n,
                                   uint64_t* first_new);
/* Edit history of a device volume (Volume::checkpoint / undo / redo, storage.cpp:358-385): roots
 * recorded and switched; no node changes, the SubDAGs are recomputed on the device. */
cqx_status cqx_volume_checkpoint(cqx_ctx* ctx, cqx_dvolume* vol);
cqx_status cqx_volume_undo(cqx_ctx* ctx, cqx_dvolume* vol);
cqx_status cqx_volume_redo(cqx_ctx* ctx, cqx_dvolume* vol);
/* Copies nodes [first, first + count) of the device store to host memory (node x 8 uint32),
 * e.g. the nodes cqx_volume_edit_spheres appended, to save the volume. */
cqx_status cqx_volume_download(cqx_ctx* ctx, const cqx_dvolume* vol, uint64_t first, uint64_t count,
                               uint32_t* nodes);
/* Device pointer to the volume's node array (node_count x 8 uint32) and its node count. */
cqx_status cqx_volume_info(const cqx_dvolume* vol, const uint32_t** dev_nodes, uint64_t* node_count,
                           uint32_t* root, cqx_subdag sub[8]);

/* Launch parameters of the traversal kernel. Zero fields take the tuned or built-in defaults
 * (variant 2, 256 threads, 32 waves per CU, regroup 8, refill 16). */
typedef struct {
    uint32_t variant;          /* 0 auto (the autotuner's choice), 1 one ray per lane, 2 persistent waves
                                  (walk steps + service passes, rays claimed in chunks), 3 = 2 with the
                                  current node's children in registers, 4 = 2 with the top nodes in
                                  LDS, 5 = 2 with two rays per lane */
    uint32_t block;            /* threads per workgroup (multiple of 64, at most 256) */
    uint32_t waves_per_cu;     /* persistent grid size: CUs x this many waves */
    uint32_t regroup_interval; /* persistent: traversal steps between service-pass checks */
    uint32_t refill_lanes;     /* persistent: run a service pass once at least this many lanes (ray
                                  slots for variant 5) wait: finished, between SubDAGs or idle */
    uint32_t max_iters;        /* per-ray iteration cap (0 = built-in 1<<22) */
    uint32_t lds_nodes;        /* variant 4: internal nodes [256, 256 + lds_nodes) read from a copy in
                                  LDS (the top of a breadth-first store; 0 = 512) */
    uint32_t bin_bits;         /* variants 2-5, cqx_intersect only: 0 = rays are taken in array order;
                                  1..4 = a prepass first bins the rays by origin into 2^(3 bin_bits)
                                  Morton-ordered cells over the origins' bounds (sampled, within the
                                  volume's SubDAG bounds), and the
                                  persistent waves take them cell by cell (each XCD a contiguous run
                                  of cells), so that rays in flight together read nearby nodes.
                                  Results are the same, in the caller's order. n < 2^32. */
    uint32_t stack_rows;       /* variants 2 and 3: 0 = a traversal stack row per node height (the
                                  volume's SubDAG height + 1, raytracing.cpp:251's stack); a power
                                  of two K >= 2 below that = a short stack of K rows per lane in LDS,
                                  reused modulo K, a pop past them descending again from the SubDAG
                                  root along the ray's position. Same records, less LDS per wave
                                  (occupancy for deep volumes). */
} cqx_launch_cfg;

/* intersectVolume over n rays, batched: rays and hits are DEVICE pointers; stream is a
 * hipStream_t (NULL = default stream). Asynchronous with respect to the host. */
cqx_status cqx_intersect(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                         uint32_t flags, float max_footprint, cqx_hit* hits,
                         const cqx_launch_cfg* cfg, void* stream);
/* cqx_intersect that allocates nothing (no counterpart in the reference: for callers that cannot
 * allocate device memory, such as code under HIP graph capture, tight serving loops and harnesses
 * that forbid hipMalloc). Same arguments, semantics and records as cqx_intersect. Instead of the
 * context's buffers, the work counters and the ray-binning buffers live in one static device
 * array per GPU (CQX_NOALLOC_SCRATCH_BYTES, held from cqx_create until the process ends), which
 * every call on that GPU shares: each call clears what must start at zero, and calls that overlap
 * on one GPU (any stream, any context) race. Ray binning (cfg->bin_bits > 0, also when the tuned
 * "auto" launch bins) takes at most CQX_NOALLOC_MAX_BINNED_RAYS rays; more fail with
 * CQX_ERR_LIMIT and nothing is launched. */
#define CQX_NOALLOC_MAX_BINNED_RAYS (1u << 22)
#define CQX_NOALLOC_SCRATCH_BYTES (1024u + 4096u * 4u + 256u * 4096u * 4u + 6u * CQX_NOALLOC_MAX_BINNED_RAYS)
cqx_status cqx_intersect_noalloc(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                                 uint32_t flags, float max_footprint, cqx_hit* hits,
                                 const cqx_launch_cfg* cfg, void* stream);
/* Launch autotuner (no reference counterpart). Times kernel variant, rays per workgroup,
 * persistent waves per CU, regroup interval and refill threshold on these rays (device pointers;
 * `hits` is overwritten), picks the fastest by coordinate descent and caches it in the context
 * for this volume depth and batch-size class; later cqx_intersect calls with cfg == NULL or
 * variant 0 use it, also on volumes and batches within 2 of the tuned depth and class (an edit
 * past the volume's bounds deepens its SubDAG by one). cqx_pathtrace and cqx_demo_trace have
 * their own "auto" launches. Optional outputs: the chosen configuration and its median time in
 * ms. */
cqx_status cqx_autotune(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                        uint32_t flags, float max_footprint, cqx_hit* hits, void* stream,
                        cqx_launch_cfg* best, float* best_ms);
cqx_status cqx_autotune_clear(cqx_ctx* ctx);

/* Many (volume, ray batch) pairs in ONE persistent launch: the batches are laid end to end in
 * one ray-id space that the persistent waves drain together, so batches of very different sizes
 * and volumes of different depths share the GPU without a launch per batch. Per-ray semantics
 * are intersectVolume's (raytracing.cpp:397-478). rays/hits are device pointers. One batched
 * launch in flight per context. (No reference counterpart: the reference raycasts one ray per
 * call.) */
typedef struct {
    const cqx_dvolume* volume;
    const cqx_ray* rays;
    cqx_hit* hits;
    uint64_t n;
    uint32_t flags;
    float max_footprint;
} cqx_batch;

cqx_status cqx_intersect_batched(cqx_ctx* ctx, const cqx_batch* batches, uint32_t n_batches,
                                 const cqx_launch_cfg* cfg, void* stream);

/* traceRayRef (src/library/raytracing.cpp:480-535 with visitVolumeNodes, utility.h:60-116): the
 * reference's brute-force raycast, one ray per GPU lane. Depth-first over every node whose box
 * dilated by 0.5 the ray pierces (FP64 slab test, entry and exit stored as float); the hit is the
 * first non-empty material node with the smallest entry among those with exit > 0. Fills hit,
 * material and distance (DBL_MAX on a miss); position and normal are 0, as the reference leaves
 * them. rays and out are device pointers. Slow by design: a checker, not the traversal. */
cqx_status cqx_trace_ref(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                         cqx_intersection* out, void* stream);
/* Same with host buffers: copies in, runs, copies out, synchronizes. */
cqx_status cqx_trace_ref_host(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                              cqx_intersection* out);

/* Same with host buffers: copies in, runs, copies out, synchronizes. */
cqx_status cqx_intersect_host(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                              uint32_t flags, float max_footprint, cqx_hit* hits);

/* ---- path tracer ------------------------------------------------------------------------------
 * Replaces the reference's per-pixel CPU tracer loop (PathtracingDemo::raytrace / traceSingleRay,
 * src/application/commands/view/pathtracing_demo.cpp:150-229) and the GLSL tracer
 * (src/application/commands/view/glsl/pathtracing.frag:723-819) with a batched GPU renderer.
 * Per path (pixel p, sample s): PCG32 with initstate (seed << 32) + p and stream s; the camera
 * ray goes through a jittered point of the pixel; at each hit the path adds beta * emission and,
 * while fewer than `bounces` scatterings were made, continues in a cosine-weighted direction about
 * the face normal from position + normal * bounce_offset with beta *= albedo; escaping rays add
 * beta * sky. The image is linear RGB (height x width x 3 floats), each pixel the mean of its
 * samples. Material ids 1..n_materials index rows 0..n_materials-1 of the tables; others are
 * black. */
typedef struct {
    uint32_t width, height, spp, bounces;
    double eye[3], target[3], fov_y_degrees;
    uint32_t n_materials;
    const float* albedo;   /* host, n_materials x 3 */
    const float* emission; /* host, n_materials x 3 */
    float sky[3];          /* radiance of escaping rays (the reference background: 0.8 0.8 1.0) */
    float bounce_offset;   /* origin offset along the face normal (reference: 0.01) */
    float max_footprint;   /* LOD (maxFootprint); CQX_MAX_FOOTPRINT_DISABLED for full detail */
    uint64_t seed;
} cqx_pt_params;

/* mode 0: wavefront (traversal kernel per bounce, captured in a hipGraph, replayed per call);
 * mode 1: megakernel (one lane per path). Both give the same image bit for bit. `image` is a
 * device pointer; synchronizes `stream` before returning. cfg NULL or variant 0: the camera
 * rays are traced one per lane (64 per workgroup unless cfg sets block), the bounce rays on
 * persistent waves (24 per CU unless cfg sets waves_per_cu); the other fields of cfg apply as
 * given. An explicit variant applies to every depth. */
cqx_status cqx_pathtrace(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_pt_params* params, uint32_t mode,
                         const cqx_launch_cfg* cfg, float* image, void* stream);

/* Gradient of a scalar image loss with respect to the material tables, by path replay: the
 * paths of cqx_pathtrace are traced again (same RNG streams, wavefront form) with each vertex's
 * material recorded, then swept backwards; geometry and traversal are discrete and exact. d_image = dLoss/dimage (device, height x width
 * x 3); d_albedo, d_emission (device, n_materials x 3) receive dLoss/dalbedo and
 * dLoss/demission. Up to 64 bounces, as cqx_pathtrace. Synchronizes `stream` before returning. */
cqx_status cqx_pathtrace_adjoint(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_pt_params* params,
                                 const cqx_launch_cfg* cfg, const float* d_image, float* d_albedo,
                                 float* d_emission, void* stream);

/* ---- reference CPU-tracer shading (compatibility mode) ---------------------------------------
 * PathtracingDemo::traceSingleRay (src/application/commands/view/pathtracing_demo.cpp:150-194)
 * for one primary ray per pixel, with surfaceColour (:44-59, positionBasedNoise :34-42),
 * gatherLighting (:82-117: a sun shadow ray toward normalize(1, -2, 10) and a sky shadow ray
 * toward normalize(n + randomPointInUnitSphere()), from position + normal * 0.001) and one
 * reflected ray from position + normal * 0.01; misses give (0.8, 0.8, 1.0); hit colours are
 * raised to `gamma`. The FP32 arithmetic is the reference's, operation for operation, except
 * powf (the device's, within 2 ulp).
 * Random points: the reference draws from ONE generator shared by all pixels in scan order
 * (pathtracing_demo.cpp:32, :61-80), which makes every pixel depend on all earlier ones. Here
 * pixel p has its own generator state, uint32(bit_mix(seed * 2^32 + p + 1)), or 17 if that is 0,
 * and uses the reference's draw function and draw order. Images match the reference's
 * statistically, not bit for bit. A draw that finds no point inside the sphere in 1024 generator
 * steps gives (0, 0, 0); the reference would loop forever. Rays the traversal flags as
 * anomalies count as misses.
 * rays: device, n x 6 floats (the camera rays of PathtracingDemo::raytrace, :214-225); out_rgb:
 * device, n x 3 floats, overwritten or, with accumulate, added to (the demo's mImage +=).
 * cfg NULL or variant 0: the primaries one per lane (64 per workgroup unless cfg sets block),
 * the secondary rays on persistent waves (16 per CU unless cfg sets waves_per_cu); the other
 * fields of cfg apply as given. An explicit variant applies to every traversal of the frame.
 * Asynchronous on `stream`; one call in flight per context. n <= 2^30. */
typedef struct {
    const float* colours; /* host, 256 x 3: the material colours (cqx_materials_colour_table) */
    float max_footprint;  /* maxFootprint of every raycast (the demo's default: 0.0035) */
    float gamma;          /* exponent of hit colours; 0 = the reference's 1/2.2 */
    uint32_t add_noise;   /* the demo's addNoise (default on) */
    uint32_t include_sun; /* includeSun (default on) */
    uint32_t include_sky; /* includeSky (default on) */
    uint32_t accumulate;  /* 1: out_rgb += colour, 0: out_rgb = colour */
    uint64_t seed;        /* per-frame seed of the per-pixel generators */
} cqx_demo_params;

cqx_status cqx_demo_trace(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n,
                          const cqx_demo_params* params, const cqx_launch_cfg* cfg, float* out_rgb, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CQX_H */
