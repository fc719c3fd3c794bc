// Launch autotuner for the traversal kernels.
//
// Tunes, on the caller's own rays and volume, the kernel variant (one ray per lane / persistent
// with refill), the ray binning grid (none, or 4^3 to 16^3 cells), rays per workgroup, persistent
// waves per CU, the regroup interval (traversal steps between refill checks) and the refill
// threshold (idle lanes that trigger a refill). The binning prepass is inside the timed span. Each
// candidate is timed with HIP events (median of 3 after a warm-up); the search is coordinate
// descent, two sweeps, from variant 2, 256 threads, 32 waves/CU, regroup 16, refill 16 (the
// built-in defaults with regroup 16: started at regroup 8, the first variant sweep could settle in
// variant 3, ~3 % slower on the 1024^3 incoherent workload, and not leave it). Results are cached in the context per
// (SubDAG height, batch-size class = floor(log2 n)) and used by every cqx_intersect whose
// configuration is left at "auto" (cfg == NULL or cfg->variant == 0), also for volumes and batches
// within 2 of a tuned height and class (tuned_cfg).
//
// The traversal stack depth is tuned too, for the persistent variants 2 and 3: the full stack
// (max SubDAG height + 1 LDS rows per lane) or a short one of 4, 8 or 16 rows, reused modulo
// its depth, whose pops past the kept rows descend again from the SubDAG root. Both visit the
// same nodes in the same order (restart descents are not node entries), so the choice changes
// LDS per wave and time only.
#include "device_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using cqx::fail;

namespace cqx {

int batch_class(uint64_t n)
{
    int c = 0;
    while ((uint64_t(1) << (c + 1)) <= n && c < 40) c++;
    return c;
}

// Entries within 2 of the SubDAG height and of the batch-size class qualify; the nearest height
// wins, then the nearest class. Heights other than the exact one matter after edits: an edit
// that reaches past the volume's bounds starts the SubDAG one level higher (and adds SubDAGs in
// other octants), and the launch tuned on the volume before the edit suits it far better than
// the built-in defaults (at 16384^3: 34 against 60 ms per 2^26 rays).
const cqx_launch_cfg* tuned_cfg(const cqx_ctx* ctx, int height, uint64_t n)
{
    const int cls = batch_class(n);
    const cqx_ctx::Tuned* best = nullptr;
    int best_d = 1 << 30;
    for (const auto& t : ctx->tuned) {
        const int dh = std::abs(t.height - height), dc = std::abs(t.cls - cls);
        if (dh > 2 || dc > 2) continue;
        const int d = dh * 8 + dc;
        if (d < best_d) {
            best_d = d;
            best = &t;
        }
    }
    return best ? &best->cfg : nullptr;
}

} // namespace cqx

namespace {

cqx_status time_cfg(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n, uint32_t flags, float fp,
                    cqx_hit* hits, const cqx_launch_cfg& c, hipStream_t s, float& ms)
{
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    std::vector<float> t;
    cqx_status st = CQX_OK;
    for (int rep = -1; rep < 3 && st == CQX_OK; rep++) {
        (void)hipEventRecord(a, s);
        st = cqx::launch_intersect(ctx, vol, rays, n, nullptr, flags, fp, hits, &c, s);
        (void)hipEventRecord(b, s);
        if (st == CQX_OK && hipEventSynchronize(b) == hipSuccess) {
            float x = 0;
            (void)hipEventElapsedTime(&x, a, b);
            if (rep >= 0) t.push_back(x);
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st != CQX_OK) return st;
    if (t.size() != 3) return fail(CQX_ERR_HIP, "timing failed");
    std::sort(t.begin(), t.end());
    ms = t[1];
    return CQX_OK;
}

} // namespace

extern "C" {

cqx_status cqx_autotune(cqx_ctx* ctx, const cqx_dvolume* vol, const cqx_ray* rays, uint64_t n, uint32_t flags,
                        float max_footprint, cqx_hit* hits, void* stream, cqx_launch_cfg* best_out, float* best_ms)
{
    const cqx::TraceRange range("cqx_autotune");
    if (!ctx || !vol || !rays || !hits) return fail(CQX_ERR_INVALID_ARG, "null argument");
    if (n == 0) return fail(CQX_ERR_INVALID_ARG, "no rays to tune on");
    if (flags & ~CQX_SURFACE_PROPERTIES) return fail(CQX_ERR_INVALID_ARG, "unknown flags");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = hipStream_t(stream);

    cqx_launch_cfg cur{};
    cur.variant = 2;
    cur.block = 256;
    cur.waves_per_cu = 32;
    cur.regroup_interval = 16;
    cur.refill_lanes = 16;
    float cur_ms = 0;
    cqx_status st = time_cfg(ctx, vol, rays, n, flags, max_footprint, hits, cur, s, cur_ms);
    if (st != CQX_OK) return st;

    const uint32_t blocks[] = { 64, 128, 256 };
    const uint32_t waves[] = { 4, 8, 12, 16, 20, 24, 28, 32 };
    const uint32_t regroups[] = { 1, 4, 8, 16, 32, 64 };
    const uint32_t refills[] = { 8, 16, 32, 48, 64 };
    auto sweep = [&](uint32_t cqx_launch_cfg::*field, const uint32_t* vals, size_t count) -> cqx_status {
        for (size_t i = 0; i < count; i++) {
            if (cur.*field == vals[i]) continue;
            cqx_launch_cfg c = cur;
            c.*field = vals[i];
            if (c.stack_rows && c.variant != 2 && c.variant != 3) continue; // short stacks: variants 2, 3
            if (c.stack_rows > uint32_t(vol->max_height)) continue;          // not shorter than the full stack
            float ms = 0;
            const cqx_status r = time_cfg(ctx, vol, rays, n, flags, max_footprint, hits, c, s, ms);
            if (r == CQX_ERR_LIMIT) continue; // e.g. the LDS stack does not fit this block size
            if (r != CQX_OK) return r;
            if (ms < cur_ms) {
                cur = c;
                cur_ms = ms;
            }
        }
        return CQX_OK;
    };
    const uint32_t variants[] = { 2, 3, 5 };
    const uint32_t bins[] = { 0, 2, 3, 4 };
    const uint32_t stacks[] = { 0, 4, 8, 16 };
    for (int pass = 0; pass < 2; pass++) {
        if ((st = sweep(&cqx_launch_cfg::variant, variants, 3)) != CQX_OK) return st;
        if ((st = sweep(&cqx_launch_cfg::bin_bits, bins, 4)) != CQX_OK) return st;
        if ((st = sweep(&cqx_launch_cfg::waves_per_cu, waves, 8)) != CQX_OK) return st;
        if ((st = sweep(&cqx_launch_cfg::regroup_interval, regroups, 6)) != CQX_OK) return st;
        if ((st = sweep(&cqx_launch_cfg::refill_lanes, refills, 5)) != CQX_OK) return st;
        if ((st = sweep(&cqx_launch_cfg::block, blocks, 3)) != CQX_OK) return st;
        if ((st = sweep(&cqx_launch_cfg::stack_rows, stacks, 4)) != CQX_OK) return st;
    }
    // the one-ray-per-lane kernel as the alternative
    for (uint32_t b : blocks) {
        cqx_launch_cfg c{};
        c.variant = 1;
        c.block = b;
        float ms = 0;
        st = time_cfg(ctx, vol, rays, n, flags, max_footprint, hits, c, s, ms);
        if (st == CQX_ERR_LIMIT) continue;
        if (st != CQX_OK) return st;
        if (ms < cur_ms) {
            cur = c;
            cur_ms = ms;
        }
    }
    const int height = vol->max_height;
    const int cls = cqx::batch_class(n);
    bool stored = false;
    for (auto& t : ctx->tuned)
        if (t.height == height && t.cls == cls) {
            t.cfg = cur;
            t.ms = cur_ms;
            stored = true;
        }
    if (!stored) ctx->tuned.push_back({ height, cls, cur, cur_ms });
    if (best_out) *best_out = cur;
    if (best_ms) *best_ms = cur_ms;
    return CQX_OK;
}

cqx_status cqx_autotune_clear(cqx_ctx* ctx)
{
    if (!ctx) return fail(CQX_ERR_INVALID_ARG, "null context");
    ctx->tuned.clear();
    return CQX_OK;
}

} // extern "C"
