// om_batch.cpp -- batched mode: many scenes (fields, sun positions) traced by one persistent
// launch from a (scene, heliostat, sample block) work queue, the decode / flux / smoothing of
// every map behind it, the whole sequence captured once in a hipGraph; and the autotuner of
// the launch parameters.
//
// No reference counterpart: the reference traces one scene at a time from a host loop over
// heliostats (recthelio_tracing.cpp:46-79).  Each map of a batch equals the map om_trace_plain
// produces for that scene alone (same Philox streams, integer accumulation).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "om_internal.h"

namespace {

constexpr size_t BATCH_PART_BYTES = (size_t)8 << 30; // budget of the packed map copies of a batch
// copies of one map at most (words): more copies spread the same-pixel atomics of small maps
// that take many hits (64 copies of a 140 x 140 map), while a 256 x 256 map keeps 16
constexpr size_t SCENE_COPY_WORDS = (size_t)5 << 18;

template <class T>
om_status dalloc(T *&p, size_t n)
{
    p = nullptr;
    if (n == 0) n = 1;
    OM_HIP(hipMalloc((void **)&p, sizeof(T) * n));
    return OM_OK;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

} // namespace

struct om_batch {
    om_ctx *ctx = nullptr;
    int n = 0;
    int sunshape = 0;
    bool poisson = false;
    int wgs = 256, rpt = 8;
    bool per_scene = false;                     // OM_BATCH_PER_SCENE: window traces scene by scene
    bool window = false;                        // OM_BATCH_WINDOW: persistent launch with LDS windows
    int n_items = 0, nblocks = 0, max_npx = 0;
    long long rays = 0;
    std::vector<om_scene *> scenes;
    std::vector<om_flux> outs;
    std::vector<float *> raw;                   // per scene, for smoothing
    om::TraceArgs *d_table = nullptr;
    // OM_BATCH_WINDOW over OM_PREC_MIXED scenes: one MixLaunch per entry, the entries' redo
    // lists and counters (4 words per entry, then a flag per scene), each entry's scene and
    // range of chunks
    bool mixed = false;
    void *d_mtable = nullptr;
    uint2 *d_redo = nullptr;
    uint32_t *d_mixc = nullptr;
    int *d_entry_scene = nullptr;
    int2 *d_entry_chunks = nullptr;
    int n_entries = 0;
    int4 *d_items = nullptr;                    // work items, or OM_BATCH_WINDOW's chunks
    unsigned int *d_queue = nullptr;            // queue block: one counter, or OM_BATCH_WINDOW's 8 queues
    om::ReduceSlot *d_slots = nullptr;
    unsigned long long *d_parts = nullptr;      // copies + carry + wrap of every scene
    float *d_raw = nullptr;
    // OM_BATCH_PER_SCENE: the batch's own trace workspace (the largest any of its scenes
    // needs) and smoothing input, so that the captured graph holds no pointer into buffers
    // of the context, which later calls may reallocate
    void *d_ws = nullptr;
    size_t ws_bytes = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void batch_free(om_batch *b)
{
    {
        // om_ctx_progress no longer reads this batch's queue block
        std::lock_guard<std::mutex> lk(b->ctx->mon_mu);
        if (b->ctx->cur_queue == b->d_queue) b->ctx->cur_queue = b->ctx->d_queue;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    dfree(b->d_table);
    if (b->d_mtable) (void)hipFree(b->d_mtable);
    dfree(b->d_redo);
    dfree(b->d_mixc);
    dfree(b->d_entry_scene);
    dfree(b->d_entry_chunks);
    dfree(b->d_items);
    dfree(b->d_queue);
    dfree(b->d_slots);
    dfree(b->d_parts);
    dfree(b->d_raw);
    if (b->d_ws) (void)hipFree(b->d_ws);
    delete b;
}

// The launch sequence of a batch: queue reset (a fill kernel, so that it can sit in a
// graph), trace, decode + flux, smoothing of each map that asks for it.
om_status batch_enqueue(om_batch *b, hipStream_t st)
{
    om_status rc;
    if (b->per_scene) {
        // om_trace_plain of each scene with the LDS-window accumulation, in the batch's
        // own workspace
        om::Workspace ws;
        ws.p = b->d_ws;
        ws.bytes = b->ws_bytes;
        for (int i = 0; i < b->n; ++i) {
            om_scene *s = b->scenes[i];
            const int keep = s->prm.accumulate;
            s->prm.accumulate = OM_ACC_WINDOW;
            rc = om::trace_into(s, 0, s->H, &b->outs[i], true, st, ws);
            s->prm.accumulate = keep;
            if (rc != OM_OK) return rc;
            if ((b->outs[i].flux || b->outs[i].flux_smooth) &&
                (rc = om::flux_finalize(s, &b->outs[i], b->d_raw, st)) != OM_OK)
                return rc;
        }
        return OM_OK;
    }
    if ((rc = om::launch_queue_reset(b->d_queue, b->n_items, b->window ? om::QUEUE_EIGHTHS : om::QUEUE_ONE, st)) !=
        OM_OK)
        return rc;
    if (b->mixed) {
        om::MixBatch m;
        m.sunshape = b->sunshape;
        m.poisson = b->poisson;
        m.table = b->d_mtable;
        m.chunks = b->d_items;
        m.n_chunks = b->n_items;
        m.queue = b->d_queue;
        m.nwg = b->nblocks;
        m.entry_scene = b->d_entry_scene;
        m.entry_chunks = b->d_entry_chunks;
        m.n_entries = b->n_entries;
        m.counters = b->d_mixc;
        m.slots = b->d_slots;
        m.n_scenes = b->n;
        if ((rc = om::launch_trace_batch_mixed(m, st)) != OM_OK) return rc;
    } else if (b->window) {
        if (b->n_items > 0 &&
            (rc = om::launch_trace_batch_window(b->sunshape, b->poisson, b->d_table, b->d_items, b->n_items,
                                                b->d_queue, b->nblocks, st)) != OM_OK)
            return rc;
    } else if (b->n_items > 0 &&
        (rc = om::launch_trace_batch(b->sunshape, b->poisson, b->wgs, b->d_table, b->d_items, b->n_items,
                                     b->d_queue, b->nblocks, st)) != OM_OK)
        return rc;
    if ((rc = om::launch_reduce_batch(b->d_slots, b->n, b->max_npx, st)) != OM_OK) return rc;
    for (int i = 0; i < b->n; ++i) {
        const om_scene *s = b->scenes[i];
        if (!b->outs[i].flux_smooth) continue;
        if ((rc = om::launch_smooth(b->raw[i], b->outs[i].flux_smooth, s->res_x, s->res_y, s->prm.smooth_radius,
                                    s->prm.smooth_ratio, s->scale, st)) != OM_OK)
            return rc;
    }
    return OM_OK;
}

om_status batch_build(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs, const om_launch *lp,
                      om_batch **out)
{
    *out = nullptr;
    if (!ctx || !scenes || !outs || n <= 0) {
        om::set_error("batch needs a context, scenes and outputs");
        return OM_ERR_INVALID;
    }
    om_batch *b = new om_batch();
    b->ctx = ctx;
    b->n = n;
    b->wgs = (lp && lp->workgroup) ? lp->workgroup : 256;
    b->rpt = (lp && lp->rays_per_thread) ? lp->rays_per_thread : 0;
    if (b->wgs != 128 && b->wgs != 256 && b->wgs != 512 && b->wgs != OM_BATCH_PER_SCENE &&
        b->wgs != OM_BATCH_WINDOW) {
        om::set_error("batch workgroup size must be 128, 256, 512, OM_BATCH_PER_SCENE or OM_BATCH_WINDOW");
        delete b;
        return OM_ERR_INVALID;
    }
    if (b->rpt < 0 || b->rpt > om::RPT_LIMIT) {
        om::set_error("rays per thread must be 1..8 (0 = automatic)");
        delete b;
        return OM_ERR_INVALID;
    }
    for (int i = 0; i < n; ++i) {
        const om_scene *s = scenes[i];
        const om_flux &f = outs[i];
        if (!s || s->ctx != ctx) {
            om::set_error("every scene of a batch must be built on the batch's context");
            delete b;
            return OM_ERR_INVALID;
        }
        if (!f.counts || !f.energy || f.capacity < (int64_t)s->res_x * s->res_y) {
            om::set_error("batch outputs need counts and energy of at least the map size");
            delete b;
            return OM_ERR_SIZE;
        }
        if (s->prm.precision != OM_PREC_FP32 && b->wgs != OM_BATCH_PER_SCENE &&
            !(s->prm.precision == OM_PREC_MIXED && b->wgs == OM_BATCH_WINDOW)) {
            om::set_error("the persistent batch launch of 128, 256 or 512 threads traces in FP32 geometry: "
                          "mixed-precision scenes need OM_BATCH_WINDOW or OM_BATCH_PER_SCENE, FP64 scenes "
                          "OM_BATCH_PER_SCENE");
            delete b;
            return OM_ERR_UNSUPPORTED;
        }
        if (i > 0 && b->wgs == OM_BATCH_WINDOW && s->prm.precision != scenes[0]->prm.precision) {
            om::set_error("scenes of a persistent batch launch must share the precision");
            delete b;
            return OM_ERR_UNSUPPORTED;
        }
        if (i > 0 && (s->prm.sunshape != scenes[0]->prm.sunshape ||
                      s->prm.micro_mode != scenes[0]->prm.micro_mode)) {
            om::set_error("scenes of a batch must share the sunshape and microsample kind");
            delete b;
            return OM_ERR_UNSUPPORTED;
        }
        b->scenes.push_back(scenes[i]);
        b->outs.push_back(f);
        b->rays += s->total_rays;
        b->max_npx = std::max(b->max_npx, s->res_x * s->res_y);
    }
    b->sunshape = scenes[0]->prm.sunshape;
    b->poisson = scenes[0]->prm.micro_mode == OM_MICRO_POISSON;
    b->window = b->wgs == OM_BATCH_WINDOW;
    b->mixed = b->window && scenes[0]->prm.precision == OM_PREC_MIXED;
    if (b->window) {
        b->rpt = 0;
    } else if (b->rpt == 0) {
        // enough items to fill 256 CUs many times over
        b->rpt = om::RPT_LIMIT;
        while (b->rpt > 1 && b->rays / ((long long)b->wgs * b->rpt) < 16384) b->rpt /= 2;
    }
    om_status rc = OM_OK;
    auto fail = [&](om_status e) {
        batch_free(b);
        return e;
    };
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(OM_ERR_HIP);
    if (b->wgs == OM_BATCH_PER_SCENE) {
        b->per_scene = true;
        b->n_items = n;
        // workspace and smoothing input of the largest scene, cleared once (each decode
        // leaves the words it reads cleared)
        for (int i = 0; i < n; ++i) {
            om_scene *s = scenes[i];
            const int keep = s->prm.accumulate;
            s->prm.accumulate = OM_ACC_WINDOW;
            const size_t need = om::trace_workspace_bytes(s);
            s->prm.accumulate = keep;
            if (need == 0) return fail(OM_ERR_UNSUPPORTED);
            b->ws_bytes = std::max(b->ws_bytes, need);
        }
        if (hipMalloc(&b->d_ws, b->ws_bytes) != hipSuccess)
            return fail(om::hip_fail(hipGetLastError(), "hipMalloc (batch workspace)"));
        if ((rc = dalloc(b->d_raw, (size_t)b->max_npx)) != OM_OK) return fail(rc);
        if ((rc = om::launch_fill_u32((uint32_t *)b->d_ws, (long long)(b->ws_bytes / 4), nullptr)) != OM_OK)
            return fail(rc);
        if (hipDeviceSynchronize() != hipSuccess) return fail(om::hip_fail(hipGetLastError(), "batch set-up"));
        *out = b;
        return OM_OK;
    }
    // work items: (table entry, heliostat, first ray) blocks of wgs * rpt rays, scene by
    // scene; a table entry is one grid of one scene (its heliostats walk that grid)
    // OM_BATCH_WINDOW: chunks (entry, heliostat, first ray, rays) of at most the window
    // kernels' chunk size, in the same order
    std::vector<int4> items;
    std::vector<int> entry_scene, entry_grid;
    std::vector<int2> entry_chunks;
    std::vector<long long> entry_rays;
    const int block = b->window ? om::window_chunk_rays() : b->wgs * b->rpt;
    for (int i = 0; i < n; ++i) {
        const om_scene *s = scenes[i];
        for (size_t g = 0; g < s->grids.size(); ++g) {
            const int e = (int)entry_scene.size();
            entry_scene.push_back(i);
            entry_grid.push_back((int)g);
            const int c0 = (int)std::min<size_t>(items.size(), INT32_MAX);
            long long rays = 0;
            for (int h = s->grids[g].h0; h < s->grids[g].h1; ++h) {
                const long long nr = (long long)s->mcount[h] * s->prm.rays_per_sample;
                rays += nr;
                for (long long r = 0; r < nr; r += block)
                    items.push_back(make_int4(e, h, (int)r, (int)std::min<long long>(block, nr - r)));
            }
            entry_chunks.push_back(make_int2(c0, (int)std::min<size_t>(items.size(), INT32_MAX)));
            entry_rays.push_back(rays);
        }
    }
    const int n_entries = (int)entry_scene.size();
    if (items.size() >= (size_t)1 << 31) return fail((om::set_error("batch has too many work items"), OM_ERR_UNSUPPORTED));
    b->n_items = (int)items.size();
    // packed map copies: 8 << shift per scene within the budget
    long long sum_px = 0;
    for (int i = 0; i < n; ++i) sum_px += (long long)scenes[i]->res_x * scenes[i]->res_y;
    int shift = 0;
    while (shift < 3 && ((size_t)16 << shift) * sum_px * sizeof(unsigned long long) <= BATCH_PART_BYTES &&
           ((size_t)16 << shift) * (size_t)b->max_npx <= SCENE_COPY_WORDS)
        ++shift;
    const int ncopy = 8 << shift;
    const size_t words = (size_t)sum_px * ncopy + (size_t)sum_px; // copies + (carry, wrap) as u32 pairs
    if ((rc = dalloc(b->d_parts, words)) != OM_OK) return fail(rc);
    if ((rc = dalloc(b->d_raw, (size_t)sum_px)) != OM_OK) return fail(rc);
    if ((rc = dalloc(b->d_table, (size_t)n_entries)) != OM_OK) return fail(rc);
    if ((rc = dalloc(b->d_items, items.size())) != OM_OK) return fail(rc);
    if ((rc = dalloc(b->d_queue, om::QUEUE_WORDS)) != OM_OK) return fail(rc);
    if ((rc = dalloc(b->d_slots, (size_t)n)) != OM_OK) return fail(rc);
    std::vector<om::TraceArgs> table(n_entries);
    std::vector<om::ReduceSlot> slots(n);
    unsigned long long *copies = b->d_parts;
    uint32_t *side = (uint32_t *)(b->d_parts + (size_t)sum_px * ncopy);
    size_t px_off = 0;
    for (int i = 0; i < n; ++i) {
        om_scene *s = scenes[i];
        const int npx = s->res_x * s->res_y;
        om::TraceArgs a = om::make_trace_args(s, 0);
        a.rays_per_thread = b->rpt;
        a.tile_rays = b->wgs * b->rpt;
        a.acc_packed = 1;
        a.part = copies + px_off * ncopy;
        a.carry = side + 2 * px_off;
        a.wrap = a.carry + npx;
        a.npx = npx;
        a.copy_shift = shift;
        a.copy_stride = npx;
        a.px_stride = 1;
        a.pack_shift = s->prm.pack_shift > 0 ? s->prm.pack_shift : 44;
        om::window_rect(s->res_x, s->res_y, a.win_x0, a.win_y0, a.win_w, a.win_h);
        for (int e = 0; e < n_entries; ++e) {
            if (entry_scene[e] != i) continue;
            // the entry's grid; maps and launch shape as above
            table[e] = a;
            om::set_trace_grid(table[e], s, entry_grid[e]);
        }
        om::ReduceSlot r;
        r.part = a.part;
        r.carry = a.carry;
        r.wrap = a.wrap;
        r.E = a.pack_shift;
        r.npx = npx;
        r.ncopy = ncopy;
        r.copy_stride = npx;
        r.px_stride = 1;
        r.counts = outs[i].counts;
        r.energy = (unsigned long long *)outs[i].energy;
        r.raw = b->d_raw + px_off;
        r.flux = outs[i].flux;
        r.scale = s->scale;
        slots[i] = r;
        b->raw.push_back(r.raw);
        px_off += (size_t)npx;
    }
    if (b->mixed) {
        // each entry's redo list: one entry per 64 rays, 2^12 to 2^24 (om_params.redo_cap
        // when set); a list that fills up sends its whole scene to the FP64 trace
        b->n_entries = n_entries;
        std::vector<size_t> redo_off(n_entries);
        size_t redo_total = 0;
        std::vector<int> caps(n_entries);
        for (int e = 0; e < n_entries; ++e) {
            const om_scene *s = scenes[entry_scene[e]];
            caps[e] = s->prm.redo_cap > 0
                          ? s->prm.redo_cap
                          : (int)std::min<long long>(std::max<long long>(entry_rays[e] / 64, 1 << 12), 1 << 24);
            redo_off[e] = redo_total;
            redo_total += (size_t)caps[e];
        }
        const size_t mlb = om::mix_launch_bytes();
        std::vector<unsigned char> mt(mlb * (size_t)n_entries);
        if ((rc = dalloc(b->d_redo, redo_total)) != OM_OK) return fail(rc);
        if ((rc = dalloc(b->d_mixc, (size_t)4 * n_entries + n)) != OM_OK) return fail(rc);
        if ((rc = dalloc(b->d_entry_scene, (size_t)n_entries)) != OM_OK) return fail(rc);
        if ((rc = dalloc(b->d_entry_chunks, (size_t)n_entries)) != OM_OK) return fail(rc);
        if (hipMalloc(&b->d_mtable, mt.size()) != hipSuccess)
            return fail(om::hip_fail(hipGetLastError(), "hipMalloc (batch table)"));
        for (int e = 0; e < n_entries; ++e)
            if ((rc = om::mix_batch_entry(scenes[entry_scene[e]], entry_grid[e], table[e], b->d_redo + redo_off[e],
                                          caps[e], b->d_mixc + 4 * (size_t)e, mt.data() + mlb * (size_t)e)) != OM_OK)
                return fail(rc);
        if (hipMemcpy(b->d_mtable, mt.data(), mt.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->d_entry_scene, entry_scene.data(), sizeof(int) * n_entries, hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(b->d_entry_chunks, entry_chunks.data(), sizeof(int2) * n_entries, hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemset(b->d_mixc, 0, sizeof(uint32_t) * ((size_t)4 * n_entries + n)) != hipSuccess)
            return fail(om::hip_fail(hipGetLastError(), "hipMemcpy (mixed batch tables)"));
    }
    if (hipMemcpy(b->d_table, table.data(), sizeof(om::TraceArgs) * n_entries, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_slots, slots.data(), sizeof(om::ReduceSlot) * n, hipMemcpyHostToDevice) != hipSuccess ||
        (!items.empty() && hipMemcpy(b->d_items, items.data(), sizeof(int4) * items.size(),
                                     hipMemcpyHostToDevice) != hipSuccess))
        return fail(om::hip_fail(hipGetLastError(), "hipMemcpy (batch tables)"));
    // the decode leaves every word cleared; clear the new buffer once
    if ((rc = om::launch_fill_u32((uint32_t *)b->d_parts, (long long)(words * 2), nullptr)) != OM_OK) return fail(rc);
    if (hipDeviceSynchronize() != hipSuccess) return fail(om::hip_fail(hipGetLastError(), "batch set-up"));
    // persistent workgroups: as many as are resident at once
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return fail(OM_ERR_HIP);
    int per_cu = b->mixed    ? om::trace_batch_window_mixed_blocks_per_cu(b->sunshape, b->poisson)
                 : b->window ? om::trace_batch_window_blocks_per_cu(b->sunshape, b->poisson)
                             : om::trace_batch_blocks_per_cu(b->sunshape, b->poisson, b->wgs);
    if (b->window && per_cu < 1)
        return fail((om::set_error("trace_batch_window_kernel cannot be resident on this device"), OM_ERR_UNSUPPORTED));
    if (per_cu < 1) per_cu = 1;
    b->nblocks = std::max(1, std::min(b->n_items, per_cu * prop.multiProcessorCount));
    *out = b;
    return OM_OK;
}

} // namespace

extern "C" om_status om_batch_create(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs,
                                     const om_launch *launch, om_batch **out)
{
    if (!out) return OM_ERR_INVALID;
    return batch_build(ctx, scenes, n, outs, launch, out);
}

extern "C" om_status om_batch_run(om_batch *b, void *stream)
{
    if (!b) return OM_ERR_INVALID;
    OM_HIP(hipSetDevice(b->ctx->device));
    if (!b->exec) {
        // capture once on a stream of the batch (the null stream cannot be captured)
        OM_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        OM_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        om_status rc = batch_enqueue(b, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (rc != OM_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return om::hip_fail(e, "hipStreamEndCapture");
        b->graph = g;
        OM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    om::set_current_queue(b->ctx, b->per_scene ? b->ctx->d_queue : b->d_queue);
    OM_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return OM_OK;
}

extern "C" om_status om_batch_info(const om_batch *b, om_launch *launch, int64_t *items, int64_t *rays)
{
    if (!b) return OM_ERR_INVALID;
    if (launch) {
        launch->workgroup = b->wgs;
        launch->rays_per_thread = b->rpt;
    }
    if (items) *items = b->n_items;
    if (rays) *rays = b->rays;
    return OM_OK;
}

extern "C" om_status om_batch_redo_count(const om_batch *b, int64_t *rays, int64_t *full_scenes)
{
    if (!b) return OM_ERR_INVALID;
    if (rays) *rays = 0;
    if (full_scenes) *full_scenes = 0;
    if (!b->mixed) return OM_OK;
    OM_HIP(hipSetDevice(b->ctx->device));
    OM_HIP(hipDeviceSynchronize());
    std::vector<uint32_t> c((size_t)4 * b->n_entries + b->n);
    OM_HIP(hipMemcpy(c.data(), b->d_mixc, sizeof(uint32_t) * c.size(), hipMemcpyDeviceToHost));
    int64_t r = 0, f = 0;
    for (int e = 0; e < b->n_entries; ++e) r += c[4 * (size_t)e];
    for (int i = 0; i < b->n; ++i) f += c[4 * (size_t)b->n_entries + i] != 0u;
    if (rays) *rays = r;
    if (full_scenes) *full_scenes = f;
    return OM_OK;
}

extern "C" om_status om_batch_destroy(om_batch *b)
{
    if (!b) return OM_ERR_INVALID;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    batch_free(b);
    return OM_OK;
}

extern "C" om_status om_trace_batched(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs,
                                      const om_launch *launch, void *stream)
{
    om_batch *b = nullptr;
    om_status rc = batch_build(ctx, scenes, n, outs, launch, &b);
    if (rc != OM_OK) return rc;
    om::set_current_queue(ctx, b->per_scene ? ctx->d_queue : b->d_queue);
    rc = batch_enqueue(b, (hipStream_t)stream);
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess && rc == OM_OK)
        rc = om::hip_fail(hipGetLastError(), "om_trace_batched");
    batch_free(b);
    return rc;
}

// Autotuner: every (workgroup size, rays per thread) candidate is timed on the batch itself
// (one warm-up run, then the best of `reps` graph replays, HIP events on a private stream);
// the fastest is returned with the whole table.  Maps are identical for every candidate,
// so only the time differs.
extern "C" om_status om_autotune(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs, int reps,
                                 om_tune *result)
{
    if (!ctx || !result) return OM_ERR_INVALID;
    std::memset(result, 0, sizeof(*result));
    if (reps <= 0) reps = 2;
    static const int wgs_list[] = {128, 256, 512, OM_BATCH_WINDOW, OM_BATCH_PER_SCENE};
    static const std::vector<int> rpt_list = {2, 4, 8}, rpt_none = {0};
    OM_HIP(hipSetDevice(ctx->device));
    hipStream_t st;
    OM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    OM_HIP(hipEventCreate(&e0));
    OM_HIP(hipEventCreate(&e1));
    om_status rc = OM_OK;
    float best = 0.0f;
    for (int w : wgs_list)
        for (int r : (w == OM_BATCH_PER_SCENE || w == OM_BATCH_WINDOW ? rpt_none : rpt_list)) {
            om_launch lp{w, r};
            om_batch *b = nullptr;
            if ((rc = batch_build(ctx, scenes, n, outs, &lp, &b)) != OM_OK) {
                // a launch form the scenes cannot use (FP64 scenes and the persistent launch)
                if (rc == OM_ERR_UNSUPPORTED && w != OM_BATCH_PER_SCENE) {
                    rc = OM_OK;
                    continue;
                }
                break;
            }
            float ms = 0.0f;
            rc = om_batch_run(b, st);
            for (int k = 0; k < reps && rc == OM_OK; ++k) {
                if (hipEventRecord(e0, st) != hipSuccess) rc = OM_ERR_HIP;
                if (rc == OM_OK) rc = om_batch_run(b, st);
                if (rc == OM_OK && hipEventRecord(e1, st) != hipSuccess) rc = OM_ERR_HIP;
                if (rc == OM_OK && hipEventSynchronize(e1) != hipSuccess) rc = OM_ERR_HIP;
                float t = 0.0f;
                if (rc == OM_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = OM_ERR_HIP;
                if (rc == OM_OK && (k == 0 || t < ms)) ms = t;
            }
            (void)hipStreamSynchronize(st);
            const long long rays = b->rays;
            batch_free(b);
            if (rc != OM_OK) break;
            int k = result->n_trials++;
            result->trials[k].workgroup = w;
            result->trials[k].rays_per_thread = r;
            result->trials[k].ms = ms;
            if (k == 0 || ms < best) {
                best = ms;
                result->best.workgroup = w;
                result->best.rays_per_thread = r;
                result->best_ms = ms;
                result->rays_per_s = ms > 0.0f ? (double)rays / (ms * 1e-3) : 0.0;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return rc;
}
