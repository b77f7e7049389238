// om_window.h -- the LDS-window accumulation of the trace kernels (OM_ACC_WINDOW), shared by
// the FP32 window kernel and the batched window kernel (om_kernels.hip) and the FP64 one
// (om_trace64.hip): packed words in LDS, each wave's buffer of hits outside the window, and
// the persistent workgroup body that takes chunks of tiles from 8 per-XCD queues.
#pragma once

#include "om_device.h"

namespace om {

// Add one hit (v = 2^E + q) to a packed word; the old value tells whether the add carried
// out of the energy field (bits 0..E-1) or wrapped the word, and both events are counted in
// the side maps of pixel px, which the decode applies.
__device__ __forceinline__ void packed_add_lds(const TraceArgs &a, unsigned long long *w, int px, uint32_t q)
{
    const int E = a.pack_shift;
    const unsigned long long lowmask = (1ull << E) - 1ull;
    const unsigned long long v = (1ull << E) + q;
    const unsigned long long old = __hip_atomic_fetch_add(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if ((old & lowmask) + q > lowmask) atomicAdd(a.carry + px, 1u);
    if (old + v < old) atomicAdd(a.wrap + px, 1u);
}

// ---------------------------------------------------------------------------------------
// OM_ACC_WINDOW: window of the packed map in LDS, persistent workgroups
// ---------------------------------------------------------------------------------------
// One workgroup per CU keeps the central block of the packed map in LDS (WIN_PX_MAX words
// with the wave buffers below: the whole 160 KiB) for every ray it traces.  Each of its
// waves traces rays: hits inside the window are LDS atomics; hits outside go to the wave's
// buffer in LDS, which reaches the packed per-XCD copies in device memory WIN_OUT / 2 or
// more hits at a time.  Vector memory operations of a wave complete in issue order, so a
// returning atomic issued per hit would hold back the wave's next loads until the memory
// side has carried it out; a flush waits for that once for up to WIN_OUT hits.
constexpr int WIN_WGS = 1024;
constexpr int WIN_WAVES = WIN_WGS / 64;
constexpr int WIN_OUT = 128;   // out-of-window buffer entries (px << 32 | q) per wave
constexpr int WIN_CHUNK = 2048; // rays a wave takes from a queue at a time
constexpr int WIN_PX_MAX = (160 * 1024 - WIN_WAVES * WIN_OUT * 8 - 256) / 8;
constexpr unsigned long long OUT_EMPTY = ~0ull;

// the per-XCD copy of this workgroup in a's packed map
__device__ __forceinline__ unsigned long long *copy_of(const TraceArgs &a)
{
    const uint32_t copy = (xcc_id() & 7u) | ((((uint32_t)blockIdx.x >> 3) & ((1u << a.copy_shift) - 1u)) << 3);
    return a.part + (size_t)copy * a.copy_stride;
}

// Adds w = (n << E) + e, one hit or a window word, to the copy of pixel px; the old value
// tells whether the add carried out of the energy field or wrapped the word (copy_check),
// which the side maps count as packed_add_lds does.
__device__ __forceinline__ unsigned long long copy_add(const TraceArgs &a, unsigned long long *part, int px,
                                                       unsigned long long w)
{
    return __hip_atomic_fetch_add(part + (size_t)px * a.px_stride, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void copy_check(const TraceArgs &a, int px, unsigned long long w, unsigned long long old)
{
    const unsigned long long lowmask = (1ull << a.pack_shift) - 1ull;
    if ((old & lowmask) + (w & lowmask) > lowmask) atomicAdd(a.carry + px, 1u);
    if (old + w < old) atomicAdd(a.wrap + px, 1u);
}

// A wave's buffer of hits outside the window; n is wave-uniform.  In the mixed trace it also
// holds the rays to trace again in FP64 (entries with REDO_TAG: the heliostat in bits 32..62,
// the ray of that heliostat in bits 0..31), which a flush appends to the redo list with one
// counter atomic for all of them: a returning atomic per flagged pass on the one counter of
// the whole launch would serialise there and hold back each wave's loads behind it.
constexpr unsigned long long REDO_TAG = 1ull << 63;
struct WaveOut {
    unsigned long long *buf; // [WIN_OUT] in LDS
    int n;
    const MixArgs *mx = nullptr; // the mixed trace's redo list (null: no REDO_TAG entries)
    int wx0 = 0, wy0 = 0;        // the origin of this workgroup's window (window_body)

    // The word of pixel (row, col) in this workgroup's window (columns modulo res_x), or -1
    // when the pixel lies outside it.
    __device__ __forceinline__ int window_word(const TraceArgs &a, int row, int col) const
    {
        const uint32_t wr = (uint32_t)(row - wy0);
        int c = col - wx0;
        if (c < 0) c += a.res_x;
        return wr < (uint32_t)a.win_h && (uint32_t)c < (uint32_t)a.win_w ? (int)wr * a.win_w + c : -1;
    }

    // Appends the REDO_TAG entries of e[] to the redo list and empties their slots; an entry
    // past the list's capacity is dropped (the launch is then traced again in FP64).
    __device__ __forceinline__ void flush_redo(unsigned long long (&e)[WIN_OUT / 64], int lane)
    {
        unsigned long long m[WIN_OUT / 64];
        uint32_t cnt = 0u;
#pragma unroll
        for (int b = 0; b < WIN_OUT / 64; ++b) {
            m[b] = __ballot(e[b] != OUT_EMPTY && (e[b] & REDO_TAG) != 0ull);
            cnt += (uint32_t)__popcll(m[b]);
        }
        if (cnt == 0u) return;
        uint32_t base = 0u;
        if (lane == 0) base = atomicAdd(mx->redo_n, cnt);
        base = __builtin_amdgcn_readfirstlane(base);
#pragma unroll
        for (int b = 0; b < WIN_OUT / 64; ++b) {
            if ((m[b] >> lane) & 1ull) {
                const uint32_t i = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m[b] >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)m[b], 0u));
                if (i < (uint32_t)mx->redo_cap)
                    mx->redo[i] = make_uint2((uint32_t)(e[b] >> 32) & 0x7fffffffu, (uint32_t)e[b]);
                e[b] = OUT_EMPTY;
            }
            base += (uint32_t)__popcll(m[b]);
        }
    }

    // Adds the buffered hits to the copies: all returning atomics first, then their checks.
    __device__ __forceinline__ void flush(const TraceArgs &a, int lane)
    {
        if (n == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); // the lanes' buffer stores are done
        unsigned long long *part = copy_of(a);
        const unsigned long long one = 1ull << a.pack_shift;
        unsigned long long e[WIN_OUT / 64], old[WIN_OUT / 64];
#pragma unroll
        for (int b = 0; b < WIN_OUT / 64; ++b) {
            const int i = b * 64 + lane;
            e[b] = i < n ? buf[i] : OUT_EMPTY;
        }
        if (mx) flush_redo(e, lane);
#pragma unroll
        for (int b = 0; b < WIN_OUT / 64; ++b) {
            old[b] = 0ull;
            if (e[b] != OUT_EMPTY) old[b] = copy_add(a, part, (int)(e[b] >> 32), one + (e[b] & 0xffffffffull));
        }
#pragma unroll
        for (int b = 0; b < WIN_OUT / 64; ++b)
            if (e[b] != OUT_EMPTY) copy_check(a, (int)(e[b] >> 32), one + (e[b] & 0xffffffffull), old[b]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); // read before the buffer is refilled
        n = 0;
    }

    // This pass's hits outside the window (out), in lane order; the whole wave takes part.
    // The buffer is flushed when the next pass could overflow it.
    __device__ __forceinline__ void push(const TraceArgs &a, bool out, int px, uint32_t q, int lane)
    {
        const unsigned long long m = __ballot(out);
        if (out) {
            const uint32_t j = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            buf[n + (int)j] = ((unsigned long long)px << 32) | q;
        }
        n += __popcll(m);
        if (n > WIN_OUT - 64) flush(a, lane);
    }

    // The mixed trace's rays to trace again in FP64 (flag), ray r of heliostat h.  The whole
    // wave takes part.
    __device__ __forceinline__ void push_redo(const TraceArgs &a, bool flag, int h, int r, int lane)
    {
        const unsigned long long m = __ballot(flag);
        if (m == 0ull) return;
        if (flag) {
            const uint32_t j = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            buf[n + (int)j] = REDO_TAG | ((unsigned long long)(uint32_t)h << 32) | (uint32_t)r;
        }
        n += __popcll(m);
        if (n > WIN_OUT - 64) flush(a, lane);
    }
};

// One wave takes the next item of a queue: lane 0 adds 1 to the counter and, in the same
// step (both in flight together), reads the abort word of om_ctx_abort, which the host
// copies into device memory; the system-scope load reads memory past the L2, where the
// copy lands.
// Returns the item index, or TAKE_ABORT when the word is set.  The whole wave is active.
constexpr uint32_t TAKE_ABORT = 0xffffffffu;
__device__ __forceinline__ uint32_t take_chunk(uint32_t *counter, const uint32_t *abort_flag, int lane)
{
    uint32_t t = 0, stop = 0;
    if (lane == 0) {
        t = atomicAdd(counter, 1u);
        stop = __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    t = __builtin_amdgcn_readfirstlane(t);
    stop = __builtin_amdgcn_readfirstlane(stop);
    return stop != 0u ? TAKE_ABORT : t;
}

// The body of a window kernel: one persistent workgroup of WIN_WGS threads per CU.  The
// window s_win is cleared, then every wave traces chunks of up to WIN_CHUNK rays of one
// tile from 8 queues, one per XCD, each over a contiguous eighth of the tile list (a wave
// drains its own XCD's queue first, then helps the others, so no wave waits for another
// until all queues are empty; the waves of an XCD take consecutive chunks, so at any time
// they trace a few neighbouring heliostats whose grid cells and blockers stay in that XCD's
// L2).  A wave stops taking chunks once the abort word is set.  The window is stored to
// this workgroup's slice of win_part at the end.
// chunk(tl, lane, nrays, out, s_win) traces nrays rays of heliostat tl.x from ray tl.y with
// one wave; finish(lane, out, s_win) ends a wave's work after its last chunk.
template <class Chunk, class Finish>
__device__ __forceinline__ void window_body(const TraceArgs &a, unsigned long long *s_win,
                                            unsigned long long (*s_out)[WIN_OUT], Chunk chunk_fn, Finish finish_fn)
{
    const int wpx = a.win_w * a.win_h; // <= WIN_PX_MAX, checked on the host
    for (int i = threadIdx.x; i < wpx; i += WIN_WGS) s_win[i] = 0ull;
    __syncthreads();
    const int lane = (int)threadIdx.x & 63;
    const uint32_t xcd = xcc_id() & 7u;
    WaveOut out{s_out[threadIdx.x >> 6], 0};
    // the window of this XCD's queue (the same for every queue unless win_slice_q is set)
    out.wx0 = a.win_slice_q ? a.win_qx0[xcd] : a.win_x0;
    out.wy0 = a.win_slice_q ? a.win_qy0[xcd] : a.win_y0;
    // tiles hold 64..2048 rays (a power of two): chunks of min(tile, WIN_CHUNK) rays
    const int chunk = min(a.tile_rays, WIN_CHUNK);
    const int parts = a.tile_rays / chunk;
    const int n = a.ntiles * parts, q = n >> 3, rmd = n & 7;
    for (uint32_t j = 0; j < 8u; ++j) {
        const int x = (int)((xcd + j) & 7u);
        const int first = x * q + min(x, rmd), cnt = q + (x < rmd ? 1 : 0);
        for (;;) {
            uint32_t t = take_chunk(a.win_queue + x, a.abort_flag, lane);
            if (t == TAKE_ABORT) j = 8u; // the host asked every persistent launch to stop
            if (t >= (uint32_t)cnt) break;
            const int c = first + (int)t;
            const int tile = c / parts, part = c - tile * parts;
            const int2 tl = a.tiles[a.tile_base + (a.win_order ? a.win_order[tile] : tile)];
            chunk_fn(make_int2(tl.x, tl.y + part * chunk), lane, chunk, out, s_win);
        }
    }
    finish_fn(lane, out, s_win);
    out.flush(a, lane);
    __syncthreads();
    unsigned long long *dst = a.win_part + (size_t)blockIdx.x * wpx;
    for (int i = threadIdx.x; i < wpx; i += WIN_WGS) dst[i] = s_win[i];
    if (a.win_slice_q && threadIdx.x == 0) a.win_slice_q[blockIdx.x] = (int)xcd;
}

template <class Chunk>
__device__ __forceinline__ void window_body(const TraceArgs &a, unsigned long long *s_win,
                                            unsigned long long (*s_out)[WIN_OUT], Chunk chunk_fn)
{
    window_body(a, s_win, s_out, chunk_fn, [](int, WaveOut &, unsigned long long *) {});
}

// ---------------------------------------------------------------------------------------
// Batched mode with LDS windows (OM_BATCH_WINDOW)
// ---------------------------------------------------------------------------------------
// One persistent workgroup of WIN_WGS threads per CU.  Its waves take chunks (table entry,
// heliostat, first ray, rays) of at most WIN_CHUNK rays from 8 per-XCD queues over
// contiguous eighths of the chunk list, which is in entry order (an entry is one grid of one
// scene), so the workgroups of an XCD work through a few scenes at a time.  The workgroup
// keeps the central block of one entry's packed map in LDS.  A wave traces the chunks it
// takes while they belong to that entry; holding a chunk of another entry, it waits at the
// workgroup barrier until every wave does (or has no chunk left).  There the entry of the
// smallest held chunk becomes the window's: if it differs, the window's nonzero words are
// added to the old entry's per-XCD copies (one returning atomic each, carries and wraps
// counted as for single hits, which the decode then applies alike) and the window is
// cleared.  Barriers come only at entry changes, so waves do not wait for one another
// between chunks.  Hits outside the window go through the wave's buffer (WaveOut), which is
// flushed before the wave reaches the barrier.
constexpr int BW_NONE = 0x7fffffff; // no chunk held

// Adds the window's nonzero words to entry a's copies and clears the window (the whole
// workgroup; barriers around it are the caller's).
__device__ __forceinline__ void flush_window(const TraceArgs &a, unsigned long long *s_win)
{
    unsigned long long *part = copy_of(a);
    const int wpx = a.win_w * a.win_h;
    for (int i = threadIdx.x; i < wpx; i += WIN_WGS) {
        const unsigned long long w = s_win[i];
        if (w == 0ull) continue;
        s_win[i] = 0ull;
        const int row = i / a.win_w, col = i - row * a.win_w;
        const int px = (a.win_y0 + row) * a.res_x + a.win_x0 + col;
        copy_check(a, px, w, copy_add(a, part, px, w));
    }
}

// The body of a batch window kernel.  table[e] holds entry e's launch arguments and
// args(table[e]) its TraceArgs; chunk_fn(table[e], chunk, lane, s_win, wo) traces one chunk
// (entry, heliostat, first ray, rays) with one wave.
template <class Entry, class Args, class Chunk>
__device__ __forceinline__ void batch_window_body(const Entry *__restrict__ table, const int4 *__restrict__ chunks,
                                                  int n_chunks, unsigned int *queue, unsigned long long *s_win,
                                                  unsigned long long (*s_out)[WIN_OUT], int *s_held, Args args,
                                                  Chunk chunk_fn)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    for (int i = threadIdx.x; i < WIN_PX_MAX; i += WIN_WGS) s_win[i] = 0ull;
    const int q = n_chunks >> 3, rmd = n_chunks & 7;
    const uint32_t xcd = xcc_id() & 7u;
    int qj = 0; // queues this wave has found empty: its own XCD's first, then the others in turn
    // the next chunk for this wave, or -1 when every queue is empty
    auto pop = [&]() -> int {
        while (qj < 8) {
            const int x = (int)((xcd + (uint32_t)qj) & 7u);
            const int cnt = q + (x < rmd ? 1 : 0);
            const uint32_t t = take_chunk(queue + x, args(table[0]).abort_flag, lane);
            if (t == TAKE_ABORT) {
                qj = 8;
                break;
            }
            if (t < (uint32_t)cnt) return x * q + min(x, rmd) + (int)t;
            ++qj;
        }
        return -1;
    };
    int held = pop();
    int he = held >= 0 ? __builtin_amdgcn_readfirstlane(chunks[held].x) : BW_NONE; // entry of the held chunk
    int cur = -1; // entry whose window the LDS holds (uniform over the workgroup)
    WaveOut wo{s_out[wave], 0};
    for (;;) {
        if (lane == 0) s_held[wave] = he;
        __syncthreads(); // every wave holds a chunk of another entry or none; the window is idle
        int ne = BW_NONE;
        for (int w = 0; w < WIN_WAVES; ++w) ne = min(ne, s_held[w]);
        ne = __builtin_amdgcn_readfirstlane(ne);
        if (ne == BW_NONE) break;
        if (ne != cur) {
            if (cur >= 0) flush_window(args(table[cur]), s_win);
            cur = ne;
            wo.wx0 = args(table[cur]).win_x0;
            wo.wy0 = args(table[cur]).win_y0;
        }
        __syncthreads(); // the window is clear, s_held read, before hits and the next round
        while (he == cur) {
            const int4 c = chunks[held];
            const int4 ck = make_int4(cur, __builtin_amdgcn_readfirstlane(c.y), __builtin_amdgcn_readfirstlane(c.z),
                                      __builtin_amdgcn_readfirstlane(c.w));
            chunk_fn(table[cur], ck, lane, s_win, wo);
            held = pop();
            he = held >= 0 ? __builtin_amdgcn_readfirstlane(chunks[held].x) : BW_NONE;
        }
        wo.flush(args(table[cur]), lane);
    }
    if (cur >= 0) flush_window(args(table[cur]), s_win);
}

} // namespace om
