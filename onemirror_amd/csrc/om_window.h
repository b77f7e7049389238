// om_window.h -- the LDS-window accumulation of the trace kernels (OM_ACC_WINDOW), shared by
// the FP32 window kernel (om_kernels.hip) and the FP64 one (om_trace64.hip): packed words in
// LDS, the ring of hits outside the window and its draining wave, and the persistent
// workgroup body that takes chunks of tiles from 8 per-XCD queues.
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
// with the ring below: the whole 160 KiB) for every ray it traces.  WIN_COMPUTE of its waves
// trace rays: hits inside the window are LDS atomics, hits outside go to a ring buffer in
// LDS.  The last wave drains the ring into the packed per-XCD copies in device memory, so
// the tracing waves never hold their loads behind a memory-side atomic.  The window is
// stored (not added) to this workgroup's slice of win_part at the end.
constexpr int WIN_WGS = 1024;
constexpr int WIN_COMPUTE = WIN_WGS / 64 - 1; // tracing waves; the last wave drains the ring
constexpr int WIN_RING = 2048;                // ring entries (px << 32 | q), a power of two
constexpr int WIN_DRAIN = 16;                 // ring entries per lane of the draining wave per round
static_assert(64 * WIN_DRAIN + 64 <= WIN_RING, "the ring holds a drain round and a pass beyond it");
constexpr int WIN_CHUNK = 2048;               // rays a tracing wave takes from a queue at a time
constexpr int WIN_PX_MAX = (160 * 1024 - WIN_RING * 8 - 256) / 8;
constexpr unsigned long long RING_EMPTY = ~0ull;

struct WinRing {
    unsigned long long *ring; // [WIN_RING]
    uint32_t *tail;           // entries reserved by the tracing waves
    uint32_t *head;           // entries drained (written by the draining wave only)
    uint32_t *done;           // tracing waves that have finished
};

// Adds this pass's hits outside the window (out) to the ring, in lane order.  Lane 0
// reserves the wave's entries with one LDS atomic and, before they are written, waits until
// the ring has room for them; head_seen caches the last head read.  The draining wave only
// waits for entries within 64 * WIN_DRAIN of the head, whose writers need no more room than
// the ring has at that head (64 * WIN_DRAIN + 64 <= WIN_RING), so neither waits for the other.
__device__ __forceinline__ void ring_push(const WinRing &R, bool out, int px, uint32_t q, int lane,
                                          uint32_t &head_seen)
{
    const unsigned long long m = __ballot(out);
    if (m == 0ull) return;
    const uint32_t n = (uint32_t)__popcll(m);
    uint32_t base = 0;
    if (lane == 0) base = __hip_atomic_fetch_add(R.tail, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    base = __builtin_amdgcn_readlane(base, 0);
    while (base + n - head_seen > (uint32_t)WIN_RING) {
        head_seen = __hip_atomic_load(R.head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (base + n - head_seen > (uint32_t)WIN_RING) __builtin_amdgcn_s_sleep(2);
    }
    if (out) {
        const uint32_t k = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        __hip_atomic_store(R.ring + ((base + k) & (uint32_t)(WIN_RING - 1)), ((unsigned long long)px << 32) | q,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// The draining wave: takes up to 64 * WIN_DRAIN entries at a time from the ring head, adds
// them to the packed per-XCD copies (one returning 64-bit atomic each, carries and wraps
// counted as in flush_slots), clears them and publishes the new head; ends when every
// tracing wave has finished and the ring is empty.  Rounds are large because each waits
// once for the memory side.
__device__ void ring_drain(const TraceArgs &a, const WinRing &R, int lane)
{
    const uint32_t copy = (xcc_id() & 7u) | ((((uint32_t)blockIdx.x >> 3) & ((1u << a.copy_shift) - 1u)) << 3);
    unsigned long long *part = a.part + (size_t)copy * a.copy_stride;
    const int E = a.pack_shift;
    const unsigned long long lowmask = (1ull << E) - 1ull;
    uint32_t head = 0;
    for (;;) {
        const uint32_t done = __hip_atomic_load(R.done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t tail = __hip_atomic_load(R.tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t avail = tail - head;
        if (avail == 0u) {
            if (done == (uint32_t)WIN_COMPUTE) break;
            __builtin_amdgcn_s_sleep(4);
            continue;
        }
        const uint32_t nb = min(avail, (uint32_t)(64 * WIN_DRAIN));
        unsigned long long e[WIN_DRAIN], old[WIN_DRAIN];
#pragma unroll
        for (int b = 0; b < WIN_DRAIN; ++b) {
            e[b] = RING_EMPTY;
            const uint32_t i = (uint32_t)(b * 64 + lane);
            if (i < nb) {
                // reserved entries are written right after their reservation
                unsigned long long *slot = R.ring + ((head + i) & (uint32_t)(WIN_RING - 1));
                do {
                    e[b] = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                } while (e[b] == RING_EMPTY);
            }
        }
#pragma unroll
        for (int b = 0; b < WIN_DRAIN; ++b) {
            old[b] = 0ull;
            if (e[b] != RING_EMPTY) {
                const uint32_t px = (uint32_t)(e[b] >> 32);
                const unsigned long long v = (1ull << E) + (e[b] & 0xffffffffull);
                old[b] = __hip_atomic_fetch_add(part + (size_t)px * a.px_stride, v, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            }
        }
#pragma unroll
        for (int b = 0; b < WIN_DRAIN; ++b) {
            if (e[b] != RING_EMPTY) {
                const uint32_t px = (uint32_t)(e[b] >> 32);
                const unsigned long long q = e[b] & 0xffffffffull;
                const unsigned long long v = (1ull << E) + q;
                if ((old[b] & lowmask) + q > lowmask) atomicAdd(a.carry + px, 1u);
                if (old[b] + v < old[b]) atomicAdd(a.wrap + px, 1u);
                __hip_atomic_store(R.ring + ((head + (uint32_t)(b * 64 + lane)) & (uint32_t)(WIN_RING - 1)),
                                   RING_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        head += nb;
        if (lane == 0) __hip_atomic_store(R.head, head, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// The body of a window kernel: one persistent workgroup of WIN_WGS threads per CU.  The
// window s_win and the ring are cleared, then WIN_COMPUTE waves trace chunks of up to
// WIN_CHUNK rays of one tile from 8 queues, one per XCD, each over a contiguous eighth of the
// tile list (a wave drains its own XCD's queue first, then helps the others, so no wave
// waits for another until all queues are empty; the waves of an XCD take consecutive
// chunks, so at any time they trace a few neighbouring heliostats whose grid cells and
// blockers stay in that XCD's L2), while the last wave drains the ring.  The window is
// stored to this workgroup's slice of win_part at the end.  chunk(tl, lane, nrays, R, s_win)
// traces nrays rays of heliostat tl.x from ray tl.y with one wave.
template <class Chunk>
__device__ __forceinline__ void window_body(const TraceArgs &a, unsigned long long *s_win,
                                            unsigned long long *s_ring, uint32_t *s_ctl, Chunk chunk_fn)
{
    const int wpx = a.win_w * a.win_h; // <= WIN_PX_MAX, checked on the host
    for (int i = threadIdx.x; i < wpx; i += WIN_WGS) s_win[i] = 0ull;
    for (int i = threadIdx.x; i < WIN_RING; i += WIN_WGS) s_ring[i] = RING_EMPTY;
    if (threadIdx.x < 3) s_ctl[threadIdx.x] = 0u;
    __syncthreads();
    const WinRing R{s_ring, s_ctl, s_ctl + 1, s_ctl + 2};
    const int lane = (int)threadIdx.x & 63;
    if ((int)(threadIdx.x >> 6) == WIN_COMPUTE) {
        ring_drain(a, R, lane);
    } else {
        // tiles hold 64..2048 rays (a power of two): chunks of min(tile, WIN_CHUNK) rays
        const int chunk = min(a.tile_rays, WIN_CHUNK);
        const int parts = a.tile_rays / chunk;
        const int n = a.ntiles * parts, q = n >> 3, rmd = n & 7;
        const uint32_t xcd = xcc_id() & 7u;
        for (uint32_t j = 0; j < 8u; ++j) {
            const int x = (int)((xcd + j) & 7u);
            const int first = x * q + min(x, rmd), cnt = q + (x < rmd ? 1 : 0);
            for (;;) {
                uint32_t t = 0;
                if (lane == 0) t = atomicAdd(a.win_queue + x, 1u);
                t = __builtin_amdgcn_readfirstlane(t); // the whole wave is active here
                if (t >= (uint32_t)cnt) break;
                const int c = first + (int)t;
                const int tile = c / parts, part = c - tile * parts;
                const int2 tl = a.tiles[a.tile_base + tile];
                chunk_fn(make_int2(tl.x, tl.y + part * chunk), lane, chunk, R, s_win);
            }
        }
        if (lane == 0) __hip_atomic_fetch_add(R.done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    unsigned long long *out = a.win_part + (size_t)blockIdx.x * wpx;
    for (int i = threadIdx.x; i < wpx; i += WIN_WGS) out[i] = s_win[i];
}

} // namespace om
