// om_kernels.hip -- gfx950 kernels of the heliostat-field ray tracer and the C ABI entry
// points that launch them.
//
//   trace_kernel    map_tracing (rectgrid_raytracing.cu:155-196) for a whole field in one
//                   launch: Philox sampling of sunshape and slope error, shading DDA,
//                   reflection, blocking DDA, receiver hit, fixed-point accumulation.
//   poisson_kernel  poisson_sample + copy_if (poisson_disk.cu:4-201, compact_sample.cu:6-19),
//                   one workgroup per heliostat, cell grid in LDS.
//   finalize_kernel fixed point -> FP32 flux (Gtest.cu:37-46).
//   smooth_kernel   trimmed_mean (image_smooth.cu:10-58) with the heaps of heap.cuh.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "om_window.h"

namespace om {

// Outcome of one ray: the shading heliostat (-1: none), the blocking heliostat (-1: none,
// -2: not looked for, the ray is shaded), the receiver pixel (row, column; row -1 for a miss
// or an obstructed ray) and the energy |sun.n'| in 2^-31 units.
struct RayOut {
    int shade, block;
    int2 rc;
    uint32_t q;
};

// Ray l of microsample m of the tile's heliostat from its origin rs, up to the receiver:
// map_tracing (rectgrid_raytracing.cu:155-196) without its last step.  The ray goes on to
// the receiver when shade and block are both -1; refl is then its reflected direction and
// q its energy.  FINE: the Buie bin comes from the split guide in device memory
// (buie_bin_split), else from the tables s_bins / s_guide (in LDS).
template <int SUN, bool FINE>
__device__ __forceinline__ RayOut trace_walks(const TraceArgs &a, const HelioTile &t, const RayStart &rs, int m,
                                              int l, const float4 *s_bins, const int *s_guide, v3 &refl)
{
    RayOut o;
    o.shade = -1;
    o.block = -2;
    o.rc = make_int2(-1, -1);
    o.q = 0u;
    const v3 orig = rs.orig;
    const cell3 &c0 = rs.c0;
    // draws: sunshape theta, sunshape phi, slope-error normal deviate, slope-error phi
    u4 x = philox_uniform_key(u4{(uint32_t)l, (uint32_t)m, (uint32_t)t.h, a.w3}, a.key0, a.key1);
    float th, st, ct, sp, cp;
    v3 nrm;
    if (SUN == OM_SUN_BUIE && FINE) {
        // buie_bin_split with its two loads issued ahead of the draws that do not depend on
        // them (the sunshape azimuth and the perturbed normal, which the reference computes
        // after the shading walk: the same values), so that the loads are in flight meanwhile
        const float v = u01(x.x) * a.buie_A;
        const int j = min(max(trunc_i(v * a.buie_gscale), 0), a.nbins - 1);
        const int2 g = a.buie_split[j];
        __builtin_amdgcn_sched_barrier(0); // the load above is issued before this work
        nrm = perturbed_normal(x, a.slope_sigma, t.hf);
        asm volatile("" ::"v"(nrm.x), "v"(nrm.y), "v"(nrm.z)); // computed here
        __builtin_amdgcn_sched_barrier(0);
        int i = min(g.x + (v >= u2f((uint32_t)g.y) ? 1 : 0), a.nbins - 1);
        float4 b = a.buie_bins[i];
        __builtin_amdgcn_sched_barrier(0);
        sincos2pi(u01(x.y), sp, cp);
        asm volatile("" ::"v"(sp), "v"(cp));
        __builtin_amdgcn_sched_barrier(0);
        if (v < a.buie_int465) {
            while (b.x > v) b = a.buie_bins[--i];
            while (b.y <= v) b = a.buie_bins[++i];
            th = (v - b.x) * b.z + b.w;
        } else {
            th = pow_<gfx>((v - a.buie_int465) * a.buie_B + a.buie_C, a.buie_inv_gp1);
        }
        th = div_rn(th, 1000.0f);
    } else {
        if (SUN == OM_SUN_PILLBOX) {
            th = a.pillbox * sqrt_rn(u01(x.x));
        } else {
            float v = u01(x.x) * a.buie_A;
            if (v < a.buie_int465) {
                const float4 b = buie_bin(s_bins, s_guide, a.buie_gscale, v, a.nbins);
                th = (v - b.x) * b.z + b.w;
            } else {
                th = pow_<gfx>((v - a.buie_int465) * a.buie_B + a.buie_C, a.buie_inv_gp1);
            }
            th = div_rn(th, 1000.0f);
        }
        sincos2pi(u01(x.y), sp, cp);
    }
    sincos_small(th, st, ct);
    const v3 to_sun = -to_world(angle2xyz(st, ct, sp, cp), t.sf);
    // the reference order: shading, perturbed normal, reflection, blocking, receiver
    o.shade = t.clear ? -1 : dda(a, orig, c0, to_sun, t.self);
    if (o.shade >= 0) return o;
    if (!(SUN == OM_SUN_BUIE && FINE)) nrm = perturbed_normal(x, a.slope_sigma, t.hf);
    refl = normalize<gfx>(reflect(-to_sun, nrm));
    o.block = dda(a, orig, c0, refl, t.self);
    if (o.block >= 0) return o;
    const float e = fabs_(dot(a.sun_n, nrm)) * 1.0f; // calEnergy with eta = 1
    o.q = (uint32_t)(e * 2147483648.0f);
    return o;
}

// The whole of map_tracing (rectgrid_raytracing.cu:155-196) for one ray: trace_walks, then
// the receiver.
template <int SUN, bool FINE>
__device__ __forceinline__ RayOut trace_from(const TraceArgs &a, const HelioTile &t, const RayStart &rs, int m,
                                             int l, const float4 *s_bins, const int *s_guide)
{
    v3 refl;
    RayOut o = trace_walks<SUN, FINE>(a, t, rs, m, l, s_bins, s_guide, refl);
    if (o.shade < 0 && o.block < 0) o.rc = receiver_rc(a, rs.orig, refl);
    return o;
}

// Adds the hits in the first nk of a thread's hit slots (pixel, or -1 for none) to the maps.
template <int WGS, int RPTS>
__device__ __forceinline__ void flush_slots(const TraceArgs &a, const int *s_hpx, const uint32_t *s_hq,
                                            const int tid, const int nk)
{
    if (!a.acc_packed) {
        for (int k = 0; k < nk; ++k) {
            const int px = s_hpx[k * WGS + tid];
            if (px < 0) continue;
            atomicAdd(a.counts + px, 1u);
            atomicAdd(a.energy + px, (unsigned long long)s_hq[k * WGS + tid]);
        }
        return;
    }
    // One 64-bit atomic per hit into a packed word: v = 2^E + q.  The old value tells
    // whether the add carried out of the energy field (bits 0..E-1) or wrapped the word;
    // both events are counted on the side so that the reduction decodes exactly.
    // copy: the XCD's own block of copies, one of them per workgroup (workgroup b runs on
    // XCD b % 8, so b / 8 counts the workgroups of one XCD)
    const uint32_t copy = (xcc_id() & 7u) | ((((uint32_t)blockIdx.x >> 3) & ((1u << a.copy_shift) - 1u)) << 3);
    unsigned long long *part = a.part + (size_t)copy * a.copy_stride;
    const int E = a.pack_shift;
    const unsigned long long lowmask = (1ull << E) - 1ull;
    unsigned long long old[RPTS];
#pragma unroll
    for (int k = 0; k < RPTS; ++k) {
        old[k] = 0;
        if (k < nk) {
            const int px = s_hpx[k * WGS + tid];
            if (px >= 0) {
                const unsigned long long v = (1ull << E) + s_hq[k * WGS + tid];
                old[k] = __hip_atomic_fetch_add(part + (size_t)px * a.px_stride, v, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RPTS; ++k) {
        if (k < nk) {
            const int px = s_hpx[k * WGS + tid];
            if (px >= 0) {
                const unsigned long long q = s_hq[k * WGS + tid];
                const unsigned long long v = (1ull << E) + q;
                if ((old[k] & lowmask) + q > lowmask) atomicAdd(a.carry + px, 1u);
                if (old[k] + v < old[k]) atomicAdd(a.wrap + px, 1u);
            }
        }
    }
}

// One tile: up to rays_per_thread rays per thread of heliostat tl.x, rays tl.y + k*WGS +
// tid among the nrays rays from tl.y.  Hits of a thread's rays are kept in LDS and added to
// the maps after the ray loop: vector memory operations complete in issue order, so an
// atomic issued inside the loop would hold back every later load of the wave until the
// memory side has carried it out.  s_hpx / s_hq: the [RPTS][WGS] hit slots of the WGS
// threads (each thread uses only its own).  RECORD: every ray's shading, blocking and pixel
// indices go to the record arrays instead (rays in the reference order).
template <int SUN, bool POISSON, bool RECORD, int WGS, bool FINE, int RPTS = RPT_MAX>
__device__ __forceinline__ void trace_tile(const TraceArgs &a, int2 tl, const int tid, const int nrays,
                                           int *s_hpx, uint32_t *s_hq, const float4 *s_bins,
                                           const int *s_guide)
{
    const HelioTile t = helio_tile(a, tl.x);
    const int L = a.L;
    const int nr = min(t.hr->mcount * L, tl.y + nrays); // end of this tile's rays
    // ray index r = m * L + l, stepped by WGS without division
    int r = tl.y + tid;
    int m = r / L;
    int l = r - m * L;
    const int dm = WGS / L, dl = WGS - (WGS / L) * L;
    int nk = 0; // rays of this thread, hit slots used
    RayStart rs{};
    int m_rs = -1; // microsample of rs: its origin is found again only when m changes
    for (int k = 0; k < a.rays_per_thread; ++k, r += WGS) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= L) {
                l -= L;
                ++m;
            }
        }
        if (r >= nr) break;
        nk = k + 1;
        asm volatile("" ::: "memory"); // arguments re-read per ray pass (see trace_window_kernel)
        if (m != m_rs) {
            rs = ray_start<POISSON>(a, t, m);
            m_rs = m;
        }
        const RayOut o = trace_from<SUN, FINE>(a, t, rs, m, l, s_bins, s_guide);
        const int px = o.rc.x >= 0 ? o.rc.x * a.res_x + o.rc.y : -1;
        if (RECORD) {
            const long long rid =
                (((long long)t.hr->ray_off_hi << 32) | (unsigned)t.hr->ray_off_lo) + r - a.rec_base;
            a.rec_shade[rid] = o.shade;
            a.rec_block[rid] = o.block;
            a.rec_pixel[rid] = px;
            continue;
        }
        s_hpx[k * WGS + tid] = px;
        s_hq[k * WGS + tid] = o.q;
    }
    if (RECORD) return;
    // Each thread adds its own hits after its last load.
    flush_slots<WGS, RPTS>(a, s_hpx, s_hq, tid, nk);
}

// One workgroup per tile of the scene's tile list (plain mode).  The arguments are read from
// the kernel-argument segment, as in trace_window_kernel.
template <int SUN, bool POISSON, bool RECORD, bool BUIE_LDS>
__global__ __launch_bounds__(WG) void trace_kernel(TraceArgs)
{
    const TraceArgs &a = *(const TraceArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    extern __shared__ float4 lds4[];
    int *s_hpx = (int *)lds4;                          // [RPT_MAX][WG] pixel or -1
    uint32_t *s_hq = (uint32_t *)(s_hpx + RPT_MAX * WG); // [RPT_MAX][WG] fixed-point energy
    const float4 *s_bins = a.buie_bins;
    const int *s_guide = a.buie_guide;
    if (SUN == OM_SUN_BUIE && BUIE_LDS) {
        float4 *lb = (float4 *)(s_hq + RPT_MAX * WG);
        int *lg = (int *)(lb + a.nbins);
        for (int i = threadIdx.x; i < a.nbins; i += WG) {
            lb[i] = a.buie_bins[i];
            lg[i] = a.buie_guide[i];
        }
        __syncthreads();
        s_bins = lb;
        s_guide = lg;
    }
    const int tile = a.tile_base + xcd_chunk(blockIdx.x, gridDim.x);
    trace_tile<SUN, POISSON, RECORD, WG, !BUIE_LDS>(a, a.tiles[tile], (int)threadIdx.x, a.tile_rays, s_hpx, s_hq, s_bins,
                                         s_guide);
}

// nrays (a multiple of 64) rays of heliostat tl.x from tl.y, traced by one wave, one ray per
// lane per pass.  A lane finds the origin of its microsample again only when the
// microsample changes: with L >= 64 rays per microsample a pass holds at most two, and most
// passes none that is new to the wave.
template <int SUN, bool POISSON>
__device__ __forceinline__ void trace_chunk_window(const TraceArgs &a, int2 tl, const int lane, const int nrays,
                                                   WaveOut &wo, unsigned long long *s_win)
{
    const HelioTile t = helio_tile(a, tl.x);
    const int L = a.L;
    const int nr = min(t.hr->mcount * L, tl.y + nrays);
    int r = tl.y + lane;
    int m = r / L;
    int l = r - m * L;
    const int dm = 64 / L, dl = 64 - (64 / L) * L;
    StartTable tab;
    for (int k = 0; k < nrays; k += 64, r += 64) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= L) {
                l -= L;
                ++m;
            }
        }
        bool out = false;
        int px = 0;
        uint32_t q = 0u;
        asm volatile("" ::: "memory"); // see trace_window_kernel
        tab.update<POISSON>(a, t, r < nr, m, lane);
        const RayStart rs = tab.get(a, m);
        if (r < nr) {
            const RayOut o = trace_from<SUN, true>(a, t, rs, m, l, a.buie_bins, a.buie_guide);
            if (o.rc.x >= 0) {
                px = o.rc.x * a.res_x + o.rc.y;
                q = o.q;
                const int wi = wo.window_word(a, o.rc.x, o.rc.y);
                if (wi >= 0)
                    packed_add_lds(a, s_win + wi, px, q);
                else
                    out = true;
            }
        }
        wo.push(a, out, px, q, lane); // the whole wave takes part: no lane has left the loop
    }
}

// Ray compaction after the blocking test (om_params.ray_compact): rays that reach the
// receiver stage are held, up to 63 of them, in lanes [0, n) of the wave until 64 fill the
// wave, so that the receiver stage runs on full waves instead of on the two thirds of each
// pass that the blocking walk leaves alive.  n is wave-uniform.
struct Pending {
    v3 orig, refl;
    uint32_t q;
    int n;
};

__device__ __forceinline__ float permute_f(int addr, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_permute(addr, __float_as_int(v)));
}

// c ? x : y per component (a select of whole structs would select their addresses)
__device__ __forceinline__ v3 sel3(bool c, const v3 &x, const v3 &y)
{
    return mk(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z);
}

// The receiver stage of the rays in the lanes where live: pixel, then the window (LDS) or
// the wave's buffer.  The whole wave takes part.
__device__ __forceinline__ void receive(const TraceArgs &a, bool live, const v3 &orig, const v3 &refl, uint32_t q,
                                        const int lane, WaveOut &wo, unsigned long long *s_win)
{
    bool out = false;
    int px = 0;
    if (live) {
        const int2 rc = receiver_rc(a, orig, refl);
        if (rc.x >= 0) {
            px = rc.x * a.res_x + rc.y;
            const int wi = wo.window_word(a, rc.x, rc.y);
            if (wi >= 0)
                packed_add_lds(a, s_win + wi, px, q);
            else
                out = true;
        }
    }
    wo.push(a, out, px, q, lane);
}

// trace_chunk_window with the receiver stage on compacted waves.  After each pass's walks,
// one forward permute (ds_permute) per ray value moves the live rays, ranked by mbcnt over
// the ballot of live lanes, behind the held ones: ranks 0..63 fill the wave, the rest wrap
// to lanes [0, n) as the new held rays; dead lanes take the other destinations, so the
// permute is one-to-one.  The caller runs the held rays at the end (finish_compact).
template <int SUN, bool POISSON>
__device__ __forceinline__ void trace_chunk_compact(const TraceArgs &a, int2 tl, const int lane, const int nrays,
                                                    WaveOut &wo, unsigned long long *s_win, Pending &P)
{
    const HelioTile t = helio_tile(a, tl.x);
    const int L = a.L;
    const int nr = min(t.hr->mcount * L, tl.y + nrays);
    int r = tl.y + lane;
    int m = r / L;
    int l = r - m * L;
    const int dm = 64 / L, dl = 64 - (64 / L) * L;
    StartTable tab;
    for (int k = 0; k < nrays; k += 64, r += 64) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= L) {
                l -= L;
                ++m;
            }
        }
        bool live = false;
        v3 orig = mk(0.0f, 0.0f, 0.0f), refl = mk(0.0f, 0.0f, 0.0f);
        uint32_t q = 0u;
        asm volatile("" ::: "memory"); // see trace_window_kernel
        tab.update<POISSON>(a, t, r < nr, m, lane);
        const RayStart rs = tab.get(a, m);
        if (r < nr) {
            const RayOut o = trace_walks<SUN, true>(a, t, rs, m, l, a.buie_bins, a.buie_guide, refl);
            live = o.shade < 0 && o.block < 0;
            orig = rs.orig;
            q = o.q;
        }
        const unsigned long long ms = __ballot(live);
        const int ns = __popcll(ms);
        const uint32_t rl = __builtin_amdgcn_mbcnt_hi((uint32_t)(ms >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ms, 0u));
        const unsigned long long md = ~ms;
        const uint32_t rd = __builtin_amdgcn_mbcnt_hi((uint32_t)(md >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)md, 0u));
        const int addr = ((live ? P.n + (int)rl : P.n + ns + (int)rd) & 63) << 2;
        const v3 ro = mk(permute_f(addr, orig.x), permute_f(addr, orig.y), permute_f(addr, orig.z));
        const v3 rr = mk(permute_f(addr, refl.x), permute_f(addr, refl.y), permute_f(addr, refl.z));
        const uint32_t rq = (uint32_t)__builtin_amdgcn_ds_permute(addr, (int)q);
        const bool held = lane < P.n;
        if (P.n + ns >= 64) {
            receive(a, true, sel3(held, P.orig, ro), sel3(held, P.refl, rr), held ? P.q : rq, lane, wo, s_win);
            P.orig = ro;
            P.refl = rr;
            P.q = rq;
            P.n += ns - 64;
        } else {
            P.orig = sel3(held, P.orig, ro);
            P.refl = sel3(held, P.refl, rr);
            P.q = held ? P.q : rq;
            P.n += ns;
        }
    }
}

__device__ __forceinline__ void finish_compact(const TraceArgs &a, const int lane, WaveOut &wo,
                                               unsigned long long *s_win, Pending &P)
{
    if (P.n > 0) receive(a, lane < P.n, P.orig, P.refl, P.q, lane, wo, s_win);
    P.n = 0;
}

// The window kernel of the FP32 geometry: window_body (om_window.h) over trace_chunk_window.
// The arguments are read from the kernel-argument segment where each use needs them (a
// compiler barrier per ray pass keeps the loads there): held in SGPRs across the ray loop,
// they exceed the SGPR file and about a hundred of them spill to VGPR lanes, each use then
// costing a v_readlane in the VALU-bound loop; scalar loads from the segment cost no VALU.
// COMPACT: ray compaction after the blocking test (trace_chunk_compact).
template <int SUN, bool POISSON, bool COMPACT>
__global__ __launch_bounds__(WIN_WGS) void trace_window_kernel(TraceArgs)
{
    const TraceArgs &a = *(const TraceArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    __shared__ unsigned long long s_win[WIN_PX_MAX];
    __shared__ unsigned long long s_out[WIN_WAVES][WIN_OUT];
    if (COMPACT) {
        Pending P{};
        window_body(
            a, s_win, s_out,
            [&](int2 tl, int lane, int nrays, WaveOut &wo, unsigned long long *w) {
                trace_chunk_compact<SUN, POISSON>(a, tl, lane, nrays, wo, w, P);
            },
            [&](int lane, WaveOut &wo, unsigned long long *w) { finish_compact(a, lane, wo, w, P); });
    } else {
        window_body(a, s_win, s_out, [&](int2 tl, int lane, int nrays, WaveOut &wo, unsigned long long *w) {
            trace_chunk_window<SUN, POISSON>(a, tl, lane, nrays, wo, w);
        });
    }
}

// Batched mode: persistent workgroups take (scene, heliostat, first ray) items from a
// queue until it is empty (or the abort word of om_ctx_abort is set).  Items of one scene are
// consecutive, so the workgroups running at any time mostly share one scene's grid and
// heliostats in their L2.  Every workgroup reads the counter until it passes n_items, so
// the grid drains.
template <int SUN, bool POISSON, int WGS>
__global__ __launch_bounds__(WGS) void trace_batch_kernel(const TraceArgs *__restrict__ table,
                                                          const int4 *__restrict__ items, int n_items,
                                                          unsigned int *queue)
{
    extern __shared__ float4 lds4[];
    int *s_hpx = (int *)lds4;
    uint32_t *s_hq = (uint32_t *)(s_hpx + RPT_MAX * WGS);
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) {
            const int it = (int)atomicAdd(queue, 1u);
            const bool stop = __hip_atomic_load(table[0].abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
            s_item = stop ? n_items : it; // om_ctx_abort: no further item
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads(); // every thread has read s_item before thread 0 overwrites it
        if (it >= n_items) break;
        const int4 w = items[it];
        const int sc = __builtin_amdgcn_readfirstlane(w.x);
        const TraceArgs &a = table[sc];
        trace_tile<SUN, POISSON, false, WGS, true>(a, make_int2(__builtin_amdgcn_readfirstlane(w.y),
                                                          __builtin_amdgcn_readfirstlane(w.z)),
                                             (int)threadIdx.x, a.tile_rays, s_hpx, s_hq, a.buie_bins,
                                             a.buie_guide);
    }
}

// nrays rays of heliostat ck.y from ray ck.z, traced by one wave: window hits as LDS atomics,
// the others through the wave's buffer, which the caller flushes before the entry changes
template <int SUN, bool POISSON>
__device__ __forceinline__ void trace_chunk_batch(const TraceArgs &a, int4 ck, const int lane,
                                                  unsigned long long *s_win, WaveOut &wo)
{
    const HelioTile t = helio_tile(a, ck.y);
    const int L = a.L;
    const int nrays = ck.w;
    const int nr = min(t.hr->mcount * L, ck.z + nrays);
    int r = ck.z + lane;
    int m = r / L;
    int l = r - m * L;
    const int dm = 64 / L, dl = 64 - (64 / L) * L;
    StartTable tab;
    for (int k = 0; k < nrays; k += 64, r += 64) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= L) {
                l -= L;
                ++m;
            }
        }
        bool out = false;
        int px = 0;
        uint32_t q = 0u;
        asm volatile("" ::: "memory");
        tab.update<POISSON>(a, t, r < nr, m, lane);
        const RayStart rs = tab.get(a, m);
        if (r < nr) {
            const RayOut o = trace_from<SUN, true>(a, t, rs, m, l, a.buie_bins, a.buie_guide);
            if (o.rc.x >= 0) {
                px = o.rc.x * a.res_x + o.rc.y;
                q = o.q;
                const uint32_t wr = (uint32_t)(o.rc.x - a.win_y0), wc = (uint32_t)(o.rc.y - a.win_x0);
                if (wr < (uint32_t)a.win_h && wc < (uint32_t)a.win_w)
                    packed_add_lds(a, s_win + wr * (uint32_t)a.win_w + wc, px, q);
                else
                    out = true;
            }
        }
        wo.push(a, out, px, q, lane);
    }
}

// OM_BATCH_WINDOW in FP32 geometry (batch_window_body, om_window.h)
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WIN_WGS) void trace_batch_window_kernel(const TraceArgs *__restrict__ table,
                                                                     const int4 *__restrict__ chunks, int n_chunks,
                                                                     unsigned int *queue)
{
    __shared__ unsigned long long s_win[WIN_PX_MAX];
    __shared__ unsigned long long s_out[WIN_WAVES][WIN_OUT];
    __shared__ int s_held[WIN_WAVES];
    batch_window_body(
        table, chunks, n_chunks, queue, s_win, s_out, s_held, [](const TraceArgs &a) -> const TraceArgs & { return a; },
        [](const TraceArgs &a, int4 ck, int lane, unsigned long long *w, WaveOut &wo) {
            trace_chunk_batch<SUN, POISSON>(a, ck, lane, w, wo);
        });
}

// Batched decode: blockIdx.y selects the map; writes (not adds) counts and energy, the raw
// float map for smoothing and the scaled flux, with the operations of reduce_packed_kernel
// and finalize_kernel.
__global__ void reduce_batch_kernel(const ReduceSlot *__restrict__ slots)
{
    const ReduceSlot &s = slots[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.npx) return;
    const int E = s.E;
    const unsigned long long lowmask = (1ull << E) - 1ull;
    unsigned long long n = 0, e = 0;
    for (int x = 0; x < s.ncopy; ++x) {
        unsigned long long &w = s.part[(size_t)x * s.copy_stride + (size_t)i * s.px_stride];
        const unsigned long long wv = w;
        n += wv >> E;
        e += wv & lowmask;
        w = 0ull;
    }
    const unsigned long long A = s.carry[i], W = s.wrap[i];
    n = n + (W << (64 - E)) - A;
    e = e + (A << E);
    s.carry[i] = 0u;
    s.wrap[i] = 0u;
    s.counts[i] = (uint32_t)n;
    s.energy[i] = e;
    const float v = (float)((double)e * 4.656612873077392578125e-10);
    s.raw[i] = v;
    if (s.flux) s.flux[i] = v * s.scale;
}

// Decode the packed per-XCD words into the output maps (which it adds to) and clear them.
// Over all copies, with A carries out of the energy field and W wraps of the word:
//   count = sum(word >> E) + W * 2^(64-E) - A,    energy = sum(word & (2^E - 1)) + A * 2^E.
// With a window (nwin workgroup slices of win_part), its pixels add the stored LDS words too.
// Everything read is cleared, the queue counters too, so the whole buffer is zero again for
// the next trace.
struct WindowRef {
    unsigned long long *part; // [nwin][w * h] window words of each workgroup
    int nwin, x0, y0, w, h, res_x;
    // the queue of each workgroup and the window origin of each queue (columns modulo res_x);
    // null: every workgroup's window is at (x0, y0)
    const int *slice_q;
    int qx0[8], qy0[8];
};

__global__ void reduce_packed_kernel(unsigned long long *part, uint32_t *carry, uint32_t *wrap, int E,
                                     uint32_t *counts, unsigned long long *energy, int npx, int ncopy,
                                     int copy_stride, int px_stride, WindowRef win)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const unsigned long long lowmask = (1ull << E) - 1ull;
    unsigned long long n = 0, e = 0;
#pragma unroll 8
    for (int x = 0; x < ncopy; ++x) {
        unsigned long long &w = part[(size_t)x * copy_stride + (size_t)i * px_stride];
        const unsigned long long wv = w;
        n += wv >> E;
        e += wv & lowmask;
        w = 0ull;
    }
    if (win.nwin > 0 && win.slice_q) {
        // the slices of the queues whose window holds the pixel
        const int row = i / win.res_x, col = i - row * win.res_x;
        uint32_t held = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int c = col - win.qx0[q];
            if (c < 0) c += win.res_x;
            if ((uint32_t)(row - win.qy0[q]) < (uint32_t)win.h && (uint32_t)c < (uint32_t)win.w) held |= 1u << q;
        }
        if (held) {
            const size_t wpx = (size_t)win.w * win.h;
            for (int g = 0; g < win.nwin; ++g) {
                const int q = win.slice_q[g];
                if (!((held >> q) & 1u)) continue;
                int c = col - win.qx0[q];
                if (c < 0) c += win.res_x;
                unsigned long long *p = win.part + (size_t)g * wpx + (size_t)(row - win.qy0[q]) * win.w + c;
                const unsigned long long wv = *p;
                n += wv >> E;
                e += wv & lowmask;
                *p = 0ull;
            }
        }
    } else if (win.nwin > 0) {
        const int row = i / win.res_x, col = i - row * win.res_x;
        const uint32_t wr = (uint32_t)(row - win.y0), wc = (uint32_t)(col - win.x0);
        if (wr < (uint32_t)win.h && wc < (uint32_t)win.w) {
            const size_t wpx = (size_t)win.w * win.h;
            unsigned long long *p = win.part + wr * (size_t)win.w + wc;
#pragma unroll 8
            for (int g = 0; g < win.nwin; ++g) {
                const unsigned long long wv = p[(size_t)g * wpx];
                n += wv >> E;
                e += wv & lowmask;
                p[(size_t)g * wpx] = 0ull;
            }
        }
    }
    const unsigned long long A = carry[i], W = wrap[i];
    n = n + (W << (64 - E)) - A;
    e = e + (A << E);
    carry[i] = 0u;
    wrap[i] = 0u;
    counts[i] += (uint32_t)n;
    energy[i] += e;
}

// ---------------------------------------------------------------------------------------
// Poisson disk
// ---------------------------------------------------------------------------------------
struct PoissonArgs {
    const HelioRec *helio;
    float r, cell;
    int k;
    int cap;
    uint32_t key0, key1, w3;
    float2 *out;
    int *count;
};

__global__ __launch_bounds__(WG) void poisson_kernel(PoissonArgs a)
{
    extern __shared__ float2 g[];
    __shared__ int s_scan[WG];
    const int h = blockIdx.x;
    const HelioRec *hr = a.helio + h;
    const float width = hr->subcol_len * (float)hr->cols;
    const float height = hr->subrow_len * (float)hr->rows;
    const float cell = a.cell, r = a.r;
    const int wi = (int)(__builtin_ceilf(width / cell) + 2.0f);
    const int hi = (int)(__builtin_ceilf(height / cell) + 2.0f);
    const int ncell = wi * hi; // <= cap, checked on the host
    const int pw = (wi % 3 != 0) ? wi / 3 + 1 : wi / 3;
    const int ph = (hi % 3 != 0) ? hi / 3 + 1 : hi / 3;
    for (int i = threadIdx.x; i < ncell; i += WG) g[i] = make_float2(-2 * cell, -2 * cell);
    __syncthreads();
    // Nine phases; cells of one phase are 3 apart and a candidate only reads cells within
    // 2 of its own, so updating in place equals reading the pre-phase snapshot.
    for (int phase = 0; phase < 9; ++phase) {
        const int xo = phase % 3, yo = phase / 3;
        for (int idx = threadIdx.x; idx < pw * ph; idx += WG) {
            const int tx = idx % pw, ty = idx / pw;
            const int my_indx = ty * wi / 3 + tx;
            const int x = tx * 3 + xo, y = ty * 3 + yo;
            if (x > wi - 1 || y > hi - 1) continue;
            if (phase == 0) {
                u4 o = philox(u4{(uint32_t)my_indx, 0u, (uint32_t)h, a.w3}, a.key0, a.key1);
                g[y * wi + x] = make_float2((u01(o.x) + (float)x - 1.0f) * cell,
                                            (u01(o.y) + (float)y - 1.0f) * cell);
                continue;
            }
            for (int i = 0; i < a.k; ++i) {
                u4 o = philox(u4{(uint32_t)(my_indx * a.k + i), (uint32_t)phase, (uint32_t)h, a.w3},
                              a.key0, a.key1);
                const float px = (u01(o.x) + (float)x - 1.0f) * cell;
                const float py = (u01(o.y) + (float)y - 1.0f) * cell;
                const int c0 = trunc_i((px - r) / cell + 1.0f), c1 = trunc_i((px + r) / cell + 1.0f);
                const int r0 = trunc_i((py - r) / cell + 1.0f), r1 = trunc_i((py + r) / cell + 1.0f);
                bool cannot = false;
                for (int rr = r0; rr <= r1; ++rr)
                    for (int cc = c0; cc <= c1; ++cc) {
                        if (cc < 0 || cc > wi - 1 || rr < 0 || rr > hi - 1) continue;
                        float2 vv = g[rr * wi + cc];
                        if (vv.x < -cell) continue;
                        float ddx = vv.x - px, ddy = vv.y - py;
                        if (sqrt_(ddx * ddx + ddy * ddy) < r) cannot = true;
                    }
                if (!cannot) {
                    g[y * wi + x] = make_float2(px, py);
                    break;
                }
            }
        }
        __syncthreads();
    }
    // stable compaction in row-major cell order (valid_element, compact_sample.cuh:13-22)
    const int chunk = (ncell + WG - 1) / WG;
    const int b0 = threadIdx.x * chunk, b1 = min(b0 + chunk, ncell);
    int cnt = 0;
    for (int i = b0; i < b1; ++i) {
        float2 p = g[i];
        cnt += (p.x > 0.0f && p.y > 0.0f && p.x < width && p.y < height) ? 1 : 0;
    }
    s_scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < WG; off <<= 1) {
        int v = (threadIdx.x >= (unsigned)off) ? s_scan[threadIdx.x - off] : 0;
        __syncthreads();
        s_scan[threadIdx.x] += v;
        __syncthreads();
    }
    int w = s_scan[threadIdx.x] - cnt;
    float2 *dst = a.out + (size_t)h * a.cap;
    for (int i = b0; i < b1; ++i) {
        float2 p = g[i];
        if (p.x > 0.0f && p.y > 0.0f && p.x < width && p.y < height) dst[w++] = p;
    }
    if (threadIdx.x == WG - 1) a.count[h] = s_scan[WG - 1];
}

// ---------------------------------------------------------------------------------------
// finalize, fill, smoothing, math self-test
// ---------------------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t *p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void fill_u64_kernel(unsigned long long *p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// fixed point -> float: exact u64 -> double, exact 2^-31 scale, one rounding to float
__global__ void finalize_kernel(const unsigned long long *energy, float *raw, float *flux, float scale,
                                int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = (float)((double)energy[i] * 4.656612873077392578125e-10);
    if (raw) raw[i] = v;
    if (flux) flux[i] = v * scale;
}

// trimmed_mean (image_smooth.cu:10-58).  Each thread keeps its max- and min-heap
// (heap.cuh:9-212, 1-based arrays) in LDS; kcap = (2r+1)^2 ratio + 1 slots each.
__global__ void smooth_kernel(const float *in, float *out, int width, int height, int radius,
                              float ratio, int kcap, float scale)
{
    extern __shared__ float heaps[];
    const int tid = threadIdx.x;
    float *mxh = heaps + (size_t)tid * 2 * (kcap + 1);
    float *mnh = mxh + (kcap + 1);
    const int idx = blockIdx.x * blockDim.x + tid;
    if (idx >= width * height) return;
    const int x = idx % width, y = idx / width;
    const int lx = (x - radius >= 0) ? x - radius : 0;
    const int ly = (y - radius >= 0) ? y - radius : 0;
    const int rx = (x + radius < width) ? x + radius : width - 1;
    const int ry = (y + radius < height) ? y + radius : height - 1;
    const int cnt = (rx - lx + 1) * (ry - ly + 1);
    const int k = (int)((float)cnt * ratio);
    int nmax = 1, nmin = 1;
    const int full = k + 1;
    float value = 0.0f;
    for (int r = ly; r <= ry; ++r)
        for (int c = lx; c <= rx; ++c) {
            const float v = in[r * width + c];
            value += v;
            if (k == 0) continue;
            if (nmax != full) {
                // insert into both heaps (Heap::insert)
                int i = nmax++;
                for (; i != 1 && v > mxh[i / 2]; i /= 2) mxh[i] = mxh[i / 2];
                mxh[i] = v;
                i = nmin++;
                for (; i != 1 && v < mnh[i / 2]; i /= 2) mnh[i] = mnh[i / 2];
                mnh[i] = v;
            } else if (v < mxh[1]) {
                // Heap::delete_and_insert on the max-heap
                int i = 1;
                while (2 * i < nmax) {
                    int cid = (2 * i + 1 >= nmax || mxh[2 * i] > mxh[2 * i + 1]) ? 2 * i : 2 * i + 1;
                    if (mxh[cid] > v) {
                        mxh[i] = mxh[cid];
                        i = cid;
                    } else {
                        break;
                    }
                }
                mxh[i] = v;
            } else if (v > mnh[1]) {
                int i = 1;
                while (2 * i < nmin) {
                    int cid = (2 * i + 1 >= nmin || mnh[2 * i] < mnh[2 * i + 1]) ? 2 * i : 2 * i + 1;
                    if (mnh[cid] < v) {
                        mnh[i] = mnh[cid];
                        i = cid;
                    } else {
                        break;
                    }
                }
                mnh[i] = v;
            }
        }
    float smax = 0.0f, smin = 0.0f;
    for (int i = 1; i < nmax; ++i) smax += mxh[i];
    for (int i = 1; i < nmin; ++i) smin += mnh[i];
    out[idx] = ((value - smax - smin) / (float)(cnt - 2 * k)) * scale;
}

// The functions as the trace kernel evaluates them (with its math policy).  fn 10 and 11
// take pairs: out[i] = in[2i] / in[2i+1] and atan2(in[2i], in[2i+1]) for i < n/2.
__global__ void math_kernel(int fn, const float *in, float *out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (fn == 10 || fn == 11) {
        if (i < n / 2) out[i] = fn == 10 ? div_rn(in[2 * i], in[2 * i + 1]) : atan2_(in[2 * i], in[2 * i + 1]);
        return;
    }
    if (i >= n) return;
    float x = in[i], s, c;
    switch (fn) {
    case 0: out[i] = u01(f2u(x)); break;
    case 1: sincos2pi(x, s, c); out[i] = s; break;
    case 2: sincos2pi(x, s, c); out[i] = c; break;
    case 3: sincos_small(x, s, c); out[i] = s; break;
    case 4: sincos_small(x, s, c); out[i] = c; break;
    case 5: out[i] = log_<gfx>(x); break;
    case 6: out[i] = exp_(x); break;
    case 7: out[i] = erfinv_<gfx>(x); break;
    case 8: out[i] = gauss_<gfx>(x); break;
    default: out[i] = sqrt_rn(x); break;
    }
}

// VALU issue probes (om_math_eval fn 12..14): a known instruction mix per lane, so that the
// kernel's time and rocprof's counters give the VALU issue rate per CU cycle.  Each thread
// runs PROBE_ITERS iterations over 8 independent chains (no dependence stalls at 4 waves
// per SIMD); per iteration and lane:
//   mix 0: 8 v_fma_f32
//   mix 1: 8 v_fma_f32 + 8 v_add_u32 (integer adds on separate chains)
//   mix 2: 8 v_fma_f32 + 2 v_exp_f32 (transcendental)
//   mix 3: 8 v_fma_f64 (the chains in double)
// The loop counter lives in scalar registers.
constexpr int PROBE_ITERS = 256;

template <int MIX>
__global__ __launch_bounds__(WG) void valu_probe_kernel(const float *in, float *out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x0 = in[i];
    float x[8];
    uint32_t k[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        x[j] = x0 + (float)j;
        k[j] = f2u(x0) + (uint32_t)j;
    }
    const float a = 0.99951171875f, b = 0.000244140625f;
    if (MIX == 3) {
        double xd[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) xd[j] = (double)x0 + (double)j;
        for (int it = 0; it < PROBE_ITERS; ++it) {
#pragma unroll
            for (int j = 0; j < 8; ++j) xd[j] = __builtin_fma(xd[j], (double)b, xd[j]); // x (1 + 2^-12)
            asm volatile("" : "+v"(xd[0]), "+v"(xd[1]), "+v"(xd[2]), "+v"(xd[3]), "+v"(xd[4]), "+v"(xd[5]),
                         "+v"(xd[6]), "+v"(xd[7]));
        }
        double sd = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) sd += xd[j];
        out[i] = (float)sd;
        return;
    }
    for (int it = 0; it < PROBE_ITERS; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = __builtin_fmaf(x[j], a, b);
        if (MIX == 1) {
#pragma unroll
            for (int j = 0; j < 8; ++j) k[j] += 0x9E3779B9u + (uint32_t)it;
        }
        if (MIX == 2) {
            x[0] = __builtin_amdgcn_exp2f(x[0]);
            x[4] = __builtin_amdgcn_exp2f(x[4]);
        }
        asm volatile("" : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]),
                     "+v"(x[6]), "+v"(x[7]));
        if (MIX == 1)
            asm volatile("" : "+v"(k[0]), "+v"(k[1]), "+v"(k[2]), "+v"(k[3]), "+v"(k[4]), "+v"(k[5]),
                         "+v"(k[6]), "+v"(k[7]));
    }
    float s = 0.0f;
    uint32_t t = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        s += x[j];
        t ^= k[j];
    }
    out[i] = s + (MIX == 1 ? (float)(t & 1u) : 0.0f);
}

// Atomic throughput probes (om_math_eval fn 17..19): ATOM_ITERS 64-bit atomic adds per
// thread at addresses spread by a multiplicative hash, so that the trace kernels' atomic
// rates can be set against what this GPU sustains in the same forms:
//   17: device memory (out, as n / 2 words), not returning;
//   18: device memory, returning (the form of the window kernels' flushes, whose carry check
//       needs the old value);
//   19: LDS, returning, over a WIN_PX_MAX-word array in one 1024-thread workgroup per CU
//       (the window kernels' hits); out[i] gets the thread's sum of the old values.
// 17 and 18 leave their sums in out's words; 18 keeps the old values alive in registers.
constexpr int ATOM_ITERS = 64;

__device__ __forceinline__ uint32_t probe_slot(uint32_t i, uint32_t j, uint32_t n)
{
    const uint32_t h = (i * (uint32_t)ATOM_ITERS + j) * 0x9E3779B1u;
    return (uint32_t)(((unsigned long long)h * n) >> 32); // in [0, n)
}

template <int KIND>
__global__ __launch_bounds__(WIN_WGS) void atomic_probe_kernel(const float *in, float *out, int n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long v = 1ull + (i < (uint32_t)n ? (unsigned long long)(f2u(in[i]) & 0xffu) : 0ull);
    if (KIND == 19) {
        __shared__ unsigned long long s[WIN_PX_MAX];
        for (int k = threadIdx.x; k < WIN_PX_MAX; k += WIN_WGS) s[k] = 0ull;
        __syncthreads();
        unsigned long long acc = 0ull;
#pragma unroll 8
        for (int j = 0; j < ATOM_ITERS; ++j)
            acc += __hip_atomic_fetch_add(s + probe_slot(i, (uint32_t)j, (uint32_t)WIN_PX_MAX), v, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
        if (i < (uint32_t)n) out[i] = (float)(acc & 0xffffffu);
        return;
    }
    if (i >= (uint32_t)n) return;
    unsigned long long *w = (unsigned long long *)out;
    const uint32_t nw = (uint32_t)n / 2u;
    unsigned long long acc = 0ull;
#pragma unroll 8
    for (int j = 0; j < ATOM_ITERS; ++j) {
        unsigned long long *p = w + probe_slot(i, (uint32_t)j, nw);
        if (KIND == 18) acc += __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (KIND == 18) asm volatile("" ::"v"(acc));
}

// ---------------------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------------------
om_status launch_poisson(om_scene *s, hipStream_t st)
{
    if (s->H <= 0) return OM_OK;
    size_t lds = sizeof(float2) * (size_t)s->poisson_cap;
    if (lds > 96 * 1024) {
        set_error("Poisson cell grid exceeds LDS (pitch too small for the heliostat)");
        return OM_ERR_UNSUPPORTED;
    }
    PoissonArgs a;
    a.helio = s->d_helio;
    a.r = s->prm.helio_pixel_length / sqrt_(1.5f);
    a.cell = a.r / sqrt_(2.0f);
    a.k = s->prm.poisson_k;
    a.cap = s->poisson_cap;
    a.key0 = (uint32_t)s->prm.seed;
    a.key1 = (uint32_t)(s->prm.seed >> 32);
    a.w3 = (s->prm.scene_index << 4) | PURPOSE_POISSON;
    a.out = s->d_poisson;
    a.count = s->d_mcount;
    hipLaunchKernelGGL(poisson_kernel, dim3(s->H), dim3(WG), lds, st, a);
    OM_HIP(hipGetLastError());
    OM_HIP(hipStreamSynchronize(st));
    return OM_OK;
}

// the Buie tables go to LDS when they fit next to the hit slots
static bool buie_in_lds(int nbins) { return (size_t)nbins * (sizeof(float4) + sizeof(int)) <= 24 * 1024; }

om_status launch_trace(om_scene *s, const TraceArgs &a, bool records, hipStream_t st)
{
    if (a.ntiles <= 0) return OM_OK;
    if (a.rays_per_thread > RPT_MAX) {
        set_error("rays per thread above the hit-slot count");
        return OM_ERR_INVALID;
    }
    const bool buie = s->prm.sunshape == OM_SUN_BUIE;
    const bool blds = buie && buie_in_lds(a.nbins);
    size_t lds = (size_t)RPT_MAX * WG * (sizeof(int) + sizeof(uint32_t));
    if (blds) lds += (sizeof(float4) + sizeof(int)) * (size_t)a.nbins;
    dim3 grid(a.ntiles), block(WG);
    const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
#define OM_TRACE_LAUNCH(SUN, BL)                                                                       \
    do {                                                                                               \
        if (poisson) {                                                                                 \
            if (records) hipLaunchKernelGGL((trace_kernel<SUN, true, true, BL>), grid, block, lds, st, a);  \
            else hipLaunchKernelGGL((trace_kernel<SUN, true, false, BL>), grid, block, lds, st, a);         \
        } else {                                                                                       \
            if (records) hipLaunchKernelGGL((trace_kernel<SUN, false, true, BL>), grid, block, lds, st, a); \
            else hipLaunchKernelGGL((trace_kernel<SUN, false, false, BL>), grid, block, lds, st, a);        \
        }                                                                                              \
    } while (0)
    if (buie) {
        if (blds) OM_TRACE_LAUNCH(OM_SUN_BUIE, true);
        else OM_TRACE_LAUNCH(OM_SUN_BUIE, false);
    } else {
        OM_TRACE_LAUNCH(OM_SUN_PILLBOX, false);
    }
#undef OM_TRACE_LAUNCH
    OM_HIP(hipGetLastError());
    return OM_OK;
}

om_status launch_finalize(const unsigned long long *energy, float *raw, float *flux, float scale, int n,
                          hipStream_t st)
{
    hipLaunchKernelGGL(finalize_kernel, dim3((n + WG - 1) / WG), dim3(WG), 0, st, energy, raw, flux,
                       scale, n);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

om_status launch_smooth(const float *in, float *out, int w, int h, int radius, float ratio, float scale,
                        hipStream_t st)
{
    int kcap = (int)((float)((2 * radius + 1) * (2 * radius + 1)) * ratio + 1);
    if (kcap < 1) kcap = 1;
    int threads = WG;
    while (threads > 64 && sizeof(float) * 2 * (size_t)(kcap + 1) * threads > 64 * 1024) threads /= 2;
    size_t lds = sizeof(float) * 2 * (size_t)(kcap + 1) * threads;
    if (lds > 64 * 1024) {
        set_error("trimmed-mean window too large");
        return OM_ERR_UNSUPPORTED;
    }
    int n = w * h;
    hipLaunchKernelGGL(smooth_kernel, dim3((n + threads - 1) / threads), dim3(threads), lds, st, in, out,
                       w, h, radius, ratio, kcap, scale);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

template <int SUN, bool POISSON>
static om_status launch_batch_wgs(int wgs, const TraceArgs *table, const int4 *items, int n_items,
                                  unsigned int *queue, int nblocks, hipStream_t st)
{
    const size_t lds = (size_t)RPT_MAX * wgs * (sizeof(int) + sizeof(uint32_t));
    switch (wgs) {
    case 128:
        hipLaunchKernelGGL((trace_batch_kernel<SUN, POISSON, 128>), dim3(nblocks), dim3(128), lds, st, table,
                           items, n_items, queue);
        break;
    case 256:
        hipLaunchKernelGGL((trace_batch_kernel<SUN, POISSON, 256>), dim3(nblocks), dim3(256), lds, st, table,
                           items, n_items, queue);
        break;
    case 512:
        hipLaunchKernelGGL((trace_batch_kernel<SUN, POISSON, 512>), dim3(nblocks), dim3(512), lds, st, table,
                           items, n_items, queue);
        break;
    default:
        set_error("batch workgroup size must be 128, 256 or 512");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipGetLastError());
    return OM_OK;
}

om_status launch_trace_batch(int sunshape, bool poisson, int wgs, const TraceArgs *table, const int4 *items,
                             int n_items, unsigned int *queue, int nblocks, hipStream_t st)
{
    if (sunshape == OM_SUN_BUIE)
        return poisson ? launch_batch_wgs<OM_SUN_BUIE, true>(wgs, table, items, n_items, queue, nblocks, st)
                       : launch_batch_wgs<OM_SUN_BUIE, false>(wgs, table, items, n_items, queue, nblocks, st);
    return poisson ? launch_batch_wgs<OM_SUN_PILLBOX, true>(wgs, table, items, n_items, queue, nblocks, st)
                   : launch_batch_wgs<OM_SUN_PILLBOX, false>(wgs, table, items, n_items, queue, nblocks, st);
}

template <int SUN, bool POISSON, int WGS>
static int batch_occupancy()
{
    int n = 0;
    const size_t lds = (size_t)RPT_MAX * WGS * (sizeof(int) + sizeof(uint32_t));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_kernel<SUN, POISSON, WGS>, WGS, lds) !=
        hipSuccess)
        return 0;
    return n;
}

template <int SUN, bool POISSON>
static int batch_occupancy_wgs(int wgs)
{
    return wgs == 128 ? batch_occupancy<SUN, POISSON, 128>()
                      : (wgs == 256 ? batch_occupancy<SUN, POISSON, 256>() : batch_occupancy<SUN, POISSON, 512>());
}

int trace_batch_blocks_per_cu(int sunshape, bool poisson, int wgs)
{
    if (sunshape == OM_SUN_BUIE)
        return poisson ? batch_occupancy_wgs<OM_SUN_BUIE, true>(wgs) : batch_occupancy_wgs<OM_SUN_BUIE, false>(wgs);
    return poisson ? batch_occupancy_wgs<OM_SUN_PILLBOX, true>(wgs) : batch_occupancy_wgs<OM_SUN_PILLBOX, false>(wgs);
}

om_status launch_trace_batch_window(int sunshape, bool poisson, const TraceArgs *table, const int4 *chunks,
                                    int n_chunks, unsigned int *queue, int nwg, hipStream_t st)
{
    const dim3 grid(nwg), block(WIN_WGS);
    if (sunshape == OM_SUN_BUIE) {
        if (poisson) hipLaunchKernelGGL((trace_batch_window_kernel<OM_SUN_BUIE, true>), grid, block, 0, st, table, chunks, n_chunks, queue);
        else hipLaunchKernelGGL((trace_batch_window_kernel<OM_SUN_BUIE, false>), grid, block, 0, st, table, chunks, n_chunks, queue);
    } else {
        if (poisson) hipLaunchKernelGGL((trace_batch_window_kernel<OM_SUN_PILLBOX, true>), grid, block, 0, st, table, chunks, n_chunks, queue);
        else hipLaunchKernelGGL((trace_batch_window_kernel<OM_SUN_PILLBOX, false>), grid, block, 0, st, table, chunks, n_chunks, queue);
    }
    OM_HIP(hipGetLastError());
    return OM_OK;
}

int trace_batch_window_blocks_per_cu(int sunshape, bool poisson)
{
    int n = 0;
    hipError_t e;
    if (sunshape == OM_SUN_BUIE)
        e = poisson ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_window_kernel<OM_SUN_BUIE, true>, WIN_WGS, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_window_kernel<OM_SUN_BUIE, false>, WIN_WGS, 0);
    else
        e = poisson ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_window_kernel<OM_SUN_PILLBOX, true>, WIN_WGS, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_window_kernel<OM_SUN_PILLBOX, false>, WIN_WGS, 0);
    return e == hipSuccess ? n : 0;
}

void window_rect(int rx, int ry, int &x0, int &y0, int &w, int &h)
{
    if ((long long)rx * ry <= WIN_PX_MAX) {
        w = rx;
        h = ry;
    } else {
        int side = 1;
        while ((side + 1) * (side + 1) <= WIN_PX_MAX) ++side;
        w = std::min(rx, side);
        h = std::min(ry, WIN_PX_MAX / w);
        w = std::min(rx, WIN_PX_MAX / h);
    }
    x0 = (rx - w) / 2;
    y0 = (ry - h) / 2;
}

int window_chunk_rays() { return WIN_CHUNK; }

om_status launch_reduce_batch(const ReduceSlot *slots, int nslot, int max_npx, hipStream_t st)
{
    hipLaunchKernelGGL(reduce_batch_kernel, dim3((max_npx + WG - 1) / WG, nslot), dim3(WG), 0, st, slots);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

__global__ void queue_reset_kernel(uint32_t *q, uint32_t n_items, uint32_t layout)
{
    const int i = threadIdx.x;
    if (i < QUEUE_WORDS) q[i] = i < QUEUE_COUNTERS ? 0u : (i == QUEUE_COUNTERS ? n_items : (i == QUEUE_COUNTERS + 1 ? layout : 0u));
}

om_status launch_queue_reset(uint32_t *q, long long n_items, int layout, hipStream_t st)
{
    if (n_items < 0 || n_items > 0x7fffffffLL) return OM_ERR_SIZE;
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, st, q, (uint32_t)n_items, (uint32_t)layout);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

void set_current_queue(om_ctx *ctx, uint32_t *q)
{
    std::lock_guard<std::mutex> lk(ctx->mon_mu);
    ctx->cur_queue = q;
}

om_status launch_fill_u32(uint32_t *p, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, p, n);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

om_status launch_math(int fn, const float *in, float *out, int n, hipStream_t st)
{
    if (fn < 0 || fn > 19) {
        set_error("math function index must be in 0..19");
        return OM_ERR_INVALID;
    }
    if (fn == 16) return launch_erfinv_ray((const uint32_t *)in, (double *)out, n, st);
    if (fn >= 17) {
        if (fn != 19 && n < 2) {
            set_error("the device-memory atomic probes need n >= 2");
            return OM_ERR_INVALID;
        }
        const dim3 ag((unsigned)((n + WIN_WGS - 1) / WIN_WGS)), ab(WIN_WGS);
        if (fn == 17) hipLaunchKernelGGL(atomic_probe_kernel<17>, ag, ab, 0, st, in, out, n);
        else if (fn == 18) hipLaunchKernelGGL(atomic_probe_kernel<18>, ag, ab, 0, st, in, out, n);
        else hipLaunchKernelGGL(atomic_probe_kernel<19>, ag, ab, 0, st, in, out, n);
        OM_HIP(hipGetLastError());
        return OM_OK;
    }
    const dim3 grid((n + WG - 1) / WG), block(WG);
    if (fn == 12) hipLaunchKernelGGL(valu_probe_kernel<0>, grid, block, 0, st, in, out, n);
    else if (fn == 13) hipLaunchKernelGGL(valu_probe_kernel<1>, grid, block, 0, st, in, out, n);
    else if (fn == 14) hipLaunchKernelGGL(valu_probe_kernel<2>, grid, block, 0, st, in, out, n);
    else if (fn == 15) hipLaunchKernelGGL(valu_probe_kernel<3>, grid, block, 0, st, in, out, n);
    else hipLaunchKernelGGL(math_kernel, grid, block, 0, st, fn, in, out, n);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

static om_status clear_maps(const om_flux *out, long long n, hipStream_t st)
{
    dim3 g((unsigned)((n + WG - 1) / WG)), b(WG);
    hipLaunchKernelGGL(fill_u32_kernel, g, b, 0, st, out->counts, n);
    hipLaunchKernelGGL(fill_u64_kernel, g, b, 0, st, (unsigned long long *)out->energy, n);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

static om_status check_flux(const om_scene *s, const om_flux *out)
{
    if (!out || !out->counts || !out->energy) {
        set_error("flux output needs counts and energy");
        return OM_ERR_INVALID;
    }
    if (out->capacity < (int64_t)s->res_x * s->res_y) {
        set_error("flux buffers smaller than the receiver map");
        return OM_ERR_SIZE;
    }
    return OM_OK;
}

// Persistent workgroups of trace_window_kernel: one per CU (its LDS takes the whole CU)
template <int SUN, bool POISSON>
static int window_blocks_per_cu()
{
    int n = 0;
    int nc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_window_kernel<SUN, POISSON, false>, WIN_WGS, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nc, trace_window_kernel<SUN, POISSON, true>, WIN_WGS, 0) !=
            hipSuccess)
        return 0;
    return std::min(n, nc); // one workgroup count for both forms
}

template <bool COMPACT>
static void launch_trace_window_as(om_scene *s, const TraceArgs &a, int nwg, hipStream_t st)
{
    const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
    dim3 grid(nwg), block(WIN_WGS);
    if (s->prm.sunshape == OM_SUN_BUIE) {
        if (poisson) hipLaunchKernelGGL((trace_window_kernel<OM_SUN_BUIE, true, COMPACT>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((trace_window_kernel<OM_SUN_BUIE, false, COMPACT>), grid, block, 0, st, a);
    } else {
        if (poisson) hipLaunchKernelGGL((trace_window_kernel<OM_SUN_PILLBOX, true, COMPACT>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((trace_window_kernel<OM_SUN_PILLBOX, false, COMPACT>), grid, block, 0, st, a);
    }
}

static om_status launch_trace_window(om_scene *s, const TraceArgs &a, int nwg, hipStream_t st)
{
    if (s->prm.ray_compact) launch_trace_window_as<true>(s, a, nwg, st);
    else launch_trace_window_as<false>(s, a, nwg, st);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

// Window of OM_ACC_WINDOW: the whole map when it fits in WIN_PX_MAX words, else the largest
// near-square block of rows and columns about the map centre (every heliostat aims at the
// receiver centre); nwg = CUs x resident workgroups per CU.  The workgroup count of a device
// is looked up once.
static om_status window_plan(om_scene *s, WindowRef &win, int &nwg)
{
    const int tile_rays = s->rays_per_thread * WG;
    if (tile_rays < 64 || (tile_rays & (tile_rays - 1)) != 0 || (tile_rays > WIN_CHUNK && tile_rays % WIN_CHUNK != 0)) {
        set_error("window accumulation needs tiles of a power of two of at least 64 rays");
        return OM_ERR_UNSUPPORTED;
    }
    if (s->ctx->win_wgs < 0) {
        int cus = 0;
        OM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->ctx->device));
        const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
        int per = s->prm.sunshape == OM_SUN_BUIE
                      ? (poisson ? window_blocks_per_cu<OM_SUN_BUIE, true>() : window_blocks_per_cu<OM_SUN_BUIE, false>())
                      : (poisson ? window_blocks_per_cu<OM_SUN_PILLBOX, true>()
                                 : window_blocks_per_cu<OM_SUN_PILLBOX, false>());
        if (per < 1) {
            set_error("trace_window_kernel cannot be resident on this device");
            return OM_ERR_UNSUPPORTED;
        }
        s->ctx->win_wgs = cus * per;
    }
    nwg = s->ctx->win_wgs;
    window_rect(s->res_x, s->res_y, win.x0, win.y0, win.w, win.h);
    win.res_x = s->res_x;
    return OM_OK;
}

// tiles of heliostats [h0, h1) are contiguous in the tile table
void tile_range(const om_scene *s, int h0, int h1, int &t0, int &t1)
{
    auto lo = std::lower_bound(s->tiles.begin(), s->tiles.end(), h0,
                               [](const int2 &t, int h) { return t.x < h; });
    auto hi = std::lower_bound(s->tiles.begin(), s->tiles.end(), h1,
                               [](const int2 &t, int h) { return t.x < h; });
    t0 = (int)(lo - s->tiles.begin());
    t1 = (int)(hi - s->tiles.begin());
}

} // namespace om

using namespace om;

namespace om {
// Layout of a trace's workspace for scene s: 8 << shift packed copies of the map (same-pixel
// atomics of concurrent workgroups then mostly land on different words), the carry and wrap
// maps, and in window mode one window slice per persistent workgroup (the queues are the
// context's).  The copies double while twice as many still fit in copy_budget bytes (at least
// one per XCD) and, with a cap on the total, halve again until the whole layout fits.
struct WsPlan {
    int shift = 0;
    size_t n16 = 0, bytes = 0, off_win = 0;
    size_t off_slice_q = 0; // window mode: the queue of each workgroup (nwg ints)
    WindowRef win{nullptr, 0, 0, 0, 0, 0, 0, nullptr, {}, {}};
    int nwg = 0;
    size_t off_redo = 0; // OM_PREC_MIXED window traces: the redo list of redo_cap entries
    int redo_cap = 0;
};

// whether a trace of s takes the mixed path (else OM_PREC_MIXED traces in FP64)
static bool mixed_window(const om_scene *s)
{
    return s->prm.precision == OM_PREC_MIXED && s->prm.accumulate == OM_ACC_WINDOW && mixed_supported(s);
}

static om_status ws_plan(om_scene *s, size_t copy_budget, size_t cap, WsPlan &p)
{
    const size_t npx = (size_t)s->res_x * s->res_y;
    p = WsPlan{};
    p.win.res_x = s->res_x;
    if (s->prm.accumulate == OM_ACC_WINDOW) {
        om_status rc = window_plan(s, p.win, p.nwg);
        if (rc != OM_OK) return rc;
    }
    int shift = 0;
    while (shift < COPY_SHIFT_MAX && ((size_t)16 << shift) * npx * sizeof(unsigned long long) <= copy_budget) ++shift;
    // the redo list of a mixed trace: one entry per 32 rays, 2^16 to 2^24 entries (a launch
    // whose list fills up is traced again wholly in FP64)
    long long want = !mixed_window(s) ? 0
                     : s->prm.redo_cap > 0 ? s->prm.redo_cap
                                           : std::min<long long>(std::max<long long>(s->total_rays / 32, 1 << 16), 1 << 24);
    for (;;) {
        for (int sh = shift;; --sh) {
            const size_t ncopy = (size_t)8 << sh;
            p.shift = sh;
            p.n16 = ncopy * npx;
            p.off_win = p.n16 * sizeof(unsigned long long) + 2 * npx * sizeof(uint32_t);
            p.bytes = p.off_win + (p.nwg > 0 ? (size_t)p.nwg * p.win.w * p.win.h * sizeof(unsigned long long) : 0);
            p.off_slice_q = p.bytes;
            if (p.nwg > 0) p.bytes += (size_t)p.nwg * sizeof(int);
            p.off_redo = (p.bytes + 255) & ~(size_t)255;
            p.redo_cap = (int)want;
            if (want > 0) p.bytes = p.off_redo + (size_t)want * sizeof(uint2);
            if (p.bytes <= cap) return OM_OK;
            if (sh == 0) break;
        }
        if (want <= 4096 || s->prm.redo_cap > 0) break;
        want /= 2;
    }
    set_error("the trace needs " + std::to_string(p.bytes) + " bytes of workspace, more than the " +
              std::to_string(cap) + " available");
    return OM_ERR_SIZE;
}

size_t trace_workspace_bytes(om_scene *s)
{
    WsPlan p;
    return ws_plan(s, PART_BYTES, ~(size_t)0, p) == OM_OK ? p.bytes : 0;
}

Workspace ctx_workspace(om_ctx *ctx)
{
    Workspace w;
    w.grow = &ctx->parts;
    return w;
}

// heliostats [h0, h1), all held by grid g, accumulated into out through the workspace plan p
static om_status trace_grid(om_scene *s, int g, int h0, int h1, const om_flux *out, hipStream_t st,
                            const WsPlan &p, char *base)
{
    TraceArgs a = make_trace_args(s, g);
    int t0, t1;
    tile_range(s, h0, h1, t0, t1);
    a.tile_base = t0;
    a.ntiles = t1 - t0;
    const int npx = s->res_x * s->res_y;
    om_status rc;
    // OM_PREC_MIXED takes the mixed window trace; without the window accumulation, or where
    // the mixed path does not apply, it traces in FP64 (the same outcome for every ray)
    const bool mixed = mixed_window(s);
    const bool fp64 = s->prm.precision == OM_PREC_FP64 || (s->prm.precision == OM_PREC_MIXED && !mixed);
    if (s->prm.accumulate == OM_ACC_PACKED || s->prm.accumulate == OM_ACC_WINDOW) {
        const int ncopy = 8 << p.shift;
        a.part = (unsigned long long *)base;
        a.carry = (uint32_t *)(a.part + p.n16);
        a.wrap = a.carry + npx;
        a.npx = npx;
        a.copy_shift = p.shift;
        a.copy_stride = npx;
        a.px_stride = 1;
        a.acc_packed = 1;
        a.pack_shift = s->prm.pack_shift > 0 ? s->prm.pack_shift : 44;
        WindowRef win = p.win;
        win.nwin = 0;
        if (fp64 && !(p.nwg > 0 && a.ntiles > 0)) {
            if ((rc = launch_trace64(s, a, trace64_args(s, g), false, st)) != OM_OK) return rc;
        } else if (p.nwg > 0 && a.ntiles > 0) {
            a.win_queue = s->ctx->d_queue;
            a.win_part = (unsigned long long *)(base + p.off_win);
            a.win_x0 = win.x0;
            a.win_y0 = win.y0;
            a.win_w = win.w;
            a.win_h = win.h;
            win.part = a.win_part;
            win.nwin = p.nwg;
            // a whole-grid trace of an azimuthal map: the queues take the grid's azimuth
            // sectors, each with its own window
            const GridState &G = s->grid_src ? s->grid_src->grids[(size_t)g] : s->grids[(size_t)g];
            if (G.win_sectors && G.d_win_order && t0 == G.t0 && t1 == G.t1) {
                a.win_order = G.d_win_order;
                a.win_slice_q = (int *)(base + p.off_slice_q);
                win.slice_q = a.win_slice_q;
                for (int q = 0; q < 8; ++q) {
                    a.win_qx0[q] = win.qx0[q] = G.win_qx0[q];
                    a.win_qy0[q] = win.qy0[q] = G.win_qy0[q];
                }
            }
            const int chunk = std::min(a.tile_rays, WIN_CHUNK);
            if ((rc = launch_queue_reset(a.win_queue, (long long)a.ntiles * (a.tile_rays / chunk), QUEUE_EIGHTHS,
                                         st)) != OM_OK)
                return rc;
            set_current_queue(s->ctx, a.win_queue);
            if (mixed) {
                MixArgs x;
                mixed_args(s, g, a, x);
                x.redo = (uint2 *)(base + p.off_redo);
                x.redo_cap = p.redo_cap;
                rc = launch_trace_mixed_window(s, a, trace64_args(s, g), x, p.nwg, (uint32_t *)base,
                                               (long long)(p.off_win / 4), st);
            } else {
                rc = fp64 ? launch_trace64_window(s, a, trace64_args(s, g), p.nwg, st) : launch_trace_window(s, a, p.nwg, st);
            }
            if (rc != OM_OK) return rc;
        } else if ((rc = launch_trace(s, a, false, st)) != OM_OK) {
            return rc;
        }
        hipLaunchKernelGGL(reduce_packed_kernel, dim3((npx + WG - 1) / WG), dim3(WG), 0, st, a.part, a.carry,
                           a.wrap, a.pack_shift, out->counts, (unsigned long long *)out->energy, npx,
                           ncopy, a.copy_stride, a.px_stride, win);
        OM_HIP(hipGetLastError());
        // the workgroups' queue numbers are read by the decode, which leaves them: cleared
        // here, since a later trace may take these words as packed copies
        if (win.slice_q) return launch_fill_u32((uint32_t *)win.slice_q, p.nwg, st);
        return OM_OK;
    }
    a.counts = out->counts;
    a.energy = (unsigned long long *)out->energy;
    return fp64 ? launch_trace64(s, a, trace64_args(s, g), false, st) : launch_trace(s, a, false, st);
}

om_status trace_into(om_scene *s, int h0, int h1, const om_flux *out, bool clear, hipStream_t st,
                     const Workspace &ws)
{
    om_status rc;
    WsPlan p;
    char *base = nullptr;
    const bool packed = s->prm.accumulate == OM_ACC_PACKED || s->prm.accumulate == OM_ACC_WINDOW;
    // a workspace that is too small fails before anything is launched
    if (packed && (rc = ws_plan(s, PART_BYTES, ws.grow ? ~(size_t)0 : ws.bytes, p)) != OM_OK) return rc;
    if (clear && (rc = clear_maps(out, s->res_x * s->res_y, st)) != OM_OK) return rc;
    if (packed) {
        size_t clear_bytes = 0;
        if (ws.grow) {
            // a new buffer is cleared once; the decode leaves the words it reads cleared
            if (ws.grow->bytes < p.bytes) clear_bytes = p.bytes;
            if ((rc = ensure_buf(*ws.grow, p.bytes)) != OM_OK) return rc;
            base = (char *)ws.grow->p;
        } else {
            base = (char *)ws.p;
            if (ws.clear_each) clear_bytes = p.off_win; // copies, carry and wrap maps
        }
        if (clear_bytes > 0 && (rc = launch_fill_u32((uint32_t *)base, (long long)(clear_bytes / 4), st)) != OM_OK)
            return rc;
    }
    // the count of rays the mixed trace sends to FP64 starts again
    if (mixed_window(s) && (rc = launch_fill_u32(s->d_mix_n + 1, 2, st)) != OM_OK) return rc;
    // one launch (and decode, which adds into out) per grid: a grid's CSR holds its own
    // heliostats only
    for (size_t g = 0; g < s->grids.size(); ++g) {
        const int a0 = std::max(h0, s->grids[g].h0), a1 = std::min(h1, s->grids[g].h1);
        if (a0 < a1 && (rc = trace_grid(s, (int)g, a0, a1, out, st, p, base)) != OM_OK) return rc;
    }
    return OM_OK;
}

om_status flux_finalize(om_scene *s, const om_flux *out, float *raw, hipStream_t st)
{
    const int n = s->res_x * s->res_y;
    om_status rc = launch_finalize((const unsigned long long *)out->energy, out->flux_smooth ? raw : nullptr,
                                   out->flux, s->scale, n, st);
    if (rc != OM_OK) return rc;
    if (out->flux_smooth)
        return launch_smooth(raw, out->flux_smooth, s->res_x, s->res_y, s->prm.smooth_radius, s->prm.smooth_ratio,
                             s->scale, st);
    return OM_OK;
}
} // namespace om

extern "C" om_status om_trace_heliostats(om_scene *s, int h0, int h1, const om_flux *out, void *stream)
{
    if (!s) return OM_ERR_INVALID;
    om_status st = check_flux(s, out);
    if (st != OM_OK) return st;
    if (h0 < 0 || h1 > s->H || h0 > h1) {
        set_error("heliostat range out of bounds");
        return OM_ERR_INVALID;
    }
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    return trace_into(s, h0, h1, out, false, (hipStream_t)stream, ctx_workspace(s->ctx));
}

extern "C" om_status om_flux_finalize(om_scene *s, const om_flux *out, void *stream)
{
    if (!s) return OM_ERR_INVALID;
    om_status st = check_flux(s, out);
    if (st != OM_OK) return st;
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    float *raw = nullptr;
    if (out->flux_smooth) {
        if ((st = ensure_buf(s->ctx->scratch, sizeof(float) * (size_t)s->res_x * s->res_y)) != OM_OK) return st;
        raw = (float *)s->ctx->scratch.p;
    }
    return flux_finalize(s, out, raw, (hipStream_t)stream);
}

extern "C" om_status om_trace_plain(om_scene *s, const om_flux *out, void *stream)
{
    if (!s) return OM_ERR_INVALID;
    om_status st = check_flux(s, out);
    if (st != OM_OK) return st;
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    if ((st = trace_into(s, 0, s->H, out, true, (hipStream_t)stream, ctx_workspace(s->ctx))) != OM_OK) return st;
    if (out->flux || out->flux_smooth) return om_flux_finalize(s, out, stream);
    return OM_OK;
}

// ---------------------------------------------------------------------------------------
// Allocation-free trace: the workspace is a static device array
// ---------------------------------------------------------------------------------------
// The map copies, side maps, window queue and slices of om_trace_plain_noalloc, and its
// smoothing input at the end.  Each device holds its own copy for the life of the process.
// In host code `scratch` names a host-side stand-in, not device memory: host code uses the
// per-device address that static_scratch looks up, never the array's name.
static __device__ __align__(256) unsigned long long scratch[OM_NOALLOC_SCRATCH_BYTES / 8];

namespace {
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_addr; // [device] device address of scratch, null before lookup
} // namespace

om_status om::static_scratch(int device, void **addr)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return OM_ERR_INVALID;
    if ((int)g_scratch_addr.size() <= device) g_scratch_addr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_addr[device] = p;
    }
    *addr = g_scratch_addr[device];
    return OM_OK;
}

extern "C" om_status om_trace_plain_noalloc(om_scene *s, const om_flux *out, void *stream)
{
    if (!s) return OM_ERR_INVALID;
    om_status st = check_flux(s, out);
    if (st != OM_OK) return st;
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    if (!s->ctx->static_scratch) {
        set_error("the context has no static scratch");
        return OM_ERR_INVALID;
    }
    const size_t npx = (size_t)s->res_x * s->res_y;
    // smoothing input at the end of the scratch, the trace workspace before it
    const size_t raw_bytes = out->flux_smooth ? ((npx * sizeof(float) + 255) & ~(size_t)255) : 0;
    if (raw_bytes >= OM_NOALLOC_SCRATCH_BYTES) {
        set_error("the smoothing input of a " + std::to_string(npx) + "-pixel map does not fit in the " +
                  std::to_string((size_t)OM_NOALLOC_SCRATCH_BYTES) + "-byte static scratch");
        return OM_ERR_SIZE;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t hs = (hipStream_t)stream;
    Workspace ws;
    ws.p = s->ctx->static_scratch;
    ws.bytes = OM_NOALLOC_SCRATCH_BYTES - raw_bytes;
    ws.clear_each = true; // every call on the device shares the array
    if ((st = trace_into(s, 0, s->H, out, true, hs, ws)) != OM_OK) return st;
    if (out->flux || out->flux_smooth)
        return flux_finalize(s, out, out->flux_smooth ? (float *)((char *)ws.p + ws.bytes) : nullptr, hs);
    return OM_OK;
}

extern "C" om_status om_autotune_accumulate(om_scene *s, const om_flux *maps, int reps, float *ms, int *best)
{
    if (!s || reps <= 0) return OM_ERR_INVALID;
    om_status st = check_flux(s, maps);
    if (st != OM_OK) return st;
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    if (s->prm.accumulate == OM_ACC_DEVICE) {
        set_error("the scene accumulates with OM_ACC_DEVICE");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    const int modes[2] = {OM_ACC_PACKED, OM_ACC_WINDOW};
    const int keep = s->prm.accumulate;
    float t[2] = {-1.0f, -1.0f};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) st = hip_fail(he, "hipEventCreate");
    for (int i = 0; i < 2 && st == OM_OK; ++i) {
        s->prm.accumulate = modes[i];
        // warm-up (buffers, first launch); a mode that cannot run on this scene is skipped
        if (trace_into(s, 0, s->H, maps, true, nullptr, ctx_workspace(s->ctx)) != OM_OK) continue;
        for (int r = 0; r < reps && st == OM_OK; ++r) {
            float ms1 = 0.0f;
            if ((he = hipEventRecord(e0, nullptr)) != hipSuccess) break;
            if ((st = trace_into(s, 0, s->H, maps, true, nullptr, ctx_workspace(s->ctx))) != OM_OK) break;
            if ((he = hipEventRecord(e1, nullptr)) != hipSuccess || (he = hipEventSynchronize(e1)) != hipSuccess ||
                (he = hipEventElapsedTime(&ms1, e0, e1)) != hipSuccess)
                break;
            if (t[i] < 0.0f || ms1 < t[i]) t[i] = ms1;
        }
        if (st == OM_OK && he != hipSuccess) st = hip_fail(he, "accumulation autotune timing");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st != OM_OK || t[0] < 0.0f) {
        s->prm.accumulate = keep;
        return st != OM_OK ? st : OM_ERR_HIP;
    }
    s->prm.accumulate = (t[1] >= 0.0f && t[1] < t[0]) ? OM_ACC_WINDOW : OM_ACC_PACKED;
    if (ms) {
        ms[0] = t[0];
        ms[1] = t[1];
    }
    if (best) *best = s->prm.accumulate;
    return OM_OK;
}

extern "C" om_status om_trace_records(om_scene *s, int h0, int h1, int32_t *shade, int32_t *block,
                                      int32_t *pixel, int64_t capacity, void *stream)
{
    if (!s || !shade || !block || !pixel) return OM_ERR_INVALID;
    if (h0 < 0 || h1 > s->H || h0 > h1) {
        set_error("heliostat range out of bounds");
        return OM_ERR_INVALID;
    }
    long long need = s->ray_offset[h1] - s->ray_offset[h0];
    if (capacity < need) {
        set_error("record buffers too small");
        return OM_ERR_SIZE;
    }
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    // records never touch the maps; any valid pointers do
    if (om_status st = ensure_buf(s->ctx->scratch2, 16 * (size_t)s->res_x * s->res_y); st != OM_OK) return st;
    const bool mixed = s->prm.precision == OM_PREC_MIXED && mixed_supported(s);
    if (mixed) {
        if (om_status st = launch_fill_u32(s->d_mix_n + 1, 2, (hipStream_t)stream); st != OM_OK) return st;
    }
    for (size_t g = 0; g < s->grids.size(); ++g) {
        const int a0 = std::max(h0, s->grids[g].h0), a1 = std::min(h1, s->grids[g].h1);
        if (a0 >= a1) continue;
        TraceArgs a = make_trace_args(s, (int)g);
        int t0, t1;
        tile_range(s, a0, a1, t0, t1);
        a.tile_base = t0;
        a.ntiles = t1 - t0;
        a.counts = (uint32_t *)s->ctx->scratch2.p;
        a.energy = (unsigned long long *)((char *)s->ctx->scratch2.p + 8 * (size_t)s->res_x * s->res_y);
        a.rec_shade = shade;
        a.rec_block = block;
        a.rec_pixel = pixel;
        a.rec_base = s->ray_offset[h0];
        om_status st;
        if (mixed) {
            MixArgs x;
            mixed_args(s, (int)g, a, x);
            st = launch_trace_mixed_records(s, a, trace64_args(s, (int)g), x, (hipStream_t)stream);
        } else if (s->prm.precision != OM_PREC_FP32) {
            st = launch_trace64(s, a, trace64_args(s, (int)g), true, (hipStream_t)stream);
        } else {
            st = launch_trace(s, a, true, (hipStream_t)stream);
        }
        if (st != OM_OK) return st;
    }
    return OM_OK;
}

extern "C" om_status om_scene_redo_count(const om_scene *s, int64_t *rays)
{
    if (!s || !rays) return OM_ERR_INVALID;
    *rays = 0;
    if (!s->ctx || !s->d_mix_n) return OM_OK;
    OM_HIP(hipSetDevice(s->ctx->device));
    OM_HIP(hipDeviceSynchronize());
    uint32_t n = 0;
    OM_HIP(hipMemcpy(&n, s->d_mix_n + 1, sizeof(n), hipMemcpyDeviceToHost));
    *rays = n;
    return OM_OK;
}

extern "C" om_status om_smooth_trimmed(om_ctx *ctx, const float *d_in, float *d_out, int width, int height,
                                       int kernel_radius, float trimmed_ratio, void *stream)
{
    if (!ctx || !d_in || !d_out || width <= 0 || height <= 0 || kernel_radius < 0 ||
        !(trimmed_ratio >= 0.0f) || trimmed_ratio >= 0.5f)
        return OM_ERR_INVALID;
    OM_HIP(hipSetDevice(ctx->device));
    hipStream_t hs = (hipStream_t)stream;
    const float *src = d_in;
    if (d_in == d_out) {
        // the reference reads a texture copy and writes the array in place
        om_status st = ensure_buf(ctx->scratch2, sizeof(float) * (size_t)width * height);
        if (st != OM_OK) return st;
        OM_HIP(hipMemcpyAsync(ctx->scratch2.p, d_in, sizeof(float) * (size_t)width * height,
                              hipMemcpyDeviceToDevice, hs));
        src = (const float *)ctx->scratch2.p;
    }
    return launch_smooth(src, d_out, width, height, kernel_radius, trimmed_ratio, 1.0f, hs);
}

extern "C" om_status om_math_eval(om_ctx *ctx, int fn, const float *d_in, float *d_out, int n, void *stream)
{
    if (!ctx || !d_in || !d_out || n < 0) return OM_ERR_INVALID;
    if (n == 0) return OM_OK;
    OM_HIP(hipSetDevice(ctx->device));
    return launch_math(fn, d_in, d_out, n, (hipStream_t)stream);
}

extern "C" om_status om_results_with_given_normal(om_scene *s, om_result *res)
{
    if (!s || !res) return OM_ERR_INVALID;
    if (!s->ctx) {
        set_error("host-only scene (created without a context) cannot be traced");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    size_t n = (size_t)s->res_x * s->res_y;
    uint32_t *cnt = nullptr;
    uint64_t *en = nullptr;
    float *fl = nullptr, *fs = nullptr;
    om_status st = OM_OK;
    hipError_t e = hipMalloc((void **)&cnt, 4 * n);
    if (e == hipSuccess) e = hipMalloc((void **)&en, 8 * n);
    if (e == hipSuccess) e = hipMalloc((void **)&fl, 4 * n);
    if (e == hipSuccess) e = hipMalloc((void **)&fs, 4 * n);
    if (e != hipSuccess) st = hip_fail(e, "hipMalloc");
    std::vector<float> hf(n), hs(n);
    if (st == OM_OK) {
        om_flux f{cnt, en, fl, fs, (int64_t)n};
        st = om_trace_plain(s, &f, nullptr);
    }
    if (st == OM_OK) {
        e = hipMemcpy(hf.data(), fl, 4 * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(hs.data(), fs, 4 * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = hip_fail(e, "hipMemcpy");
    }
    if (cnt) (void)hipFree(cnt);
    if (en) (void)hipFree(en);
    if (fl) (void)hipFree(fl);
    if (fs) (void)hipFree(fs);
    if (st != OM_OK) return st;
    res->ratio = om_intercept_ratio(hf.data(), s->res_x, s->res_y, s->total_energy);
    res->ratio_smooth = om_intercept_ratio(hs.data(), s->res_x, s->res_y, s->total_energy);
    res->total_energy = s->total_energy;
    res->rays = s->total_rays;
    if (res->flux) std::memcpy(res->flux, hf.data(), 4 * n);
    if (res->flux_smooth) std::memcpy(res->flux_smooth, hs.data(), 4 * n);
    return OM_OK;
}
