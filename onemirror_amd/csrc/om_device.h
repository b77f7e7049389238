// om_device.h -- device helpers shared by the trace kernels of om_kernels.hip and
// om_diff.hip: launch constants, the Buie bin lookup, the short exact division and square
// root, the grid walk, the slope-error normal and the receiver lookup.
#pragma once

#include <hip/hip_runtime.h>

#include "om_internal.h"

namespace om {

constexpr int WG = 256;
constexpr int RPT_MAX = RPT_LIMIT; // rays per thread of one trace workgroup (hit slots in LDS)
constexpr int COPY_SHIFT_MAX = 3;             // up to 64 copies of the packed map
constexpr size_t PART_BYTES = 64u << 20;      // memory budget of the copies

// Workgroup b of an n-workgroup launch runs on XCD b % 8.  Give each XCD a contiguous
// range of tiles (neighbouring heliostats share blockers and grid cells in its L2).
__device__ __forceinline__ int xcd_chunk(int b, int n)
{
    int q = n >> 3, rmd = n & 7;
    int x = b & 7, i = b >> 3;
    return x * q + min(x, rmd) + i;
}

// Buie bin of value v: the largest i with cdf_y[i] <= v, which is what
// samplelights::max_less_index (scene_instance_process.cu:158-173) returns for a strictly
// increasing CDF.  A guide entry starts the walk next to it.
__device__ __forceinline__ float4 buie_bin(const float4 *bins, const int *guide, float gscale, float v,
                                           int n)
{
    int j = min(max(trunc_i(v * gscale), 0), n - 1);
    int i = guide[j];
    float4 b = bins[i];
    while (b.x > v) b = bins[--i];
    while (b.y <= v) b = bins[++i];
    return b;
}

// The same bin from the guide and the upper edge of its bin, held side by side: a value
// beyond that edge starts at the next bin, so the walk of buie_bin takes no step when the
// guide entry's value range meets at most one bin edge (the usual case), and the bin is the
// second load of the lookup, never a third.
__device__ __forceinline__ float4 buie_bin_split(const TraceArgs &a, float v)
{
    const int j = min(max(trunc_i(v * a.buie_gscale), 0), a.nbins - 1);
    const int2 g = a.buie_split[j];
    int i = g.x + (v >= u2f((uint32_t)g.y) ? 1 : 0); // at most nbins - 1: v < int465
    float4 b = a.buie_bins[i];
    while (b.x > v) b = a.buie_bins[--i];
    while (b.y <= v) b = a.buie_bins[++i];
    return b;
}

// ---------------------------------------------------------------------------------------
// Short forms of the correctly rounded division and square root
// ---------------------------------------------------------------------------------------
// The correctly rounded f32 division on gfx9 is
//   div_scale(den), div_scale(num), rcp, fma, fma, mul, fma, fma, fma, div_fmas, div_fixup.
// div_scale rescales an operand only when the denominator, its reciprocal or the quotient is
// denormal, the numerator is within 2^23 ulp of the denormal range or the exponents differ by
// 96 or more; div_fmas is then a plain fma, and div_fixup changes nothing unless an operand
// is zero, infinite or NaN.  For 2^-60 <= |b| <= 2^60 and a == 0 or 2^-90 <= |a| <= 2^60 the
// steps below therefore give the same bits (a zero numerator gives a zero of either sign).
__device__ __forceinline__ float div_rn(float a, float b)
{
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    const float y = __builtin_fmaf(e, y0, y0);
    float q = a * y;
    float r = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

// The correctly rounded f32 square root on gfx9 takes v_sqrt (within 1 ulp) and keeps it or
// one of its two neighbours by the signs of the residuals; inputs below 2^-96 are scaled
// first and zero and infinity pass through.  For 2^-90 <= x <= 2^90 these steps are the same.
__device__ __forceinline__ float sqrt_rn(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = u2f(f2u(s) - 1u), hi = u2f(f2u(s) + 1u);
    const float rlo = __builtin_fmaf(-lo, s, x), rhi = __builtin_fmaf(-hi, s, x);
    const float r = (rlo <= 0.0f) ? lo : s;
    return (rhi > 0.0f) ? hi : r;
}

// Math policy of the trace kernel.  Its divisions and square roots have operands inside the
// ranges above: scenes keep coordinates and sizes within 2^20 m, cell edges within
// [2^-20, 2^20] m and every ray origin strictly inside the grid (checked when the scene is
// built), so
//   walk set-up    |k*edge - cur| / |dir|: numerator 0 or in [2^-44, 2^22], |dir| in (1e-6, 1]
//                  edge / |dir|
//   intersection   1 / det with 1e-6 <= |det| <= |E1| |E2| < 2^42
//   normalize      1 / sqrt(v.v) for |v| = 1 up to rounding
//   log            (m - 1) / (m + 1) with m in [0.7, 1.42)
//   erfinv tail    sqrt(w), w in [5, 90]
//   Buie angle     theta[mrad] / 1000 with theta > 2^-60
struct gfx {
    __device__ static float div(float a, float b) { return div_rn(a, b); }
    __device__ static float sqrt(float x) { return sqrt_rn(x); }
};

// ---------------------------------------------------------------------------------------
// Grid walk: NotCollision + Intersect (rectgrid_raytracing.cu:21-117)
// ---------------------------------------------------------------------------------------
// Amanatides-Woo walk from the cell of the origin (which is never tested); the result is the
// first candidate in CSR order whose parallelogram the ray hits.  The ray's own heliostat is
// skipped (the reference rejects it through t < 1e-6, which holds in exact arithmetic).  The
// walk also ends, with the same result, once the cell index on an axis has passed the
// occupied index range in the direction of travel: every later cell is empty.

// Cell of a ray origin, shared by the two walks of a ray: make_int3((o - pos) / interval)
// (rectgrid_raytracing.cu:43).  With power-of-two cell edges x / 2^k == x * 2^-k.
struct cell3 {
    float rx, ry, rz; // o - grid pos
    int gx, gy, gz;
    int addr;         // (gx * gny + gy) * gnz + gz
};
__device__ __forceinline__ cell3 start_cell(const TraceArgs &a, v3 o)
{
    cell3 c;
    c.rx = o.x - a.gpx;
    c.ry = o.y - a.gpy;
    c.rz = o.z - a.gpz;
    if (a.grid_pow2) {
        c.gx = trunc_i(c.rx * a.inv_gix);
        c.gy = trunc_i(c.ry * a.inv_giy);
        c.gz = trunc_i(c.rz * a.inv_giz);
    } else {
        c.gx = trunc_i(div_rn(c.rx, a.gix));
        c.gy = trunc_i(div_rn(c.ry, a.giy));
        c.gz = trunc_i(div_rn(c.rz, a.giz));
    }
    c.addr = (c.gx * a.gny + c.gy) * a.gnz + c.gz;
    return c;
}

// tMax of one axis, absDivide(calTMax(...), dir) (rectgrid_raytracing.cu:3-17): calTMax is
// (idx + 1) * edge - cur for dir >= 0 and cur - idx * edge otherwise, i.e. +-(k * edge - cur)
// with k = idx + (dir >= 0); IEEE subtraction and division are odd in each operand, so the
// absolute quotient is the same value.  absDivide returns float(INT_MAX) for |dir| <= 1e-6.
__device__ __forceinline__ float first_crossing(float dir, bool fwd, float edge, int idx, float cur)
{
    const float t = (float)(idx + (fwd ? 1 : 0)) * edge - cur;
    const float q = fabs_(div_rn(t, dir));
    return fabs_(dir) <= EPS_LE ? INT_MAX_F : q;
}
// tDelta of one axis, absDivide(interval, dir)
__device__ __forceinline__ float cell_crossing(float dir, float edge)
{
    const float q = fabs_(div_rn(edge, dir));
    return fabs_(dir) <= EPS_LE ? INT_MAX_F : q;
}

__device__ int dda(const TraceArgs &a, v3 o, const cell3 &c, v3 d, int self)
{
    const bool fx = d.x >= 0.0f, fy = d.y >= 0.0f, fz = d.z >= 0.0f;
    // Steps left on each axis before its index leaves the admissible window: the grid,
    // cut at the occupied range in the direction of travel.  The origin lies inside the
    // grid, so only that end can be crossed.
    int nx = fx ? a.occ_hi[0] - c.gx : c.gx - a.occ_lo[0];
    int ny = fy ? a.occ_hi[1] - c.gy : c.gy - a.occ_lo[1];
    int nz = fz ? a.occ_hi[2] - c.gz : c.gz - a.occ_lo[2];
    if ((nx | ny | nz) < 0) return -1;
    // no step left on any axis: the first step, whichever axis it takes, ends the walk
    // before any cell is tested
    if ((nx | ny | nz) == 0) return -1;
    float mx = first_crossing(d.x, fx, a.gix, c.gx, c.rx);
    float my = first_crossing(d.y, fy, a.giy, c.gy, c.ry);
    float mz = first_crossing(d.z, fz, a.giz, c.gz, c.rz);
    const float dx = cell_crossing(d.x, a.gix), dy = cell_crossing(d.y, a.giy),
                dz = cell_crossing(d.z, a.giz);
    // cell index strides of one step on each axis
    const int yz = a.gny * a.gnz;
    const int ax = fx ? yz : -yz, ay = fy ? a.gnz : -a.gnz, az = fz ? 1 : -1;
    // One step: the address of the next cell, false once the walk has left the window.
    int addr = c.addr;
#define OM_WALK_STEP(more)                                                                   \
    do {                                                                                   \
        /* axis of the next crossing, ties as in the reference: x only if strictly first */ \
        const bool sx = mx < my && mx < mz;                                                \
        const bool sy = !(mx < my) && my < mz;                                             \
        const bool sz = !sx && !sy;                                                        \
        nx -= sx ? 1 : 0;                                                                  \
        ny -= sy ? 1 : 0;                                                                  \
        nz -= sz ? 1 : 0;                                                                  \
        more = (nx | ny | nz) >= 0;                                                        \
        addr += sx ? ax : (sy ? ay : az);                                                  \
        mx = sx ? mx + dx : mx;                                                            \
        my = sy ? my + dy : my;                                                            \
        mz = sz ? mz + dz : mz;                                                            \
    } while (0)
    bool first;
    OM_WALK_STEP(first);
    if (!first) return -1;
    // the candidate range of a cell: one address for both ends (the second load takes an offset)
    const int *cs = a.cell_start + (uint32_t)addr;
    int i0 = cs[0], i1 = cs[1];
    for (;;) {
        // the next cell is found, and its range requested, before this cell's candidates are
        // tested, so that the range load of one step is in flight during the previous one
        bool more;
        OM_WALK_STEP(more);
        int n0 = 0, n1 = 0;
        if (more) {
            const int *ns = a.cell_start + (uint32_t)addr;
            n0 = ns[0];
            n1 = ns[1];
        }
        for (int i = i0; i < i1; ++i) {
            const float4 *g = a.cell_geo + 3u * (uint32_t)i;
            const float4 A = g[0];
            const int h = (int)f2u(A.w);
            if (h == self) continue;
            const float4 E1 = g[1], E2 = g[2];
            float t, u, v;
            if (parallelogram<gfx>(o, d, mk(A.x, A.y, A.z), mk(E1.x, E1.y, E1.z),
                                   mk(E2.x, E2.y, E2.z), t, u, v))
                return h;
        }
        if (!more) return -1;
        i0 = n0;
        i1 = n1;
    }
#undef OM_WALK_STEP
}

// perturbed mirror normal from the slope-error draws (map_turbulance,
// scene_instance_process.cu:14-23; rectgrid_raytracing.cu:182-186)
__device__ __forceinline__ v3 perturbed_normal(const u4 &x, float sigma, const frame &hf)
{
    float st, ct, sp, cp;
    const float ths = gauss_<gfx>(u01(x.z)) * sigma;
    sincos_small(ths, st, ct);
    sincos2pi(u01(x.w), sp, cp);
    return normalize<gfx>(to_world(angle2xyz(st, ct, sp, cp), hf));
}

// receiver_drawing (rectgrid_raytracing.cu:135-153): (row, column) of the hit, row -1 for a
// miss; u or v equal to 1 lands on the last row or column
__device__ __forceinline__ int2 receiver_rc(const TraceArgs &a, v3 orig, v3 refl)
{
    if (a.recv_type == 1) return cylinder_rc(orig, refl, a.cyl_base, a.cyl_r2, a.cyl_h, a.res_x, a.res_y);
    if (a.recv_type == 2)
        return cone_rc(orig, refl, a.cyl_base, a.cone_r0, a.cone_k, a.cyl_h, a.cone_inv_area, a.res_x, a.res_y);
    float t, u, v;
    if (!parallelogram<gfx>(orig, refl, a.rA, a.rE1, a.rE2, t, u, v)) return make_int2(-1, -1);
    return make_int2(min(trunc_i(u * (float)a.res_y), a.res_y - 1), min(trunc_i(v * (float)a.res_x), a.res_x - 1));
}
__device__ __forceinline__ int receiver_pixel(const TraceArgs &a, v3 orig, v3 refl)
{
    const int2 rc = receiver_rc(a, orig, refl);
    return rc.x >= 0 ? rc.x * a.res_x + rc.y : -1;
}

// Philox4x32-10 of om_common.h for kernels with a uniform key (the same bits).  The key of
// each round is bumped by a scalar add in place: as plain code the ten round keys are
// loop-invariant for every ray of a kernel, so the compiler precomputes all twenty, runs out
// of SGPRs and re-reads them from VGPR lanes (two v_readlane and a hazard wait per round);
// an asm add cannot be hoisted, so each round key lives in one SGPR pair for one round.
__device__ __forceinline__ u4 philox_uniform_key(u4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        if (r < 9) {
            asm volatile("s_add_u32 %0, %0, 0x9e3779b9" : "+s"(k0) : : "scc");
            asm volatile("s_add_u32 %0, %0, 0xbb67ae85" : "+s"(k1) : : "scc");
        }
    }
    return c;
}

// Per-tile state of one heliostat: its record, frame and position, and the sun frame.
struct HelioTile {
    const HelioRec *hr;
    int h;
    frame hf, sf;
    v3 hpos;
    bool one_row, one_col;
    bool clear; // the shading walk finds nothing for any ray of this heliostat
    int self;   // the id a walk skips: the heliostat's own, or -1 (none) for the literal rule
};

__device__ __forceinline__ HelioTile helio_tile(const TraceArgs &a, int h)
{
    HelioTile t;
    t.h = h;
    t.hr = a.helio + h;
    const HelioRec *hr = t.hr;
    t.hf.n = mk(hr->n[0], hr->n[1], hr->n[2]);
    t.hf.u = mk(hr->fu[0], hr->fu[1], hr->fu[2]);
    t.hf.v = mk(hr->fv[0], hr->fv[1], hr->fv[2]);
    t.hf.identity = hr->ident;
    t.sf.n = a.sun_n;
    t.sf.u = a.sun_u;
    t.sf.v = a.sun_v;
    t.sf.identity = a.sun_ident;
    t.hpos = mk(hr->pos[0], hr->pos[1], hr->pos[2]);
    t.one_row = hr->rows == 1;
    t.one_col = hr->cols == 1;
    // under the literal rule a walk may stop at the ray's own heliostat, which the clear
    // test (other heliostats only) does not cover
    t.clear = hr->clear != 0 && !a.literal_self;
    t.self = a.literal_self ? -1 : h;
    return t;
}

// Origin of the rays of microsample m, shared by its L rays: the microsample centre in world
// coordinates and the grid cell that holds it.
struct RayStart {
    v3 orig;
    cell3 c0;
};

// t: any per-tile state with the heliostat record hr, its index h, one_row, one_col, the
// frame hf and the position hpos (HelioTile, or the differentiable kernels' tile)
template <bool POISSON, class Tile>
__device__ __forceinline__ RayStart ray_start(const TraceArgs &a, const Tile &t, int m)
{
    const HelioRec *hr = t.hr;
    // microsample centre (steps_for_raytracing.cu:4-38, :53-65)
    v3 loc;
    if (POISSON) {
        float2 pt = a.poisson[(size_t)t.h * a.poisson_cap + m];
        // valid points lie inside (0, width) x (0, height): one facet -> block 0
        int br = t.one_row ? 0 : trunc_i(pt.y / hr->subrow_len);
        int bc = t.one_col ? 0 : trunc_i(pt.x / hr->subcol_len);
        loc = mk(pt.x + (float)bc * hr->gx - hr->sx2, hr->sy2, pt.y + (float)br * hr->gy - hr->sz2);
    } else {
        int row = m / hr->cols_all, col = m - row * hr->cols_all;
        int br = row / hr->sub_rows, bc = col / hr->sub_cols;
        loc = mk((float)col * a.pitch + (float)bc * hr->gx + a.pitch2 - hr->sx2, hr->sy2,
                 (float)row * a.pitch + (float)br * hr->gy + a.pitch2 - hr->sz2);
    }
    RayStart st;
    st.orig = to_world(loc, t.hf) + t.hpos;
    st.c0 = start_cell(a, st.orig);
    return st;
}

// The starts of 64 consecutive microsamples mt, mt + 1, ... of a tile's heliostat, one per
// lane of a wave (the origin and its cell).  A wave's pass of 64 rays spans at most 64
// microsamples, so the table is refilled only when a pass reaches past it: once per L
// passes, once per chunk of 2048 rays for L >= 32.  Each pass reads its rays' starts from
// the lanes that hold them (ds_bpermute) instead of finding them again: with fewer than 64
// rays per microsample a lane's microsample changes at every pass.  The whole wave calls
// both members.
struct StartTable {
    int mt = -(1 << 30);
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    int gx = 0, gy = 0, gz = 0;

    template <bool POISSON, class Tile>
    __device__ __forceinline__ void update(const TraceArgs &a, const Tile &t, bool active, int m, int lane)
    {
        if (__ballot(active && m - mt >= 64) == 0ull) return;
        mt = __builtin_amdgcn_readfirstlane(m); // lane 0's: the pass's first microsample
        const int mj = mt + lane;
        if (mj < t.hr->mcount) {
            const RayStart st = ray_start<POISSON>(a, t, mj);
            ox = st.orig.x;
            oy = st.orig.y;
            oz = st.orig.z;
            gx = st.c0.gx;
            gy = st.c0.gy;
            gz = st.c0.gz;
        }
    }
    // the start of microsample m, mt <= m < mt + 64, as ray_start gives it
    __device__ __forceinline__ RayStart get(const TraceArgs &a, int m) const
    {
        const int src = min(max(m - mt, 0), 63) << 2;
        RayStart st;
        st.orig = mk(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(ox))),
                     __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(oy))),
                     __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(oz))));
        st.c0.gx = __builtin_amdgcn_ds_bpermute(src, gx);
        st.c0.gy = __builtin_amdgcn_ds_bpermute(src, gy);
        st.c0.gz = __builtin_amdgcn_ds_bpermute(src, gz);
        // as start_cell forms them
        st.c0.rx = st.orig.x - a.gpx;
        st.c0.ry = st.orig.y - a.gpy;
        st.c0.rz = st.orig.z - a.gpz;
        st.c0.addr = (st.c0.gx * a.gny + st.c0.gy) * a.gnz + st.c0.gz;
        return st;
    }
};

// XCD (accelerator complex die) that runs the calling wave.  The packed map words are kept
// one copy per XCD to spread same-pixel atomics; the atomics themselves are agent-scope, so
// the choice of copy affects contention only, never the result.
__device__ __forceinline__ uint32_t xcc_id()
{
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & 15u;
}

} // namespace om
