// om_trace64.hip -- the trace in FP64 geometry (om_params.precision = OM_PREC_FP64).
//
// The same rays as the FP32 path: the Philox draws of each ray (the uniforms converted to
// double exactly) and the same microsamples, sampled and traced in double: sunshape angle
// (Buie table and power-law tail, or pillbox), slope-error normal (inverse normal CDF),
// local2world frames, both grid walks (NotCollision with its quirks), the parallelogram test
// and the receiver.  This is the FP64 model the FP32 path's error is measured against: per
// ray it reaches the same outcome except where an FP32 rounding moves a ray across a pixel,
// cell or mirror edge.  Energies |sun.n| go into the same exact fixed-point accumulation as
// the FP32 path (2^-31 units, one truncation per ray).
//
// Reference: map_tracing and NotCollision (rectgrid_raytracing.cu:3-196), the samplers of
// scene_instance_process.cu:6-226, receiver_drawing (:135-153).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "om_window.h"

namespace om {
namespace {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mkd(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 ld(const double *p) { return d3{p[0], p[1], p[2]}; }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 neg(d3 a) { return d3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ d3 scl(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b)
{
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// 1 / d and 1 / sqrt(x) within about an ulp: the hardware estimates refined by two Newton
// steps each.  The FP64 model rounds both correctly; a last-bit difference in a direction or
// a crossing parameter moves a ray by some 1e-16 of the scene, far below the geometry's
// tolerances.
__device__ __forceinline__ double rcp64(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double rsq64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}
__device__ __forceinline__ d3 normalized(d3 v) { return scl(v, rsq64(dot(v, v))); }
__device__ __forceinline__ d3 reflected(d3 i, d3 n) { return sub(i, scl(n, 2.0 * dot(n, i))); }

// The double transcendentals out of line: inlined, their general argument reduction holds
// about a hundred registers across the whole ray; as calls they hold them only inside.
__device__ __noinline__ double2 sincos64(double x)
{
    double s, c;
    sincos(x, &s, &c);
    return make_double2(s, c);
}

// (sin x, cos x) for the angles of a ray, |x| <= 8: the sunshape and slope-error polar angles
// (milliradians) and the azimuths 2 pi u.  Cody-Waite reduction by pi/2 split in three parts
// (P1 and P2 with 21 trailing zero bits, so q P1 and q P2 are exact for the quadrants q that
// occur) to |r| <= pi/4, then the Taylor series of sin to r^17 and of cos to r^18, whose
// next terms are below 2^-60 of the result on that interval.  Larger arguments take the
// library's general reduction.
__device__ __forceinline__ double2 sincos_ray(double x)
{
    if (!(fabs(x) <= 8.0)) return sincos64(x);
    constexpr double P1 = 1.5707963267341256, P2 = 6.077100506303966e-11, P3 = 2.0222662487959506e-21;
    const double q = rint(x * 0.6366197723675814);
    const double r = ((x - q * P1) - q * P2) - q * P3;
    const double r2 = r * r;
    const double s = r + r * r2 * (-0.16666666666666666 + r2 * (0.008333333333333333 + r2 * (-0.0001984126984126984 +
                     r2 * (2.7557319223985893e-06 + r2 * (-2.505210838544172e-08 + r2 * (1.6059043836821613e-10 +
                     r2 * (-7.647163731819816e-13 + r2 * 2.8114572543455206e-15)))))));
    const double c = 1.0 + r2 * (-0.5 + r2 * (0.041666666666666664 + r2 * (-0.001388888888888889 +
                     r2 * (2.48015873015873e-05 + r2 * (-2.755731922398589e-07 + r2 * (2.08767569878681e-09 +
                     r2 * (-1.1470745597729725e-11 + r2 * (4.779477332387385e-14 + r2 * -1.5619206968586225e-16))))))));
    switch ((int)q & 3) {
    case 0: return make_double2(s, c);
    case 1: return make_double2(c, -s);
    case 2: return make_double2(-s, -c);
    default: return make_double2(-c, s);
    }
}
// (sin x, cos x) of a polar angle of a ray, the sunshape's or the slope error's, which lie
// within 0.125 rad (the Buie profile ends at 43.6 mrad): the Taylor series of sin to x^11
// and of cos to x^12, whose next terms are below 2^-66 of the result there; no reduction.
// Larger angles take sincos_ray.
__device__ __forceinline__ double2 sincos_polar(double x)
{
    if (!(fabs(x) <= 0.125)) return sincos_ray(x);
    const double x2 = x * x;
    const double s = x + x * x2 * (-0.16666666666666666 + x2 * (0.008333333333333333 + x2 * (-0.0001984126984126984 +
                     x2 * (2.7557319223985893e-06 + x2 * -2.505210838544172e-08))));
    const double c = 1.0 + x2 * (-0.5 + x2 * (0.041666666666666664 + x2 * (-0.001388888888888889 +
                     x2 * (2.48015873015873e-05 + x2 * (-2.755731922398589e-07 + x2 * 2.08767569878681e-09)))));
    return make_double2(s, c);
}
__device__ __noinline__ double pow64(double x, double y) { return pow(x, y); }
__device__ __noinline__ double erfinv64(double x) { return erfinv(x); }

// erfinv(x) of a ray's slope-error draw, x = 2u - 1 with u = (2k + 1) 2^-24: the form of
// M. Giles, "Approximating the erfinv function" (GPU Computing Gems, 2010), x p(w - 3.125)
// with w = -log((1 - x)(1 + x)), on w < 6.25 (99.9 % of the draws); (1 - x)(1 + x) is
// exact in double for these x.  p is this build's own degree-22 least-squares fit to
// erfinv(x) / x on Chebyshev nodes of [0, 6.25]; over all 2^23 draws of the region it is
// within 1.8e-15 relative of erfinv.  The draws beyond take the library's erfinv.  Out of
// line, as the library's transcendentals: inlined, its log and polynomial spill the ray's
// registers.
__device__ __constant__ double ERFINV_P[23] = {
    1.6536545626831025,
    0.24015818242558906,
    -0.006033670871433301,
    -0.0007407025341673867,
    0.00018673420804426685,
    -1.38825233585321e-05,
    -1.365469213819447e-06,
    4.234787740289138e-07,
    -2.9070360058863695e-08,
    -4.112631858218718e-09,
    1.0512079981016051e-09,
    -5.415473598366289e-11,
    -1.2973959923550662e-11,
    2.6336180516038526e-12,
    -8.17295193661494e-14,
    -4.0557333294761386e-14,
    6.662194463545632e-15,
    2.169483066444032e-17,
    -1.351253836906662e-16,
    1.1135525520949027e-17,
    1.3621871100405752e-18,
    -1.6827858204529238e-19,
    -5.05530203750712e-21};
__device__ __noinline__ double erfinv_ray(double x)
{
    const double t = (1.0 - x) * (1.0 + x);
    if (!(t > 0.0019304541362277093)) return erfinv64(x); // w >= 6.25
    const double s = -log(t) - 3.125;
    double p = ERFINV_P[22];
#pragma unroll
    for (int i = 21; i >= 0; --i) p = fma(p, s, ERFINV_P[i]);
    return p * x;
}
__device__ __noinline__ double atan2_64(double y, double x) { return atan2(y, x); }

struct dframe {
    d3 n, u, v;
    bool ident;
};
// local2world (global_function.cuh:14-46) with the frame of n held
__device__ __forceinline__ d3 world(d3 d, const dframe &f)
{
    if (f.ident) return d;
    return mkd(d.x * f.u.x + d.y * f.n.x + d.z * f.v.x, d.x * f.u.y + d.y * f.n.y + d.z * f.v.y,
               d.x * f.u.z + d.y * f.n.z + d.z * f.v.z);
}

// rayParallelogramIntersect (global_function.cuh:78-105) with E1 = B - A, E2 = C - A held
__device__ __forceinline__ bool parallelogram64(d3 orig, d3 dir, d3 A, d3 E1, d3 E2, double &t, double &u, double &v)
{
    const d3 pvec = cross(dir, E2);
    const double det = dot(E1, pvec);
    if (fabs(det) < 1e-6) return false;
    const double invDet = rcp64(det);
    const d3 T = sub(orig, A);
    u = dot(T, pvec) * invDet;
    if (u < 0 || u > 1) return false;
    const d3 qvec = cross(T, E1);
    v = dot(dir, qvec) * invDet;
    if (v < 0 || v > 1) return false;
    t = dot(E2, qvec) * invDet;
    if (t < 1e-6) return false;
    return true;
}

// The grid cell of a ray origin, shared by both walks of the rays of one microsample
struct Cell64 {
    int g[3];
};
__device__ __forceinline__ Cell64 cell64(const Trace64Args &b, d3 o)
{
    const d3 rel = sub(o, ld(b.gpos));
    return Cell64{{(int)(rel.x / b.gint[0]), (int)(rel.y / b.gint[1]), (int)(rel.z / b.gint[2])}};
}

// NotCollision + Intersect (rectgrid_raytracing.cu:21-117) in double: start cell by
// truncation and never tested, absDivide's INT_MAX for |dir| <= 1e-6, the reference's
// tie-break, first candidate in CSR order; the walk ends once an axis has passed the occupied
// index range (every later cell is empty).  The FP64 model divides by each direction
// component once (absDivide's quotients as products with 1 / dir).  self: the id skipped,
// or -1.
// c: the cell of o (cell64)
__device__ int walk64(const TraceArgs &a, const Trace64Args &b, d3 o, const Cell64 &c, d3 d, int self)
{
    const d3 rel = sub(o, ld(b.gpos));
    int g[3] = {c.g[0], c.g[1], c.g[2]};
    const double dd[3] = {d.x, d.y, d.z}, cur[3] = {rel.x, rel.y, rel.z};
    const int gn[3] = {a.gnx, a.gny, a.gnz};
    int st[3];
    double tm[3], td[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        st[k] = dd[k] >= 0 ? 1 : -1;
        const double t = dd[k] >= 0 ? (g[k] + 1) * b.gint[k] - cur[k] : cur[k] - g[k] * b.gint[k];
        const double inv = rcp64(dd[k]); // used where |dd[k]| > 1e-6 only
        tm[k] = (fabs(dd[k]) <= 1e-6) ? 2147483647.0 : fabs(t * inv);
        td[k] = (fabs(dd[k]) <= 1e-6) ? 2147483647.0 : fabs(b.gint[k] * inv);
    }
    for (;;) {
        int k;
        if (tm[0] < tm[1]) k = (tm[0] < tm[2]) ? 0 : 2;
        else k = (tm[1] < tm[2]) ? 1 : 2;
        g[k] += st[k];
        if (g[k] >= gn[k] || g[k] < 0) return -1;
        tm[k] += td[k];
        if ((st[0] > 0 ? g[0] > a.occ_hi[0] : g[0] < a.occ_lo[0]) ||
            (st[1] > 0 ? g[1] > a.occ_hi[1] : g[1] < a.occ_lo[1]) ||
            (st[2] > 0 ? g[2] > a.occ_hi[2] : g[2] < a.occ_lo[2]))
            return -1;
        const int addr = (g[0] * a.gny + g[1]) * a.gnz + g[2];
        const int i1 = a.cell_start[addr + 1];
        for (int i = a.cell_start[addr]; i < i1; ++i) {
            const Geo64 &G = b.cell_geo[i];
            const int h = G.id;
            if (h == self) continue;
            double t, u, v;
            if (parallelogram64(o, d, ld(G.A), ld(G.E1), ld(G.E2), t, u, v)) return h;
        }
    }
}

// receiver_drawing in double; cylinder and cone as the new receivers of the FP32 path:
// (row, column) of the hit, row -1 for a miss
__device__ int2 receiver64(const TraceArgs &a, const Trace64Args &b, d3 o, d3 d)
{
    const int rx = a.res_x, ry = a.res_y;
    if (a.recv_type == 1) {
        const double ox = o.x - b.cyl_base[0], oz = o.z - b.cyl_base[2];
        const double qa = d.x * d.x + d.z * d.z, qb = 2.0 * (ox * d.x + oz * d.z);
        const double qc = ox * ox + oz * oz - b.cyl_r * b.cyl_r;
        const double disc = qb * qb - 4.0 * qa * qc;
        if (qa < 1e-6 || disc <= 0.0) return make_int2(-1, -1);
        const double sq = sqrt(disc);
        const double q0 = -0.5 * (qb + (qb >= 0 ? sq : -sq));
        const double t1 = q0 / qa, t2 = qc / q0;
        const double t = t1 < t2 ? t1 : t2;
        if (t < 1e-6) return make_int2(-1, -1);
        const double hx = ox + t * d.x, hz = oz + t * d.z, hy = o.y + t * d.y - b.cyl_base[1];
        if (hy < 0.0 || hy > b.cyl_h) return make_int2(-1, -1);
        double th = atan2_64(hz, hx);
        if (th < 0.0) th += 6.283185307179586;
        const int col = min((int)(th / 6.283185307179586 * rx), rx - 1);
        const int row = min((int)(hy / b.cyl_h * ry), ry - 1);
        return make_int2(row, col);
    }
    if (a.recv_type == 2) {
        const double h = b.cyl_h, r0 = b.cone_r0, r1 = b.cone_r1, k = (r1 - r0) / h;
        const double ox = o.x - b.cyl_base[0], oy = o.y - b.cyl_base[1], oz = o.z - b.cyl_base[2];
        const double w0 = r0 + k * oy, wd = k * d.y;
        const double qa = d.x * d.x + d.z * d.z - wd * wd;
        const double qb = 2.0 * (ox * d.x + oz * d.z - w0 * wd);
        const double qc = ox * ox + oz * oz - w0 * w0;
        const double disc = qb * qb - 4.0 * qa * qc;
        if (d.x * d.x + d.z * d.z < 1e-6 || disc < 0.0 || qa == 0.0) return make_int2(-1, -1);
        const double sq = sqrt(disc);
        const double q0 = -0.5 * (qb + (qb >= 0 ? sq : -sq));
        const double ta = q0 / qa, tb = qc / q0;
        const double t0 = ta < tb ? ta : tb, t1 = ta < tb ? tb : ta;
        double t = -1.0, hy = 0.0;
        for (int i = 0; i < 2; ++i) {
            const double ti = i == 0 ? t0 : t1;
            if (ti < 1e-6) continue;
            const double yi = oy + ti * d.y;
            if (yi < 0.0 || yi > h) continue;
            t = ti;
            hy = yi;
            break;
        }
        if (t < 0.0) return make_int2(-1, -1);
        const double hx = ox + t * d.x, hz = oz + t * d.z;
        if (!(d.x * hx + d.z * hz - d.y * k * (r0 + k * hy) < 0.0)) return make_int2(-1, -1);
        double th = atan2_64(hz, hx);
        if (th < 0.0) th += 6.283185307179586;
        const int col = min((int)(th / 6.283185307179586 * rx), rx - 1);
        const int row = min((int)(hy * (2.0 * r0 + k * hy) / (h * (r0 + r1)) * ry), ry - 1);
        return make_int2(row, col);
    }
    double t, u, v;
    if (!parallelogram64(o, d, ld(b.rA), ld(b.rE1), ld(b.rE2), t, u, v)) return make_int2(-1, -1);
    return make_int2(min((int)(u * ry), ry - 1), min((int)(v * rx), rx - 1));
}

// microsample centre in the heliostat frame (steps_for_raytracing.cu:4-38, :53-65)
template <bool POISSON>
__device__ __forceinline__ d3 micro_local64(const TraceArgs &a, const Trace64Args &b, const HelioRec *hr,
                                            const Helio64 &h64, int h, int m)
{
    if (POISSON) {
        const float2 pt = a.poisson[(size_t)h * a.poisson_cap + m];
        const int block_row = (int)((double)pt.y / h64.subrow);
        const int block_col = (int)((double)pt.x / h64.subcol);
        return mkd(pt.x + block_col * h64.gx - h64.sz[0] / 2, h64.sz[1] / 2, pt.y + block_row * h64.gy - h64.sz[2] / 2);
    }
    const double p = b.pitch;
    const int row = m / hr->cols_all, col = m - row * hr->cols_all;
    const int block_row = row / hr->sub_rows, block_col = col / hr->sub_cols;
    return mkd(col * p + block_col * h64.gx + p / 2 - h64.sz[0] / 2, h64.sz[1] / 2,
               row * p + block_row * h64.gy + p / 2 - h64.sz[2] / 2);
}

struct Ray64 {
    int shade, block, px; // px: receiver pixel, -1 for none
    int2 rc;              // its (row, column)
    uint32_t q;
};


// one ray (map_tracing, rectgrid_raytracing.cu:155-196) in double
template <int SUN>
__device__ Ray64 trace_ray64(const TraceArgs &a, const Trace64Args &b, const dframe &hf, const dframe &sf, d3 orig,
                             const Cell64 &c0,
                             int h, int m, int l, int self, bool clear)
{
    Ray64 r{-1, -2, -1, make_int2(-1, -1), 0u};
    // the key words read as wave-uniform values (a's fields can arrive through a vector
    // address where the caller's argument struct is not the kernel's)
    const u4 x = philox_uniform_key(u4{(uint32_t)l, (uint32_t)m, (uint32_t)h, a.w3},
                                    __builtin_amdgcn_readfirstlane(a.key0), __builtin_amdgcn_readfirstlane(a.key1));
    const double u0 = u01(x.x), u1 = u01(x.y), u2 = u01(x.z), u3 = u01(x.w);
    // the slope-error normal does not depend on the sunshape draw: with the Buie sunshape it
    // is formed while the guide entry's load is in flight
    const double v = u0 * b.A;
    const bool head = SUN != OM_SUN_PILLBOX && v < b.int465;
    int lo = 0;
    if (head) lo = b.guide[min(max((int)(v * b.gscale), 0), b.nbins - 1)];
    __builtin_amdgcn_sched_barrier(0); // the load above is issued before this work
    const double ths = erfinv_ray(2.0 * u2 - 1.0) * 1.4142135623730951 * b.sigma;
    const double phs = 6.283185307179586 * u3;
    const double2 sts = sincos_polar(ths), sps = sincos_ray(phs);
    const d3 nrm = normalized(world(mkd(sts.x * sps.y, sts.y, sts.x * sps.x), hf));
    __builtin_amdgcn_sched_barrier(0);
    double th;
    if (SUN == OM_SUN_PILLBOX) {
        th = b.pillbox * sqrt(u0);
    } else {
        if (head) {
            // the largest lo with cdf_y[lo] <= v (max_less_index, scene_instance_process.cu:
            // 158-173), from the guide entry of v
            while (lo > 0 && b.cdf_y[lo] > v) --lo;
            while (b.cdf_y[lo + 1] <= v) ++lo;
            th = (v - b.cdf_y[lo]) * b.slope[lo] + b.cdf_x[lo];
        } else {
            // the tail: the table holds this pow for each draw that reaches here
            th = b.tail[(x.x >> 9) - b.tail_k0];
        }
        th /= 1000.0;
    }
    const double ph = 6.283185307179586 * u1;
    const double2 st = sincos_polar(th), sp = sincos_ray(ph);
    const d3 to_sun = neg(world(mkd(st.x * sp.y, st.y, st.x * sp.x), sf));
    r.shade = clear ? -1 : walk64(a, b, orig, c0, to_sun, self);
    if (r.shade >= 0) return r;
    const d3 refl = normalized(reflected(neg(to_sun), nrm));
    r.block = walk64(a, b, orig, c0, refl, self);
    if (r.block >= 0) return r;
    r.rc = receiver64(a, b, orig, refl);
    if (r.rc.x >= 0) {
        r.px = r.rc.x * a.res_x + r.rc.y;
        r.q = (uint32_t)(fabs(dot(ld(b.sun_n), nrm)) * 2147483648.0);
    }
    return r;
}

// One workgroup per tile of the scene's tile list, as trace_kernel: up to rays_per_thread
// rays per thread, hits kept in LDS slots and added after the ray loop (packed per-XCD
// copies, or two device atomics per hit), or per-ray records.
struct Trace64Launch {
    TraceArgs a;
    Trace64Args b;
};

// The arguments are read from the kernel-argument segment where they are used, as in the FP32
// trace kernels: held in registers across the ray they exceed the SGPR file.
template <int SUN, bool POISSON, bool RECORD>
__global__ __launch_bounds__(WG) void trace64_kernel(Trace64Launch)
{
    const Trace64Launch &args = *(const Trace64Launch *)__builtin_amdgcn_kernarg_segment_ptr();
    const TraceArgs &a = args.a;
    const Trace64Args &b = args.b;
    __shared__ int s_hpx[RPT_MAX * WG];
    __shared__ uint32_t s_hq[RPT_MAX * WG];
    const int tid = (int)threadIdx.x;
    const int2 tl = a.tiles[a.tile_base + xcd_chunk(blockIdx.x, gridDim.x)];
    const int h = tl.x;
    const HelioRec *hr = a.helio + h;
    const Helio64 &h64 = b.helio[h];
    dframe hf, sf;
    hf.n = ld(h64.n);
    hf.u = ld(h64.u);
    hf.v = ld(h64.v);
    hf.ident = h64.ident != 0;
    sf.n = ld(b.sun_n);
    sf.u = ld(b.sun_u);
    sf.v = ld(b.sun_v);
    sf.ident = b.sun_ident != 0;
    const d3 hpos = ld(h64.pos);
    const int self = a.literal_self ? -1 : h;
    const bool clear = hr->clear != 0 && !a.literal_self;
    const int L = a.L;
    const int nr = min(hr->mcount * L, tl.y + a.tile_rays);
    int nk = 0;
    int m_o = -1; // microsample of orig: its origin is found again only when m changes
    d3 orig{0.0, 0.0, 0.0};
    Cell64 c0{};
    for (int k = 0; k < a.rays_per_thread; ++k) {
        const int r = tl.y + tid + k * WG;
        if (r >= nr) break;
        nk = k + 1;
        asm volatile("" ::: "memory"); // arguments re-read per ray pass
        const int m = r / L, l = r - m * L;
        if (m != m_o) {
            orig = add(world(micro_local64<POISSON>(a, b, hr, h64, h, m), hf), hpos);
            c0 = cell64(b, orig);
            m_o = m;
        }
        const Ray64 o = trace_ray64<SUN>(a, b, hf, sf, orig, c0, h, m, l, self, clear);
        if (RECORD) {
            const long long rid = (((long long)hr->ray_off_hi << 32) | (unsigned)hr->ray_off_lo) + r - a.rec_base;
            a.rec_shade[rid] = o.shade;
            a.rec_block[rid] = o.block;
            a.rec_pixel[rid] = o.px;
            continue;
        }
        s_hpx[k * WG + tid] = o.px;
        s_hq[k * WG + tid] = o.q;
    }
    if (RECORD) return;
    if (!a.acc_packed) {
        for (int k = 0; k < nk; ++k) {
            const int px = s_hpx[k * WG + tid];
            if (px < 0) continue;
            atomicAdd(a.counts + px, 1u);
            atomicAdd(a.energy + px, (unsigned long long)s_hq[k * WG + tid]);
        }
        return;
    }
    const uint32_t copy = (xcc_id() & 7u) | ((((uint32_t)blockIdx.x >> 3) & ((1u << a.copy_shift) - 1u)) << 3);
    unsigned long long *part = a.part + (size_t)copy * a.copy_stride;
    const int E = a.pack_shift;
    const unsigned long long lowmask = (1ull << E) - 1ull;
    for (int k = 0; k < nk; ++k) {
        const int px = s_hpx[k * WG + tid];
        if (px < 0) continue;
        const unsigned long long q = s_hq[k * WG + tid];
        const unsigned long long v = (1ull << E) + q;
        const unsigned long long old = __hip_atomic_fetch_add(part + (size_t)px * a.px_stride, v, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
        if ((old & lowmask) + q > lowmask) atomicAdd(a.carry + px, 1u);
        if (old + v < old) atomicAdd(a.wrap + px, 1u);
    }
}

// nrays (a multiple of 64) rays of heliostat tl.x from tl.y traced by one wave of a window
// workgroup, one ray per lane per pass: hits inside the LDS window are LDS atomics, the
// others go to the wave's buffer (window_body, om_window.h).
// v of the lane src / 4 (ds_bpermute of both halves)
__device__ __forceinline__ double bperm_d(int src, double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int SUN, bool POISSON>
__device__ __forceinline__ void trace_chunk_window64(const TraceArgs &a, const Trace64Args &b, int2 tl, const int lane,
                                                     const int nrays, WaveOut &wo, unsigned long long *s_win)
{
    const int h = tl.x;
    const HelioRec *hr = a.helio + h;
    const Helio64 &h64 = b.helio[h];
    dframe hf, sf;
    hf.n = ld(h64.n);
    hf.u = ld(h64.u);
    hf.v = ld(h64.v);
    hf.ident = h64.ident != 0;
    sf.n = ld(b.sun_n);
    sf.u = ld(b.sun_u);
    sf.v = ld(b.sun_v);
    sf.ident = b.sun_ident != 0;
    const d3 hpos = ld(h64.pos);
    const int self = a.literal_self ? -1 : h;
    const bool clear = hr->clear != 0 && !a.literal_self;
    const int L = a.L;
    const int nr = min(hr->mcount * L, tl.y + nrays);
    int r = tl.y + lane;
    int m = r / L;
    int l = r - m * L;
    const int dm = 64 / L, dl = 64 - (64 / L) * L;
    // The origins and cells of 64 consecutive microsamples mt, mt + 1, ..., one per lane,
    // refilled when a pass reaches past them (once per chunk for L >= 32); each pass reads
    // its rays' from the lanes that hold them, as the FP32 and mixed window traces do.
    int mt = -(1 << 30);
    d3 to{0.0, 0.0, 0.0};
    int tg0 = 0, tg1 = 0, tg2 = 0;
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
        asm volatile("" ::: "memory"); // arguments re-read per ray pass
        if (__ballot(r < nr && m - mt >= 64) != 0ull) { // the whole wave takes part
            mt = __builtin_amdgcn_readfirstlane(m);     // lane 0's: the pass's first microsample
            const int mj = mt + lane;
            if (mj < hr->mcount) {
                to = add(world(micro_local64<POISSON>(a, b, hr, h64, h, mj), hf), hpos);
                const Cell64 c = cell64(b, to);
                tg0 = c.g[0];
                tg1 = c.g[1];
                tg2 = c.g[2];
            }
        }
        const int src = min(max(m - mt, 0), 63) << 2;
        const d3 orig{bperm_d(src, to.x), bperm_d(src, to.y), bperm_d(src, to.z)};
        const Cell64 c0{{__builtin_amdgcn_ds_bpermute(src, tg0), __builtin_amdgcn_ds_bpermute(src, tg1),
                         __builtin_amdgcn_ds_bpermute(src, tg2)}};
        if (r < nr) {
            const Ray64 o = trace_ray64<SUN>(a, b, hf, sf, orig, c0, h, m, l, self, clear);
            if (o.px >= 0) {
                px = o.px;
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

template <int SUN, bool POISSON>
__global__ __launch_bounds__(WIN_WGS) void trace_window64_kernel(Trace64Launch)
{
    const Trace64Launch &args = *(const Trace64Launch *)__builtin_amdgcn_kernarg_segment_ptr();
    // the FP64 trace of a mixed launch whose redo list filled up: nothing to do otherwise
    if (args.b.gate && *args.b.gate <= args.b.gate_cap) return;
    __shared__ unsigned long long s_win[WIN_PX_MAX];
    __shared__ unsigned long long s_out[WIN_WAVES][WIN_OUT];
    window_body(args.a, s_win, s_out,
                [&](int2 tl, int lane, int nrays, WaveOut &wo, unsigned long long *w) {
                    trace_chunk_window64<SUN, POISSON>(args.a, args.b, tl, lane, nrays, wo, w);
                });
}

// ---------------------------------------------------------------------------------------
// OM_PREC_MIXED: the FP64 model's outcome of every ray at close to FP32 cost
// ---------------------------------------------------------------------------------------
// A ray is sampled in FP32 (the same Philox draws).  Its direction is the model's nominal
// reflection of the heliostat (computed in double at upload, HelioMix) plus a deviation
// formed in FP32 from the small sunshape and slope-error angles, and its receiver hit is
// found in FP32 relative to the receiver's reference point, from the model's nominal miss
// vector and the deviation.  Both walks run in FP32 over float images of the model's
// vertices, from the origin's offset inside its cell (formed in double).  Each FP32
// decision is taken only when it holds for every value within its error bound (MixArgs); a
// ray with a decision inside its bound is put on a list and traced by the FP64 trace
// afterwards.  So every ray's shading heliostat, blocking heliostat and
// pixel are the model's; its energy |sun.n| is the model's to the FP32 draws' error.

constexpr int MIX_UNSURE = -3;
constexpr float U24 = 5.9604644775390625e-08f; // 2^-24
constexpr int REDO_BLOCKS = 1024;              // workgroups of the redo list's FP64 trace

__device__ __forceinline__ float norm1(v3 a) { return fabs_(a.x) + fabs_(a.y) + fabs_(a.z); }

struct MixLaunch {
    TraceArgs a; // its FP32 scene fields hold the float images of the FP64 scene (mixed_args)
    Trace64Args b;
    MixArgs x;
};

__device__ __forceinline__ v3 f3(d3 v) { return mk((float)v.x, (float)v.y, (float)v.z); }

// rayParallelogramIntersect (global_function.cuh:78-105) in FP32, certified against the same
// test in double on the model's ray: 1 hit, 0 miss, -1 when a tested quantity lies within its
// error bound of its threshold.  With T = o - A, c0 = e_pos + |T| (e_dir + 11 2^-24) (the
// error of T and of the ray's position along T) and c1 = e_dir + 8 2^-24 (the direction's):
//   |det - det_model| <= |E1||E2| c1
//   |u - u_model| <= (|E2| c0 + |u| |E1||E2| c1) / |det| + 4 2^-24 |u|
//   |v - v_model| <= (|E1| c0 + |v| |E1||E2| c1) / |det| + 4 2^-24 |v|
//   |t - t_model| <= (|E1||E2| c0 + |t| |E1||E2| c1) / |det| + 4 2^-24 |t|
// (first order; the reciprocal is the hardware estimate, within 2^-23).  |T| is bounded by
// its 1-norm and |E1|, |E2| (rounded up) ride in the w components of the vertex records.
__device__ __forceinline__ int parallelogram_cert(const MixArgs &x, v3 o, v3 d, const float4 &A, const float4 &E1,
                                                  const float4 &E2)
{
    const v3 e1 = mk(E1.x, E1.y, E1.z), e2 = mk(E2.x, E2.y, E2.z);
    const v3 p = cross(d, e2);
    const float det = dot(e1, p);
    const float inv = __builtin_amdgcn_rcpf(det);
    const v3 T = o - mk(A.x, A.y, A.z);
    const v3 q = cross(T, e1);
    const float u = dot(T, p) * inv, v = dot(d, q) * inv, t = dot(e2, q) * inv;
    const float tn = fabs_(T.x) + fabs_(T.y) + fabs_(T.z);
    const float c0 = fma_(tn, x.e_dir + 11.0f * U24, x.e_pos);
    const float mdet = E1.w * E2.w * (x.e_dir + 8.0f * U24);
    const float ai = fabs_(inv), adet = fabs_(det);
    const float au = fabs_(u), av = fabs_(v), at = fabs_(t);
    // (|u| mdet + |E2| c0) / |det| + 4 2^-24 |u|, and likewise, with the factors shared by the
    // three formed once (four roundings each, as in the direct form)
    const float k4 = fma_(mdet, ai, 4.0f * U24), ci = c0 * ai;
    const float mu = fma_(au, k4, E2.w * ci);
    const float mv = fma_(av, k4, E1.w * ci);
    const float mt = fma_(at, k4, E1.w * E2.w * ci);
    const bool miss = adet < 1e-6f - mdet || u < -mu || u > 1.0f + mu || v < -mv || v > 1.0f + mv || t < 1e-6f - mt;
    const bool hit = adet >= 1e-6f + mdet && u >= mu && u <= 1.0f - mu && v >= mv && v <= 1.0f - mv && t >= 1e-6f + mt;
    return hit ? 1 : (miss ? 0 : -1);
}

// The origin of a microsample's rays: the float origin o = pos + w with w = o - pos formed in
// float from the microsample's local position and the model's frame (within ew = 2^-21 |w|_1
// of the model's w), the grid cell of the origin and the origin's offset inside it.  pos and
// w are floats, so pos + w - gpos is formed exactly in double, once per microsample: the
// cell is the model's truncation of (o - gpos) / edge unless the quotient lies within ew /
// edge of an integer, and the offset inside the cell, rounded to float, carries the walk's
// crossings with an error of ew + 2^-23 edge instead of the rounding of field coordinates.
// A microsample's facet is certain likewise (the quotient of the Poisson point by the facet
// length); ok is false when either is not, and every ray of the microsample goes to FP64.
struct MixStart {
    v3 of, w;
    cell3 c; // rx, ry, rz: the offset of the origin inside its cell
    float ew;
    bool ok;
};
template <bool POISSON>
__device__ __forceinline__ MixStart mix_start(const TraceArgs &a, const MixArgs &x, const HelioRec *hr,
                                              const HelioMix &hm, int h, int m)
{
    MixStart s;
    s.ok = true;
    v3 loc;
    if (POISSON) {
        const float2 pt = a.poisson[(size_t)h * a.poisson_cap + m];
        int br = 0, bc = 0;
        // the facet of the point: quotients by the hardware reciprocal (within 1.5 ulp of
        // the correctly rounded ones, inside the margin's 4 2^-24 |q|)
        if (hr->rows > 1) {
            const float q = pt.y * __builtin_amdgcn_rcpf(hr->subrow_len), g = fabs_(q - rintf(q));
            br = trunc_i(q);
            s.ok = s.ok && g > fma_(fabs_(q), 4.0f * U24, 1e-6f);
        }
        if (hr->cols > 1) {
            const float q = pt.x * __builtin_amdgcn_rcpf(hr->subcol_len), g = fabs_(q - rintf(q));
            bc = trunc_i(q);
            s.ok = s.ok && g > fma_(fabs_(q), 4.0f * U24, 1e-6f);
        }
        loc = mk(pt.x + (float)bc * hr->gx - hr->sx2, hr->sy2, pt.y + (float)br * hr->gy - hr->sz2);
    } else {
        const int row = m / hr->cols_all, col = m - row * hr->cols_all;
        const int br = row / hr->sub_rows, bc = col / hr->sub_cols;
        loc = mk((float)col * a.pitch + (float)bc * hr->gx + a.pitch2 - hr->sx2, hr->sy2,
                 (float)row * a.pitch + (float)br * hr->gy + a.pitch2 - hr->sz2);
    }
    const v3 u = mk(hm.u[0], hm.u[1], hm.u[2]), n = mk(hm.n[0], hm.n[1], hm.n[2]), v = mk(hm.v[0], hm.v[1], hm.v[2]);
    s.w = loc.x * u + loc.y * n + loc.z * v;
    const v3 pos = mk(hr->pos[0], hr->pos[1], hr->pos[2]);
    s.of = pos + s.w;
    s.ew = fma_(norm1(s.w), 4.7683716e-07f, 1e-12f);
    const double rx = ((double)pos.x + (double)s.w.x) - (double)a.gpx;
    const double ry = ((double)pos.y + (double)s.w.y) - (double)a.gpy;
    const double rz = ((double)pos.z + (double)s.w.z) - (double)a.gpz;
    // the quotients by the rounded reciprocals: within 2^-51 |q| of r / gi, which the margins
    // below cover with their 2^-30 (|q| < 2^16 cells)
    const double qx = rx * x.rgi[0], qy = ry * x.rgi[1], qz = rz * x.rgi[2];
    s.c.gx = (int)qx;
    s.c.gy = (int)qy;
    s.c.gz = (int)qz;
    s.c.addr = (s.c.gx * a.gny + s.c.gy) * a.gnz + s.c.gz;
    s.c.rx = (float)(rx - (double)s.c.gx * (double)a.gix);
    s.c.ry = (float)(ry - (double)s.c.gy * (double)a.giy);
    s.c.rz = (float)(rz - (double)s.c.gz * (double)a.giz);
    const double gx = qx - (double)s.c.gx, gy = qy - (double)s.c.gy, gz = qz - (double)s.c.gz;
    const float mx = fma_(s.ew * a.inv_gix, 1.01f, 9.3132257e-10f), my = fma_(s.ew * a.inv_giy, 1.01f, 9.3132257e-10f),
                mz = fma_(s.ew * a.inv_giz, 1.01f, 9.3132257e-10f);
    s.ok = s.ok && (float)fmin(gx, 1.0 - gx) > mx && (float)fmin(gy, 1.0 - gy) > my && (float)fmin(gz, 1.0 - gz) > mz;
    return s;
}

// NotCollision (rectgrid_raytracing.cu:38-117) in FP32 with certified decisions, in the form
// of dda (om_device.h): the first candidate hit, -1 for none, MIX_UNSURE.  c: the start cell
// with the origin's offset inside it (mix_start), so that the first crossing numerators,
// edge - r or r, are within e_num = ew + 2^-22 edge of the model's.  The crossing parameters
// m_k = numerator / |d_k| are formed with the hardware reciprocal (within 1 ulp); each is
// within e_k = e_num / |d_k| + m_k (e_dir / |d_k| + (n + 8) 2^-24) of the model's after n
// steps (the direction's error, the reciprocal and product, the n additions).  A step's axis
// is certain when the chosen m_c lies below the runner-up m_2 by more than
// 2 e_num r + (m_c + m_2) (e_dir r + (n + 8) 2^-24), r the largest 1 / |d_k|, which bounds
// e_c + e_2.  A direction component within 2 e_dir + 1e-6 of zero could differ from the
// model's in sign or in absDivide's 1e-6 test: such a ray is unsure.
__device__ int walk_cert(const TraceArgs &a, const MixArgs &x, v3 o, const cell3 &c, float ew, v3 d, int self)
{
    const bool fx = d.x >= 0.0f, fy = d.y >= 0.0f, fz = d.z >= 0.0f;
    int nx = fx ? a.occ_hi[0] - c.gx : c.gx - a.occ_lo[0];
    int ny = fy ? a.occ_hi[1] - c.gy : c.gy - a.occ_lo[1];
    int nz = fz ? a.occ_hi[2] - c.gz : c.gz - a.occ_lo[2];
    const float near0 = 2.0f * x.e_dir + 1.0001e-6f;
    if (fabs_(d.x) <= near0 || fabs_(d.y) <= near0 || fabs_(d.z) <= near0) return MIX_UNSURE;
    if ((nx | ny | nz) < 0) return -1;
    if ((nx | ny | nz) == 0) return -1;
    const float rx = fabs_(__builtin_amdgcn_rcpf(d.x)), ry = fabs_(__builtin_amdgcn_rcpf(d.y)),
                rz = fabs_(__builtin_amdgcn_rcpf(d.z));
    float mx = (fx ? a.gix - c.rx : c.rx) * rx;
    float my = (fy ? a.giy - c.ry : c.ry) * ry;
    float mz = (fz ? a.giz - c.rz : c.rz) * rz;
    const float dx = a.gix * rx, dy = a.giy * ry, dz = a.giz * rz;
    const float rm = fmaxf(fmaxf(rx, ry), rz) * 1.000001f;
    const float ea = 2.0f * fma_(x.gi_max, 2.3841858e-07f, ew) * rm; // 2 (ew + 2^-22 edge) r
    float eb = fma_(x.e_dir, rm, 8.0f * U24);                          // e_dir r + (n + 8) 2^-24
    const int yz = a.gny * a.gnz;
    const int sax = fx ? yz : -yz, say = fy ? a.gnz : -a.gnz, saz = fz ? 1 : -1;
    int addr = c.addr;
#define OM_CERT_STEP(more, sure)                                                                   \
    do {                                                                                         \
        const bool sx = mx < my && mx < mz;                                                      \
        const bool sy = !(mx < my) && my < mz;                                                   \
        const bool sz = !sx && !sy;                                                              \
        /* the chosen axis holds the smallest value (ties too), m2 the second smallest */       \
        const float mc = fminf(fminf(mx, my), mz);                                               \
        const float m2 = __builtin_amdgcn_fmed3f(mx, my, mz);                                    \
        sure = m2 - mc > fma_(mc + m2, eb, ea);                                                  \
        eb += U24;                                                                               \
        nx -= sx ? 1 : 0;                                                                        \
        ny -= sy ? 1 : 0;                                                                        \
        nz -= sz ? 1 : 0;                                                                        \
        more = (nx | ny | nz) >= 0;                                                              \
        addr += sx ? sax : (sy ? say : saz);                                                     \
        mx = sx ? mx + dx : mx;                                                                  \
        my = sy ? my + dy : my;                                                                  \
        mz = sz ? mz + dz : mz;                                                                  \
    } while (0)
    bool first, sure;
    OM_CERT_STEP(first, sure);
    if (!sure) return MIX_UNSURE;
    if (!first) return -1;
    const int *cs = a.cell_start + (uint32_t)addr;
    int i0 = cs[0], i1 = cs[1];
    for (;;) {
        // the next step and its range request before this cell's candidates (as dda); the
        // step's certainty matters only when the walk goes on past this cell
        bool more;
        OM_CERT_STEP(more, sure);
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
            const int r = parallelogram_cert(x, o, d, A, g[1], g[2]);
            if (r != 0) return r > 0 ? h : MIX_UNSURE;
        }
        if (!sure) return MIX_UNSURE;
        if (!more) return -1;
        i0 = n0;
        i1 = n1;
    }
#undef OM_CERT_STEP
}

// distance of x from the nearest integer
__device__ __forceinline__ float int_gapf(float x) { return fabs_(x - rintf(x)); }

// sin x and cos x - 1 of a polar angle |x| < 0.25 (the series of sincos_small with the 1 of
// the cosine left out, so that the deviation keeps its relative precision)
__device__ __forceinline__ void sincos_m1(float x, float &s, float &cm1)
{
    const float x2 = x * x;
    float ps = fma_(x2, -1.98412698e-04f, 8.33333333e-03f);
    ps = fma_(x2, ps, -1.66666667e-01f);
    s = fma_(x * x2, ps, x);
    float pc = fma_(x2, 2.48015873e-05f, -1.38888889e-03f);
    pc = fma_(x2, pc, 4.16666667e-02f);
    pc = fma_(x2, pc, -0.5f);
    cm1 = x2 * pc;
}

// receiver_drawing (rectgrid_raytracing.cu:135-153) for the rectangle and the cylinder, from
// the receiver's reference point F: the ray meets the receiver at F + c + tau d.  (row,
// column), row -1 for a miss, row MIX_UNSURE when the hit or the pixel lies within its bound
// of an edge.  bc bounds the error of c against the model's ray (its point nearest F); a
// point tau along the ray is within bc + |tau| e_dir of the model's, and the hit on the
// surface within that over the cosine of incidence.
__device__ __forceinline__ int2 receiver_local(const TraceArgs &a, const MixArgs &x, v3 c, v3 d, float D, float bc)
{
    const int2 miss = make_int2(-1, -1), unsure = make_int2(MIX_UNSURE, 0);
    const float rx = (float)a.res_x, ry = (float)a.res_y;
    const float cn1 = norm1(c);
    if (a.recv_type == 1) {
        const float r = x.cyl_r;
        const float A2 = d.x * d.x + d.z * d.z, B2 = c.x * d.x + c.z * d.z, C2 = c.x * c.x + c.z * c.z - r * r;
        if (fabs_(A2 - 1e-6f) <= 4.0f * x.e_dir) return unsure;
        if (A2 < 1e-6f) return miss;
        // the line's distance from the axis, rho^2 = r^2 - disc / A2, against r
        const float disc = B2 * B2 - A2 * C2;
        const float ia = __builtin_amdgcn_rcpf(A2);
        const float tc = fabs_(B2) * ia; // the parameter of the closest approach
        const float dl = bc + tc * x.e_dir + fma_(cn1 + tc, 4.0f * U24, 0.0f) + r * x.e_dir * __builtin_amdgcn_rsqf(A2);
        const float rho2 = r * r - disc * ia;
        const float lo = r > dl ? (r - dl) * (r - dl) : 0.0f, hi = (r + dl) * (r + dl);
        if (fabs_(rho2 - r * r) <= 8.0f * U24 * (r * r + fabs_(disc * ia)) || (rho2 >= lo && rho2 <= hi))
            return unsure;
        if (disc <= 0.0f) return miss;
        const float sq = sqrt_rn(disc);
        const float q = -(B2 + (B2 >= 0.0f ? sq : -sq));
        const float ta = q * ia, tb = div_rn(C2, q);
        const float tau = fminf(ta, tb);
        const v3 W = c + tau * d;
        const float t = D + tau;
        const float at = fabs_(tau);
        const float ci = fabs_(d.x * W.x + d.z * W.z) / r; // cosine of incidence (|d| = 1)
        const float mw = div_rn(bc + at * x.e_dir, fmaxf(ci, 1e-30f)) + (cn1 + at) * 4.0f * U24;
        if (fabs_(t - 1e-6f) <= mw + fabs_(t) * 4.0f * U24) return unsure;
        if (t < 1e-6f) return miss;
        const float Y = (W.y + 0.5f * x.cyl_h) / x.cyl_h * ry, mY = (mw + fabs_(W.y) * 4.0f * U24) / x.cyl_h * ry + ry * 4.0f * U24;
        if (Y < -mY || Y - ry > mY) return miss; // Y - ry is exact near ry (ry + mY would round)
        float th = atan2_(W.z, W.x);
        if (th < 0.0f) th += 6.28318548f;
        const float X = th * 0.159154937f * rx;
        const float mX = (mw / r + 8.0f * U24 * 6.2831853f) * 0.159154937f * rx + rx * 4.0f * U24;
        if (int_gapf(Y) <= mY || int_gapf(X) <= mX || X >= rx - mX) return unsure;
        return make_int2(min(trunc_i(Y), a.res_y - 1), min(trunc_i(X), a.res_x - 1));
    }
    const float dn = dot(d, x.nn), ad = fabs_(dn);
    const float det = dot(d, x.nr);
    if (fabs_(fabs_(det) - 1e-6f) <= x.nrn * (x.e_dir + 4.0f * U24)) return unsure;
    if (fabs_(det) < 1e-6f) return miss;
    const float tau = -dot(c, x.nn) * __builtin_amdgcn_rcpf(dn);
    const v3 P = c + tau * d;
    const float at = fabs_(tau);
    const float mp = (bc + at * x.e_dir) * __builtin_amdgcn_rcpf(ad) * 1.0001f + (cn1 + at) * 8.0f * U24;
    const float t = D + tau;
    if (fabs_(t - 1e-6f) <= mp + fabs_(t) * 4.0f * U24) return unsure;
    const float pn = norm1(P);
    const float U = (x.uF + dot(P, x.g1)) * ry, V = (x.vF + dot(P, x.g2)) * rx;
    const float mU = ((mp + pn * 8.0f * U24) * x.g1n + 4.0f * U24) * ry + ry * 4.0f * U24;
    const float mV = ((mp + pn * 8.0f * U24) * x.g2n + 4.0f * U24) * rx + rx * 4.0f * U24;
    // the edges: U - ry is exact near ry, where ry + mU would round to float and leave a sliver
    // (ry + mU, fl(ry + mU)] that neither test catches; nearer an integer than the bound
    // (an edge included) the pixel is unsure; else the hit is certain
    if (U < -mU || U - ry > mU || V < -mV || V - rx > mV || t < 1e-6f) return miss;
    if (int_gapf(U) <= mU || int_gapf(V) <= mV) return unsure;
    return make_int2(min(trunc_i(U), a.res_y - 1), min(trunc_i(V), a.res_x - 1));
}

struct MixRay {
    int shade, block; // as Ray64; shade MIX_UNSURE: the ray goes to the FP64 trace
    int2 rc;
    uint32_t q;
};

// One ray (map_tracing, rectgrid_raytracing.cu:155-196) of the mixed trace, in FP32.  The
// draws are those of trace_ray64: the sunshape angle from the float image of the model's Buie
// table (or, for the tail draws k >= tail_k0, the power law in float), the slope-error angle
// from the FP32 erfinv.  Directions are the nominal ones plus deviations:
// the light i = i0 + di, the normal n = n0 + dn, and the reflection
//   r = i - 2 (n.i) n = a + dr,   dr = di - 2 c1 n0 - 2 (c0 + c1) dn,   c1 = n0.di + dn.i0 + dn.di
// (c0 = n0.i0, a = i0 - 2 c0 n0 from the model); |r| = 1 to the FP32 rounding of the draws,
// as the model's normalisations give.
template <int SUN>
__device__ __forceinline__ MixRay trace_ray_mixed(const TraceArgs &a, const Trace64Args &b, const MixArgs &x,
                                                  const HelioMix &hm, const MixStart &ms, int h, int m, int l,
                                                  int self, bool clear)
{
    MixRay r{-1, -2, make_int2(-1, -1), 0u};
    if (!ms.ok) {
        r.shade = MIX_UNSURE;
        return r;
    }
    const u4 xr = philox_uniform_key(u4{(uint32_t)l, (uint32_t)m, (uint32_t)h, a.w3}, a.key0, a.key1);
    const v3 n0 = mk(hm.n[0], hm.n[1], hm.n[2]);
    float th, ths, st, cm, sp, cp, sts, cms, sps, cps;
    v3 dn;
    if (SUN == OM_SUN_BUIE) {
        // The two loads of the Buie lookup (buie_bin_split) are issued ahead of the work that
        // does not depend on them (the slope-error normal and the sunshape azimuth), which
        // then runs while they are in flight.  The tail draws (k >= tail_k0, the model's
        // choice) take the power law in float, from the value's excess over int465 formed
        // without cancellation (MixArgs::tail_du), within 2^-20 relative of the model's angle.
        const uint32_t k = xr.x >> 9;
        const float v = fminf(u01(xr.x) * a.buie_A, x.buie_vmax);
        const int j = min(max(trunc_i(v * a.buie_gscale), 0), a.nbins - 1);
        const int2 g = a.buie_split[j];
        __builtin_amdgcn_sched_barrier(0);
        ths = gauss_<gfx>(u01(xr.z)) * a.slope_sigma;
        sincos_m1(ths, sts, cms);
        sincos2pi(u01(xr.w), sps, cps);
        dn = cms * n0 + sts * (cps * mk(hm.u[0], hm.u[1], hm.u[2]) + sps * mk(hm.v[0], hm.v[1], hm.v[2]));
        asm volatile("" ::"v"(dn.x), "v"(dn.y), "v"(dn.z)); // computed here
        __builtin_amdgcn_sched_barrier(0);
        int i = min(g.x + (v >= u2f((uint32_t)g.y) ? 1 : 0), a.nbins - 1);
        float4 bb = a.buie_bins[i];
        __builtin_amdgcn_sched_barrier(0);
        sincos2pi(u01(xr.y), sp, cp);
        asm volatile("" ::"v"(sp), "v"(cp));
        __builtin_amdgcn_sched_barrier(0);
        if (k < b.tail_k0) {
            while (bb.x > v) bb = a.buie_bins[--i];
            while (bb.y <= v) bb = a.buie_bins[++i];
            th = ((v - bb.x) * bb.z + bb.w) * 1e-3f;
        } else {
            const float dv = fma_((float)(k - b.tail_k0), x.tail_du, x.tail_dv0);
            th = pow_<gfx>(dv * a.buie_B + a.buie_C, a.buie_inv_gp1) * 1e-3f;
        }
    } else {
        th = a.pillbox * __builtin_amdgcn_sqrtf(u01(xr.x));
        ths = gauss_<gfx>(u01(xr.z)) * a.slope_sigma;
        sincos_m1(ths, sts, cms);
        sincos2pi(u01(xr.w), sps, cps);
        dn = cms * n0 + sts * (cps * mk(hm.u[0], hm.u[1], hm.u[2]) + sps * mk(hm.v[0], hm.v[1], hm.v[2]));
        sincos2pi(u01(xr.y), sp, cp);
    }
    if (!(fabs_(th) < 0.25f && fabs_(ths) < 0.25f)) { // the series' range
        r.shade = MIX_UNSURE;
        return r;
    }
    sincos_m1(th, st, cm);
    const v3 di = cm * x.s_n + st * (cp * x.s_u + sp * x.s_v);
    if (!clear) {
        r.shade = walk_cert(a, x, ms.of, ms.c, ms.ew, -(x.s_n + di), self);
        if (r.shade != -1) return r; // a blocker, or MIX_UNSURE
    }
    const float c1 = dot(n0, di) + dot(dn, x.s_n) + dot(dn, di);
    const v3 dr = di - (2.0f * c1) * n0 - (2.0f * (hm.c0 + c1)) * dn;
    const v3 d = mk(hm.a[0], hm.a[1], hm.a[2]) + dr;
    r.block = walk_cert(a, x, ms.of, ms.c, ms.ew, d, self);
    if (r.block != -1) {
        if (r.block == MIX_UNSURE) r.shade = MIX_UNSURE;
        return r;
    }
    // c = qf + w + D dr with its error bound: the model's qf and D rounded, w and D dr formed
    // in float, dr within e_dr of the model's
    const v3 qf = mk(hm.qf[0], hm.qf[1], hm.qf[2]);
    const v3 c = qf + ms.w + hm.D * dr;
    // dr is within 2^-19 (theta + 4 theta_s) + 2^-26 of the model's deviation: the draws'
    // relative errors (a few 2^-24; the FP32 erfinv's and the float Buie tail's, at most 2^-21
    // and 2^-20 over all 2^23 draws) scale with the angles and
    // the slope error's count twice in the reflection; the float Buie lookup's absolute error
    // (at most about 3e-9 rad, where the table's slope is steepest) does not.  The scene bound
    // e_dr is this at the largest angles.
    const float edr = fminf(fma_(fabs_(th) + 4.0f * fabs_(ths), 1.9073486e-06f, 1.4901161e-08f), x.e_dr);
    const float bc = (norm1(qf) + norm1(ms.w) + hm.D * norm1(dr)) * 8.0f * U24 + hm.D * edr;
    r.rc = receiver_local(a, x, c, d, hm.D, bc);
    if (r.rc.x == MIX_UNSURE) {
        r.shade = MIX_UNSURE;
    } else if (r.rc.x >= 0) {
        // |sun.n| = |sun.n0 + sun.dn|
        r.q = (uint32_t)(fabs_(hm.e0 + dot(x.sun, dn)) * 2147483648.0f);
    }
    return r;
}

// Ray r of heliostat h traced by trace_ray64 from nothing but its indices: the rays of the
// redo list, and the unsure rays of the records kernel.
template <int SUN, bool POISSON>
__device__ __forceinline__ Ray64 ray64_at(const TraceArgs &a, const Trace64Args &b, int h, int r)
{
    const HelioRec *hr = a.helio + h;
    const Helio64 &h64 = b.helio[h];
    dframe hf, sf;
    hf.n = ld(h64.n);
    hf.u = ld(h64.u);
    hf.v = ld(h64.v);
    hf.ident = h64.ident != 0;
    sf.n = ld(b.sun_n);
    sf.u = ld(b.sun_u);
    sf.v = ld(b.sun_v);
    sf.ident = b.sun_ident != 0;
    const int m = r / a.L, l = r - m * a.L;
    const d3 orig = add(world(micro_local64<POISSON>(a, b, hr, h64, h, m), hf), ld(h64.pos));
    return trace_ray64<SUN>(a, b, hf, sf, orig, cell64(b, orig), h, m, l, h, hr->clear != 0);
}

// A hit of pixel (row, col) into the workgroup's LDS window, or marked for the wave's buffer
// (out)
__device__ __forceinline__ void window_hit(const TraceArgs &a, const WaveOut &wo, unsigned long long *s_win, int2 rc,
                                           int px, uint32_t q, bool &out)
{
    const int wi = wo.window_word(a, rc.x, rc.y);
    if (wi >= 0) packed_add_lds(a, s_win + wi, px, q);
    else out = true;
}

// nrays (a multiple of 64) rays of heliostat tl.x from tl.y traced by one wave of a window
// workgroup, one ray per lane per pass (as trace_chunk_window, om_kernels.hip)
template <int SUN, bool POISSON>
__device__ __forceinline__ void trace_chunk_mixed(const MixLaunch &L, int2 tl, const int lane, const int nrays,
                                                  WaveOut &wo, unsigned long long *s_win)
{
    const TraceArgs &a = L.a;
    wo.mx = &L.x;
    const int h = tl.x;
    const HelioRec *hr = a.helio + h;
    const HelioMix hm = L.x.hmix[h];
    const bool clear = hr->clear != 0;
    const int Lr = a.L;
    const int nr = min(hr->mcount * Lr, tl.y + nrays);
    int r = tl.y + lane;
    int m = r / Lr;
    int l = r - m * Lr;
    const int dm = 64 / Lr, dl = 64 - (64 / Lr) * Lr;
    // The starts of 64 consecutive microsamples mt, mt + 1, ..., one per lane: w, the cell
    // and the offset inside it (rx NaN: not certain).  A pass spans at most 64 microsamples,
    // so the table is refilled only when a pass reaches past it: once per L passes, once per
    // chunk for L >= 32.  Each pass reads its rays' starts from the lanes that hold them.
    int mt = -(1 << 30);
    float tw0 = 0.0f, tw1 = 0.0f, tw2 = 0.0f, tr0 = 0.0f, tr1 = 0.0f, tr2 = 0.0f;
    int tg0 = 0, tg1 = 0, tg2 = 0;
    const v3 pos = mk(hr->pos[0], hr->pos[1], hr->pos[2]);
    for (int k = 0; k < nrays; k += 64, r += 64) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= Lr) {
                l -= Lr;
                ++m;
            }
        }
        bool out = false, flag = false;
        int px = 0;
        uint32_t q = 0u;
        asm volatile("" ::: "memory"); // arguments re-read per ray pass
        // the whole wave takes part in the refill and the reads
        if (__ballot(r < nr && m - mt >= 64) != 0ull) {
            mt = __builtin_amdgcn_readfirstlane(m); // lane 0's: the pass's first microsample
            const int mj = mt + lane;
            if (mj < hr->mcount) {
                const MixStart t = mix_start<POISSON>(a, L.x, hr, hm, h, mj);
                tw0 = t.w.x;
                tw1 = t.w.y;
                tw2 = t.w.z;
                tg0 = t.c.gx;
                tg1 = t.c.gy;
                tg2 = t.c.gz;
                tr0 = t.ok ? t.c.rx : __builtin_nanf("");
                tr1 = t.c.ry;
                tr2 = t.c.rz;
            }
        }
        const int src = min(max(m - mt, 0), 63) << 2;
        MixStart ms;
        ms.w = mk(__int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(tw0))),
                  __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(tw1))),
                  __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(tw2))));
        ms.c.gx = __builtin_amdgcn_ds_bpermute(src, tg0);
        ms.c.gy = __builtin_amdgcn_ds_bpermute(src, tg1);
        ms.c.gz = __builtin_amdgcn_ds_bpermute(src, tg2);
        ms.c.rx = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(tr0)));
        ms.c.ry = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(tr1)));
        ms.c.rz = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(tr2)));
        if (r < nr) {
            // as mix_start forms them
            ms.of = pos + ms.w;
            ms.ew = fma_(norm1(ms.w), 4.7683716e-07f, 1e-12f);
            ms.c.addr = (ms.c.gx * a.gny + ms.c.gy) * a.gnz + ms.c.gz;
            ms.ok = ms.c.rx == ms.c.rx;
            const MixRay o = trace_ray_mixed<SUN>(a, L.b, L.x, hm, ms, h, m, l, h, clear);
            if (o.shade == MIX_UNSURE) {
                flag = true;
            } else if (o.rc.x >= 0) {
                px = o.rc.x * a.res_x + o.rc.y;
                q = o.q;
                window_hit(a, wo, s_win, o.rc, px, q, out);
            }
        }
        // the whole wave takes part: no lane has left the loop.  A ray that finds the redo
        // list full is dropped: the whole launch is then traced again in FP64
        // (launch_trace_mixed_window).
        wo.push_redo(a, flag, h, r, lane);
        wo.push(a, out, px, q, lane);
    }
}

template <int SUN, bool POISSON>
__global__ __launch_bounds__(WIN_WGS) void trace_window_mixed_kernel(MixLaunch)
{
    const MixLaunch &L = *(const MixLaunch *)__builtin_amdgcn_kernarg_segment_ptr();
    __shared__ unsigned long long s_win[WIN_PX_MAX];
    __shared__ unsigned long long s_out[WIN_WAVES][WIN_OUT];
    window_body(L.a, s_win, s_out, [&](int2 tl, int lane, int nrays, WaveOut &wo, unsigned long long *w) {
        trace_chunk_mixed<SUN, POISSON>(L, tl, lane, nrays, wo, w);
    });
}

// The rays of the redo list traced by the FP64 trace into the packed per-XCD copies; thread
// 0 adds the launch's count to the scene's running total (and counts a full list in
// redo_n[2]).  A full list leaves everything to the FP64 trace of the whole launch.
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WG) void trace_redo_kernel(MixLaunch)
{
    const MixLaunch &L = *(const MixLaunch *)__builtin_amdgcn_kernarg_segment_ptr();
    const TraceArgs &a = L.a;
    const uint32_t cnt = *L.x.redo_n;
    const bool full = cnt > (uint32_t)L.x.redo_cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(L.x.redo_n + 1, cnt);
        if (full) atomicAdd(L.x.redo_n + 2, 1u);
    }
    // The list lies in the trace's workspace, whose words a later trace may take as packed
    // copies, which it adds to: every entry written is cleared once read.
    if (full) {
        for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < (uint32_t)L.x.redo_cap; i += gridDim.x * WG)
            L.x.redo[i] = make_uint2(0u, 0u);
        return;
    }
    const uint32_t n = cnt;
    unsigned long long *part = copy_of(a);
    const unsigned long long one = 1ull << a.pack_shift;
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG) {
        const uint2 e = L.x.redo[i];
        L.x.redo[i] = make_uint2(0u, 0u);
        const Ray64 o = ray64_at<SUN, POISSON>(a, L.b, (int)e.x, (int)e.y);
        if (o.px >= 0) {
            const unsigned long long w = one + o.q;
            copy_check(a, o.px, w, copy_add(a, part, o.px, w));
        }
    }
}

// Per-ray records of the mixed trace (one workgroup per tile, as trace64_kernel): a ray with
// an unsure decision is traced by the FP64 trace in place and counted in redo_n[1].
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WG) void trace_mixed_records_kernel(MixLaunch)
{
    const MixLaunch &L = *(const MixLaunch *)__builtin_amdgcn_kernarg_segment_ptr();
    const TraceArgs &a = L.a;
    const int tid = (int)threadIdx.x;
    const int2 tl = a.tiles[a.tile_base + xcd_chunk(blockIdx.x, gridDim.x)];
    const int h = tl.x;
    const HelioRec *hr = a.helio + h;
    const HelioMix hm = L.x.hmix[h];
    const bool clear = hr->clear != 0;
    const int Lr = a.L;
    const int nr = min(hr->mcount * Lr, tl.y + a.tile_rays);
    for (int k = 0; k < a.rays_per_thread; ++k) {
        const int r = tl.y + tid + k * WG;
        if (r >= nr) break;
        asm volatile("" ::: "memory");
        const int m = r / Lr, l = r - m * Lr;
        const MixStart ms = mix_start<POISSON>(a, L.x, hr, hm, h, m);
        const MixRay o = trace_ray_mixed<SUN>(a, L.b, L.x, hm, ms, h, m, l, h, clear);
        int shade = o.shade, block = o.block, px = o.rc.x >= 0 ? o.rc.x * a.res_x + o.rc.y : -1;
        if (o.shade == MIX_UNSURE) {
            atomicAdd(L.x.redo_n + 1, 1u);
            const Ray64 f = ray64_at<SUN, POISSON>(a, L.b, h, r);
            shade = f.shade;
            block = f.block;
            px = f.px;
        }
        const long long rid = (((long long)hr->ray_off_hi << 32) | (unsigned)hr->ray_off_lo) + r - a.rec_base;
        a.rec_shade[rid] = shade;
        a.rec_block[rid] = block;
        a.rec_pixel[rid] = px;
    }
}


// ---------------------------------------------------------------------------------------
// OM_BATCH_WINDOW with OM_PREC_MIXED: the persistent batch launch of mixed scenes
// ---------------------------------------------------------------------------------------
// The table holds one MixLaunch per entry (one grid of one scene), with the batch's packed
// copies of the scene's map, the entry's own redo list and its counters (redo_n[0]: rays
// listed by this run, redo_n[2]: set when the list filled up).  The persistent kernel is the
// FP32 batch window kernel's body with the mixed trace's chunks; a wave's buffer sends its
// redo entries to the current entry's list before the window moves on to another entry.
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WIN_WGS) void trace_batch_window_mixed_kernel(const MixLaunch *__restrict__ table,
                                                                           const int4 *__restrict__ chunks,
                                                                           int n_chunks, unsigned int *queue)
{
    __shared__ unsigned long long s_win[WIN_PX_MAX];
    __shared__ unsigned long long s_out[WIN_WAVES][WIN_OUT];
    __shared__ int s_held[WIN_WAVES];
    batch_window_body(
        table, chunks, n_chunks, queue, s_win, s_out, s_held, [](const MixLaunch &L) -> const TraceArgs & { return L.a; },
        [](const MixLaunch &L, int4 ck, int lane, unsigned long long *w, WaveOut &wo) {
            trace_chunk_mixed<SUN, POISSON>(L, make_int2(ck.y, ck.z), lane, ck.w, wo, w);
        });
}

// Every entry's redo list traced in FP64 into its scene's copies.  Each workgroup visits
// every entry in turn and takes a strided share of its list, starting at a workgroup that
// moves with the entry, so that the rays of all lists spread over the whole grid: a grid
// per entry left the lists of the large scenes to a few workgroups each.  A list that
// filled up marks the entry (redo_n[2]) and its scene (scene_full), whose maps are then
// cleared and traced again wholly in FP64 (trace64_batch_full_kernel).
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WG) void trace_redo_batch_kernel(const MixLaunch *__restrict__ table,
                                                              const int *__restrict__ entry_scene,
                                                              uint32_t *scene_full, int n_entries)
{
    const uint32_t nthreads = gridDim.x * WG;
    for (int en = 0; en < n_entries; ++en) {
        const MixLaunch &L = table[en];
        const TraceArgs &a = L.a;
        const uint32_t cnt = *L.x.redo_n;
        if (cnt > (uint32_t)L.x.redo_cap) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                L.x.redo_n[2] = 1u;
                atomicOr(scene_full + entry_scene[en], 1u);
            }
            continue;
        }
        const uint32_t first = ((blockIdx.x + (uint32_t)en * 97u) % gridDim.x) * WG + threadIdx.x;
        if (first >= cnt) continue;
        unsigned long long *part = copy_of(a);
        const unsigned long long one = 1ull << a.pack_shift;
        for (uint32_t i = first; i < cnt; i += nthreads) {
            const uint2 e = L.x.redo[i];
            const Ray64 o = ray64_at<SUN, POISSON>(a, L.b, (int)e.x, (int)e.y);
            if (o.px >= 0) {
                const unsigned long long w = one + o.q;
                copy_check(a, o.px, w, copy_add(a, part, o.px, w));
            }
        }
    }
}

// The scenes whose redo list filled up (blockIdx.y: the scene): their packed copies and side
// maps cleared for the FP64 trace of the whole scene.
__global__ void clear_full_scenes_kernel(const ReduceSlot *__restrict__ slots, const uint32_t *__restrict__ scene_full)
{
    if (scene_full[blockIdx.y] == 0u) return;
    const ReduceSlot &s = slots[blockIdx.y];
    const long long nw = (long long)s.ncopy * s.copy_stride, step = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nw; i += step) s.part[i] = 0ull;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < s.npx; i += step) {
        s.carry[i] = 0u;
        s.wrap[i] = 0u;
    }
}

// The FP64 trace of every chunk of the scenes whose redo list filled up, one ray per thread
// from its indices (ray64_at), into the scene's copies; entry_chunks[e] is the range of
// entry e's chunks.  Scenes that fitted are skipped without reading their chunks.
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WG) void trace64_batch_full_kernel(const MixLaunch *__restrict__ table,
                                                                const int4 *__restrict__ chunks,
                                                                const int2 *__restrict__ entry_chunks,
                                                                const int *__restrict__ entry_scene,
                                                                const uint32_t *__restrict__ scene_full, int n_entries)
{
    for (int e = 0; e < n_entries; ++e) {
        if (scene_full[entry_scene[e]] == 0u) continue;
        const MixLaunch &L = table[e];
        const TraceArgs &a = L.a;
        unsigned long long *part = copy_of(a);
        const unsigned long long one = 1ull << a.pack_shift;
        const int2 cr = entry_chunks[e];
        for (int c = cr.x + (int)blockIdx.x; c < cr.y; c += (int)gridDim.x) {
            const int4 ck = chunks[c];
            for (int i = (int)threadIdx.x; i < ck.w; i += WG) {
                const Ray64 o = ray64_at<SUN, POISSON>(a, L.b, ck.y, ck.z + i);
                if (o.px >= 0) {
                    const unsigned long long w = one + o.q;
                    copy_check(a, o.px, w, copy_add(a, part, o.px, w));
                }
            }
        }
    }
}

// Each run starts with empty lists: redo_n[0] and redo_n[2] of every entry (4 words per
// entry at c) and the scene flags after them.
__global__ void batch_mix_reset_kernel(uint32_t *c, int n_entries, int n_scenes)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_entries + n_scenes; i += gridDim.x * blockDim.x) {
        if (i < n_entries) {
            c[4 * i] = 0u;
            c[4 * i + 2] = 0u;
        } else {
            c[4 * n_entries + (i - n_entries)] = 0u;
        }
    }
}

} // namespace

// When the mixed launch's redo list filled up (*gate > cap): clears n words at p and the
// chunk counters of the launch's queue (whose item count and layout the FP64 window trace
// shares), so that the FP64 trace takes every chunk again.  Otherwise the queue keeps the
// mixed launch's counts, which om_ctx_progress reports.
__global__ void clear_if_full_kernel(const uint32_t *gate, uint32_t cap, uint32_t *p, long long n, uint32_t *queue)
{
    if (*gate <= cap) return;
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)QUEUE_COUNTERS) queue[threadIdx.x] = 0u;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// om_math_eval 16: erfinv_ray(2u - 1) of the draw words in[i], as the FP64 trace evaluates it
__global__ void erfinv_ray_kernel(const uint32_t *in, double *out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = erfinv_ray(2.0 * (double)u01(in[i]) - 1.0);
}

om_status launch_erfinv_ray(const uint32_t *in, double *out, int n, hipStream_t st)
{
    if (n <= 0) return OM_OK;
    hipLaunchKernelGGL(erfinv_ray_kernel, dim3((n + 255) / 256), dim3(256), 0, st, in, out, n);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

__global__ void buie_tail64_kernel(double *out, uint32_t k0, uint32_t n, double A, double int465, double B,
                                   double C, double inv_gp1)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the draw and the expression of trace_ray64, term for term
    const double u0 = u01((k0 + i) << 9);
    const double v = u0 * A;
    out[i] = pow64((v - int465) * B + C, inv_gp1);
}

om_status build_buie_tail64(double *out, uint32_t k0, uint32_t n, double A, double int465, double B, double C,
                            double inv_gp1)
{
    if (n == 0) return OM_OK;
    hipLaunchKernelGGL(buie_tail64_kernel, dim3((n + 255) / 256), dim3(256), 0, nullptr, out, k0, n, A, int465, B, C,
                       inv_gp1);
    OM_HIP(hipGetLastError());
    OM_HIP(hipDeviceSynchronize());
    return OM_OK;
}

om_status launch_trace64_window(om_scene *s, const TraceArgs &a, const Trace64Args &b, int nwg, hipStream_t st)
{
    if (a.ntiles <= 0) return OM_OK;
    if (!b.helio || !b.cell_geo) {
        set_error("scene has no FP64 geometry");
        return OM_ERR_INVALID;
    }
    const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
    const dim3 grid(nwg), block(WIN_WGS);
    const Trace64Launch ab{a, b};
    if (s->prm.sunshape == OM_SUN_BUIE) {
        if (poisson) hipLaunchKernelGGL((trace_window64_kernel<OM_SUN_BUIE, true>), grid, block, 0, st, ab);
        else hipLaunchKernelGGL((trace_window64_kernel<OM_SUN_BUIE, false>), grid, block, 0, st, ab);
    } else {
        if (poisson) hipLaunchKernelGGL((trace_window64_kernel<OM_SUN_PILLBOX, true>), grid, block, 0, st, ab);
        else hipLaunchKernelGGL((trace_window64_kernel<OM_SUN_PILLBOX, false>), grid, block, 0, st, ab);
    }
    OM_HIP(hipGetLastError());
    return OM_OK;
}

bool mixed_supported(const om_scene *s)
{
    // the cone receiver and the literal self-intersection rule trace in FP64
    return s->recv_type != 2 && !s->prm.literal_self && s->d_mix_n != nullptr && s->d_hmix != nullptr;
}

void mixed_args(const om_scene *s, int g, TraceArgs &a, MixArgs &x)
{
    a.cell_geo = g < (int)s->d_cell_geo_m.size() ? s->d_cell_geo_m[g] : nullptr;
    a.buie_bins = s->d_buie_bins_m;
    a.buie_split = s->d_buie_split_m;
    a.buie_A = s->buie_A_m;
    a.buie_gscale = s->buie_gscale_m;
    // the tail's power law from the model's constants
    a.buie_int465 = (float)s->a64.int465;
    a.buie_B = (float)s->a64.B;
    a.buie_C = (float)s->a64.C;
    a.buie_inv_gp1 = (float)s->a64.inv_gp1;
    x = s->mix;
    x.tail_du = (float)std::ldexp(s->a64.A, -23);
    x.tail_dv0 = s->a64.tail_k0 < (1u << 23)
                     ? (float)((double)u01(s->a64.tail_k0 << 9) * s->a64.A - s->a64.int465)
                     : 0.0f;
    const GridState &G = s->grids[g];
    x.gi_max = std::max(std::max(G.interval[0], G.interval[1]), G.interval[2]);
    for (int k = 0; k < 3; ++k) x.rgi[k] = 1.0 / (double)G.interval[k];
    x.buie_vmax = s->buie_vmax_m;
    x.redo_n = s->d_mix_n;
}

om_status launch_trace_mixed_window(om_scene *s, const TraceArgs &a, const Trace64Args &b, const MixArgs &x, int nwg,
                                    uint32_t *acc, long long acc_words, hipStream_t st)
{
    if (a.ntiles <= 0) return OM_OK;
    if (!b.helio || !b.cell_geo || !a.cell_geo || !x.hmix || !x.redo_n || !x.redo || x.redo_cap <= 0) {
        set_error("scene has no mixed-precision geometry or the trace no redo list");
        return OM_ERR_INVALID;
    }
    om_status rc = launch_fill_u32(x.redo_n, 1, st); // this launch's list is empty
    if (rc != OM_OK) return rc;
    const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
    const MixLaunch ml{a, b, x};
    const dim3 grid(nwg), block(WIN_WGS), rgrid(REDO_BLOCKS), rblock(WG);
#define OM_MIX(SUN, P)                                                                     \
    do {                                                                                   \
        hipLaunchKernelGGL((trace_window_mixed_kernel<SUN, P>), grid, block, 0, st, ml);   \
        hipLaunchKernelGGL((trace_redo_kernel<SUN, P>), rgrid, rblock, 0, st, ml);        \
    } while (0)
    if (s->prm.sunshape == OM_SUN_BUIE) {
        if (poisson) OM_MIX(OM_SUN_BUIE, true);
        else OM_MIX(OM_SUN_BUIE, false);
    } else {
        if (poisson) OM_MIX(OM_SUN_PILLBOX, true);
        else OM_MIX(OM_SUN_PILLBOX, false);
    }
#undef OM_MIX
    OM_HIP(hipGetLastError());
    // A full list: the packed copies and side maps are cleared, the queue's chunk counters
    // zeroed, and the whole launch is traced again by the FP64 window trace, whose workgroups
    // overwrite their window slices.  Both kernels read the list's count and return at once
    // when it fitted.
    hipLaunchKernelGGL(clear_if_full_kernel, dim3(1024), dim3(WG), 0, st, x.redo_n, (uint32_t)x.redo_cap, acc,
                       acc_words, a.win_queue);
    OM_HIP(hipGetLastError());
    Trace64Args bg = b;
    bg.gate = x.redo_n;
    bg.gate_cap = (uint32_t)x.redo_cap;
    return launch_trace64_window(s, a, bg, nwg, st);
}

om_status launch_trace_mixed_records(om_scene *s, const TraceArgs &a, const Trace64Args &b, const MixArgs &x,
                                     hipStream_t st)
{
    if (a.ntiles <= 0) return OM_OK;
    if (a.rays_per_thread > RPT_MAX) {
        set_error("rays per thread above the hit-slot count");
        return OM_ERR_INVALID;
    }
    if (!b.helio || !b.cell_geo || !a.cell_geo || !x.hmix || !x.redo_n) {
        set_error("scene has no mixed-precision geometry");
        return OM_ERR_INVALID;
    }
    const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
    const MixLaunch ml{a, b, x};
    const dim3 grid(a.ntiles), block(WG);
    if (s->prm.sunshape == OM_SUN_BUIE) {
        if (poisson) hipLaunchKernelGGL((trace_mixed_records_kernel<OM_SUN_BUIE, true>), grid, block, 0, st, ml);
        else hipLaunchKernelGGL((trace_mixed_records_kernel<OM_SUN_BUIE, false>), grid, block, 0, st, ml);
    } else {
        if (poisson) hipLaunchKernelGGL((trace_mixed_records_kernel<OM_SUN_PILLBOX, true>), grid, block, 0, st, ml);
        else hipLaunchKernelGGL((trace_mixed_records_kernel<OM_SUN_PILLBOX, false>), grid, block, 0, st, ml);
    }
    OM_HIP(hipGetLastError());
    return OM_OK;
}

size_t mix_launch_bytes() { return sizeof(MixLaunch); }

om_status mix_batch_entry(const om_scene *s, int g, const TraceArgs &a, uint2 *redo, int redo_cap,
                          uint32_t *redo_n, void *dst)
{
    if (!mixed_supported(s)) {
        set_error("a scene of the batch cannot take the mixed path (cone receiver or literal self-intersection): "
                  "OM_BATCH_PER_SCENE traces it in FP64");
        return OM_ERR_UNSUPPORTED;
    }
    MixLaunch L{a, trace64_args(s, g), MixArgs{}};
    mixed_args(s, g, L.a, L.x);
    if (!L.b.helio || !L.b.cell_geo || !L.a.cell_geo || !L.x.hmix) {
        set_error("scene has no mixed-precision geometry");
        return OM_ERR_INVALID;
    }
    L.x.redo = redo;
    L.x.redo_cap = redo_cap;
    L.x.redo_n = redo_n;
    std::memcpy(dst, &L, sizeof(L));
    return OM_OK;
}

om_status launch_trace_batch_mixed(const MixBatch &m, hipStream_t st)
{
    if (m.n_chunks <= 0 || m.n_entries <= 0) return OM_OK;
    const MixLaunch *table = (const MixLaunch *)m.table;
    hipLaunchKernelGGL(batch_mix_reset_kernel, dim3((m.n_entries + m.n_scenes + WG - 1) / WG), dim3(WG), 0, st,
                       m.counters, m.n_entries, m.n_scenes);
    uint32_t *full = m.counters + 4 * (size_t)m.n_entries;
    const dim3 grid(m.nwg), block(WIN_WGS), rgrid(REDO_BLOCKS), wblock(WG);
#define OM_MIXB(SUN, P)                                                                                             \
    do {                                                                                                            \
        hipLaunchKernelGGL((trace_batch_window_mixed_kernel<SUN, P>), grid, block, 0, st, table, m.chunks,           \
                           m.n_chunks, m.queue);                                                                    \
        hipLaunchKernelGGL((trace_redo_batch_kernel<SUN, P>), rgrid, wblock, 0, st, table, m.entry_scene, full,      \
                           m.n_entries);                                                                            \
        hipLaunchKernelGGL(clear_full_scenes_kernel, dim3(256, m.n_scenes), wblock, 0, st, m.slots, full);           \
        hipLaunchKernelGGL((trace64_batch_full_kernel<SUN, P>), dim3(REDO_BLOCKS), wblock, 0, st, table, m.chunks,   \
                           m.entry_chunks, m.entry_scene, full, m.n_entries);                                       \
    } while (0)
    if (m.sunshape == OM_SUN_BUIE) {
        if (m.poisson) OM_MIXB(OM_SUN_BUIE, true);
        else OM_MIXB(OM_SUN_BUIE, false);
    } else {
        if (m.poisson) OM_MIXB(OM_SUN_PILLBOX, true);
        else OM_MIXB(OM_SUN_PILLBOX, false);
    }
#undef OM_MIXB
    OM_HIP(hipGetLastError());
    return OM_OK;
}

int trace_batch_window_mixed_blocks_per_cu(int sunshape, bool poisson)
{
    int n = 0;
    hipError_t e;
    if (sunshape == OM_SUN_BUIE)
        e = poisson ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_window_mixed_kernel<OM_SUN_BUIE, true>,
                                                                   WIN_WGS, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_batch_window_mixed_kernel<OM_SUN_BUIE, false>,
                                                                   WIN_WGS, 0);
    else
        e = poisson ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &n, trace_batch_window_mixed_kernel<OM_SUN_PILLBOX, true>, WIN_WGS, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &n, trace_batch_window_mixed_kernel<OM_SUN_PILLBOX, false>, WIN_WGS, 0);
    return e == hipSuccess ? n : 0;
}

Trace64Args trace64_args(const om_scene *s, int g)
{
    Trace64Args b = s->a64;
    b.cell_geo = g < (int)s->d_cell_geo64.size() ? s->d_cell_geo64[g] : nullptr;
    const GridState &G = s->grids[g];
    for (int k = 0; k < 3; ++k) {
        b.gpos[k] = G.pos[k];
        b.gint[k] = G.interval[k];
    }
    return b;
}

om_status launch_trace64(om_scene *s, const TraceArgs &a, const Trace64Args &b, bool records, hipStream_t st)
{
    if (a.ntiles <= 0) return OM_OK;
    if (a.rays_per_thread > RPT_MAX) {
        set_error("rays per thread above the hit-slot count");
        return OM_ERR_INVALID;
    }
    if (!b.helio || !b.cell_geo) {
        set_error("scene has no FP64 geometry");
        return OM_ERR_INVALID;
    }
    const bool poisson = s->prm.micro_mode == OM_MICRO_POISSON;
    const dim3 grid(a.ntiles), block(WG);
    const Trace64Launch ab{a, b};
#define OM_T64(SUN)                                                                                    \
    do {                                                                                               \
        if (poisson) {                                                                                 \
            if (records) hipLaunchKernelGGL((trace64_kernel<SUN, true, true>), grid, block, 0, st, ab);  \
            else hipLaunchKernelGGL((trace64_kernel<SUN, true, false>), grid, block, 0, st, ab);         \
        } else {                                                                                       \
            if (records) hipLaunchKernelGGL((trace64_kernel<SUN, false, true>), grid, block, 0, st, ab); \
            else hipLaunchKernelGGL((trace64_kernel<SUN, false, false>), grid, block, 0, st, ab);        \
        }                                                                                              \
    } while (0)
    if (s->prm.sunshape == OM_SUN_BUIE) OM_T64(OM_SUN_BUIE);
    else OM_T64(OM_SUN_PILLBOX);
#undef OM_T64
    OM_HIP(hipGetLastError());
    return OM_OK;
}

} // namespace om
