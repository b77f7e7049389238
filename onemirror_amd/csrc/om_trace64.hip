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
    const u4 x = philox_uniform_key(u4{(uint32_t)l, (uint32_t)m, (uint32_t)h, a.w3}, a.key0, a.key1);
    const double u0 = u01(x.x), u1 = u01(x.y), u2 = u01(x.z), u3 = u01(x.w);
    double th;
    if (SUN == OM_SUN_PILLBOX) {
        th = b.pillbox * sqrt(u0);
    } else {
        const double v = u0 * b.A;
        if (v < b.int465) {
            // the largest lo with cdf_y[lo] <= v (max_less_index, scene_instance_process.cu:
            // 158-173), from the guide entry of v
            int lo = b.guide[min(max((int)(v * b.gscale), 0), b.nbins - 1)];
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
    const double ths = erfinv_ray(2.0 * u2 - 1.0) * 1.4142135623730951 * b.sigma;
    const double phs = 6.283185307179586 * u3;
    const double2 sts = sincos_polar(ths), sps = sincos_ray(phs);
    const d3 nrm = normalized(world(mkd(sts.x * sps.y, sts.y, sts.x * sps.x), hf));
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
    int m_o = -1; // microsample of orig
    d3 orig{0.0, 0.0, 0.0};
    Cell64 c0{};
    for (int k = 0; k < nrays; k += 64) {
        const int r = tl.y + k + lane;
        bool out = false;
        int px = 0;
        uint32_t q = 0u;
        asm volatile("" ::: "memory"); // arguments re-read per ray pass
        if (r < nr) {
            const int m = r / L, l = r - m * L;
            if (m != m_o) {
                orig = add(world(micro_local64<POISSON>(a, b, hr, h64, h, m), hf), hpos);
                c0 = cell64(b, orig);
                m_o = m;
            }
            const Ray64 o = trace_ray64<SUN>(a, b, hf, sf, orig, c0, h, m, l, self, clear);
            if (o.px >= 0) {
                px = o.px;
                q = o.q;
                const uint32_t wr = (uint32_t)(o.rc.x - a.win_y0), wc = (uint32_t)(o.rc.y - a.win_x0);
                if (wr < (uint32_t)a.win_h && wc < (uint32_t)a.win_w)
                    packed_add_lds(a, s_win + wr * (uint32_t)a.win_w + wc, px, q);
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
    __shared__ unsigned long long s_win[WIN_PX_MAX];
    __shared__ unsigned long long s_out[WIN_WAVES][WIN_OUT];
    window_body(args.a, s_win, s_out,
                [&](int2 tl, int lane, int nrays, WaveOut &wo, unsigned long long *w) {
                    trace_chunk_window64<SUN, POISSON>(args.a, args.b, tl, lane, nrays, wo, w);
                });
}

} // namespace

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
