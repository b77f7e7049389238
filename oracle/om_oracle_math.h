/*
 * om_oracle_math.h -- internal helpers of the CPU oracle (test infrastructure).
 *
 * float3 algebra with the host semantics of the reference's vector header
 * (vector_arithmetic.cuh:51-54 rsqrtf = 1/sqrtf; :345-349 float3/float multiplies by the
 * reciprocal; :383-413 dot, cross, normalize, reflect; :571-574 float3 -> int3 truncates),
 * the geometry of global_function.cuh:14-105, and the project-written transcendental
 * functions that the HIP kernels evaluate with the same operation sequence.
 *
 * Build with -ffp-contract=off: every a*b+c below must round twice, as on the GPU.
 */
#ifndef OM_ORACLE_MATH_H
#define OM_ORACLE_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } f3;
typedef struct { double x, y, z; } d3;

static inline f3 mk3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline f3 scl3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline f3 sclf3(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* vector_arithmetic.cuh:397-401 with the host rsqrtf of :51-54 */
static inline f3 normalize3(f3 v)
{
    float inv = 1.0f / sqrtf(dot3(v, v));
    return scl3(v, inv);
}
/* vector_arithmetic.cuh:410-413: i - 2*n*dot(n,i), evaluated as (2*n)*d */
static inline f3 reflect3(f3 i, f3 n)
{
    float d = dot3(n, i);
    f3 tn = sclf3(2.0f, n);
    return sub3(i, scl3(tn, d));
}

static inline d3 mkd(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 addd(d3 a, d3 b) { return mkd(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 subd(d3 a, d3 b) { return mkd(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 negd(d3 a) { return mkd(-a.x, -a.y, -a.z); }
static inline d3 scld(d3 a, double s) { return mkd(a.x * s, a.y * s, a.z * s); }
static inline double dotd(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 crossd(d3 a, d3 b)
{
    return mkd(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline d3 normalized(d3 v) { return scld(v, 1.0 / sqrt(dotd(v, v))); }
static inline d3 reflectd(d3 i, d3 n) { return subd(i, scld(n, 2.0 * dotd(n, i))); }
static inline d3 f2d(f3 a) { return mkd(a.x, a.y, a.z); }

/* The reference's Epsilon is the double literal 1e-6 (global_constant.h:6); every test
 * against it compares a float promoted to double, which C reproduces as written. */
#define OC_EPS 1e-6

/* global_function.cuh:14-52 */
static inline f3 local2world3(f3 d, f3 n)
{
    if (fabsf(n.x) < OC_EPS && fabsf(n.z) < OC_EPS)
        return d;
    f3 u = normalize3(cross3(mk3(0.0f, 1.0f, 0.0f), n));
    f3 v = normalize3(cross3(u, n));
    return mk3(d.x * u.x + d.y * n.x + d.z * v.x,
               d.x * u.y + d.y * n.y + d.z * v.y,
               d.x * u.z + d.y * n.z + d.z * v.z);
}
static inline d3 local2worldd(d3 d, d3 n)
{
    if (fabs(n.x) < OC_EPS && fabs(n.z) < OC_EPS)
        return d;
    d3 u = normalized(crossd(mkd(0.0, 1.0, 0.0), n));
    d3 v = normalized(crossd(u, n));
    return mkd(d.x * u.x + d.y * n.x + d.z * v.x,
               d.x * u.y + d.y * n.y + d.z * v.y,
               d.x * u.z + d.y * n.z + d.z * v.z);
}

/* global_function.cuh:61-74 */
static inline f3 rotate_y3(f3 o, f3 old_dir, f3 new_dir)
{
    int dir = (cross3(old_dir, new_dir).y > 0) ? 1 : -1;
    float c = dot3(old_dir, new_dir);
    float s = (float)dir * sqrtf(1 - c * c);
    return mk3(c * o.x + s * o.z, o.y, -s * o.x + c * o.z);
}

/* global_function.cuh:78-105 */
static inline int ray_parallelogram3(f3 orig, f3 dir, f3 A, f3 B, f3 C,
                                     float *t, float *u, float *v)
{
    f3 E1 = sub3(B, A);
    f3 E2 = sub3(C, A);
    f3 pvec = cross3(dir, E2);
    float det = dot3(E1, pvec);
    if (fabsf(det) < OC_EPS) return 0;
    float invDet = 1 / det;
    f3 T = sub3(orig, A);
    *u = dot3(T, pvec) * invDet;
    if (*u < 0 || *u > 1) return 0;
    f3 qvec = cross3(T, E1);
    *v = dot3(dir, qvec) * invDet;
    if (*v < 0 || *v > 1) return 0;
    *t = dot3(E2, qvec) * invDet;
    if (*t < OC_EPS) return 0;
    return 1;
}
static inline int ray_parallelogramd(d3 orig, d3 dir, d3 A, d3 B, d3 C,
                                     double *t, double *u, double *v)
{
    d3 E1 = subd(B, A);
    d3 E2 = subd(C, A);
    d3 pvec = crossd(dir, E2);
    double det = dotd(E1, pvec);
    if (fabs(det) < OC_EPS) return 0;
    double invDet = 1 / det;
    d3 T = subd(orig, A);
    *u = dotd(T, pvec) * invDet;
    if (*u < 0 || *u > 1) return 0;
    d3 qvec = crossd(T, E1);
    *v = dotd(dir, qvec) * invDet;
    if (*v < 0 || *v > 1) return 0;
    *t = dotd(E2, qvec) * invDet;
    if (*t < OC_EPS) return 0;
    return 1;
}

/* float -> int truncation of make_int3(float3) (vector_arithmetic.cuh:571-574); values are
 * clamped to +-2^30 first so that the conversion is defined on every platform. */
static inline int trunc_i(float x)
{
    if (!(x > -1073741824.0f)) x = -1073741824.0f;
    if (x > 1073741824.0f) x = 1073741824.0f;
    return (int)x;
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* atan2(y, x) in (-pi, pi] (the project's float function, no reference counterpart):
 * t = min/max of |x|, |y|; above tan(pi/8) atan(t) = pi/4 + atan((t - 1)/(t + 1)), so the
 * alternating series to s^23 runs on |s| <= 0.4143; then the octant. */
static inline float oc_atan2f(float y, float x)
{
    float ax = fabsf(x), ay = fabsf(y);
    float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    float t = mn / mx;
    int big = t > 0.414213562f;
    float s = big ? (t - 1.0f) / (t + 1.0f) : t;
    float z = s * s;
    float p = -0.0434782609f;
    p = fmaf(p, z, 0.0476190476f);
    p = fmaf(p, z, -0.0526315789f);
    p = fmaf(p, z, 0.0588235294f);
    p = fmaf(p, z, -0.0666666667f);
    p = fmaf(p, z, 0.0769230769f);
    p = fmaf(p, z, -0.0909090909f);
    p = fmaf(p, z, 0.111111111f);
    p = fmaf(p, z, -0.142857143f);
    p = fmaf(p, z, 0.2f);
    p = fmaf(p, z, -0.333333333f);
    float r = fmaf(p * z, s, s);
    if (big) r = r + 0.785398163f;
    if (ay > ax) r = 1.57079633f - r;
    if (x < 0.0f) r = 3.14159265f - r;
    return y < 0.0f ? -r : r;
}

/* Cylinder receiver pixel (new; CylinderReceiver::GIntersect is an empty stub,
 * receiver.cuh:103-122): vertical axis through base, radius^2 r2, height h, outer surface
 * only.  Entering root from the closest approach tc: t = tc - sqrt((r2 - d2) / a); column
 * = azimuth atan2(z, x) in [0, 2 pi) x res_x, row = (y - base.y) / h x res_y. */
static inline int cylinder_pixel3(f3 o, f3 d, f3 base, float r2, float h, int res_x, int res_y)
{
    float ox = o.x - base.x, oz = o.z - base.z;
    float a = d.x * d.x + d.z * d.z;
    if (!(a >= OC_EPS)) return -1;
    float tc = -(ox * d.x + oz * d.z) / a;
    float px = ox + tc * d.x, pz = oz + tc * d.z;
    float d2 = px * px + pz * pz;
    if (!(d2 < r2)) return -1;
    float t = tc - sqrtf((r2 - d2) / a);
    if (t < OC_EPS) return -1;
    float hx = ox + t * d.x, hz = oz + t * d.z;
    float hy = (o.y + t * d.y) - base.y;
    if (hy < 0.0f || hy > h) return -1;
    float th = oc_atan2f(hz, hx);
    if (th < 0.0f) th = th + 6.28318531f;
    int col = trunc_i(th * 0.159154943f * (float)res_x);
    int row = trunc_i(hy / h * (float)res_y);
    if (col > res_x - 1) col = res_x - 1;
    if (row > res_y - 1) row = res_y - 1;
    return row * res_x + col;
}

/* Truncated-cone receiver (the project's; CircularTruncatedConeReceiver::GIntersect is an
 * empty stub, receiver.cuh:125-140): vertical axis through base, radius r0 + k y at height
 * y in [0, h]; outer surface only.  The quadratic |p_xz|^2 = r(p_y)^2 is solved from the
 * ray's closest approach to the axis in the xz plane (small coefficients); the nearer root
 * ahead within the height range must enter the surface.  Column = azimuth as for the
 * cylinder, row = lateral area below the hit over the whole: y (2 r0 + k y) inv_area. */
static inline int cone_pixel3(f3 o, f3 d, f3 base, float r0, float k, float h, float inv_area, int res_x,
                              int res_y)
{
    float ox = o.x - base.x, oy = o.y - base.y, oz = o.z - base.z;
    float axy = d.x * d.x + d.z * d.z;
    if (!(axy >= OC_EPS)) return -1;
    float tc = -(ox * d.x + oz * d.z) / axy;
    float px = ox + tc * d.x, py = oy + tc * d.y, pz = oz + tc * d.z;
    float d2 = px * px + pz * pz;
    float w0 = r0 + k * py, wd = k * d.y;
    float A = axy - wd * wd;
    float disc = axy * (w0 * w0 - d2) + wd * wd * d2;
    if (!(disc >= 0.0f) || A == 0.0f) return -1;
    float sq = sqrtf(disc);
    float ta = tc + (w0 * wd - sq) / A, tb = tc + (w0 * wd + sq) / A;
    float t0 = ta < tb ? ta : tb, t1 = ta < tb ? tb : ta;
    float t = -1.0f, hy = 0.0f;
    for (int i = 0; i < 2; ++i) {
        float ti = i == 0 ? t0 : t1;
        if (ti < OC_EPS) continue;
        float yi = oy + ti * d.y;
        if (yi < 0.0f || yi > h) continue;
        t = ti;
        hy = yi;
        break;
    }
    if (t < 0.0f) return -1;
    float hx = ox + t * d.x, hz = oz + t * d.z;
    if (!(d.x * hx + d.z * hz - d.y * (k * (r0 + k * hy)) < 0.0f)) return -1;
    float th = oc_atan2f(hz, hx);
    if (th < 0.0f) th = th + 6.28318531f;
    int col = trunc_i(th * 0.159154943f * (float)res_x);
    int row = trunc_i(hy * (2.0f * r0 + k * hy) * inv_area * (float)res_y);
    if (col > res_x - 1) col = res_x - 1;
    if (row > res_y - 1) row = res_y - 1;
    return row * res_x + col;
}

#endif
