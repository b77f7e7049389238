// om_common.h -- single-precision geometry, sampling and RNG shared by the gfx950 kernels
// and the host-side scene build.  Every function is one fixed IEEE operation sequence:
// the library is compiled with -ffp-contract=off, divisions and square roots are
// correctly rounded, and transcendental functions are project-written polynomials, so a
// value computed here is the same bits on the host and on the GPU.
//
// Reference semantics (Hu_smooth/ClassedSunRay):
//   vector algebra            vector_arithmetic.cuh:341-413 (host rsqrtf = 1/sqrtf, :51-54)
//   frames and intersection   global_function.cuh:14-105, :175-180
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OM_HD __host__ __device__ __forceinline__
#else
#define OM_HD inline
#endif

namespace om {

// ---- the reference's Epsilon is the double 1e-6 (global_constant.h:6); a float x
// compared against it is promoted to double.  Equivalent float thresholds:
//   x <  1e-6  <=>  x <  EPS_LT        x <= 1e-6  <=>  x <= EPS_LE
constexpr float EPS_LT = 0x1.0c6f7cp-20f;  // smallest float above 1e-6
constexpr float EPS_LE = 0x1.0c6f7ap-20f;  // largest float below 1e-6
constexpr float INT_MAX_F = 2147483648.0f; // float(INT_MAX), absDivide's sentinel

OM_HD float fabs_(float x) { return __builtin_fabsf(x); }
OM_HD float sqrt_(float x) { return __builtin_sqrtf(x); }
OM_HD float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
OM_HD float rint_(float x) { return __builtin_rintf(x); }
OM_HD float floor_(float x) { return __builtin_floorf(x); }
OM_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
OM_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

struct v3 {
    float x, y, z;
};
OM_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
OM_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
OM_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
OM_HD v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
OM_HD v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
OM_HD v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
OM_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OM_HD v3 cross(v3 a, v3 b)
{
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Division and square root of the functions below come from a policy M.  ieee is the plain
// correctly rounded operation; the trace kernel passes a policy with shorter sequences that
// give the same bits for the operand ranges it uses.
struct ieee {
    OM_HD static float div(float a, float b) { return a / b; }
    OM_HD static float sqrt(float x) { return sqrt_(x); }
};

// normalize with the host rsqrtf: v * (1 / sqrtf(v.v))
template <class M = ieee>
OM_HD v3 normalize(v3 v)
{
    float inv = M::div(1.0f, M::sqrt(dot(v, v)));
    return v * inv;
}
// reflect(i, n) = i - (2 n) dot(n, i)
OM_HD v3 reflect(v3 i, v3 n)
{
    float d = dot(n, i);
    return i - (2.0f * n) * d;
}

// float -> int truncation, clamped to +-2^30 so that every platform agrees (NaN -> -2^30)
OM_HD int trunc_i(float x)
{
    if (!(x > -1073741824.0f)) x = -1073741824.0f;
    if (x > 1073741824.0f) x = 1073741824.0f;
    return (int)x;
}

// ---- orthonormal frame of local2world (global_function.cuh:14-52): u = norm((0,1,0) x n),
// v = norm(u x n); identity when |n.x| and |n.z| are below Epsilon.
struct frame {
    v3 u, n, v;
    int identity;
};
OM_HD frame make_frame(v3 n)
{
    frame f;
    f.n = n;
    f.identity = (fabs_(n.x) < EPS_LT && fabs_(n.z) < EPS_LT) ? 1 : 0;
    if (f.identity) {
        f.u = mk(1.0f, 0.0f, 0.0f);
        f.v = mk(0.0f, 0.0f, 1.0f);
        return f;
    }
    f.u = normalize(cross(mk(0.0f, 1.0f, 0.0f), n));
    f.v = normalize(cross(f.u, n));
    return f;
}
OM_HD v3 to_world(v3 d, const frame &f)
{
    if (f.identity) return d;
    return v3{d.x * f.u.x + d.y * f.n.x + d.z * f.v.x,
              d.x * f.u.y + d.y * f.n.y + d.z * f.v.y,
              d.x * f.u.z + d.y * f.n.z + d.z * f.v.z};
}

// ---- rotateY (global_function.cuh:61-74)
OM_HD v3 rotate_y(v3 o, v3 old_dir, v3 new_dir)
{
    int dir = (cross(old_dir, new_dir).y > 0) ? 1 : -1;
    float c = dot(old_dir, new_dir);
    float s = (float)dir * sqrt_(1.0f - c * c);
    return v3{c * o.x + s * o.z, o.y, -s * o.x + c * o.z};
}

// ---- rayParallelogramIntersect (global_function.cuh:78-105) with E1 = B - A and
// E2 = C - A precomputed by the caller.
template <class M = ieee>
OM_HD bool parallelogram(v3 orig, v3 dir, v3 A, v3 E1, v3 E2, float &t, float &u, float &v)
{
    v3 pvec = cross(dir, E2);
    float det = dot(E1, pvec);
    if (fabs_(det) < EPS_LT) return false;
    float invDet = M::div(1.0f, det);
    v3 T = orig - A;
    u = dot(T, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    v3 qvec = cross(T, E1);
    v = dot(dir, qvec) * invDet;
    if (v < 0.0f || v > 1.0f) return false;
    t = dot(E2, qvec) * invDet;
    if (t < EPS_LT) return false;
    return true;
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
struct u4 {
    uint32_t x, y, z, w;
};
OM_HD u4 philox(u4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
// counter word 3: scene index << 4 | purpose
enum : uint32_t { PURPOSE_RAY = 0, PURPOSE_POISSON = 1 };

// uniform on (0,1): (2 (x >> 9) + 1) 2^-24, exact in float
OM_HD float u01(uint32_t x) { return (float)((x >> 9) * 2u + 1u) * 5.9604644775390625e-08f; }

// sin and cos of 2 pi u: quadrant of t = 4u, remainder in [-1/2, 1/2] (both exact), angle
// a = r pi/2, Taylor series of sin to a^9 and of cos to a^10.
OM_HD void sincos2pi(float u, float &s, float &c)
{
    float t = u * 4.0f;
    float q = rint_(t);
    float r = t - q;
    float a = r * 1.57079637f;
    float a2 = a * a;
    float ps = fma_(a2, 2.75573192e-06f, -1.98412698e-04f);
    ps = fma_(a2, ps, 8.33333333e-03f);
    ps = fma_(a2, ps, -1.66666667e-01f);
    float sn = fma_(a * a2, ps, a);
    float pc = fma_(a2, -2.75573192e-07f, 2.48015873e-05f);
    pc = fma_(a2, pc, -1.38888889e-03f);
    pc = fma_(a2, pc, 4.16666667e-02f);
    pc = fma_(a2, pc, -0.5f);
    float cs = fma_(a2, pc, 1.0f);
    int qi = (int)q & 3;
    float s0 = (qi & 1) ? cs : sn;
    float c0 = (qi & 1) ? sn : cs;
    s = (qi & 2) ? -s0 : s0;
    c = ((qi + 1) & 2) ? -c0 : c0;
}

// sin and cos of a small polar angle: Taylor series below 0.25 rad, else sincos2pi
OM_HD void sincos_small(float x, float &s, float &c)
{
    if (fabs_(x) < 0.25f) {
        float x2 = x * x;
        float ps = fma_(x2, -1.98412698e-04f, 8.33333333e-03f);
        ps = fma_(x2, ps, -1.66666667e-01f);
        s = fma_(x * x2, ps, x);
        float pc = fma_(x2, 2.48015873e-05f, -1.38888889e-03f);
        pc = fma_(x2, pc, 4.16666667e-02f);
        pc = fma_(x2, pc, -0.5f);
        c = fma_(x2, pc, 1.0f);
    } else {
        float u = x * 0.159154943f;
        u = u - floor_(u);
        sincos2pi(u, s, c);
    }
}

// natural log of a positive normal float: m in [sqrt(1/2), sqrt(2)), 2 atanh((m-1)/(m+1))
// as an odd series to s^11, plus e ln2 in two parts.
template <class M = ieee>
OM_HD float log_(float x)
{
    uint32_t ix = f2u(x);
    int e = (int)(ix >> 23) - 127;
    float m = u2f((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float f = m - 1.0f;
    float s = M::div(f, m + 1.0f);
    float s2 = s * s;
    float p = fma_(s2, 0.0909090909f, 0.111111111f);
    p = fma_(s2, p, 0.142857143f);
    p = fma_(s2, p, 0.2f);
    p = fma_(s2, p, 0.333333333f);
    float at = fma_(s * s2, p, s);
    float lm = at * 2.0f;
    float fe = (float)e;
    return fma_(fe, 0.693145752f, fma_(fe, 1.42860677e-06f, lm));
}

// exp for |x| < 87: 2^k exp(r), r = x - k ln2 (two parts), Taylor to r^8
OM_HD float exp_(float x)
{
    float k = rint_(x * 1.44269504f);
    float r = fma_(-k, 0.693145752f, x);
    r = fma_(-k, 1.42860677e-06f, r);
    float p = fma_(r, 2.48015873e-05f, 1.98412698e-04f);
    p = fma_(r, p, 1.38888889e-03f);
    p = fma_(r, p, 8.33333333e-03f);
    p = fma_(r, p, 4.16666667e-02f);
    p = fma_(r, p, 1.66666667e-01f);
    p = fma_(r, p, 0.5f);
    p = fma_(r, p, 1.0f);
    p = fma_(r, p, 1.0f);
    int ki = (int)k;
    return p * u2f((uint32_t)(ki + 127) << 23);
}

template <class M = ieee>
OM_HD float pow_(float x, float y) { return exp_(y * log_<M>(x)); }

// erfinv in single precision: M. Giles, "Approximating the erfinv function",
// GPU Computing Gems Jade Edition (2011): polynomial in w = -log(1 - x^2), two branches.
template <class M = ieee>
OM_HD float erfinv_(float x)
{
    float w = -log_<M>((1.0f - x) * (1.0f + x));
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = fma_(p, w, 3.43273939e-07f);
        p = fma_(p, w, -3.5233877e-06f);
        p = fma_(p, w, -4.39150654e-06f);
        p = fma_(p, w, 0.00021858087f);
        p = fma_(p, w, -0.00125372503f);
        p = fma_(p, w, -0.00417768164f);
        p = fma_(p, w, 0.246640727f);
        p = fma_(p, w, 1.50140941f);
    } else {
        w = M::sqrt(w) - 3.0f;
        p = -0.000200214257f;
        p = fma_(p, w, 0.000100950558f);
        p = fma_(p, w, 0.00134934322f);
        p = fma_(p, w, -0.00367342844f);
        p = fma_(p, w, 0.00573950773f);
        p = fma_(p, w, -0.0076224613f);
        p = fma_(p, w, 0.00943887047f);
        p = fma_(p, w, 1.00167406f);
        p = fma_(p, w, 2.83297682f);
    }
    return p * x;
}

// standard normal deviate: sqrt(2) erfinv(2u - 1)
template <class M = ieee>
OM_HD float gauss_(float u) { return erfinv_<M>(2.0f * u - 1.0f) * 1.41421356f; }

// angle2xyz (global_function.cuh:175-180) from the sines and cosines
OM_HD v3 angle2xyz(float st, float ct, float sp, float cp) { return v3{st * cp, ct, st * sp}; }

// atan2(y, x) in (-pi, pi]: t = min/max of |x|, |y| in [0, 1]; above tan(pi/8) the
// identity atan(t) = pi/4 + atan((t - 1)/(t + 1)) brings the argument to |s| <= 0.4143,
// where the alternating series s - s^3/3 + ... - s^23/23 is within 2e-11; then the octant.
OM_HD float atan2_(float y, float x)
{
    const float ax = fabs_(x), ay = fabs_(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    const float t = mn / mx;
    const bool big = t > 0.414213562f;
    const float s = big ? (t - 1.0f) / (t + 1.0f) : t;
    const float z = s * s;
    float p = -0.0434782609f;             // -1/23
    p = fma_(p, z, 0.0476190476f);        //  1/21
    p = fma_(p, z, -0.0526315789f);       // -1/19
    p = fma_(p, z, 0.0588235294f);        //  1/17
    p = fma_(p, z, -0.0666666667f);       // -1/15
    p = fma_(p, z, 0.0769230769f);        //  1/13
    p = fma_(p, z, -0.0909090909f);       // -1/11
    p = fma_(p, z, 0.111111111f);         //  1/9
    p = fma_(p, z, -0.142857143f);        // -1/7
    p = fma_(p, z, 0.2f);                 //  1/5
    p = fma_(p, z, -0.333333333f);        // -1/3
    float r = fma_(p * z, s, s);
    if (big) r = r + 0.785398163f;
    if (ay > ax) r = 1.57079633f - r;
    if (x < 0.0f) r = 3.14159265f - r;
    return y < 0.0f ? -r : r;
}

// Cylinder receiver (new: the reference's CylinderReceiver::GIntersect is an empty stub,
// receiver.cuh:103-122).  Vertical axis through base = (cx, y0, cz), radius^2 r2, height h;
// only the outer surface receives: the entering root of |(o + t d - base)_xz| = r taken
// from the closest approach tc (well conditioned far from the axis), 0 <= y - y0 <= h.
// Pixel: column = azimuth atan2(z, x) over [0, 2 pi) x res_x, row = (y - y0) / h x res_y.
// Pixel (row, column) of the cylinder hit; row < 0 for a miss.
OM_HD int2 cylinder_rc(v3 o, v3 d, v3 base, float r2, float h, int res_x, int res_y)
{
    const int2 miss = make_int2(-1, -1);
    const float ox = o.x - base.x, oz = o.z - base.z;
    const float a = d.x * d.x + d.z * d.z;
    if (!(a >= EPS_LT)) return miss;
    const float tc = -(ox * d.x + oz * d.z) / a;
    const float px = ox + tc * d.x, pz = oz + tc * d.z;
    const float d2 = px * px + pz * pz;
    if (!(d2 < r2)) return miss;
    const float t = tc - sqrt_((r2 - d2) / a);
    if (t < EPS_LT) return miss;
    const float hx = ox + t * d.x, hz = oz + t * d.z;
    const float hy = (o.y + t * d.y) - base.y;
    if (hy < 0.0f || hy > h) return miss;
    float th = atan2_(hz, hx);
    if (th < 0.0f) th = th + 6.28318531f;
    int col = trunc_i(th * 0.159154943f * (float)res_x);
    int row = trunc_i(hy / h * (float)res_y);
    col = col < res_x - 1 ? col : res_x - 1;
    row = row < res_y - 1 ? row : res_y - 1;
    return make_int2(row, col);
}

OM_HD int cylinder_pixel(v3 o, v3 d, v3 base, float r2, float h, int res_x, int res_y)
{
    const int2 rc = cylinder_rc(o, d, base, r2, h, res_x, res_y);
    return rc.x >= 0 ? rc.x * res_x + rc.y : -1;
}

// Truncated-cone receiver (new: the reference's CircularTruncatedConeReceiver::GIntersect is
// an empty stub, receiver.cuh:125-140).  Vertical axis through base = (cx, y0, cz); radius
// r(y) = r0 + k y at height y in [0, h] above the base (r0 the bottom radius, k = (r1 - r0)
// / h); only the outer surface receives.  The ray is first moved to its closest approach to
// the axis in the xz plane (tc), where the quadratic |p_xz|^2 = r(p_y)^2 has small
// coefficients: with w0 = r(py), wd = k dy, A = dx^2 + dz^2 - wd^2, its discriminant is
// (dx^2 + dz^2)(w0^2 - d2) + wd^2 d2 (d2 = |p_xz|^2) and its roots tc + (w0 wd -+ sqrt(disc))
// / A.  The nearer root ahead of the ray within the height range must enter the surface
// (the ray against the outward normal (x, -k r(y), z)); otherwise the ray misses.
// Pixel: column = azimuth as for the cylinder; row = the lateral area below the hit as a
// fraction of the whole, y (2 r0 + k y) / (h (r0 + r1)) = y (2 r0 + k y) inv_area, so that
// every pixel covers the same area.
OM_HD int2 cone_rc(v3 o, v3 d, v3 base, float r0, float k, float h, float inv_area, int res_x, int res_y)
{
    const int2 miss = make_int2(-1, -1);
    const float ox = o.x - base.x, oy = o.y - base.y, oz = o.z - base.z;
    const float axy = d.x * d.x + d.z * d.z;
    if (!(axy >= EPS_LT)) return miss;
    const float tc = -(ox * d.x + oz * d.z) / axy;
    const float px = ox + tc * d.x, py = oy + tc * d.y, pz = oz + tc * d.z;
    const float d2 = px * px + pz * pz;
    const float w0 = r0 + k * py, wd = k * d.y;
    const float A = axy - wd * wd;
    const float disc = axy * (w0 * w0 - d2) + wd * wd * d2;
    if (!(disc >= 0.0f) || A == 0.0f) return miss;
    const float sq = sqrt_(disc);
    const float ta = tc + (w0 * wd - sq) / A, tb = tc + (w0 * wd + sq) / A;
    const float t0 = ta < tb ? ta : tb, t1 = ta < tb ? tb : ta;
    float t = -1.0f, hy = 0.0f;
    for (int i = 0; i < 2; ++i) {
        const float ti = i == 0 ? t0 : t1;
        if (ti < EPS_LT) continue;
        const float yi = oy + ti * d.y;
        if (yi < 0.0f || yi > h) continue;
        t = ti;
        hy = yi;
        break;
    }
    if (t < 0.0f) return miss;
    const float hx = ox + t * d.x, hz = oz + t * d.z;
    // entering: d . (x, -k r(y), z) < 0
    if (!(d.x * hx + d.z * hz - d.y * (k * (r0 + k * hy)) < 0.0f)) return miss;
    float th = atan2_(hz, hx);
    if (th < 0.0f) th = th + 6.28318531f;
    int col = trunc_i(th * 0.159154943f * (float)res_x);
    int row = trunc_i(hy * (2.0f * r0 + k * hy) * inv_area * (float)res_y);
    col = col < res_x - 1 ? col : res_x - 1;
    row = row < res_y - 1 ? row : res_y - 1;
    return make_int2(row, col);
}

OM_HD int cone_pixel(v3 o, v3 d, v3 base, float r0, float k, float h, float inv_area, int res_x, int res_y)
{
    const int2 rc = cone_rc(o, d, base, r0, k, h, inv_area, res_x, res_y);
    return rc.x >= 0 ? rc.x * res_x + rc.y : -1;
}

} // namespace om
