/*
 * om_oracle.c -- CPU oracle, FP32 reference semantics (test infrastructure only; see
 * om_oracle.h).  Every routine restates a reference routine of Hu_smooth/ClassedSunRay
 * and cites it; the Philox counter layout and the project-written transcendental
 * functions are the ones the HIP kernels use, so that results are bit-identical.
 */
#include "om_oracle.h"
#include "om_oracle_math.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------ */
/* Philox4x32-10: Salmon et al., "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11.   */
/* Multipliers 0xD2511F53, 0xCD9E8D57; Weyl key increments 0x9E3779B9, 0xBB67AE85.       */
/* ------------------------------------------------------------------------------------ */
void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Counter layout (shared with the kernels):
 *   ray draws     ctr = {l, m, h, scene<<4 | 0}  -> {sun theta, sun phi, slope N, slope phi}
 *   Poisson draws ctr = {idx, phase, h, scene<<4 | 1} -> {x, y}
 *   key = {seed low word, seed high word}                                               */
enum { OC_PURPOSE_RAY = 0, OC_PURPOSE_POISSON = 1, OC_PURPOSE_POOL = 2 };

/* uniform on (0,1): (2*(x>>9)+1) * 2^-24, 23-bit resolution, exact in float */
float oc_u01(uint32_t x) { return (float)((x >> 9) * 2u + 1u) * 5.9604644775390625e-08f; }

/* sin(2 pi u), cos(2 pi u): quadrant from t = 4u (exact), remainder r = t - rint(t)
 * (exact), angle a = r*pi/2 in [-pi/4, pi/4], Taylor series of sin to a^9 and cos to
 * a^10 (truncation below 2e-9). */
void oc_sincos2pi(float u, float *s, float *c)
{
    float t = u * 4.0f;
    float q = rintf(t);
    float r = t - q;
    float a = r * 1.57079637f;
    float a2 = a * a;
    float ps = fmaf(a2, 2.75573192e-06f, -1.98412698e-04f);
    ps = fmaf(a2, ps, 8.33333333e-03f);
    ps = fmaf(a2, ps, -1.66666667e-01f);
    float sn = fmaf(a * a2, ps, a);
    float pc = fmaf(a2, -2.75573192e-07f, 2.48015873e-05f);
    pc = fmaf(a2, pc, -1.38888889e-03f);
    pc = fmaf(a2, pc, 4.16666667e-02f);
    pc = fmaf(a2, pc, -0.5f);
    float cs = fmaf(a2, pc, 1.0f);
    switch ((int)q & 3) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

/* sin, cos of a small angle (sunshape and slope-error polar angles): Taylor series for
 * |x| < 0.25 (truncation below 1.1e-11), otherwise reduced through sincos2pi. */
void oc_sincos_small(float x, float *s, float *c)
{
    if (fabsf(x) < 0.25f) {
        float x2 = x * x;
        float ps = fmaf(x2, -1.98412698e-04f, 8.33333333e-03f);
        ps = fmaf(x2, ps, -1.66666667e-01f);
        *s = fmaf(x * x2, ps, x);
        float pc = fmaf(x2, 2.48015873e-05f, -1.38888889e-03f);
        pc = fmaf(x2, pc, 4.16666667e-02f);
        pc = fmaf(x2, pc, -0.5f);
        *c = fmaf(x2, pc, 1.0f);
    } else {
        float u = x * 0.159154943f;
        u = u - floorf(u);
        oc_sincos2pi(u, s, c);
    }
}

/* natural log of a positive normal float: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * log m = 2 atanh(s), s = (m-1)/(m+1), series to s^11; e*ln2 in two parts. */
float oc_logf(float x)
{
    uint32_t ix = f2u(x);
    int e = (int)(ix >> 23) - 127;
    float m = u2f((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;
    float s = f / (m + 1.0f);
    float s2 = s * s;
    float p = fmaf(s2, 0.0909090909f, 0.111111111f);
    p = fmaf(s2, p, 0.142857143f);
    p = fmaf(s2, p, 0.2f);
    p = fmaf(s2, p, 0.333333333f);
    float at = fmaf(s * s2, p, s);
    float lm = at * 2.0f;
    float fe = (float)e;
    return fmaf(fe, 0.693145752f, fmaf(fe, 1.42860677e-06f, lm));
}

/* exp for |x| < 87: k = rint(x/ln2), r = x - k ln2 (two parts), Taylor to r^8, 2^k by
 * exponent construction. */
float oc_expf(float x)
{
    float k = rintf(x * 1.44269504f);
    float r = fmaf(-k, 0.693145752f, x);
    r = fmaf(-k, 1.42860677e-06f, r);
    float p = fmaf(r, 2.48015873e-05f, 1.98412698e-04f);
    p = fmaf(r, p, 1.38888889e-03f);
    p = fmaf(r, p, 8.33333333e-03f);
    p = fmaf(r, p, 4.16666667e-02f);
    p = fmaf(r, p, 1.66666667e-01f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    int ki = (int)k;
    return p * u2f((uint32_t)(ki + 127) << 23);
}

/* x^y for x > 0 */
float oc_powf(float x, float y) { return oc_expf(y * oc_logf(x)); }

/* erfinv, single precision: M. Giles, "Approximating the erfinv function",
 * GPU Computing Gems Jade Edition (2011), section 10.2 (two-branch polynomial in
 * w = -log(1-x^2)). */
float oc_erfinvf(float x)
{
    float w = -oc_logf((1.0f - x) * (1.0f + x));
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = fmaf(p, w, 3.43273939e-07f);
        p = fmaf(p, w, -3.5233877e-06f);
        p = fmaf(p, w, -4.39150654e-06f);
        p = fmaf(p, w, 0.00021858087f);
        p = fmaf(p, w, -0.00125372503f);
        p = fmaf(p, w, -0.00417768164f);
        p = fmaf(p, w, 0.246640727f);
        p = fmaf(p, w, 1.50140941f);
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = fmaf(p, w, 0.000100950558f);
        p = fmaf(p, w, 0.00134934322f);
        p = fmaf(p, w, -0.00367342844f);
        p = fmaf(p, w, 0.00573950773f);
        p = fmaf(p, w, -0.0076224613f);
        p = fmaf(p, w, 0.00943887047f);
        p = fmaf(p, w, 1.00167406f);
        p = fmaf(p, w, 2.83297682f);
    }
    return p * x;
}

/* standard normal deviate from one uniform: sqrt(2) erfinv(2u - 1) */
float oc_gauss_from_u(float u) { return oc_erfinvf(2.0f * u - 1.0f) * 1.41421356f; }

void oc_math_batch(int fn, const float *in, float *out, int n)
{
    if (fn == 10 || fn == 11) {
        /* pairs: division and atan2 of (in[2i], in[2i+1]) */
        for (int i = 0; i < n / 2; ++i)
            out[i] = fn == 10 ? in[2 * i] / in[2 * i + 1] : oc_atan2f(in[2 * i], in[2 * i + 1]);
        return;
    }
    for (int i = 0; i < n; ++i) {
        float s, c, x = in[i];
        switch (fn) {
        case 0: out[i] = oc_u01(f2u(x)); break;
        case 1: oc_sincos2pi(x, &s, &c); out[i] = s; break;
        case 2: oc_sincos2pi(x, &s, &c); out[i] = c; break;
        case 3: oc_sincos_small(x, &s, &c); out[i] = s; break;
        case 4: oc_sincos_small(x, &s, &c); out[i] = c; break;
        case 5: out[i] = oc_logf(x); break;
        case 6: out[i] = oc_expf(x); break;
        case 7: out[i] = oc_erfinvf(x); break;
        case 8: out[i] = oc_gauss_from_u(x); break;
        default: out[i] = sqrtf(x); break;
        }
    }
}

void oc_philox_batch(const uint32_t *ctr, const uint32_t key[2], uint32_t *out, int n)
{
    for (int i = 0; i < n; ++i) oc_philox4x32_10(ctr + 4 * i, key, out + 4 * i);
}

/* ------------------------------------------------------------------------------------ */
/* scene                                                                                 */
/* ------------------------------------------------------------------------------------ */
/* one grid (grid.cpp:3-82) over the heliostats [h0, h1) */
typedef struct {
    f3 pos, interval;
    int h0, h1;
    int gn[3];
    int64_t cells, nnz;
    int32_t *cell_start, *cell_items;
    int occ_lo[3], occ_hi[3];
} oc_grid;

struct oc_scene {
    oc_scene_desc d;
    oc_params p;
    int H;
    /* grids (every grid of the description; a heliostat's rays walk its own) */
    int ngrids;
    oc_grid *grids;
    int *helio_grid;          /* [H] index of the grid holding each heliostat */
    /* heliostats (heliostat.cpp:16-62) */
    f3 *pos, *size, *normal, *v0, *v1, *v2, *v3;
    float *gap;   /* [H][2] */
    int *matrix;  /* [H][2] */
    /* receiver (receiver.cpp:24-117) */
    f3 rv[4], recv_normal, focus;
    float recv_pixel;
    int geo, res_x, res_y;
    float srec;               /* pixel area */
    f3 cyl_base;              /* cylinder: base of the axis, radius^2, height */
    float cyl_r2, cyl_h;
    float cone_r0, cone_k, cone_inv_area; /* cone: bottom radius, dr/dy, 1 / (h (r0 + r1)) */
    /* sun (scene_instance_process.cu:6-11, :163-226) */
    f3 sun;
    float buie_k, buie_gamma, buie_int465, buie_A, buie_B, buie_C, buie_inv_gp1;
    int nbins;
    float *cdf_x, *cdf_y, *slope;
    /* microsamples (steps_for_raytracing.cu:67-158) */
    int32_t *mcount;
    int64_t *ray_offset;
    int64_t total_micro, total_rays;
    int max_micro;
    int poisson_cap;
    float *poisson_pts; /* [H][cap][2] */
    /* per-heliostat lattice constants */
    float *subrow_len, *subcol_len; /* subhelio_rowlength_collength x, y */
    int *sub_rows, *sub_cols;
    /* flux scale (Gtest.cu:29,37-46) */
    float scale, total_energy;
    int32_t *tag;             /* [H] or NULL: heliostats counted in oc_stats' tag_* fields */
    int literal_self;         /* 1: a walk tests the ray's own heliostat like any candidate */
};

static f3 ld3(const float *p) { return mk3(p[0], p[1], p[2]); }

/* SceneProcessor::set_samplelights host part (scene_instance_process.cu:116-156,
 * :223-260): Buie constants and trapezoid CDF with nbins bins on [0, 4.65] mrad. */
static void build_buie(oc_scene *s)
{
    int num_group = s->p.buie_bins > 0 ? s->p.buie_bins : 1024;
    float csr = s->p.csr;
    float upper_lm = 4.65f;
    float k = 0.9f * logf(13.5f * csr) * powf(csr, -0.3f);
    float gamma = 2.2f * logf(0.52f * csr) * powf(csr, 0.43f) - 0.1f;
    /* inter_larger_465_intensity (scene_instance_process.cu:122-126) */
    float i93 = expf(k) / (gamma + 1) * powf(9.3f, gamma + 1);
    float i465 = expf(k) / (gamma + 1) * powf(upper_lm, gamma + 1);
    float value_465_930 = i93 - i465;
    s->nbins = num_group;
    s->cdf_x = (float *)malloc(sizeof(float) * (num_group + 1));
    s->cdf_y = (float *)malloc(sizeof(float) * (num_group + 1));
    s->slope = (float *)malloc(sizeof(float) * num_group);
    /* parameters_generate (scene_instance_process.cu:133-156) */
    float length_interval = upper_lm / (float)num_group;
    float x = 0.0f;
    s->cdf_x[0] = 0.0f;
    s->cdf_y[0] = 0.0f;
    float hist_pre = cosf((float)(0.326 * x)) / cosf((float)(0.308 * x));
    for (int i = 1; i <= num_group; ++i) {
        x += length_interval;
        float hist_cur = cosf((float)(0.326 * x)) / cosf((float)(0.308 * x));
        float subarea = (hist_pre + hist_cur) * length_interval / 2.0f;
        s->cdf_x[i] = x;
        s->cdf_y[i] = subarea + s->cdf_y[i - 1];
        s->slope[i - 1] = length_interval / subarea;
        hist_pre = hist_cur;
    }
    float value_less_465 = s->cdf_y[num_group];
    s->buie_k = k;
    s->buie_gamma = gamma;
    s->buie_int465 = value_less_465;
    s->buie_A = value_less_465 + value_465_930;
    s->buie_B = (gamma + 1) / expf(k);
    s->buie_C = powf(4.65f, gamma + 1);
    s->buie_inv_gp1 = 1 / (gamma + 1);
}

/* samplelights::max_less_index (scene_instance_process.cu:158-173) */
static int max_less_index(const float *cdf_y, float value, int n)
{
    int left = 0, right = n, mid;
    while (left <= right) {
        mid = (left + right) >> 1;
        if (value > cdf_y[mid]) left = mid + 1;
        else if (value < cdf_y[mid]) right = mid - 1;
        else return mid;
    }
    return right;
}

/* RectGrid::Cinit + CGridHelioMatch + boxIntersect (grid.cpp:3-82) of one grid over its
 * heliostats [g->h0, g->h1), with global ids (grid.cpp:53).  Cell ranges are clamped to the
 * grid (out-of-grid binning is undefined behaviour in the reference). */
static int build_grid(oc_scene *s, oc_grid *g, f3 gpos, f3 gsize, f3 gint)
{
    g->pos = gpos;
    g->interval = gint;
    g->gn[0] = (int)ceilf(gsize.x / gint.x);
    g->gn[1] = (int)ceilf(gsize.y / gint.y);
    g->gn[2] = (int)ceilf(gsize.z / gint.z);
    if (g->gn[0] <= 0 || g->gn[1] <= 0 || g->gn[2] <= 0) return -1;
    g->cells = (int64_t)g->gn[0] * g->gn[1] * g->gn[2];
    int32_t *cnt = (int32_t *)calloc((size_t)g->cells, sizeof(int32_t));
    int nh = g->h1 - g->h0;
    int (*rng)[6] = malloc(sizeof(int[6]) * (size_t)(nh > 0 ? nh : 1));
    int64_t nnz = 0;
    for (int i = 0; i < 3; ++i) { g->occ_lo[i] = 1 << 30; g->occ_hi[i] = -1; }
    for (int i = g->h0; i < g->h1; ++i) {
        int *r = rng[i - g->h0];
        float diag = sqrtf(dot3(s->size[i], s->size[i]));
        float radius = sqrtf(diag) / 2;
        f3 mn = mk3(s->pos[i].x - radius, s->pos[i].y - radius, s->pos[i].z - radius);
        f3 mx = mk3(s->pos[i].x + radius, s->pos[i].y + radius, s->pos[i].z + radius);
        int lo[3] = { trunc_i(sub3(mn, gpos).x / gint.x), trunc_i(sub3(mn, gpos).y / gint.y),
                      trunc_i(sub3(mn, gpos).z / gint.z) };
        int hi[3] = { trunc_i(sub3(mx, gpos).x / gint.x), trunc_i(sub3(mx, gpos).y / gint.y),
                      trunc_i(sub3(mx, gpos).z / gint.z) };
        for (int a = 0; a < 3; ++a) {
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > g->gn[a] - 1) hi[a] = g->gn[a] - 1;
            r[a] = lo[a];
            r[3 + a] = hi[a];
        }
        for (int x = lo[0]; x <= hi[0]; ++x)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    int64_t c = (int64_t)x * g->gn[1] * g->gn[2] + (int64_t)y * g->gn[2] + z;
                    cnt[c]++;
                    nnz++;
                }
        if (lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2])
            for (int a = 0; a < 3; ++a) {
                if (lo[a] < g->occ_lo[a]) g->occ_lo[a] = lo[a];
                if (hi[a] > g->occ_hi[a]) g->occ_hi[a] = hi[a];
            }
    }
    g->nnz = nnz;
    g->cell_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(g->cells + 1));
    g->cell_items = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    g->cell_start[0] = 0;
    for (int64_t c = 0; c < g->cells; ++c) g->cell_start[c + 1] = g->cell_start[c] + cnt[c];
    memset(cnt, 0, sizeof(int32_t) * (size_t)g->cells);
    /* items in push order: heliostat index ascending within each cell */
    for (int i = g->h0; i < g->h1; ++i) {
        const int *r = rng[i - g->h0];
        for (int x = r[0]; x <= r[3]; ++x)
            for (int y = r[1]; y <= r[4]; ++y)
                for (int z = r[2]; z <= r[5]; ++z) {
                    int64_t c = (int64_t)x * g->gn[1] * g->gn[2] + (int64_t)y * g->gn[2] + z;
                    g->cell_items[g->cell_start[c] + cnt[c]++] = i;
                }
    }
    free(cnt);
    free(rng);
    return 0;
}

/* every grid of the description, or the one grid of grid_* holding every heliostat */
static int build_grids(oc_scene *s)
{
    const oc_scene_desc *d = &s->d;
    int n = d->n_grids > 0 ? d->n_grids : 1;
    s->ngrids = n;
    s->grids = (oc_grid *)calloc((size_t)n, sizeof(oc_grid));
    s->helio_grid = (int *)malloc(sizeof(int) * (size_t)(s->H + 1));
    int next = 0;
    for (int k = 0; k < n; ++k) {
        const float *gp = d->n_grids > 0 ? d->grids[k].pos : d->grid_pos;
        const float *gs = d->n_grids > 0 ? d->grids[k].size : d->grid_size;
        const float *gi = d->n_grids > 0 ? d->grids[k].interval : d->grid_interval;
        int first = d->n_grids > 0 ? d->grids[k].first_helio : 0;
        int cnt = d->n_grids > 0 ? d->grids[k].n_helio : s->H;
        if (first != next || cnt < 0 || first + cnt > s->H) return -6;
        s->grids[k].h0 = first;
        s->grids[k].h1 = first + cnt;
        for (int i = first; i < first + cnt; ++i) s->helio_grid[i] = k;
        next += cnt;
        int rc = build_grid(s, &s->grids[k], ld3(gp), ld3(gs), ld3(gi));
        if (rc) return rc;
    }
    return next == s->H ? 0 : -6;
}

/* Cylinder receiver (new; CylinderReceiver is an empty stub, receiver.cuh:103-122): axis
 * through recv_pos at mid-height, radius size.x, height size.y; the map unrolls the outer
 * surface (res_x columns over the azimuth, res_y rows over the height), by default from
 * geo = int(1 / pixel length).  Heliostats aim at the axis point. */
static int build_cylinder(oc_scene *s, int geo)
{
    float r = s->d.recv_size[0], h = s->d.recv_size[1];
    if (!(r > 0.0f && h > 0.0f)) return -5;
    int rx = s->d.recv_res_x, ry = s->d.recv_res_y;
    if (rx <= 0 || ry <= 0) {
        rx = (int)(6.28318531f * r * (float)geo);
        ry = (int)(h * (float)geo);
    }
    if (rx <= 0 || ry <= 0) return -4;
    s->res_x = rx;
    s->res_y = ry;
    s->cyl_base = mk3(s->d.recv_pos[0], s->d.recv_pos[1] - h / 2, s->d.recv_pos[2]);
    s->cyl_r2 = r * r;
    s->cyl_h = h;
    s->recv_pixel = h / (float)ry;
    s->srec = (6.28318531f * r / (float)rx) * (h / (float)ry);
    s->focus = ld3(s->d.recv_pos);
    return 0;
}

/* Truncated-cone receiver (the project's; the reference's is an empty stub,
 * receiver.cuh:125-140): axis through recv_pos at mid-height, top radius size.x, bottom
 * radius size.y, height size.z (the order of topradius_bottomradius_hight_); the map unrolls
 * the lateral surface, res_x columns over the azimuth, res_y rows of equal area, by default
 * the mean circumference and the slant height in pixels. */
static int build_cone(oc_scene *s, int geo)
{
    float r1 = s->d.recv_size[0], r0 = s->d.recv_size[1], h = s->d.recv_size[2];
    if (!(r0 >= 0.0f && r1 >= 0.0f && r0 + r1 > 0.0f && h > 0.0f)) return -5;
    float slant = sqrtf(h * h + (r1 - r0) * (r1 - r0));
    int rx = s->d.recv_res_x, ry = s->d.recv_res_y;
    if (rx <= 0 || ry <= 0) {
        rx = (int)(3.14159265f * (r0 + r1) * (float)geo);
        ry = (int)(slant * (float)geo);
    }
    if (rx <= 0 || ry <= 0) return -4;
    s->res_x = rx;
    s->res_y = ry;
    s->cyl_base = mk3(s->d.recv_pos[0], s->d.recv_pos[1] - h / 2, s->d.recv_pos[2]);
    s->cyl_h = h;
    s->cone_r0 = r0;
    s->cone_k = (r1 - r0) / h;
    s->cone_inv_area = 1.0f / (h * (r0 + r1));
    s->recv_pixel = slant / (float)ry;
    s->srec = 3.14159265f * (r0 + r1) * slant / ((float)rx * (float)ry);
    s->focus = ld3(s->d.recv_pos);
    return 0;
}

/* RectangleReceiver::CInit (receiver.cpp:24-117) */
static int build_receiver(oc_scene *s)
{
    if (s->d.recv_type < 0 || s->d.recv_type > 2) return -2;
    int geo = (int)(1.0f / s->d.recv_pixel_length);
    s->geo = geo;
    if (s->d.recv_type == 1) return build_cylinder(s, geo);
    if (s->d.recv_type == 2) return build_cone(s, geo);
    s->recv_pixel = 1.0f / (float)geo;
    s->srec = s->recv_pixel * s->recv_pixel;
    f3 size = ld3(s->d.recv_size), pos = ld3(s->d.recv_pos);
    f3 ln, v[4];
    switch (s->d.recv_face) {
    case 0: ln = mk3(0.0f, 0.0f, 1.0f);
        v[0] = mk3(-size.x / 2, -size.y / 2, size.z / 2);
        v[1] = mk3(-size.x / 2, size.y / 2, size.z / 2);
        v[2] = mk3(size.x / 2, size.y / 2, size.z / 2);
        v[3] = mk3(size.x / 2, -size.y / 2, size.z / 2);
        break;
    case 1: ln = mk3(1.0f, 0.0f, 0.0f);
        v[0] = mk3(size.x / 2, -size.y / 2, size.z / 2);
        v[1] = mk3(size.x / 2, size.y / 2, size.z / 2);
        v[2] = mk3(size.x / 2, size.y / 2, -size.z / 2);
        v[3] = mk3(size.x / 2, -size.y / 2, -size.z / 2);
        break;
    case 2: ln = mk3(0.0f, 0.0f, -1.0f);
        v[0] = mk3(size.x / 2, -size.y / 2, -size.z / 2);
        v[1] = mk3(size.x / 2, size.y / 2, -size.z / 2);
        v[2] = mk3(-size.x / 2, size.y / 2, -size.z / 2);
        v[3] = mk3(-size.x / 2, -size.y / 2, -size.z / 2);
        break;
    case 3: ln = mk3(-1.0f, 0.0f, 0.0f);
        v[0] = mk3(-size.x / 2, -size.y / 2, -size.z / 2);
        v[1] = mk3(-size.x / 2, size.y / 2, -size.z / 2);
        v[2] = mk3(-size.x / 2, size.y / 2, size.z / 2);
        v[3] = mk3(-size.x / 2, -size.y / 2, size.z / 2);
        break;
    default: return -3;
    }
    f3 n = normalize3(ld3(s->d.recv_normal));
    s->recv_normal = n;
    for (int i = 0; i < 4; ++i) s->rv[i] = add3(rotate_y3(v[i], ln, n), pos);
    s->focus = scl3(add3(s->rv[0], s->rv[2]), 1.0f / 2);
    s->res_x = (int)(size.x * (float)geo);
    s->res_y = (int)(size.y * (float)geo);
    if (s->res_x <= 0 || s->res_y <= 0) return -4;
    return 0;
}

/* RectangleHelio::CRotate(focus, sun) (heliostat.cpp:16-62) */
/* RectangleHelio::CRotate(focus, sun) (heliostat.cpp:16-62), or CRotate(dir) on the given
 * normals of InitContent(normals) (heliostat.cpp:22-27,73-76) */
static void build_heliostats(oc_scene *s)
{
    for (int i = 0; i < s->H; ++i) {
        f3 n;
        if (s->d.helio_normal) {
            n = normalize3(ld3(s->d.helio_normal + 3 * i));
        } else {
            f3 rd = normalize3(sub3(s->focus, s->pos[i]));
            n = normalize3(sub3(rd, s->sun));
        }
        s->normal[i] = n;
        f3 sz = s->size[i];
        f3 l0 = mk3(-sz.x / 2, sz.y / 2, -sz.z / 2);
        f3 l1 = mk3(l0.x, l0.y, l0.z + sz.z);
        f3 l2 = mk3(l1.x + sz.x, l1.y, l1.z);
        f3 l3 = mk3(l0.x + sz.x, l0.y, l0.z);
        s->v0[i] = add3(local2world3(l0, n), s->pos[i]);
        s->v1[i] = add3(local2world3(l1, n), s->pos[i]);
        s->v2[i] = add3(local2world3(l2, n), s->pos[i]);
        s->v3[i] = add3(local2world3(l3, n), s->pos[i]);
    }
}

/* Poisson-disk dart throwing of poisson_sample (poisson_disk.cu:4-201), with the phase
 * snapshot of cudaMemcpyToArray (:226) kept explicitly, followed by the stable
 * compaction of generate_microhelio_origin / valid_element (compact_sample.cu:6-19,
 * compact_sample.cuh:13-22).  Returns the number of accepted points written to pts. */
static int poisson_one(const oc_scene *s, int h, float width, float height, float r, int k,
                       float *pts, int cap, float *grid_out)
{
    float cell = r / sqrtf(2.0f);
    int width_int = (int)(ceilf(width / cell) + 2);
    int height_int = (int)(ceilf(height / cell) + 2);
    int phase_length = (int)(ceilf(sqrtf(2.0f)) + 1);
    int pw = (width_int % phase_length != 0) ? width_int / phase_length + 1 : width_int / phase_length;
    int ph = (height_int % phase_length != 0) ? height_int / phase_length + 1 : height_int / phase_length;
    int ncell = width_int * height_int;
    float *g = (float *)malloc(sizeof(float) * 2 * ncell);
    float *snap = (float *)malloc(sizeof(float) * 2 * ncell);
    for (int i = 0; i < ncell; ++i) { g[2 * i] = -2 * cell; g[2 * i + 1] = -2 * cell; }
    uint32_t key[2] = { (uint32_t)s->p.seed, (uint32_t)(s->p.seed >> 32) };
    uint32_t w3 = (s->p.scene_index << 4) | OC_PURPOSE_POISSON;
    for (int phase = 0; phase < phase_length * phase_length; ++phase) {
        int xo = phase % phase_length, yo = phase / phase_length;
        memcpy(snap, g, sizeof(float) * 2 * ncell);
        for (int ty = 0; ty < ph; ++ty)
            for (int tx = 0; tx < pw; ++tx) {
                int my_indx = ty * width_int / phase_length + tx;
                int x = tx * phase_length + xo;
                int y = ty * phase_length + yo;
                if (x > width_int - 1 || y > height_int - 1) continue;
                if (phase == 0) {
                    uint32_t c[4] = { (uint32_t)my_indx, 0u, (uint32_t)h, w3 }, o[4];
                    oc_philox4x32_10(c, key, o);
                    g[2 * (y * width_int + x)] = (oc_u01(o[0]) + (float)x - 1) * cell;
                    g[2 * (y * width_int + x) + 1] = (oc_u01(o[1]) + (float)y - 1) * cell;
                    continue;
                }
                for (int i = 0; i < k; ++i) {
                    uint32_t c[4] = { (uint32_t)(my_indx * k + i), (uint32_t)phase, (uint32_t)h, w3 }, o[4];
                    oc_philox4x32_10(c, key, o);
                    float px = (oc_u01(o[0]) + (float)x - 1) * cell;
                    float py = (oc_u01(o[1]) + (float)y - 1) * cell;
                    int c0 = trunc_i((px - r) / cell + 1), c1 = trunc_i((px + r) / cell + 1);
                    int r0 = trunc_i((py - r) / cell + 1), r1 = trunc_i((py + r) / cell + 1);
                    int cannot = 0;
                    for (int rr = r0; rr <= r1; ++rr)
                        for (int cc = c0; cc <= c1; ++cc) {
                            if (cc < 0 || cc > width_int - 1 || rr < 0 || rr > height_int - 1) continue;
                            float vx = snap[2 * (rr * width_int + cc)];
                            float vy = snap[2 * (rr * width_int + cc) + 1];
                            if (vx < -cell) continue;
                            float dx = vx - px, dy = vy - py;
                            if (sqrtf(dx * dx + dy * dy) < r) cannot = 1;
                        }
                    if (!cannot) {
                        g[2 * (y * width_int + x)] = px;
                        g[2 * (y * width_int + x) + 1] = py;
                        break;
                    }
                }
            }
    }
    if (grid_out) memcpy(grid_out, g, sizeof(float) * 2 * ncell);
    int n = 0;
    for (int i = 0; i < ncell; ++i) {
        float x = g[2 * i], y = g[2 * i + 1];
        if ((x > 0) && (y > 0) && (x < width) && (y < height)) {
            if (n < cap) { pts[2 * n] = x; pts[2 * n + 1] = y; }
            ++n;
        }
    }
    free(g);
    free(snap);
    return n;
}

/* microsample counts (steps_for_raytracing.cu:67-82, :105-125) */
static int build_micro(oc_scene *s)
{
    int H = s->H;
    float pitch = s->p.helio_pixel_length;
    s->mcount = (int32_t *)malloc(sizeof(int32_t) * H);
    s->ray_offset = (int64_t *)malloc(sizeof(int64_t) * (H + 1));
    s->subrow_len = (float *)malloc(sizeof(float) * H);
    s->subcol_len = (float *)malloc(sizeof(float) * H);
    s->sub_rows = (int *)malloc(sizeof(int) * H);
    s->sub_cols = (int *)malloc(sizeof(int) * H);
    s->poisson_cap = 0;
    s->poisson_pts = NULL;
    for (int i = 0; i < H; ++i) {
        int rows = s->matrix[2 * i], cols = s->matrix[2 * i + 1];
        float gx = s->gap[2 * i], gy = s->gap[2 * i + 1];
        s->subrow_len[i] = (s->size[i].z - gy * (float)(rows - 1)) / (float)rows;
        s->subcol_len[i] = (s->size[i].x - gx * (float)(cols - 1)) / (float)cols;
        s->sub_rows[i] = trunc_i(s->subrow_len[i] / pitch);
        s->sub_cols[i] = trunc_i(s->subcol_len[i] / pitch);
    }
    if (s->p.micro_mode == OC_MICRO_POISSON) {
        /* capacity = padded cell count of the largest heliostat */
        int cap = 0;
        float r = pitch / sqrtf(1.5f);
        float cell = r / sqrtf(2.0f);
        for (int i = 0; i < H; ++i) {
            float width = s->subcol_len[i] * (float)s->matrix[2 * i + 1];
            float height = s->subrow_len[i] * (float)s->matrix[2 * i];
            int wi = (int)(ceilf(width / cell) + 2), hi = (int)(ceilf(height / cell) + 2);
            if (wi * hi > cap) cap = wi * hi;
        }
        s->poisson_cap = cap;
        s->poisson_pts = (float *)malloc(sizeof(float) * 2 * (size_t)cap * H);
        #pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < H; ++i) {
            float width = s->subcol_len[i] * (float)s->matrix[2 * i + 1];
            float height = s->subrow_len[i] * (float)s->matrix[2 * i];
            s->mcount[i] = poisson_one(s, i, width, height, r, s->p.poisson_k,
                                       s->poisson_pts + 2 * (size_t)cap * i, cap, NULL);
        }
    } else {
        for (int i = 0; i < H; ++i)
            s->mcount[i] = s->sub_rows[i] * s->sub_cols[i] * s->matrix[2 * i] * s->matrix[2 * i + 1];
    }
    s->ray_offset[0] = 0;
    s->total_micro = 0;
    s->max_micro = 0;
    for (int i = 0; i < H; ++i) {
        s->ray_offset[i + 1] = s->ray_offset[i] + (int64_t)s->mcount[i] * s->p.rays_per_sample;
        s->total_micro += s->mcount[i];
        if (s->mcount[i] > s->max_micro) s->max_micro = s->mcount[i];
    }
    s->total_rays = s->ray_offset[H];
    return 0;
}

/* local microsample centre (steps_for_raytracing.cu:4-38) */
static f3 micro_local(const oc_scene *s, int h, int m)
{
    f3 sz = s->size[h];
    float gx = s->gap[2 * h], gy = s->gap[2 * h + 1];
    if (s->p.micro_mode == OC_MICRO_POISSON) {
        const float *pt = s->poisson_pts + 2 * ((size_t)s->poisson_cap * h + m);
        int block_row = trunc_i(pt[1] / s->subrow_len[h]);
        int block_col = trunc_i(pt[0] / s->subcol_len[h]);
        return mk3(pt[0] + (float)block_col * gx - sz.x / 2, sz.y / 2,
                   pt[1] + (float)block_row * gy - sz.z / 2);
    }
    float p = s->p.helio_pixel_length;
    int cols_all = s->matrix[2 * h + 1] * s->sub_cols[h];
    int row = m / cols_all, col = m % cols_all;
    int block_row = row / s->sub_rows[h], block_col = col / s->sub_cols[h];
    return mk3((float)col * p + (float)block_col * gx + p / 2 - sz.x / 2, sz.y / 2,
               (float)row * p + (float)block_row * gy + p / 2 - sz.z / 2);
}

static void flux_scale(oc_scene *s)
{
    /* results_with_given_normal (Gtest.cu:14,29,37-46) */
    float Id = s->d.dni;
    float total = 0.0f;
    for (int i = 0; i < s->H; ++i)
        total += fabsf(dot3(s->normal[i], s->sun) * Id * s->size[i].x * s->size[i].z);
    float Ssub = s->p.helio_pixel_length * s->p.helio_pixel_length;
    float rou = s->p.reflectivity;
    int Nc = s->p.rays_per_sample;
    float Srec = s->srec;
    total *= rou / Srec;
    s->total_energy = total;
    s->scale = Id * Ssub * rou / (float)Nc / Srec;
}

int oc_scene_build(const oc_scene_desc *desc, const oc_params *prm, oc_scene **out)
{
    *out = NULL;
    if (!desc || !prm || desc->n_helio < 0 || prm->rays_per_sample <= 0) return -1;
    oc_scene *s = (oc_scene *)calloc(1, sizeof(oc_scene));
    s->d = *desc;
    s->p = *prm;
    int H = desc->n_helio;
    s->H = H;
    s->pos = malloc(sizeof(f3) * (H + 1));
    s->size = malloc(sizeof(f3) * (H + 1));
    s->normal = malloc(sizeof(f3) * (H + 1));
    s->v0 = malloc(sizeof(f3) * (H + 1));
    s->v1 = malloc(sizeof(f3) * (H + 1));
    s->v2 = malloc(sizeof(f3) * (H + 1));
    s->v3 = malloc(sizeof(f3) * (H + 1));
    s->gap = malloc(sizeof(float) * 2 * (H + 1));
    s->matrix = malloc(sizeof(int) * 2 * (H + 1));
    for (int i = 0; i < H; ++i) {
        s->pos[i] = ld3(desc->helio_pos + 3 * i);
        s->size[i] = ld3(desc->helio_size + 3 * i);
        s->gap[2 * i] = desc->helio_gap[2 * i];
        s->gap[2 * i + 1] = desc->helio_gap[2 * i + 1];
        s->matrix[2 * i] = desc->helio_matrix[2 * i];
        s->matrix[2 * i + 1] = desc->helio_matrix[2 * i + 1];
    }
    /* InitContent order (solar_scene.cpp:83-99): sun, grid, receiver, heliostats */
    s->sun = normalize3(ld3(desc->sun_dir));
    build_buie(s);
    int rc = build_grids(s);
    if (!rc) rc = build_receiver(s);
    if (rc) { oc_scene_free(s); return rc; }
    build_heliostats(s);
    build_micro(s);
    flux_scale(s);
    *out = s;
    return 0;
}

void oc_scene_free(oc_scene *s)
{
    if (!s) return;
    free(s->pos); free(s->size); free(s->normal);
    free(s->v0); free(s->v1); free(s->v2); free(s->v3);
    free(s->gap); free(s->matrix);
    for (int k = 0; k < s->ngrids && s->grids; ++k) {
        free(s->grids[k].cell_start);
        free(s->grids[k].cell_items);
    }
    free(s->grids); free(s->helio_grid);
    free(s->cdf_x); free(s->cdf_y); free(s->slope);
    free(s->mcount); free(s->ray_offset); free(s->poisson_pts);
    free(s->subrow_len); free(s->subcol_len); free(s->sub_rows); free(s->sub_cols);
    free(s->tag);
    free(s);
}

int oc_scene_set_tags(oc_scene *s, const int32_t *tags)
{
    free(s->tag);
    s->tag = NULL;
    if (!tags) return 0;
    s->tag = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->H > 0 ? s->H : 1));
    if (!s->tag) return -1;
    memcpy(s->tag, tags, sizeof(int32_t) * (size_t)s->H);
    return 0;
}

int oc_scene_counts(const oc_scene *s, int64_t out[8])
{
    int64_t cells = 0, nnz = 0;
    for (int k = 0; k < s->ngrids; ++k) { cells += s->grids[k].cells; nnz += s->grids[k].nnz; }
    out[0] = s->H; out[1] = cells; out[2] = nnz; out[3] = s->res_x; out[4] = s->res_y;
    out[5] = s->total_micro; out[6] = s->total_rays; out[7] = s->max_micro;
    return 0;
}

void oc_scene_helio(const oc_scene *s, float *normal, float *v0, float *v1, float *v3)
{
    for (int i = 0; i < s->H; ++i) {
        if (normal) memcpy(normal + 3 * i, &s->normal[i], 12);
        if (v0) memcpy(v0 + 3 * i, &s->v0[i], 12);
        if (v1) memcpy(v1 + 3 * i, &s->v1[i], 12);
        if (v3) memcpy(v3 + 3 * i, &s->v3[i], 12);
    }
}

int oc_scene_grid_k(const oc_scene *s, int k, int32_t *cell_start, int32_t *cell_items, int32_t grid_num[3],
                    int64_t *nnz)
{
    if (k < 0 || k >= s->ngrids) return -1;
    const oc_grid *g = &s->grids[k];
    if (cell_start) memcpy(cell_start, g->cell_start, sizeof(int32_t) * (size_t)(g->cells + 1));
    if (cell_items) memcpy(cell_items, g->cell_items, sizeof(int32_t) * (size_t)g->nnz);
    if (grid_num) { grid_num[0] = g->gn[0]; grid_num[1] = g->gn[1]; grid_num[2] = g->gn[2]; }
    if (nnz) *nnz = g->nnz;
    return 0;
}

void oc_scene_grid(const oc_scene *s, int32_t *cell_start, int32_t *cell_items, int32_t grid_num[3])
{
    oc_scene_grid_k(s, 0, cell_start, cell_items, grid_num, NULL);
}

void oc_scene_receiver(const oc_scene *s, float *v4x3, float focus[3], float *pixel_length)
{
    if (v4x3) memcpy(v4x3, s->rv, 48);
    if (focus) memcpy(focus, &s->focus, 12);
    if (pixel_length) *pixel_length = s->recv_pixel;
}

void oc_scene_sun(const oc_scene *s, float sun_dir[3], float *c)
{
    if (sun_dir) memcpy(sun_dir, &s->sun, 12);
    if (c) {
        c[0] = s->buie_int465; c[1] = s->buie_A; c[2] = s->buie_B;
        c[3] = s->buie_C; c[4] = s->buie_inv_gp1; c[5] = s->buie_gamma;
    }
}

int oc_scene_buie_table(const oc_scene *s, float *cdf_x, float *cdf_y, float *slope)
{
    if (cdf_x) memcpy(cdf_x, s->cdf_x, sizeof(float) * (s->nbins + 1));
    if (cdf_y) memcpy(cdf_y, s->cdf_y, sizeof(float) * (s->nbins + 1));
    if (slope) memcpy(slope, s->slope, sizeof(float) * s->nbins);
    return s->nbins;
}

void oc_scene_micro(const oc_scene *s, int32_t *count, int64_t *ray_offset)
{
    if (count) memcpy(count, s->mcount, sizeof(int32_t) * s->H);
    if (ray_offset) memcpy(ray_offset, s->ray_offset, sizeof(int64_t) * (s->H + 1));
}

int oc_scene_poisson_grid(const oc_scene *s, int h, float *grid, int cap, int dims[2], float rc[2])
{
    if (h < 0 || h >= s->H) return -1;
    float pitch = s->p.helio_pixel_length;
    float r = pitch / sqrtf(1.5f);
    float cell = r / sqrtf(2.0f);
    float width = s->subcol_len[h] * (float)s->matrix[2 * h + 1];
    float height = s->subrow_len[h] * (float)s->matrix[2 * h];
    int wi = (int)(ceilf(width / cell) + 2), hi = (int)(ceilf(height / cell) + 2);
    dims[0] = wi;
    dims[1] = hi;
    rc[0] = r;
    rc[1] = cell;
    if (!grid || cap < wi * hi) return wi * hi;
    float *pts = (float *)malloc(sizeof(float) * 2 * (size_t)wi * hi);
    int n = poisson_one(s, h, width, height, r, s->p.poisson_k, pts, wi * hi, grid);
    free(pts);
    return n;
}

int oc_scene_poisson(const oc_scene *s, int h, float *pts, int cap)
{
    if (s->p.micro_mode != OC_MICRO_POISSON || h < 0 || h >= s->H) return -1;
    int n = s->mcount[h] < cap ? s->mcount[h] : cap;
    memcpy(pts, s->poisson_pts + 2 * (size_t)s->poisson_cap * h, sizeof(float) * 2 * n);
    return s->mcount[h];
}

void oc_scene_flux_scale(const oc_scene *s, float *scale, float *total_energy)
{
    if (scale) *scale = s->scale;
    if (total_energy) *total_energy = s->total_energy;
}

/* ------------------------------------------------------------------------------------ */
/* per-ray sampling                                                                      */
/* ------------------------------------------------------------------------------------ */

/* sunshape polar angle in rad from one uniform.  Buie: samplelights::linear_interpolate
 * (scene_instance_process.cu:175-197) then theta/1000 (:205).  Pillbox (no reference
 * code): uniform over the solar disk, theta = theta_max sqrt(u). */
static float sun_theta(const oc_scene *s, float u)
{
    if (s->p.sunshape == OC_SUN_PILLBOX)
        return s->p.pillbox_half_angle * sqrtf(u);
    float v = u * s->buie_A;
    float th;
    if (v < s->buie_int465) {
        int id = max_less_index(s->cdf_y, v, s->nbins);
        th = (v - s->cdf_y[id]) * s->slope[id] + s->cdf_x[id];
    } else {
        th = oc_powf((v - s->buie_int465) * s->buie_B + s->buie_C, s->buie_inv_gp1);
    }
    return th / 1000.0f;
}

/* draws for ray (h, m, l): direction towards the sun (map_tracing steps at
 * rectgrid_raytracing.cu:171-174) and the perturbed mirror normal (:182-186, with the
 * slope error of map_turbulance, scene_instance_process.cu:14-23). */
static void sample_ray(const oc_scene *s, int h, int m, int l, f3 *to_sun, f3 *nrm)
{
    uint32_t key[2] = { (uint32_t)s->p.seed, (uint32_t)(s->p.seed >> 32) };
    uint32_t c[4] = { (uint32_t)l, (uint32_t)m, (uint32_t)h, (s->p.scene_index << 4) | OC_PURPOSE_RAY };
    uint32_t o[4];
    oc_philox4x32_10(c, key, o);
    float th = sun_theta(s, oc_u01(o[0]));
    float st, ct, sp, cp;
    oc_sincos_small(th, &st, &ct);
    oc_sincos2pi(oc_u01(o[1]), &sp, &cp);
    f3 smp = mk3(st * cp, ct, st * sp);            /* angle2xyz (global_function.cuh:175-180) */
    *to_sun = neg3(local2world3(smp, s->sun));
    float ths = oc_gauss_from_u(oc_u01(o[2])) * s->p.slope_sigma;
    float sts, cts, sps, cps;
    oc_sincos_small(ths, &sts, &cts);
    oc_sincos2pi(oc_u01(o[3]), &sps, &cps);
    f3 turb = mk3(sts * cps, cts, sts * sps);
    *nrm = normalize3(local2world3(turb, s->normal[h]));
}

void oc_sample_ray(const oc_scene *s, int h, int m, int l, float sun_to[3], float normal[3])
{
    f3 a, b;
    sample_ray(s, h, m, l, &a, &b);
    memcpy(sun_to, &a, 12);
    memcpy(normal, &b, 12);
}

/* ------------------------------------------------------------------------------------ */
/* 3D-DDA                                                                                */
/* ------------------------------------------------------------------------------------ */
typedef struct { int64_t steps, steps_eff, tests, self_skips, self_hits; } dda_ctr;

/* calTMax (rectgrid_raytracing.cu:3-9) */
static inline float cal_tmax(float dir, float interval, int idx, float cur)
{
    if (dir >= 0) return (float)(idx + 1) * interval - cur;
    return cur - (float)idx * interval;
}
/* absDivide (rectgrid_raytracing.cu:11-17) */
static inline float abs_divide(float den, float num)
{
    if (num <= OC_EPS && num >= -OC_EPS) return (float)2147483647;
    return fabsf(den / num);
}

/* Intersect (rectgrid_raytracing.cu:21-36): first candidate of the cell's CSR list hit by
 * the ray.  By default the ray's own heliostat is skipped (exact-arithmetic reading of the
 * t >= 1e-6 guard of rayParallelogramIntersect, global_function.cuh:102: the origin lies on
 * the mirror plane, so t = 0); with literal_self it is tested like any candidate, as the
 * reference's loop does, and a pass is counted in self_hits. */
static int cell_hit(const oc_scene *s, const oc_grid *g, f3 o, f3 d, int64_t addr, int self, dda_ctr *k)
{
    for (int32_t i = g->cell_start[addr]; i < g->cell_start[addr + 1]; ++i) {
        int h = g->cell_items[i];
        if (h == self && !s->literal_self) { k->self_skips++; continue; }
        k->tests++;
        float t, u, v;
        if (ray_parallelogram3(o, d, s->v0[h], s->v1[h], s->v3[h], &t, &u, &v)) {
            if (h == self) k->self_hits++;
            return h;
        }
    }
    return -1;
}

int oc_scene_set_ray_stream(oc_scene *s, uint32_t scene_index)
{
    s->p.scene_index = scene_index;
    return 0;
}

int oc_scene_set_literal_self(oc_scene *s, int on)
{
    s->literal_self = on ? 1 : 0;
    return 0;
}

/* NotCollision (rectgrid_raytracing.cu:38-117).  Returns the blocking heliostat, or -1
 * when the ray leaves the grid.  With early_exit, the walk stops as soon as the cell index
 * on some axis has passed the occupied index range in the direction of travel: every later
 * cell is then empty, so the result is unchanged. */
static const oc_grid *grid_of(const oc_scene *s, int h)
{
    return &s->grids[(h >= 0 && h < s->H) ? s->helio_grid[h] : 0];
}

static int dda(const oc_scene *s, f3 o, f3 d, int self, int early_exit, dda_ctr *k,
               int32_t *trace, int cap, int *ntrace)
{
    /* the grid of the ray's heliostat (grid 0 for a probe without one) */
    const oc_grid *G = grid_of(s, self);
    f3 gpos = G->pos, gint = G->interval;
    f3 rel = sub3(o, gpos);
    int px = trunc_i(rel.x / gint.x), py = trunc_i(rel.y / gint.y), pz = trunc_i(rel.z / gint.z);
    int sx = (d.x >= 0) ? 1 : -1, sy = (d.y >= 0) ? 1 : -1, sz = (d.z >= 0) ? 1 : -1;
    float tx = cal_tmax(d.x, gint.x, px, o.x - gpos.x);
    float ty = cal_tmax(d.y, gint.y, py, o.y - gpos.y);
    float tz = cal_tmax(d.z, gint.z, pz, o.z - gpos.z);
    float mx = abs_divide(tx, d.x), my = abs_divide(ty, d.y), mz = abs_divide(tz, d.z);
    float dx = abs_divide(gint.x, d.x), dy = abs_divide(gint.y, d.y), dz = abs_divide(gint.z, d.z);
    int gx = px, gy = py, gz = pz;
    int eff_done = 0;
    int n = 0;
    for (;;) {
        int axis;
        if (mx < my) axis = (mx < mz) ? 0 : 2;
        else axis = (my < mz) ? 1 : 2;
        if (axis == 0) {
            gx += sx;
            if (gx >= G->gn[0] || gx < 0) return -1;
            mx += dx;
        } else if (axis == 1) {
            gy += sy;
            if (gy >= G->gn[1] || gy < 0) return -1;
            my += dy;
        } else {
            gz += sz;
            if (gz >= G->gn[2] || gz < 0) return -1;
            mz += dz;
        }
        k->steps++;
        int past = (sx > 0 ? gx > G->occ_hi[0] : gx < G->occ_lo[0]) ||
                   (sy > 0 ? gy > G->occ_hi[1] : gy < G->occ_lo[1]) ||
                   (sz > 0 ? gz > G->occ_hi[2] : gz < G->occ_lo[2]);
        if (past) {
            if (!eff_done) { eff_done = 1; k->steps_eff++; }
            if (early_exit) return -1;
        } else if (!eff_done) {
            k->steps_eff++;
        }
        int64_t addr = (int64_t)gx * G->gn[1] * G->gn[2] + (int64_t)gy * G->gn[2] + gz;
        if (trace && n < cap) trace[n] = (int32_t)addr;
        n++;
        if (ntrace) *ntrace = n;
        int h = cell_hit(s, G, o, d, addr, self, k);
        if (h >= 0) return h;
    }
}

int oc_dda_probe(const oc_scene *s, const float orig[3], const float dir[3], int self,
                 int32_t *cells, int cap, int *hit_helio)
{
    dda_ctr k = {0, 0, 0, 0, 0};
    int n = 0;
    int h = dda(s, ld3(orig), ld3(dir), self, 0, &k, cells, cap, &n);
    if (hit_helio) *hit_helio = h;
    return n;
}

/* ------------------------------------------------------------------------------------ */
/* field trace                                                                           */
/* ------------------------------------------------------------------------------------ */

/* receiver_drawing (rectgrid_raytracing.cu:135-153); the row/col of a hit at u or v == 1
 * is clamped to the last pixel (the reference writes past the image there). */
static int receiver_pixel(const oc_scene *s, f3 o, f3 d)
{
    if (s->d.recv_type == 1)
        return cylinder_pixel3(o, d, s->cyl_base, s->cyl_r2, s->cyl_h, s->res_x, s->res_y);
    if (s->d.recv_type == 2)
        return cone_pixel3(o, d, s->cyl_base, s->cone_r0, s->cone_k, s->cyl_h, s->cone_inv_area, s->res_x,
                           s->res_y);
    float t, u, v;
    if (!ray_parallelogram3(o, d, s->rv[0], s->rv[1], s->rv[3], &t, &u, &v)) return -1;
    int row = trunc_i(u * (float)s->res_y);
    int col = trunc_i(v * (float)s->res_x);
    if (row > s->res_y - 1) row = s->res_y - 1;
    if (col > s->res_x - 1) col = s->res_x - 1;
    return row * s->res_x + col;
}

int oc_trace_f32(const oc_scene *s, int h0, int h1,
                 uint32_t *counts, uint64_t *energy,
                 int32_t *rec_shade, int32_t *rec_block, int32_t *rec_pixel,
                 int early_exit, int nthreads, oc_stats *stats)
{
    if (h0 < 0) h0 = 0;
    if (h1 > s->H) h1 = s->H;
    int npx = s->res_x * s->res_y;
    int L = s->p.rays_per_sample;
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    oc_stats tot;
    memset(&tot, 0, sizeof(tot));
    #pragma omp parallel num_threads(nt)
    {
        uint32_t *lc = (uint32_t *)calloc(npx, sizeof(uint32_t));
        uint64_t *le = (uint64_t *)calloc(npx, sizeof(uint64_t));
        oc_stats st;
        memset(&st, 0, sizeof(st));
        dda_ctr k = {0, 0, 0, 0, 0}, kt = {0, 0, 0, 0, 0};
        #pragma omp for schedule(dynamic, 1)
        for (int h = h0; h < h1; ++h) {
            f3 n = s->normal[h];
            const int tagged = s->tag && s->tag[h];
            if (tagged) st.tag_rays += (int64_t)s->mcount[h] * L;
            for (int m = 0; m < s->mcount[h]; ++m) {
                f3 orig = add3(local2world3(micro_local(s, h, m), n), s->pos[h]);
                for (int l = 0; l < L; ++l) {
                    int64_t rid = s->ray_offset[h] - s->ray_offset[h0] + (int64_t)m * L + l;
                    st.rays++;
                    f3 to_sun, nrm;
                    sample_ray(s, h, m, l, &to_sun, &nrm);
                    int sh = dda(s, orig, to_sun, h, early_exit, tagged ? &kt : &k, NULL, 0, NULL);
                    if (rec_shade) rec_shade[rid] = sh;
                    if (sh >= 0) {
                        st.shaded++;
                        if (rec_block) rec_block[rid] = -2;
                        if (rec_pixel) rec_pixel[rid] = -1;
                        continue;
                    }
                    f3 refl = normalize3(reflect3(neg3(to_sun), nrm));
                    int bl = dda(s, orig, refl, h, early_exit, &k, NULL, 0, NULL);
                    if (rec_block) rec_block[rid] = bl;
                    if (bl >= 0) {
                        st.blocked++;
                        if (rec_pixel) rec_pixel[rid] = -1;
                        continue;
                    }
                    int px = receiver_pixel(s, orig, refl);
                    if (rec_pixel) rec_pixel[rid] = px;
                    if (px < 0) continue;
                    st.hit++;
                    float e = fabsf(dot3(s->sun, nrm)) * 1.0f;   /* calEnergy, eta = 1 */
                    lc[px] += 1u;
                    le[px] += (uint64_t)(uint32_t)(e * 2147483648.0f);
                }
            }
        }
        #pragma omp critical
        {
            if (counts) for (int i = 0; i < npx; ++i) counts[i] += lc[i];
            if (energy) for (int i = 0; i < npx; ++i) energy[i] += le[i];
            tot.rays += st.rays; tot.shaded += st.shaded; tot.blocked += st.blocked;
            tot.hit += st.hit; tot.dda_steps += k.steps + kt.steps;
            tot.dda_steps_eff += k.steps_eff + kt.steps_eff;
            tot.cand_tests += k.tests + kt.tests; tot.self_skips += k.self_skips + kt.self_skips;
            tot.self_hits += k.self_hits + kt.self_hits;
            tot.tag_rays += st.tag_rays; tot.tag_shade_steps_eff += kt.steps_eff;
            tot.tag_shade_tests += kt.tests; tot.tag_shade_self_skips += kt.self_skips;
        }
        free(lc);
        free(le);
    }
    if (stats) *stats = tot;
    return 0;
}

/* fixed point (2^-31 units) to FP32: exact u64 -> double, exact power-of-two scale, one
 * rounding to float. */
void oc_energy_to_float(const uint64_t *energy, int n, float *out)
{
    for (int i = 0; i < n; ++i) out[i] = (float)((double)energy[i] * 4.656612873077392578125e-10);
}

/* ------------------------------------------------------------------------------------ */
/* FP64 evaluation of the same model                                                      */
/* ------------------------------------------------------------------------------------ */

/* exact inverse of the standard normal CDF for the FP64 model: Newton on libm's erf
 * started from the single-precision approximation. */
static double erfinv_d(double x)
{
    double y = (double)oc_erfinvf((float)x);
    for (int i = 0; i < 4; ++i) {
        double err = erf(y) - x;
        y -= err / (1.1283791670955126 * exp(-y * y));
    }
    return y;
}

/* the FP64 model's erfinv element-wise (tests) */
void oc_erfinv_d_batch(const double *in, double *out, int n)
{
    for (int i = 0; i < n; ++i) out[i] = erfinv_d(in[i]);
}

typedef struct {
    const oc_scene *s;
    d3 sun, rv[4];
    d3 *normal, *v0, *v1, *v3, *pos;
    double *cdf_x, *cdf_y, *slope;
    double int465, A, B, C, inv_gp1;
    d3 cyl_base;
    double cyl_r, cyl_h;
    double cone_r0, cone_r1;
    d3 focus;
} oc_scene64;

/* the truncated cone in double: the roots of |p_xz|^2 = (r0 + k p_y)^2 from the original
 * origin, the nearer one ahead within [0, h] entering the surface; rows of equal area */
static int cone_pixeld(d3 o, d3 d, const oc_scene64 *q, int res_x, int res_y)
{
    double h = q->cyl_h, r0 = q->cone_r0, r1 = q->cone_r1, k = (r1 - r0) / h;
    double ox = o.x - q->cyl_base.x, oy = o.y - q->cyl_base.y, oz = o.z - q->cyl_base.z;
    double w0 = r0 + k * oy, wd = k * d.y;
    double a = d.x * d.x + d.z * d.z - wd * wd;
    double b = 2.0 * (ox * d.x + oz * d.z - w0 * wd);
    double c = ox * ox + oz * oz - w0 * w0;
    double disc = b * b - 4.0 * a * c;
    if (d.x * d.x + d.z * d.z < 1e-6 || disc < 0.0 || a == 0.0) return -1;
    double sq = sqrt(disc);
    double q0 = -0.5 * (b + (b >= 0 ? sq : -sq));
    double ta = q0 / a, tb = c / q0;
    double t0 = ta < tb ? ta : tb, t1 = ta < tb ? tb : ta;
    double t = -1.0, hy = 0.0;
    for (int i = 0; i < 2; ++i) {
        double ti = i == 0 ? t0 : t1;
        if (ti < 1e-6) continue;
        double yi = oy + ti * d.y;
        if (yi < 0.0 || yi > h) continue;
        t = ti;
        hy = yi;
        break;
    }
    if (t < 0.0) return -1;
    double hx = ox + t * d.x, hz = oz + t * d.z;
    if (!(d.x * hx + d.z * hz - d.y * k * (r0 + k * hy) < 0.0)) return -1;
    double th = atan2(hz, hx);
    if (th < 0.0) th += 6.283185307179586;
    int col = (int)(th / 6.283185307179586 * res_x);
    int row = (int)(hy * (2.0 * r0 + k * hy) / (h * (r0 + r1)) * res_y);
    if (col > res_x - 1) col = res_x - 1;
    if (row > res_y - 1) row = res_y - 1;
    return row * res_x + col;
}

/* the cylinder receiver in double: the entering root of the quadratic in t */
static int cylinder_pixeld(d3 o, d3 d, const oc_scene64 *q, int res_x, int res_y)
{
    double ox = o.x - q->cyl_base.x, oz = o.z - q->cyl_base.z;
    double a = d.x * d.x + d.z * d.z, b = 2.0 * (ox * d.x + oz * d.z);
    double c = ox * ox + oz * oz - q->cyl_r * q->cyl_r;
    double disc = b * b - 4.0 * a * c;
    if (a < 1e-6 || disc <= 0.0) return -1;
    double sq = sqrt(disc);
    /* stable roots: q0 = -(b + sign(b) sq) / 2, t = q0 / a or c / q0 */
    double q0 = -0.5 * (b + (b >= 0 ? sq : -sq));
    double t1 = q0 / a, t2 = c / q0;
    double t = t1 < t2 ? t1 : t2;
    if (t < 1e-6) return -1;
    double hx = ox + t * d.x, hz = oz + t * d.z, hy = o.y + t * d.y - q->cyl_base.y;
    if (hy < 0.0 || hy > q->cyl_h) return -1;
    double th = atan2(hz, hx);
    if (th < 0.0) th += 6.283185307179586;
    int col = (int)(th / 6.283185307179586 * res_x), row = (int)(hy / q->cyl_h * res_y);
    if (col > res_x - 1) col = res_x - 1;
    if (row > res_y - 1) row = res_y - 1;
    return row * res_x + col;
}

/* The FP64 model's Buie table (inter_larger_465_intensity and parameters_generate,
 * scene_instance_process.cu:54-226, in double): the trapezoid CDF of the profile
 * cos(0.326 x) / cos(0.308 x) on ng bins of [0, 4.65] mrad, cx, cy [ng + 1] and the slopes
 * [ng]; c = {int465, A, B, C, 1 / (gamma + 1)} of the power-law tail to 9.3 mrad. */
void oc_buie_table_d(double csr, int ng, double *cx, double *cy, double *sl, double c[5])
{
    double upper = 4.65;
    double k = 0.9 * log(13.5 * csr) * pow(csr, -0.3);
    double g = 2.2 * log(0.52 * csr) * pow(csr, 0.43) - 0.1;
    double i93 = exp(k) / (g + 1) * pow(9.3, g + 1), i465 = exp(k) / (g + 1) * pow(upper, g + 1);
    double li = upper / ng;
    cx[0] = 0; cy[0] = 0;
    double hp = 1.0;
    for (int i = 1; i <= ng; ++i) {
        double x = li * i;
        double hc = cos(0.326 * x) / cos(0.308 * x);
        double sa = (hp + hc) * li / 2;
        cx[i] = x;
        cy[i] = sa + cy[i - 1];
        sl[i - 1] = li / sa;
        hp = hc;
    }
    c[0] = cy[ng];
    c[1] = c[0] + (i93 - i465);
    c[2] = (g + 1) / exp(k);
    c[3] = pow(upper, g + 1);
    c[4] = 1 / (g + 1);
}

/* The FP64 model's sunshape angle in mrad of the draw u (sun_theta in double): the value
 * v = u A below int465 interpolated in the table's bin (the largest lo with cy[lo] <= v),
 * else the power-law tail. */
static double buie_theta_d(const oc_scene64 *q, int ng, double u)
{
    double v = u * q->A;
    if (v < q->int465) {
        int lo = 0, hi = ng;
        while (hi - lo > 1) { int md = (lo + hi) / 2; if (q->cdf_y[md] <= v) lo = md; else hi = md; }
        return (v - q->cdf_y[lo]) * q->slope[lo] + q->cdf_x[lo];
    }
    return pow((v - q->int465) * q->B + q->C, q->inv_gp1);
}

/* buie_theta_d element-wise on the table of (csr, ng) (tests) */
void oc_buie_theta_d_batch(double csr, int ng, const double *u, double *th, int n)
{
    oc_scene64 q;
    memset(&q, 0, sizeof q);
    q.cdf_x = malloc(sizeof(double) * (ng + 1));
    q.cdf_y = malloc(sizeof(double) * (ng + 1));
    q.slope = malloc(sizeof(double) * ng);
    double c[5];
    oc_buie_table_d(csr, ng, q.cdf_x, q.cdf_y, q.slope, c);
    q.int465 = c[0]; q.A = c[1]; q.B = c[2]; q.C = c[3]; q.inv_gp1 = c[4];
    for (int i = 0; i < n; ++i) th[i] = buie_theta_d(&q, ng, u[i]);
    free(q.cdf_x); free(q.cdf_y); free(q.slope);
}

static void build64(const oc_scene *s, oc_scene64 *q)
{
    int H = s->H;
    q->s = s;
    d3 sun = normalized(mkd(s->d.sun_dir[0], s->d.sun_dir[1], s->d.sun_dir[2]));
    q->sun = sun;
    /* receiver: same construction in double */
    f3 size = ld3(s->d.recv_size);
    d3 sz = f2d(size), pos = mkd(s->d.recv_pos[0], s->d.recv_pos[1], s->d.recv_pos[2]);
    d3 ln, v[4];
    switch (s->d.recv_face) {
    case 0: ln = mkd(0, 0, 1);
        v[0] = mkd(-sz.x / 2, -sz.y / 2, sz.z / 2); v[1] = mkd(-sz.x / 2, sz.y / 2, sz.z / 2);
        v[2] = mkd(sz.x / 2, sz.y / 2, sz.z / 2); v[3] = mkd(sz.x / 2, -sz.y / 2, sz.z / 2); break;
    case 1: ln = mkd(1, 0, 0);
        v[0] = mkd(sz.x / 2, -sz.y / 2, sz.z / 2); v[1] = mkd(sz.x / 2, sz.y / 2, sz.z / 2);
        v[2] = mkd(sz.x / 2, sz.y / 2, -sz.z / 2); v[3] = mkd(sz.x / 2, -sz.y / 2, -sz.z / 2); break;
    case 2: ln = mkd(0, 0, -1);
        v[0] = mkd(sz.x / 2, -sz.y / 2, -sz.z / 2); v[1] = mkd(sz.x / 2, sz.y / 2, -sz.z / 2);
        v[2] = mkd(-sz.x / 2, sz.y / 2, -sz.z / 2); v[3] = mkd(-sz.x / 2, -sz.y / 2, -sz.z / 2); break;
    default: ln = mkd(-1, 0, 0);
        v[0] = mkd(-sz.x / 2, -sz.y / 2, -sz.z / 2); v[1] = mkd(-sz.x / 2, sz.y / 2, -sz.z / 2);
        v[2] = mkd(-sz.x / 2, sz.y / 2, sz.z / 2); v[3] = mkd(-sz.x / 2, -sz.y / 2, sz.z / 2); break;
    }
    d3 n = normalized(mkd(s->d.recv_normal[0], s->d.recv_normal[1], s->d.recv_normal[2]));
    for (int i = 0; i < 4; ++i) {
        int dir = (crossd(ln, n).y > 0) ? 1 : -1;
        double c = dotd(ln, n);
        double sn = dir * sqrt(fmax(0.0, 1 - c * c));
        d3 o = v[i];
        q->rv[i] = addd(mkd(c * o.x + sn * o.z, o.y, -sn * o.x + c * o.z), pos);
    }
    d3 focus = scld(addd(q->rv[0], q->rv[2]), 0.5);
    if (s->d.recv_type == 1) {
        double r = s->d.recv_size[0], h = s->d.recv_size[1];
        focus = pos;
        q->cyl_base = mkd(pos.x, pos.y - h / 2, pos.z);
        q->cyl_r = r;
        q->cyl_h = h;
    } else if (s->d.recv_type == 2) {
        double h = s->d.recv_size[2];
        focus = pos;
        q->cyl_base = mkd(pos.x, pos.y - h / 2, pos.z);
        q->cone_r1 = s->d.recv_size[0];
        q->cone_r0 = s->d.recv_size[1];
        q->cyl_h = h;
    }
    q->focus = focus;
    q->normal = malloc(sizeof(d3) * (H + 1));
    q->v0 = malloc(sizeof(d3) * (H + 1));
    q->v1 = malloc(sizeof(d3) * (H + 1));
    q->v3 = malloc(sizeof(d3) * (H + 1));
    q->pos = malloc(sizeof(d3) * (H + 1));
    for (int i = 0; i < H; ++i) {
        d3 p = f2d(s->pos[i]), z = f2d(s->size[i]);
        q->pos[i] = p;
        d3 nn = s->d.helio_normal
                    ? normalized(mkd(s->d.helio_normal[3 * i], s->d.helio_normal[3 * i + 1], s->d.helio_normal[3 * i + 2]))
                    : normalized(subd(normalized(subd(focus, p)), sun));
        q->normal[i] = nn;
        d3 l0 = mkd(-z.x / 2, z.y / 2, -z.z / 2);
        d3 l1 = mkd(l0.x, l0.y, l0.z + z.z);
        d3 l3 = mkd(l0.x + z.x, l0.y, l0.z);
        q->v0[i] = addd(local2worldd(l0, nn), p);
        q->v1[i] = addd(local2worldd(l1, nn), p);
        q->v3[i] = addd(local2worldd(l3, nn), p);
    }
    /* Buie table in double */
    int ng = s->nbins;
    q->cdf_x = malloc(sizeof(double) * (ng + 1));
    q->cdf_y = malloc(sizeof(double) * (ng + 1));
    q->slope = malloc(sizeof(double) * ng);
    double c[5];
    oc_buie_table_d(s->p.csr, ng, q->cdf_x, q->cdf_y, q->slope, c);
    q->int465 = c[0]; q->A = c[1]; q->B = c[2]; q->C = c[3]; q->inv_gp1 = c[4];
}

static void free64(oc_scene64 *q)
{
    free(q->normal); free(q->v0); free(q->v1); free(q->v3); free(q->pos);
    free(q->cdf_x); free(q->cdf_y); free(q->slope);
}

static int dda64(const oc_scene *s, const oc_scene64 *q, d3 o, d3 d, int self)
{
    const oc_grid *G = grid_of(s, self);
    d3 gpos = f2d(G->pos);
    d3 gint = f2d(G->interval);
    d3 rel = subd(o, gpos);
    int g[3] = { (int)(rel.x / gint.x), (int)(rel.y / gint.y), (int)(rel.z / gint.z) };
    double dd[3] = { d.x, d.y, d.z }, gi[3] = { gint.x, gint.y, gint.z }, cur[3] = { rel.x, rel.y, rel.z };
    int st[3];
    double tm[3], td[3];
    /* absDivide's quotients as products with 1 / dir (one division per axis: the FP64
     * model's form; either form is exact to an ulp) */
    for (int a = 0; a < 3; ++a) {
        st[a] = dd[a] >= 0 ? 1 : -1;
        double t = dd[a] >= 0 ? (g[a] + 1) * gi[a] - cur[a] : cur[a] - g[a] * gi[a];
        double inv = 1.0 / dd[a];
        tm[a] = (fabs(dd[a]) <= OC_EPS) ? 2147483647.0 : fabs(t * inv);
        td[a] = (fabs(dd[a]) <= OC_EPS) ? 2147483647.0 : fabs(gi[a] * inv);
    }
    for (;;) {
        int a;
        if (tm[0] < tm[1]) a = (tm[0] < tm[2]) ? 0 : 2;
        else a = (tm[1] < tm[2]) ? 1 : 2;
        g[a] += st[a];
        if (g[a] >= G->gn[a] || g[a] < 0) return -1;
        tm[a] += td[a];
        if ((st[0] > 0 ? g[0] > G->occ_hi[0] : g[0] < G->occ_lo[0]) ||
            (st[1] > 0 ? g[1] > G->occ_hi[1] : g[1] < G->occ_lo[1]) ||
            (st[2] > 0 ? g[2] > G->occ_hi[2] : g[2] < G->occ_lo[2]))
            return -1;
        int64_t addr = (int64_t)g[0] * G->gn[1] * G->gn[2] + (int64_t)g[1] * G->gn[2] + g[2];
        for (int32_t i = G->cell_start[addr]; i < G->cell_start[addr + 1]; ++i) {
            int h = G->cell_items[i];
            if (h == self) continue;
            double t, u, v;
            if (ray_parallelogramd(o, d, q->v0[h], q->v1[h], q->v3[h], &t, &u, &v)) return h;
        }
    }
}

static d3 micro_local_d(const oc_scene *s, int h, int m)
{
    d3 sz = f2d(s->size[h]);
    double gx = s->gap[2 * h], gy = s->gap[2 * h + 1];
    double subrow = (sz.z - gy * (s->matrix[2 * h] - 1)) / s->matrix[2 * h];
    double subcol = (sz.x - gx * (s->matrix[2 * h + 1] - 1)) / s->matrix[2 * h + 1];
    if (s->p.micro_mode == OC_MICRO_POISSON) {
        const float *pt = s->poisson_pts + 2 * ((size_t)s->poisson_cap * h + m);
        int block_row = (int)(pt[1] / subrow);
        int block_col = (int)(pt[0] / subcol);
        return mkd(pt[0] + block_col * gx - sz.x / 2, sz.y / 2, pt[1] + block_row * gy - sz.z / 2);
    }
    double p = s->p.helio_pixel_length;
    int cols_all = s->matrix[2 * h + 1] * s->sub_cols[h];
    int row = m / cols_all, col = m % cols_all;
    int block_row = row / s->sub_rows[h], block_col = col / s->sub_cols[h];
    return mkd(col * p + block_col * gx + p / 2 - sz.x / 2, sz.y / 2,
               row * p + block_row * gy + p / 2 - sz.z / 2);
}

int oc_trace_f64(const oc_scene *s, int h0, int h1,
                 uint32_t *counts, double *energy64,
                 int32_t *rec_shade, int32_t *rec_block, int32_t *rec_pixel,
                 int nthreads, oc_stats *stats)
{
    if (h0 < 0) h0 = 0;
    if (h1 > s->H) h1 = s->H;
    oc_scene64 q;
    build64(s, &q);
    int npx = s->res_x * s->res_y;
    int L = s->p.rays_per_sample;
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    oc_stats tot;
    memset(&tot, 0, sizeof(tot));
    uint32_t key[2] = { (uint32_t)s->p.seed, (uint32_t)(s->p.seed >> 32) };
    #pragma omp parallel num_threads(nt)
    {
        uint32_t *lc = (uint32_t *)calloc(npx, sizeof(uint32_t));
        double *le = (double *)calloc(npx, sizeof(double));
        oc_stats st;
        memset(&st, 0, sizeof(st));
        #pragma omp for schedule(dynamic, 1)
        for (int h = h0; h < h1; ++h) {
            d3 n = q.normal[h];
            for (int m = 0; m < s->mcount[h]; ++m) {
                d3 orig = addd(local2worldd(micro_local_d(s, h, m), n), q.pos[h]);
                for (int l = 0; l < L; ++l) {
                    int64_t rid = s->ray_offset[h] - s->ray_offset[h0] + (int64_t)m * L + l;
                    st.rays++;
                    uint32_t c[4] = { (uint32_t)l, (uint32_t)m, (uint32_t)h,
                                      (s->p.scene_index << 4) | OC_PURPOSE_RAY }, o[4];
                    oc_philox4x32_10(c, key, o);
                    double u0 = oc_u01(o[0]), u1 = oc_u01(o[1]), u2 = oc_u01(o[2]), u3 = oc_u01(o[3]);
                    double th;
                    if (s->p.sunshape == OC_SUN_PILLBOX) {
                        th = (double)s->p.pillbox_half_angle * sqrt(u0);
                    } else {
                        th = buie_theta_d(&q, s->nbins, u0) / 1000.0;
                    }
                    double ph = 6.283185307179586 * u1;
                    d3 smp = mkd(sin(th) * cos(ph), cos(th), sin(th) * sin(ph));
                    d3 to_sun = negd(local2worldd(smp, q.sun));
                    double ths = erfinv_d(2.0 * u2 - 1.0) * sqrt(2.0) * (double)s->p.slope_sigma;
                    double phs = 6.283185307179586 * u3;
                    d3 turb = mkd(sin(ths) * cos(phs), cos(ths), sin(ths) * sin(phs));
                    d3 nrm = normalized(local2worldd(turb, n));
                    int sh = dda64(s, &q, orig, to_sun, h);
                    if (rec_shade) rec_shade[rid] = sh;
                    if (sh >= 0) {
                        st.shaded++;
                        if (rec_block) rec_block[rid] = -2;
                        if (rec_pixel) rec_pixel[rid] = -1;
                        continue;
                    }
                    d3 refl = normalized(reflectd(negd(to_sun), nrm));
                    int bl = dda64(s, &q, orig, refl, h);
                    if (rec_block) rec_block[rid] = bl;
                    if (bl >= 0) {
                        st.blocked++;
                        if (rec_pixel) rec_pixel[rid] = -1;
                        continue;
                    }
                    double t, u, v;
                    int px = -1;
                    if (s->d.recv_type == 1) {
                        px = cylinder_pixeld(orig, refl, &q, s->res_x, s->res_y);
                    } else if (s->d.recv_type == 2) {
                        px = cone_pixeld(orig, refl, &q, s->res_x, s->res_y);
                    } else if (ray_parallelogramd(orig, refl, q.rv[0], q.rv[1], q.rv[3], &t, &u, &v)) {
                        int row = (int)(u * s->res_y), col = (int)(v * s->res_x);
                        if (row > s->res_y - 1) row = s->res_y - 1;
                        if (col > s->res_x - 1) col = s->res_x - 1;
                        px = row * s->res_x + col;
                    }
                    if (rec_pixel) rec_pixel[rid] = px;
                    if (px < 0) continue;
                    st.hit++;
                    lc[px] += 1u;
                    le[px] += fabs(dotd(q.sun, nrm));
                }
            }
        }
        #pragma omp critical
        {
            if (counts) for (int i = 0; i < npx; ++i) counts[i] += lc[i];
            if (energy64) for (int i = 0; i < npx; ++i) energy64[i] += le[i];
            tot.rays += st.rays; tot.shaded += st.shaded; tot.blocked += st.blocked; tot.hit += st.hit;
        }
        free(lc);
        free(le);
    }
    free64(&q);
    if (stats) *stats = tot;
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* pooled sampling comparator (reference RNG scheme)                                     */
/* ------------------------------------------------------------------------------------ */

/* set_perturbation + set_samplelights pools (scene_instance_process.cu:25-51, :163-226)
 * rebuilt for every heliostat (Gtest.cu:27), random group starts per microsample
 * (steps_for_raytracing.cu:181-187) and the pool indices of map_tracing
 * (rectgrid_raytracing.cu:171, :183-185, including the m-1 start of the perturbation). */
int oc_trace_pooled(const oc_scene *s, int h0, int h1, int groups, int lights,
                    uint64_t pool_seed, double *e_helio, int nthreads)
{
    if (h0 < 0) h0 = 0;
    if (h1 > s->H) h1 = s->H;
    int pool = groups * lights;
    int L = s->p.rays_per_sample;
    if (L > pool) return -1;
    uint32_t key[2] = { (uint32_t)pool_seed, (uint32_t)(pool_seed >> 32) };
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    #pragma omp parallel num_threads(nt)
    {
        f3 *light = (f3 *)malloc(sizeof(f3) * pool);
        f3 *turb = (f3 *)malloc(sizeof(f3) * pool);
        int *grp = (int *)malloc(sizeof(int) * (s->max_micro + 1));
        dda_ctr k = {0, 0, 0, 0, 0};
        #pragma omp for schedule(dynamic, 1)
        for (int h = h0; h < h1; ++h) {
            for (int i = 0; i < pool; ++i) {
                uint32_t c[4] = { (uint32_t)i, 0u, (uint32_t)h, OC_PURPOSE_POOL }, o[4];
                oc_philox4x32_10(c, key, o);
                float th = sun_theta(s, oc_u01(o[0]));
                float st, ct, sp, cp;
                oc_sincos_small(th, &st, &ct);
                oc_sincos2pi(oc_u01(o[1]), &sp, &cp);
                light[i] = mk3(st * cp, ct, st * sp);
                float ths = oc_gauss_from_u(oc_u01(o[2])) * s->p.slope_sigma;
                oc_sincos_small(ths, &st, &ct);
                oc_sincos2pi(oc_u01(o[3]), &sp, &cp);
                turb[i] = mk3(st * cp, ct, st * sp);
            }
            for (int m = 0; m < s->mcount[h]; ++m) {
                uint32_t c[4] = { (uint32_t)m, 1u, (uint32_t)h, OC_PURPOSE_POOL }, o[4];
                oc_philox4x32_10(c, key, o);
                grp[m] = (int)((uint64_t)o[0] * (uint64_t)pool >> 32);
            }
            double acc = 0.0;
            f3 n = s->normal[h];
            for (int m = 0; m < s->mcount[h]; ++m) {
                f3 orig = add3(local2world3(micro_local(s, h, m), n), s->pos[h]);
                int gm = grp[m], gp = grp[m > 0 ? m - 1 : 0];
                for (int l = 0; l < L; ++l) {
                    f3 to_sun = neg3(local2world3(light[(gm + l) % pool], s->sun));
                    if (dda(s, orig, to_sun, h, 1, &k, NULL, 0, NULL) >= 0) continue;
                    f3 nrm = normalize3(local2world3(turb[(gp + l) % pool], n));
                    f3 refl = normalize3(reflect3(neg3(to_sun), nrm));
                    if (dda(s, orig, refl, h, 1, &k, NULL, 0, NULL) >= 0) continue;
                    if (receiver_pixel(s, orig, refl) < 0) continue;
                    acc += fabsf(dot3(s->sun, nrm));
                }
            }
            e_helio[h - h0] = acc;
        }
        free(light);
        free(turb);
        free(grp);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* trimmed mean (image_smooth.cu:10-58, heap.cuh:9-212)                                  */
/* ------------------------------------------------------------------------------------ */
typedef struct { int num, max_size, is_max; float *el; } oc_heap;

static inline int heap_cmp(const oc_heap *hp, float a, float b) { return hp->is_max ? a > b : a < b; }

static void heap_insert(oc_heap *hp, float e)
{
    if (hp->num == hp->max_size) return;
    int i = hp->num++;
    for (; i != 1 && heap_cmp(hp, e, hp->el[i / 2]); i /= 2) hp->el[i] = hp->el[i / 2];
    hp->el[i] = e;
}

static void heap_replace_top(oc_heap *hp, float e)
{
    if (hp->num == 1) return;
    int i = 1, cid;
    for (; 2 * i < hp->num;) {
        cid = (2 * i + 1 >= hp->num || heap_cmp(hp, hp->el[2 * i], hp->el[2 * i + 1])) ? 2 * i : 2 * i + 1;
        if (heap_cmp(hp, hp->el[cid], e)) { hp->el[i] = hp->el[cid]; i = cid; }
        else break;
    }
    hp->el[i] = e;
}

static float heap_sum(const oc_heap *hp)
{
    float s = 0.0f;
    for (int i = 1; i < hp->num; ++i) s += hp->el[i];
    return s;
}

int oc_trimmed_mean(const float *in, float *out, int width, int height,
                    int kernel_radius, float trimmed_ratio)
{
    int kcap = (int)((float)((2 * kernel_radius + 1) * (2 * kernel_radius + 1)) * trimmed_ratio + 1);
    float *mx = (float *)malloc(sizeof(float) * (kcap + 2));
    float *mn = (float *)malloc(sizeof(float) * (kcap + 2));
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            int lx = (x - kernel_radius >= 0) ? x - kernel_radius : 0;
            int ly = (y - kernel_radius >= 0) ? y - kernel_radius : 0;
            int rx = (x + kernel_radius < width) ? x + kernel_radius : width - 1;
            int ry = (y + kernel_radius < height) ? y + kernel_radius : height - 1;
            int cnt = (rx - lx + 1) * (ry - ly + 1);
            int k = (int)((float)cnt * trimmed_ratio);
            oc_heap hmax = { 1, k + 1, 1, mx }, hmin = { 1, k + 1, 0, mn };
            float value = 0.0f;
            for (int r = ly; r <= ry; ++r)
                for (int c = lx; c <= rx; ++c) {
                    float v = in[r * width + c];
                    value += v;
                    if (k == 0) continue;   /* both heaps full and empty: no update */
                    if (hmax.num != hmax.max_size) {
                        heap_insert(&hmax, v);
                        heap_insert(&hmin, v);
                    } else if (v < hmax.el[1]) {
                        heap_replace_top(&hmax, v);
                    } else if (v > hmin.el[1]) {
                        heap_replace_top(&hmin, v);
                    }
                }
            out[y * width + x] = (value - heap_sum(&hmax) - heap_sum(&hmin)) / (float)(cnt - 2 * k);
        }
    free(mx);
    free(mn);
    return 0;
}

/* ImageSaver::savetxt(..., float totalEnergy) (image_save.cpp:26-47): rows are scanned
 * from the last to the first. */
float oc_intercept_ratio(const float *flux, int width, int height, float total_energy)
{
    float in = 0.0f;
    for (int r = 0; r < height; ++r)
        for (int c = 0; c < width; ++c) {
            float v = flux[(height - 1 - r) * width + c];
            if (v < OC_EPS) continue;
            in += v;
        }
    return in / total_energy;
}

/* ------------------------------------------------------------------------------------ */
/* geometry probes                                                                       */
/* ------------------------------------------------------------------------------------ */
void oc_local2world(const float d[3], const float n[3], float out[3])
{
    f3 r = local2world3(ld3(d), ld3(n));
    memcpy(out, &r, 12);
}

void oc_rotate_y(const float o[3], const float old_dir[3], const float new_dir[3], float out[3])
{
    f3 r = rotate_y3(ld3(o), ld3(old_dir), ld3(new_dir));
    memcpy(out, &r, 12);
}

int oc_ray_parallelogram(const float orig[3], const float dir[3], const float A[3],
                         const float B[3], const float C[3], float tuv[3])
{
    float t = 0, u = 0, v = 0;
    int r = ray_parallelogram3(ld3(orig), ld3(dir), ld3(A), ld3(B), ld3(C), &t, &u, &v);
    tuv[0] = t; tuv[1] = u; tuv[2] = v;
    return r;
}

int oc_cylinder_pixel_batch(const float *orig, const float *dir, int n, const float base[3], float r2,
                            float h, int res_x, int res_y, int32_t *out)
{
    for (int i = 0; i < n; ++i)
        out[i] = cylinder_pixel3(ld3(orig + 3 * i), ld3(dir + 3 * i), ld3(base), r2, h, res_x, res_y);
    return 0;
}

int oc_cone_pixel_batch(const float *orig, const float *dir, int n, const float base[3], float r0, float k,
                        float h, float inv_area, int res_x, int res_y, int32_t *out)
{
    for (int i = 0; i < n; ++i)
        out[i] = cone_pixel3(ld3(orig + 3 * i), ld3(dir + 3 * i), ld3(base), r0, k, h, inv_area, res_x, res_y);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* differentiable model in FP64 (no reference counterpart): the same Philox draws and     */
/* microsample centres as oc_trace_f64; per heliostat theta = (a_u, a_v, t_u, t_v):       */
/*   aim = focus + a_u e1 + a_v e2,  n_a = normalize(normalize(aim - pos) - sun),          */
/*   n = normalize(n_a + 1e-3 (t_u fu + t_v fv))  (fu, fv: frame of the nominal normal);  */
/* per visible ray: nrm = normalize(n + (nrm0 - n0)), r = normalize(reflect(-to_sun, nrm)),*/
/* receiver-plane hit -> pixel coordinates (Y, X) = (u res_y, v res_x) - 1/2, energy       */
/* e = |sun.nrm|, splat e K(x - X) K(y - Y) with the tapered Gaussian K of radius 3 px.   */
/* Visibility is given per ray (vis, in ray order of [h0, h1)).                           */
/* ------------------------------------------------------------------------------------ */
static double diff_kern(double d, double eR, double R, double iZ)
{
    return (exp(-0.5 * d * d) - eR * (1.0 + 0.5 * (R * R - d * d))) * iZ;
}

int oc_diff_splat64(const oc_scene *s, const double *theta, int h0, int h1, const uint8_t *vis,
                    double *S, double *E, int nthreads)
{
    if (s->d.recv_type != 0) return -1;
    if (h0 < 0) h0 = 0;
    if (h1 > s->H) h1 = s->H;
    oc_scene64 q;
    build64(s, &q);
    const int npx = s->res_x * s->res_y, L = s->p.rays_per_sample;
    const d3 A = q.rv[0], E1 = subd(q.rv[1], q.rv[0]), E2 = subd(q.rv[3], q.rv[0]);
    const d3 N = normalized(crossd(E1, E2)), e1 = normalized(E1), e2 = normalized(E2);
    const double iE1 = 1.0 / dotd(E1, E1), iE2 = 1.0 / dotd(E2, E2);
    const double R = 3.0, eR = exp(-0.5 * R * R);
    const double iZ = 1.0 / (sqrt(2.0 * 3.14159265358979323846) * erf(R / sqrt(2.0)) - eR * (2.0 * R + 2.0 * R * R * R / 3.0));
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    double *maps = (double *)calloc((size_t)nt * npx, sizeof(double));
    double *esum = (double *)calloc((size_t)nt, sizeof(double));
    uint32_t key[2] = { (uint32_t)s->p.seed, (uint32_t)(s->p.seed >> 32) };
    #pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *Sm = maps + (size_t)tid * npx;
        #pragma omp for schedule(static)
        for (int h = h0; h < h1; ++h) {
            const d3 n0 = q.normal[h], pos = q.pos[h];
            d3 fu = mkd(1, 0, 0), fv = mkd(0, 0, 1);
            if (!(fabs(n0.x) < OC_EPS && fabs(n0.z) < OC_EPS)) {
                fu = normalized(crossd(mkd(0.0, 1.0, 0.0), n0));
                fv = normalized(crossd(fu, n0));
            }
            const double *th = theta + 4 * (size_t)h;
            d3 aim = addd(addd(q.focus, scld(e1, th[0])), scld(e2, th[1]));
            d3 na = normalized(subd(normalized(subd(aim, pos)), q.sun));
            d3 nt_ = normalized(addd(na, scld(addd(scld(fu, th[2]), scld(fv, th[3])), 1e-3)));
            for (int m = 0; m < s->mcount[h]; ++m) {
                d3 orig = addd(local2worldd(micro_local_d(s, h, m), n0), pos);
                for (int l = 0; l < L; ++l) {
                    int64_t rid = s->ray_offset[h] - s->ray_offset[h0] + (int64_t)m * L + l;
                    if (!vis[rid]) continue;
                    uint32_t c[4] = { (uint32_t)l, (uint32_t)m, (uint32_t)h,
                                      (s->p.scene_index << 4) | OC_PURPOSE_RAY }, o[4];
                    oc_philox4x32_10(c, key, o);
                    double u0 = oc_u01(o[0]), u1 = oc_u01(o[1]), u2 = oc_u01(o[2]), u3 = oc_u01(o[3]);
                    double thv;
                    if (s->p.sunshape == OC_SUN_PILLBOX) {
                        thv = (double)s->p.pillbox_half_angle * sqrt(u0);
                    } else {
                        thv = buie_theta_d(&q, s->nbins, u0) / 1000.0;
                    }
                    double ph = 6.283185307179586 * u1;
                    d3 smp = mkd(sin(thv) * cos(ph), cos(thv), sin(thv) * sin(ph));
                    d3 to_sun = negd(local2worldd(smp, q.sun));
                    double ths = erfinv_d(2.0 * u2 - 1.0) * sqrt(2.0) * (double)s->p.slope_sigma;
                    double phs = 6.283185307179586 * u3;
                    d3 turb = mkd(sin(ths) * cos(phs), cos(ths), sin(ths) * sin(phs));
                    d3 nrm0 = normalized(local2worldd(turb, n0));
                    d3 nrm = normalized(addd(nt_, subd(nrm0, n0)));
                    d3 r = normalized(reflectd(negd(to_sun), nrm));
                    double den = dotd(r, N);
                    if (fabs(den) < OC_EPS) continue;
                    double t = dotd(subd(A, orig), N) / den;
                    if (t < OC_EPS) continue;
                    d3 pp = addd(orig, scld(r, t));
                    double u = dotd(subd(pp, A), E1) * iE1, v = dotd(subd(pp, A), E2) * iE2;
                    double Y = u * s->res_y - 0.5, X = v * s->res_x - 0.5;
                    double e = fabs(dotd(q.sun, nrm));
                    esum[tid] += e;
                    int x0 = (int)ceil(X - R), x1 = (int)floor(X + R), y0 = (int)ceil(Y - R), y1 = (int)floor(Y + R);
                    if (x0 < 0) x0 = 0;
                    if (y0 < 0) y0 = 0;
                    if (x1 > s->res_x - 1) x1 = s->res_x - 1;
                    if (y1 > s->res_y - 1) y1 = s->res_y - 1;
                    double kx[8];
                    for (int xx = x0; xx <= x1 && xx - x0 < 8; ++xx) kx[xx - x0] = diff_kern(xx - X, eR, R, iZ);
                    for (int yy = y0; yy <= y1; ++yy) {
                        double ky = e * diff_kern(yy - Y, eR, R, iZ);
                        for (int xx = x0; xx <= x1 && xx - x0 < 8; ++xx) {
                            double w = ky * kx[xx - x0];
                            if (w > 0.0) Sm[(size_t)yy * s->res_x + xx] += w;
                        }
                    }
                }
            }
        }
    }
    for (int t = 0; t < nt; ++t) {
        for (int i = 0; i < npx; ++i) S[i] += maps[(size_t)t * npx + i];
        *E += esum[t];
    }
    free(maps);
    free(esum);
    free64(&q);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* the faceted heliostat of the annual driver                                            */
/* ------------------------------------------------------------------------------------ */

/* main.cpp:84-91.  receiverH (main.cpp:8) is a double literal, so the height terms are
 * evaluated in double; the rest in float as declared there. */
int oc_master_normal(float X, float Y, float Z, const float sun_dir[3], float out[3])
{
    const double receiverH = 13.01;
    float a = sun_dir[0], b = sun_dir[1], c = sun_dir[2];
    float distance = (float)sqrt(X * X + Z * Z + (receiverH - Y) * (receiverH - Y));
    if (!(distance > 0.0f)) return -1;
    float rox = -X / distance;
    float roz = -Z / distance;
    float roy = (float)((receiverH - Y) / distance);
    float NR = sqrtf((-a + rox) * (-a + rox) + (-b + roy) * (-b + roy) + (-c + roz) * (-c + roz));
    if (!(NR > 0.0f)) return -1;
    out[0] = (-a + rox) / NR;
    out[1] = (-b + roy) / NR;
    out[2] = (-c + roz) / NR;
    return 0;
}

/* SceneFileProc::matrixMulti (scene_file_proc.cpp:536-543) */
static void matrix_multi(float A[3][3], const float *b, float *c)
{
    for (int i = 0; i < 3; i++) {
        c[i] = 0;
        for (int j = 0; j < 3; j++) c[i] += A[i][j] * b[j];
    }
}

/* SceneFileProc::SceneFileRead(SolarScene*, a, b, mirrorNormal, X, Y, Z), facet part
 * (scene_file_proc.cpp:289-339) and facet sizes (:478-481). */
int oc_facet_heliostat(float a, float b, const float mirror_normal[3], float X, float Y, float Z,
                       float pos[27], float normal[27], float size[27])
{
    float centerX[9] = { -sqrt(2.0) / 3 - 0.1, 0, sqrt(2.0) / 3 + 0.1, -sqrt(2.0) / 3 - 0.1, 0,
                         sqrt(2.0) / 3 + 0.1, -sqrt(2.0) / 3 - 0.1, 0, sqrt(2.0) / 3 + 0.1 };
    float centerZ[9] = { -sqrt(2.0) / 3 - 0.1, -sqrt(2.0) / 3 - 0.1, -sqrt(2.0) / 3 - 0.1, 0, 0, 0,
                         sqrt(2.0) / 3 + 0.1, sqrt(2.0) / 3 + 0.1, sqrt(2.0) / 3 + 0.1 };
    float centerY[9];
    float m[3][3];
    const float *nm = mirror_normal;
    if (a == 0.0f || b == 0.0f) return -1;
    for (int i = 0; i < 9; i++)
        centerY[i] = centerX[i] * centerX[i] / a / a + centerZ[i] * centerZ[i] / b / b;
    float temp = sqrtf(nm[0] * nm[0] + nm[1] * nm[1]);
    if (nm[2] > 0) {
        m[0][0] = nm[2] / temp; m[2][0] = -nm[0] / temp; m[1][0] = 0;
    } else if (nm[2] < 0) {
        m[0][0] = -nm[2] / temp; m[2][0] = nm[0] / temp; m[1][0] = 0;
    } else {
        m[0][0] = 0; m[1][0] = 0; m[2][0] = 1;
    }
    for (int i = 0; i < 3; i++) m[i][1] = nm[i];
    m[0][2] = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    m[1][2] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
    m[2][2] = m[0][0] * m[1][1] - m[1][0] * m[0][1];
    for (int i = 0; i < 9; i++) {
        float center[3] = { centerX[i], centerY[i], centerZ[i] };
        float nrm[3] = { -center[0] * 2 / a / a, 1, -center[2] * 2 / b / b };
        float ct[3], nt[3];
        matrix_multi(m, center, ct);
        matrix_multi(m, nrm, nt);
        pos[3 * i] = ct[0] + X;
        pos[3 * i + 1] = ct[1] + Y;
        pos[3 * i + 2] = ct[2] + Z;
        normal[3 * i] = nt[0];
        normal[3 * i + 1] = nt[1];
        normal[3 * i + 2] = nt[2];
        size[3 * i] = (float)(1.4142136 / 3);
        size[3 * i + 1] = 0.1f;
        size[3 * i + 2] = (float)(1.4142136 / 3);
    }
    for (int i = 0; i < 27; ++i)
        if (!isfinite(pos[i]) || !isfinite(normal[i])) return -1;
    return 0;
}
