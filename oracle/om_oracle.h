/*
 * om_oracle.h -- CPU oracle for the heliostat-field Monte-Carlo ray tracer.
 *
 * TEST INFRASTRUCTURE, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker or the
 * host baseline.  The product (onemirror_amd/) never links, loads or falls back to it.
 *
 * It is a plain-C restatement of the reference routines of Hu_smooth/ClassedSunRay
 * (every function cites the reference file:line it follows), fed by Philox4x32-10
 * draws at a fixed counter layout instead of cuRAND pools.  Two variants exist:
 *   - FP32 ("reference semantics"): the same IEEE single-precision operation sequence
 *     as the HIP kernels, so per-ray indices and integer pixel counts must be bit-exact;
 *   - FP64: the same model evaluated in double precision (libm transcendentals, exact
 *     inverse normal), used as the accuracy yardstick for the FP32 flux.
 *
 * Citations use ClassedSunRay/ as the root of the reference sources.
 */
#ifndef OM_ORACLE_H
#define OM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- one grid of a field (RectGrid, grid.h:25-49) and its heliostats
 *      [first_helio, first_helio + n_helio) (start_helio_pos_, num_helios_) ---- */
typedef struct {
    float pos[3];
    float size[3];
    float interval[3];
    int   first_helio;
    int   n_helio;
} oc_grid_desc;

/* ---- scene description: the inputs a .scn file carries (scene_file_proc.cpp:41-286)
 *      plus the sun of SunRay (sunray.h:4-60). All arrays are host memory. ---- */
typedef struct {
    int          n_helio;
    const float *helio_pos;    /* [n][3] centre, world (x east, y up, z south)          */
    const float *helio_size;   /* [n][3] x = width, y = thickness, z = height           */
    const float *helio_gap;    /* [n][2] gap between facets (x, z)                      */
    const int   *helio_matrix; /* [n][2] facet rows, facet cols                         */
    int   recv_type;           /* 0 = rectangle (receiver.cpp), 1 = cylinder, 2 = cone (new) */
    float recv_pos[3];
    float recv_size[3];        /* rectangle: box size; cylinder: (radius, height, -);
                                  cone: (top radius, bottom radius, height)              */
                               /* cylinder recv_pos: axis point at mid-height           */
    float recv_normal[3];
    int   recv_face;           /* rectangle face 0..3 (receiver.cpp:30-48)              */
    float recv_pixel_length;   /* solarenergy::receiver_pixel_length (common_var.cpp:12)*/
    float grid_pos[3];
    float grid_size[3];
    float grid_interval[3];
    float sun_dir[3];          /* direction light travels (main.cpp:71-73)              */
    float dni;
    int   recv_res_x, recv_res_y; /* cylinder columns (azimuth), rows; 0 = from pixel length */
    int   n_grids;             /* 0: one grid (grid_* above) holding every heliostat;    */
    const oc_grid_desc *grids; /* else grids[k] hold consecutive ranges covering [0, n),
                                  each heliostat's rays walking its own grid over global
                                  ids (recthelio_tracing.cpp:46-79, grid.cpp:53)          */
    const float *helio_normal; /* [n][3] given normals (InitContent(normals),
                                  solar_scene.cpp:101-115), or NULL: aim at the focus     */
} oc_scene_desc;

/* The faceted heliostat (scene_file_proc.cpp:288-339,464-487) and the master normal of the
 * annual driver (main.cpp:84-91): see om_facet_heliostat / om_master_normal. */
int oc_facet_heliostat(float a, float b, const float mirror_normal[3], float X, float Y, float Z,
                       float pos[27], float normal[27], float size[27]);
int oc_master_normal(float X, float Y, float Z, const float sun_dir[3], float out[3]);

/* ---- tracing parameters: the solarenergy:: globals (common_var.cpp:9-27) plus the
 *      per-ray sampling choices that replace the cuRAND pools. ---- */
enum { OC_SUN_PILLBOX = 0, OC_SUN_BUIE = 1 };
enum { OC_MICRO_LATTICE = 0, OC_MICRO_POISSON = 1 };

typedef struct {
    uint64_t seed;
    uint32_t scene_index;      /* Philox counter word 3 (high bits)                     */
    int   sunshape;            /* OC_SUN_*                                              */
    float pillbox_half_angle;  /* rad                                                   */
    float csr;                 /* Buie circumsolar ratio                                */
    int   buie_bins;           /* num_group of set_samplelights (1024 in the reference) */
    float slope_sigma;         /* rad, disturb_std                                      */
    int   micro_mode;          /* OC_MICRO_*                                            */
    float helio_pixel_length;  /* microsample pitch                                     */
    int   poisson_k;           /* trials per cell (steps_for_raytracing.h:11, k = 8)    */
    int   rays_per_sample;     /* Nc = lights per microsample                           */
    float reflectivity;        /* reflected_rate                                        */
} oc_params;

typedef struct oc_scene oc_scene;

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- project-written single-precision math shared (bit-for-bit) with the kernels ---- */
float oc_u01(uint32_t x);
void  oc_sincos2pi(float u, float *s, float *c);
void  oc_sincos_small(float x, float *s, float *c);
float oc_logf(float x);
float oc_expf(float x);
float oc_powf(float x, float y);
float oc_erfinvf(float x);
float oc_gauss_from_u(float u);
/* batch evaluation for tests: fn 0 u01(bits), 1 sin2pi, 2 cos2pi, 3 sin_small, 4 cos_small,
   5 log, 6 exp, 7 erfinv, 8 gauss_from_u, 9 sqrt */
/* fn 10 and 11 take pairs (in[2i], in[2i+1]): division and atan2, n/2 outputs */
void  oc_math_batch(int fn, const float *in, float *out, int n);
/* the FP64 model's erfinv (Newton on libm's erf from oc_erfinvf), element-wise */
void  oc_erfinv_d_batch(const double *in, double *out, int n);
/* the FP64 model's Buie table: cx, cy [ng + 1], slopes [ng], c = {int465, A, B, C, 1/(gamma+1)};
   and its sunshape angle (mrad) of draws u, element-wise */
void  oc_buie_table_d(double csr, int ng, double *cx, double *cy, double *sl, double c[5]);
void  oc_buie_theta_d_batch(double csr, int ng, const double *u, double *th, int n);
void  oc_philox_batch(const uint32_t *ctr /*[n][4]*/, const uint32_t key[2], uint32_t *out, int n);

/* ---- scene construction: SolarScene::InitContent (solar_scene.cpp:83-99) ---- */
int  oc_scene_build(const oc_scene_desc *desc, const oc_params *prm, oc_scene **out);
void oc_scene_free(oc_scene *s);

/* accessors used by parity tests against the product's scene arrays */
int  oc_scene_counts(const oc_scene *s, int64_t out[8]);
/* out[0]=n_helio out[1]=grid cells out[2]=nnz out[3]=res_x out[4]=res_y
   out[5]=total microsamples out[6]=total rays out[7]=max microsamples per heliostat */
void oc_scene_helio(const oc_scene *s, float *normal, float *v0, float *v1, float *v3);
void oc_scene_grid(const oc_scene *s, int32_t *cell_start, int32_t *cell_items, int32_t grid_num[3]);
/* grid k: CSR, cells per axis, nnz; returns -1 for no such grid */
int  oc_scene_grid_k(const oc_scene *s, int k, int32_t *cell_start, int32_t *cell_items, int32_t grid_num[3],
                     int64_t *nnz);
void oc_scene_receiver(const oc_scene *s, float *v4x3, float focus[3], float *pixel_length);
void oc_scene_sun(const oc_scene *s, float sun_dir[3], float *buie_consts /*[6]*/);
int  oc_scene_buie_table(const oc_scene *s, float *cdf_x, float *cdf_y, float *slope);
void oc_scene_micro(const oc_scene *s, int32_t *count, int64_t *ray_offset);
int  oc_scene_poisson(const oc_scene *s, int h, float *pts /*[2*cap]*/, int cap);
/* Heliostat h's Poisson cell grid before compaction (poisson_sample's array after the last
 * phase, poisson_disk.cu:93-201): dims = (cells per row, rows) including the padding ring,
 * rc = (r, cell edge), grid[2 * (y * dims[0] + x)] the point of cell (x, y) or -2 cell for
 * none.  Returns the accepted point count, or the cell count when grid is NULL or cap is
 * smaller than it. */
int  oc_scene_poisson_grid(const oc_scene *s, int h, float *grid, int cap, int dims[2], float rc[2]);
void oc_scene_flux_scale(const oc_scene *s, float *scale, float *total_energy);

/* ---- tracing ---- */
typedef struct {
    int64_t rays;          /* rays traced                                                */
    int64_t shaded;        /* stopped by the shading test                                */
    int64_t blocked;       /* stopped by the blocking test                               */
    int64_t hit;           /* reached the receiver                                       */
    int64_t dda_steps;     /* cells entered by both DDAs (reference walk to grid exit)   */
    int64_t dda_steps_eff; /* cells entered up to the occupied-band early exit           */
    int64_t cand_tests;    /* candidate parallelogram tests                              */
    int64_t self_skips;    /* candidate == own heliostat, skipped                        */
    /* the rays of tagged heliostats (oc_scene_set_tags) and their shading walks' share of
     * dda_steps_eff, cand_tests and self_skips (a work count only: walks and results are
     * the same for every heliostat)                                                      */
    int64_t tag_rays, tag_shade_steps_eff, tag_shade_tests, tag_shade_self_skips;
    int64_t self_hits;     /* walks stopped by the ray's own heliostat (literal_self only)   */
} oc_stats;

/* Literal self-intersection semantics of Intersect (rectgrid_raytracing.cu:21-36): on = 1
 * tests the ray's own heliostat like every other candidate, so that an FP32 t >= 1e-6 at
 * the ray's own mirror (global_function.cuh:102) stops the walk; on = 0 (default) skips it. */
int oc_scene_set_literal_self(oc_scene *s, int on);

/* Philox stream (scene_index) of the rays from now on; the microsamples keep the stream the
 * scene was built with.  This is how a scene fork draws (om_scene_fork: the root's
 * microsamples, the fork's rays). */
int oc_scene_set_ray_stream(oc_scene *s, uint32_t scene_index);

/* Tags heliostats for oc_stats' tag_* counts (tags[H], nonzero = tagged; NULL clears). */
int oc_scene_set_tags(oc_scene *s, const int32_t *tags);

/* FP32 reference-semantics trace of heliostats [h0, h1).
 * counts[res_y*res_x] (u32) and energy[res_y*res_x] (u64 fixed point, 2^-31 units) are
 * accumulated into (not cleared).  rec_* (may be NULL) are indexed by the ray id
 * ray_offset[h] - ray_offset[h0] + m*Nc + l and hold -1 when empty: shade = blocker of the
 * sun ray, block =
 * blocker of the reflected ray (-2 when not evaluated), pixel = receiver pixel (-1 miss).
 * early_exit selects the occupied-band early exit (results are identical either way).
 * nthreads <= 0 uses OpenMP's default. */
int oc_trace_f32(const oc_scene *s, int h0, int h1,
                 uint32_t *counts, uint64_t *energy,
                 int32_t *rec_shade, int32_t *rec_block, int32_t *rec_pixel,
                 int early_exit, int nthreads, oc_stats *stats);

/* FP64 evaluation of the same model for heliostats [h0, h1): energy64 accumulated per
 * pixel (double), counts per pixel, optional records as above. */
int oc_trace_f64(const oc_scene *s, int h0, int h1,
                 uint32_t *counts, double *energy64,
                 int32_t *rec_shade, int32_t *rec_block, int32_t *rec_pixel,
                 int nthreads, oc_stats *stats);

/* Statistical comparator: the reference's pooled sampling scheme (two pools of
 * groups*lights directions rebuilt per heliostat, random group offsets, perturbation
 * index of microsample m-1) driven by an independent Philox key.  Returns intercepted
 * energy per heliostat (sum of |s.n| over receiver hits) into e_helio[h1-h0]. */
int oc_trace_pooled(const oc_scene *s, int h0, int h1, int groups, int lights,
                    uint64_t pool_seed, double *e_helio, int nthreads);

/* fixed-point energy map -> FP32 raw map (same ops as the finalize kernel) */
void oc_energy_to_float(const uint64_t *energy, int n, float *out);

/* trimmed-mean smoothing: ImageSmoother::image_smooth (image_smooth.cu:10-115) with the
 * two-heap scheme of heap.cuh:9-212.  in and out may not alias. */
int oc_trimmed_mean(const float *in, float *out, int width, int height,
                    int kernel_radius, float trimmed_ratio);

/* intercept ratio: ImageSaver::savetxt(..., float total) (image_save.cpp:26-47) */
float oc_intercept_ratio(const float *flux, int width, int height, float total_energy);

/* DDA probe for unit tests: visits of NotCollision (rectgrid_raytracing.cu:38-117).
 * Writes up to cap visited cell addresses (after the start cell) and returns the number
 * visited until exit or hit; *hit_helio = blocker or -1.  self = heliostat to skip. */
int oc_dda_probe(const oc_scene *s, const float orig[3], const float dir[3], int self,
                 int32_t *cells, int cap, int *hit_helio);

/* geometry probes for unit tests (global_function.cuh:14-105, receiver.cuh:79-82) */
void oc_local2world(const float d[3], const float n[3], float out[3]);
void oc_rotate_y(const float o[3], const float old_dir[3], const float new_dir[3], float out[3]);
int  oc_ray_parallelogram(const float orig[3], const float dir[3], const float A[3],
                          const float B[3], const float C[3], float tuv[3]);
void oc_sample_ray(const oc_scene *s, int h, int m, int l, float sun_to[3], float normal[3]);
/* Differentiable model in FP64 (no reference counterpart): splat map S[res_y*res_x] and
 * energy E of the visible rays of heliostats [h0, h1) at parameters theta[H][4] =
 * (aim offset along e1, e2 in m; tracking deltas along fu, fv in mrad), accumulated into
 * S and *E.  vis: one byte per ray of [h0, h1) in ray order (1 = visible). */
int  oc_diff_splat64(const oc_scene *s, const double *theta, int h0, int h1, const uint8_t *vis,
                     double *S, double *E, int nthreads);
/* cylinder receiver pixel of n rays (new component; no reference counterpart) */
int  oc_cylinder_pixel_batch(const float *orig, const float *dir, int n, const float base[3], float r2,
                             float h, int res_x, int res_y, int32_t *out);
/* truncated-cone receiver pixels (cone_pixel3) of n rays */
int  oc_cone_pixel_batch(const float *orig, const float *dir, int n, const float base[3], float r0, float k,
                         float h, float inv_area, int res_x, int res_y, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif
