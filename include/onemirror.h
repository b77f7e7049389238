/*
 * onemirror.h -- C ABI of the MI355X (gfx950) heliostat-field Monte-Carlo ray tracer.
 *
 * The reference (Hu_smooth/ClassedSunRay) exposes its tracer as C++ free functions and
 * static methods over a singleton scene.  Each entry point below replaces one of them;
 * the comment names the reference interface (file:line, root ClassedSunRay/).
 *
 * Conventions
 *   - Every function returns an om_status; nothing calls exit().  om_last_error() gives a
 *     message for the last failure on the calling thread.
 *   - Device pointers are plain addresses in the memory of the context's GPU; sizes are
 *     element counts.  Stream arguments are hipStream_t passed as void* (NULL = default).
 *   - Trace and smoothing calls are ordered on the given stream and do not synchronise;
 *     om_results_with_given_normal is synchronous, like the reference driver.
 *   - Handles are thread-compatible: one thread per scene at a time.
 */
#ifndef ONEMIRROR_H
#define ONEMIRROR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    OM_OK = 0,
    OM_ERR_INVALID = 1,     /* bad argument                                     */
    OM_ERR_IO = 2,          /* file could not be opened                          */
    OM_ERR_PARSE = 3,       /* malformed scene file                              */
    OM_ERR_HIP = 4,         /* HIP runtime error                                 */
    OM_ERR_NOMEM = 5,       /* device or host allocation failed                  */
    OM_ERR_UNSUPPORTED = 6, /* configuration outside what the library supports   */
    OM_ERR_SIZE = 7         /* caller buffer too small                           */
} om_status;

typedef struct om_ctx om_ctx;
typedef struct om_scene om_scene;

/* One grid of a field (RectGrid pos_/size_/interval_, grid.h:25-49; the Grid block of a
 * .scn file, scene_file_proc.cpp:110-138,185-219) and the heliostats it holds,
 * [first_helio, first_helio + n_helio) of the scene's list (start_helio_pos_, num_helios_). */
typedef struct {
    float pos[3];
    float size[3];
    float interval[3];
    int   first_helio;
    int   n_helio;
} om_grid_desc;

/* Scene description: what a .scn file carries (scene_file_proc.cpp:41-286) plus the sun
 * of SunRay (sunray.h:4-60).  World frame x = east, y = up, z = south (main.cpp:71-73).
 * Grids: with n_grids == 0, grid_pos/size/interval describe the one grid, which holds
 * every heliostat; with n_grids > 0, grids[k] hold consecutive heliostat ranges that cover
 * [0, n_helio) in order (a .scn file lists each grid's heliostats after its block) and
 * grid_pos/size/interval are not read.  Which grids a scene traces: om_params.grid_select. */
typedef struct {
    int          n_helio;
    const float *helio_pos;    /* [n][3] heliostat centre                                 */
    const float *helio_size;   /* [n][3] width (x), thickness (y), height (z)             */
    const float *helio_gap;    /* [n][2] gap between facets                               */
    const int   *helio_matrix; /* [n][2] facet rows, facet columns                        */
    int   recv_type;           /* 0 rectangle (receiver.cuh:66-101); 1 cylinder and 2
                                  truncated cone (new: both are empty stubs in the
                                  reference, receiver.cuh:103-140)                        */
    float recv_pos[3];         /* rectangle centre; cylinder, cone: axis point at mid-height */
    float recv_size[3];        /* rectangle box size; cylinder (radius, height, unused);
                                  cone (top radius, bottom radius, height), the order of
                                  topradius_bottomradius_hight_ (receiver.cuh:138)         */
    float recv_normal[3];
    int   recv_face;           /* 0..3 (receiver.cpp:30-48)                               */
    float recv_pixel_length;   /* solarenergy::receiver_pixel_length (common_var.cpp:12)  */
    float grid_pos[3];         /* RectGrid pos_/size_/interval_ (grid.h:25-49)            */
    float grid_size[3];
    float grid_interval[3];
    float sun_dir[3];          /* direction the light travels                             */
    float dni;                 /* W/m^2                                                   */
    int   recv_res_x, recv_res_y; /* cylinder / cone map columns (azimuth) and rows
                                     (height; the cone's rows cover equal areas);
                                     0 = from recv_pixel_length                           */
    int   n_grids;             /* 0: the grid_* fields above; else the grids array        */
    const om_grid_desc *grids; /* [n_grids]                                               */
    float ground_length, ground_width; /* "ground" line (SolarScene ground_length_/width_;
                                          not used in tracing)                            */
    const float *helio_normal; /* [n][3] given heliostat normals (InitContent(normals),
                                  solar_scene.cpp:101-115; CRotate(dir), heliostat.cpp:22-27),
                                  normalised and kept for every sun; NULL: each heliostat is
                                  aimed at the receiver focus (CRotate(focus, sun))       */
} om_scene_desc;

enum { OM_SUN_PILLBOX = 0, OM_SUN_BUIE = 1 };
enum { OM_MICRO_LATTICE = 0, OM_MICRO_POISSON = 1 };

/* Tracing parameters: the solarenergy:: globals (common_var.cpp:9-27) plus the per-ray
 * Philox sampling that replaces the cuRAND pools (scene_instance_process.cu:6-226). */
typedef struct {
    uint64_t seed;             /* Philox key                                              */
    uint32_t scene_index;      /* Philox stream of this scene (batched runs)              */
    int   sunshape;            /* OM_SUN_*                                                */
    float pillbox_half_angle;  /* rad                                                     */
    float csr;                 /* Buie circumsolar ratio                                  */
    int   buie_bins;           /* CDF bins on [0, 4.65] mrad (reference: 1024)            */
    float slope_sigma;         /* rad (disturb_std)                                       */
    int   micro_mode;          /* OM_MICRO_*; the reference default is the lattice        */
    float helio_pixel_length;  /* microsample pitch (helio_pixel_length)                  */
    int   poisson_k;           /* dart trials per cell (steps_for_raytracing.h:11)        */
    int   rays_per_sample;     /* Nc, sun rays per microsample                            */
    float reflectivity;        /* reflected_rate                                          */
    int   smooth_radius;       /* trimmed-mean window radius (Gtest.cu:50)                */
    float smooth_ratio;        /* trimmed fraction (Gtest.cu:51)                          */
    int   accumulate;          /* OM_ACC_*: how hits reach the maps (results identical)   */
    int   pack_shift;          /* OM_ACC_PACKED: energy-field bits E (32..62, 0 = 44)     */
    int   grid_select;         /* k >= 0: the scene is grid k and its heliostats alone, as
                                  the reference traces grid 0 (Gtest.cu:12,21-23; default
                                  0); OM_GRIDS_ALL: every grid, each heliostat's rays
                                  walking the grid that holds it                          */
    int   literal_self;        /* 0 (default): a walk skips the ray's own heliostat, the
                                  exact-arithmetic reading of the t >= 1e-6 test of
                                  rayParallelogramIntersect (global_function.cuh:102), since
                                  the ray starts on that mirror's plane; 1: the own heliostat
                                  is tested like any candidate, as the reference's Intersect
                                  loop does (rectgrid_raytracing.cu:21-36), so that an FP32
                                  rounding of t to >= 1e-6 stops the ray at its own mirror.
                                  Not supported by the differentiable mode.              */
    int   precision;           /* OM_PREC_FP32 (default): the reference's FP32 ray geometry,
                                  bit-exact against the host restatement; OM_PREC_FP64: the
                                  same Philox draws with sampling, frames, walks,
                                  intersections and receiver in double (the FP64 model the
                                  FP32 error is measured against), energies summed exactly
                                  in the same fixed point.  Plain traces, per-heliostat
                                  traces, records and per-scene batches; not the persistent
                                  batch launch or the differentiable mode.
                                  OM_PREC_MIXED: the FP64 model's outcome of every ray at
                                  close to FP32 cost.  Sampling, both walks and the receiver
                                  hit in FP32, each direction an FP32 deviation from the
                                  model's nominal reflection of its heliostat (computed in
                                  double); every FP32 decision (start cell, walk step,
                                  candidate test, pixel) is taken only outside its error
                                  bound, and the rays with a decision inside it are traced
                                  again in FP64.  Shading, blocking and pixel of every ray
                                  are the FP64 model's; energies differ from it by the FP32
                                  draws (about 1e-8 relative).  Rectangle and cylinder receivers with
                                  OM_ACC_WINDOW; scenes with the cone receiver or
                                  literal_self, and other accumulation modes, trace in
                                  FP64.  Same entry points as OM_PREC_FP64.              */
    int   ray_compact;         /* OM_ACC_WINDOW, FP32 geometry: 1 packs the rays that pass
                                  the blocking test into full waves (ballot, mbcnt and one
                                  ds_permute per ray value) before the receiver stage;
                                  0 (default): each wave runs the receiver on the lanes its
                                  walks leave alive.  Results are identical.              */
    int   redo_cap;            /* OM_PREC_MIXED: entries of a launch's list of rays for the
                                  FP64 trace (0, the default: one per 32 rays, 2^16 to
                                  2^24).  A launch whose list fills up is traced again
                                  wholly in FP64: the same results, more time.           */
} om_params;

enum { OM_PREC_FP32 = 0, OM_PREC_FP64 = 1, OM_PREC_MIXED = 2 };

enum { OM_GRIDS_ALL = -1 };

/* OM_ACC_PACKED: one 64-bit atomic per hit into a per-XCD word holding (count << E) +
 * energy; carries out of the energy field and wraps of the word are detected from the
 * returned old value and counted on the side, and a reduction kernel decodes the exact
 * counts and energies.  OM_ACC_DEVICE: two device-scope atomics (count, energy) per hit
 * straight into the output maps.  OM_ACC_WINDOW: packed words as OM_ACC_PACKED, but one
 * persistent workgroup per CU keeps the central block of the map (up to 18,400 pixels) in
 * LDS for all the tiles it traces, so only hits outside it reach device memory. */
enum { OM_ACC_PACKED = 0, OM_ACC_DEVICE = 1, OM_ACC_WINDOW = 2 };

/* Output maps of one receiver, row-major [res_y][res_x], all device memory. */
typedef struct {
    uint32_t *counts;          /* rays per pixel                                          */
    uint64_t *energy;          /* sum of |s.n| per pixel, fixed point in 2^-31 units      */
    float    *flux;            /* W/m^2, NULL to skip                                     */
    float    *flux_smooth;     /* trimmed-mean smoothed W/m^2, NULL to skip               */
    int64_t   capacity;        /* elements available in each non-NULL array               */
} om_flux;

typedef struct {
    int     n_helio;
    int     res_x, res_y;
    int     grid_num[3];
    int64_t grid_cells, grid_nnz;
    int64_t total_micro;       /* microsamples over all heliostats                        */
    int64_t total_rays;        /* rays one full trace draws                               */
    int     max_micro;
    float   pixel_length;      /* receiver pixel edge                                     */
    float   flux_scale;        /* DNI * Ssub * rho / Nc / Srec (Gtest.cu:37-46)           */
    float   total_energy;      /* sum |n.s| DNI A rho / Srec (Gtest.cu:29,44)             */
    int     n_grids;           /* grids traced; grid_num is grid 0's, cells and nnz sums  */
    float   pixel_area;        /* Srec, the area of one receiver pixel (m^2)              */
} om_scene_info;

/* Host-side result of the full driver. */
typedef struct {
    float *flux;               /* host [res_y*res_x], may be NULL                         */
    float *flux_smooth;        /* host [res_y*res_x], may be NULL                         */
    float  ratio;              /* intercept ratio, unsmoothed (Gtest.cu:47)               */
    float  ratio_smooth;       /* intercept ratio, smoothed (Gtest.cu:59)                 */
    float  total_energy;
    int64_t rays;
} om_result;

const char *om_status_string(om_status s);
const char *om_last_error(void);
void        om_default_params(om_params *p);

/* Context on one GPU. (Replaces the process-global cuRAND generator set in main.cpp:51-54
 * and the implicit current device.) */
om_status om_ctx_create(int device, om_ctx **out);
om_status om_ctx_destroy(om_ctx *ctx);

/* Abort and progress of the persistent launches of a context: the window traces
 * (OM_ACC_WINDOW, every precision) and both persistent batch launches.  No reference
 * counterpart (the reference exits the process on any error, utils.h:13-22); the caller's
 * watchdog uses them from any thread while a launch runs.
 * om_ctx_abort(ctx, 1) copies a nonzero word into device memory on a stream of the
 * context's own; every wave reads it each time it takes a chunk of work (one system-scope
 * load beside the queue atomic).  The running launch, and each later one until
 * om_ctx_abort(ctx, 0), stops within one chunk per wave.  Its maps then hold only the
 * chunks traced so far. */
om_status om_ctx_abort(om_ctx *ctx, int on);
/* Chunks of work taken so far by the persistent launch of this context that runs now, or
 * that ran last, and its total.  The launch writes both into a queue block on its own
 * stream, and this call copies the block on a stream of the context's own, so the two
 * always belong to one launch.  An aborted launch counts at most one chunk per wave that
 * it took but did not trace.  Do not call it while om_batch_destroy runs on the context. */
om_status om_ctx_progress(om_ctx *ctx, int64_t *taken, int64_t *total);

/* Scene from a .scn file + SolarScene::InitContent: SceneFileProc::SceneFileRead
 * (scene_file_proc.cpp:41) then sun, grid, receiver, heliostat aiming
 * (solar_scene.cpp:83-99).  sun_dir and dni are the solarenergy globals. */
om_status om_scene_from_scn(om_ctx *ctx, const char *path, const float sun_dir[3], float dni,
                            const om_params *params, om_scene **out);

/* Parse a .scn file into a description whose arrays live until om_scene_desc_free.  Every
 * Grid block becomes an entry of desc->grids (grid 0 is also copied to grid_pos/size/
 * interval); the first receiver is kept (the reference traces receivers[0], Gtest.cu:13).
 * Where the reference only prints a warning and goes on into undefined behaviour, parsing
 * fails with OM_ERR_PARSE: a grid whose "n" differs from the heliostats listed under it
 * (scene_file_proc.cpp:112-115,250-252), heliostat type other than 0 (:253), a receiver type
 * other than 0..2 (:96-101) or a grid type other than 0 (:123-126, :220-222). */
om_status om_scn_parse(const char *path, om_scene_desc *desc);
void      om_scene_desc_free(om_scene_desc *desc);

/* Write a description as a .scn file that SceneFileProc::SceneFileRead reads back to the
 * same description (numbers with 9 significant digits, which round-trip every float): the
 * ground, receiver and grid blocks, and each grid's heliostats with "gap" and "matrix"
 * lines where they change.  No reference counterpart (the reference only reads scenes). */
om_status om_scn_write(const char *path, const om_scene_desc *desc);

/* Scene from arrays (generated fields) + InitContent (solar_scene.cpp:83-99).  With
 * ctx == NULL the scene is built on the host only (aiming, grid, receiver, sun tables) for
 * inspection through om_scene_export/om_scene_get_info; it cannot be traced, and its
 * Poisson microsample counts are 0 because those points are drawn on the GPU. */
om_status om_scene_create(om_ctx *ctx, const om_scene_desc *desc, const om_params *params,
                          om_scene **out);
om_status om_scene_destroy(om_scene *scene);
om_status om_scene_get_info(const om_scene *scene, om_scene_info *info);

/* New sun position: renormalise the sun and re-aim every heliostat at the receiver focus
 * (set_sunray_content + set_helio_content, scene_instance_process.cpp:21-28); heliostats
 * with given normals (om_scene_desc.helio_normal) keep them. */
om_status om_scene_set_sun(om_scene *scene, const float sun_dir[3], float dni);

/* ---- Annual time steps (main.cpp:56-116 builds a new scene for every sun position of the
 * year on the same field).  A fork is that scene for another sun direction, DNI and Philox
 * stream (scene_index): it shares the source's grid, microsamples, tiles and Buie tables on
 * the device and owns its re-aimed heliostats, so that hundreds of time steps of one field
 * fit side by side (for a batch).  The source (or the root it was forked from) must outlive
 * its forks.  The microsamples are the root's (Poisson points drawn with the root's scene
 * index); the rays use the fork's scene_index.  With the root's scene index, tracing a fork
 * gives what a scene built from scratch with that sun gives.  Given heliostat normals are
 * kept, as by om_scene_set_sun. */
om_status om_scene_fork(const om_scene *src, const float sun_dir[3], float dni, uint32_t scene_index,
                        om_scene **out);
/* Light travel direction of the sun at zenith and azimuth (degrees, azimuth clockwise from
 * north) as main.cpp:71-73 sets sun_dir: (-sin z sin az, -cos z, sin z cos az). */
om_status om_sun_direction(float zenith_deg, float azimuth_deg, float out[3]);
/* Clear-sky DNI of getDNI (SPA.cpp:1174-1189), W/m^2: (1353 + 45 cos(2 pi (n + 10) / 365))
 * (a0 + a1 exp(-k / cos z)), n the day of a 365-day year from month (1-12) and day, and
 * Hottel's a0, a1, k for a site at elevation_m.  NaN for a month out of range. */
float om_dni_hottel(float zenith_deg, int month, int day, float elevation_m);

/* ---- The faceted heliostat of the annual driver (main.cpp:56-116): nine flat facets on a
 * 3 x 3 lattice of pitch sqrt(2)/3 + 0.1 m, lifted onto the paraboloid y = x^2/a^2 + z^2/b^2,
 * with the paraboloid's normals, turned by the matrix whose column 1 is the master normal
 * and placed at (X, Y, Z) (SceneFileProc::SceneFileRead(SolarScene*, a, b, mirrorNormal, X,
 * Y, Z), scene_file_proc.cpp:288-339,464-487).  The matrix's horizontal column is scaled by
 * sqrt(n0^2 + n1^2) as the reference writes it (:302), which is a unit vector only for n1 = 0.
 * Outputs pos, normal (not normalised) and size, [9][3] each. */
om_status om_facet_heliostat(float a, float b, const float mirror_normal[3], float X, float Y, float Z,
                             float pos[27], float normal[27], float size[27]);
/* Master normal of the annual driver (main.cpp:84-91): the bisector of the direction from
 * (X, Y, Z) to the point (0, 13.01, 0) and the direction to the sun (-sun_dir). */
om_status om_master_normal(float X, float Y, float Z, const float sun_dir[3], float out[3]);
/* The whole faceted scene as the reference builds it (scene_file_proc.cpp:288-534): the
 * nine facets with their normals given, gap 0, matrix 1 x 1, a 1.4 x 1.4 m receiver at
 * (0, 13, -1.5) facing +z (face 0), one grid of 110 m cells at (-5.5, -5.5, 94.5) of size
 * (110, 110, 220), pixel length 0.01, sun straight down and DNI 1000 (set sun_dir and dni
 * before building).  The description points into the caller's arrays. */
om_status om_facet_scene_desc(float a, float b, const float mirror_normal[3], float X, float Y, float Z,
                              float pos[27], float normal[27], float size[27], float gap[18], int matrix[18],
                              om_scene_desc *desc);

/* Copy scene arrays to host buffers for inspection (parity tests).  what:
 *  "helio_normal" [n][3] f32   "helio_v0" / "helio_v1" / "helio_v3" [n][3] f32
 *  "grid_start" [cells+1] i32  "grid_items" [nnz] i32 (grid 0; "grid_start.K",
 *  "grid_items.K" for grid K, "grid_num.K" i32[3] its cells per axis, "grid_helio.K" i32[2]
 *  its first heliostat and count)                        "receiver" [4][3] f32 + focus[3]
 *  "buie" cdf_x[b+1], cdf_y[b+1], slope[b] f32         "micro_count" [n] i32
 *  "ray_offset" [n+1] i64      "poisson" [n][max_micro][2] f32 (zero padded)
 *  "sun" sun[3] + buie constants int465, A, B, C, 1/(gamma+1)
 *  "helio_clear" [n] i32: 1 where no ray toward the sun from that mirror can meet another
 *  heliostat (its shading walks are skipped; they would all return -1)   */
om_status om_scene_export(const om_scene *scene, const char *what, void *host, size_t bytes);

/* Full trace of the field into caller maps (cleared first), then flux scaling and, when
 * flux_smooth is set, trimmed-mean smoothing with params.smooth_radius/ratio.  This is the
 * device part of results_with_given_normal (Gtest.cu:10-63): the per-heliostat loop of
 * recthelio_ray_tracing (recthelio_tracing.cpp:46-79) -> ray_tracing (ray_tracing.cu:3-32)
 * -> map_tracing (rectgrid_raytracing.cu:155-196) in one launch. */
om_status om_trace_plain(om_scene *scene, const om_flux *out, void *stream);

/* Allocation-free variant of om_trace_plain, with no counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same outputs as om_trace_plain, bit for bit; it
 * calls no hipMalloc or hipFree.  Its workspace is a static device array of
 * OM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: the packed map copies (8 to
 * 64 of 8 bytes per pixel, fewer when space is short), 8 bytes per pixel of side maps, in
 * OM_ACC_WINDOW mode one window slice of up to 147,200 bytes per persistent workgroup (one
 * per CU), and with flux_smooth 4 bytes per pixel of smoothing input.  A scene whose
 * workspace does not fit gets OM_ERR_SIZE before anything is launched (with 256 CUs: maps up
 * to about 880,000 pixels in OM_ACC_PACKED mode, 400,000 in OM_ACC_WINDOW mode).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race and give wrong maps. */
#define OM_NOALLOC_SCRATCH_BYTES (64u << 20)
om_status om_trace_plain_noalloc(om_scene *scene, const om_flux *out, void *stream);

/* Trace heliostats [h0, h1) and accumulate (no clear) into counts/energy: the per-heliostat
 * entry recthelio_ray_tracing (recthelio_tracing.cpp:46-79; recthelio_tracing.h:16). */
om_status om_trace_heliostats(om_scene *scene, int h0, int h1, const om_flux *out, void *stream);

/* Per-ray records for parity checks: for every ray of heliostats [h0, h1) (indexed by
 * ray_offset[h] - ray_offset[h0] + m*Nc + l) the blocking heliostat of the sun ray (shade),
 * of the reflected ray (block; -2 when not traced) and the receiver pixel (-1 for none).
 * Device int32 arrays of at least ray_offset[h1]-ray_offset[h0] elements. */
om_status om_trace_records(om_scene *scene, int h0, int h1, int32_t *shade, int32_t *block,
                           int32_t *pixel, int64_t capacity, void *stream);

/* OM_PREC_MIXED: the number of rays of the scene's last trace (or records call) that a
 * decision within its error bound sent to the FP64 trace.  Waits for the device.  No
 * counterpart in the reference (its traces are FP32 only). */
om_status om_scene_redo_count(const om_scene *scene, int64_t *rays);

/* Energy/fixed point -> flux maps (the scaling loop of Gtest.cu:37-46) for a scene. */
om_status om_flux_finalize(om_scene *scene, const om_flux *out, void *stream);

/* ImageSmoother::image_smooth (image_smooth.cuh:20, image_smooth.cu:60-115): trimmed
 * mean over a (2r+1)^2 clipped window with the two-heap update of heap.cuh.  d_in and
 * d_out may be the same buffer (the reference works in place). */
om_status om_smooth_trimmed(om_ctx *ctx, const float *d_in, float *d_out, int width, int height,
                            int kernel_radius, float trimmed_ratio, void *stream);

/* results_with_given_normal (Gtest.cu:10-63; Gtest.cuh:8): trace, scale, intercept ratio,
 * smooth, scale, ratio.  Synchronous; host outputs. */
om_status om_results_with_given_normal(om_scene *scene, om_result *result);

/* Accumulation autotuner, no reference counterpart: traces the whole scene into maps with
 * OM_ACC_PACKED and with OM_ACC_WINDOW (identical maps; a mode the scene cannot use is
 * skipped, time -1), best of reps timed traces each after a warm-up, and keeps the faster
 * mode for the scene's later traces.  ms[2] (may be NULL) receives the two times, best the
 * mode kept.  Synchronous, on the null stream; the maps are overwritten. */
om_status om_autotune_accumulate(om_scene *scene, const om_flux *maps, int reps, float *ms, int *best);

/* Intercept ratio of a host flux map (ImageSaver::savetxt, image_save.cpp:26-47). */
float om_intercept_ratio(const float *flux, int width, int height, float total_energy);

/* ---- Text outputs of the reference driver (both append to the file, as it does). */
/* ImageSaver::savetxt(filename, w, h, h_data, precision) (image_save.cpp:5-24): the host map
 * [height][width] written last row first, one line per row, each pixel "0," below 1e-6 and
 * "%.<precision>f," otherwise (the reference's default precision is 2). */
om_status om_savetxt_map(const char *path, int width, int height, const float *map, int precision);
/* ImageSaver::savetxt(filename, w, h, h_data, totalEnergy) (image_save.cpp:26-47), what
 * results_with_given_normal writes per trace (Gtest.cu:47,59): one line holding the
 * intercept ratio (om_intercept_ratio) in default float formatting ("%g"). */
om_status om_savetxt_ratio(const char *path, int width, int height, const float *map, float total_energy);
/* load_noraml (scene_normals.cpp:3-13): a count N then N normals "x y z" into out[3N] (the
 * om_scene_desc.helio_normal of InitContent(normals)).  *n receives N; out == NULL only
 * counts; fewer readable normals than N is OM_ERR_PARSE. */
om_status om_load_normals(const char *path, float *out, int capacity, int *n);

/* ---- Batched mode: no reference counterpart (the reference traces one scene at a time
 * from a host loop over heliostats, recthelio_tracing.cpp:46-79).  n scenes (fields, sun
 * positions: annual time steps) are traced into n sets of maps by one persistent launch that
 * takes (scene, heliostat, sample block) items from a work queue; the decode, flux scaling
 * and smoothing of every map follow.  Each map equals what om_trace_plain gives for that
 * scene alone.  Scenes of a batch share the context, the sunshape and the microsample kind. */
typedef struct om_batch om_batch;

typedef struct {
    int workgroup;             /* threads per workgroup: 128, 256 or 512 (0 = 256), or
                                  OM_BATCH_PER_SCENE or OM_BATCH_WINDOW                  */
    int rays_per_thread;       /* 1..8, rays per work item = workgroup x this (0 = auto)  */
} om_launch;
/* om_launch.workgroup: no persistent batch launch; each scene is traced as om_trace_plain
 * traces it with OM_ACC_WINDOW (the LDS-window kernel of 1024 threads), one after another
 * in the batch's graph.  Large scenes run faster this way; the maps are the same. */
enum { OM_BATCH_PER_SCENE = 1024 };
/* om_launch.workgroup: one persistent launch of 1024-thread workgroups, one per CU, that
 * keeps the central block of the current scene's packed map in LDS (as OM_ACC_WINDOW does
 * for one scene).  Waves take chunks of at most 2048 rays of one heliostat from 8 per-XCD
 * queues in scene order; a workgroup adds its window to the scene's maps in device memory
 * when its waves move on to another scene (or grid).  rays_per_thread is ignored.  The maps
 * are the same.  It also takes OM_PREC_MIXED scenes (all scenes of the batch then mixed):
 * the mixed trace's rays with an uncertain decision are listed per scene and grid and traced
 * in FP64 after the launch, and a scene whose list fills up is traced again wholly in FP64. */
enum { OM_BATCH_WINDOW = 2048 };

/* Batch over scenes[0..n) with outputs outs[0..n) (counts and energy required, flux and
 * flux_smooth optional; all overwritten by each run).  The scenes and outputs must outlive
 * the batch; re-aiming a scene (om_scene_set_sun) needs a new batch. */
om_status om_batch_create(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs,
                          const om_launch *launch, om_batch **out);
/* Run the batch on a stream: the first run captures the launch sequence in a hipGraph,
 * every run replays it.  Calls on one batch must not overlap. */
om_status om_batch_run(om_batch *batch, void *stream);
om_status om_batch_info(const om_batch *batch, om_launch *launch, int64_t *items, int64_t *rays);
/* OM_BATCH_WINDOW over OM_PREC_MIXED scenes: the rays of the batch's last run that a decision
 * within its error bound sent to the FP64 trace, and the scenes whose redo list filled up
 * (each then traced wholly in FP64).  Zero for other batches.  Waits for the device.  No
 * counterpart in the reference (its traces are FP32 only). */
om_status om_batch_redo_count(const om_batch *batch, int64_t *rays, int64_t *full_scenes);
om_status om_batch_destroy(om_batch *batch);
/* One-shot batch without a graph; synchronous (waits for the stream). */
om_status om_trace_batched(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs,
                           const om_launch *launch, void *stream);

/* Autotuner of the batch launch: times every workgroup size in {128, 256, 512} x rays per
 * thread in {2, 4, 8}, OM_BATCH_WINDOW and OM_BATCH_PER_SCENE, on the given batch (best of reps graph
 * replays after a warm-up) and returns the fastest with the whole table.  Results do not
 * depend on these parameters. */
typedef struct {
    om_launch best;
    float     best_ms;
    double    rays_per_s;
    int       n_trials;
    struct { int workgroup, rays_per_thread; float ms; } trials[12];
} om_tune;
om_status om_autotune(om_ctx *ctx, om_scene *const *scenes, int n, const om_flux *outs, int reps,
                      om_tune *result);

/* ---- Differentiable mode: no reference counterpart (the reference traces fixed aims).
 * Parameters per heliostat, float[H][4]: aim-point offset (m) along the receiver's two unit
 * edges e1 = (v1-v0)/|v1-v0| and e2 = (v3-v0)/|v3-v0|, and tracking deltas (mrad) that tilt
 * the aimed normal toward the heliostat frame axes fu, fv:
 *   n = normalize(normalize(normalize(focus + a_u e1 + a_v e2 - pos) - sun)
 *                 + 1e-3 (t_u fu + t_v fv)).
 * Per ray the slope-error deviation of the plain path is added to n and renormalised; the
 * reflected ray's receiver-plane point, in pixel coordinates, is splatted with energy
 * |sun.n'| by a Gaussian of sigma = 1 px tapered to zero value and slope at 3 px.  Rectangle
 * receivers only. */
/* Words (uint64) of the visibility ballots a forward pass writes. */
om_status om_diff_mask_words(const om_scene *scene, int64_t *words);
/* Forward: d_splat double[res_y*res_x] and d_energy double[1] (energy of the visible rays)
 * at d_params; d_mask receives one ballot bit per ray (visible = not shaded, not blocked,
 * receiver plane hit ahead), which the backward pass holds fixed. */
om_status om_diff_forward(om_scene *scene, const float *d_params, uint64_t *d_mask, double *d_splat,
                          double *d_energy, void *stream);
/* Backward (the adjoint kernel): d_grad float[H][4] = dL/dparams from d_dsplat = dL/dS and
 * d_denergy = dL/dE (device), recomputing the visible rays from their Philox counters. */
om_status om_diff_backward(om_scene *scene, const float *d_params, const uint64_t *d_mask,
                           const double *d_dsplat, const double *d_denergy, float *d_grad, void *stream);
/* Visibility of the rays of heliostats [h0, h1) from the ballots, one byte per ray in ray order. */
om_status om_diff_visibility(om_scene *scene, const uint64_t *d_mask, int h0, int h1, uint8_t *d_vis,
                             int64_t capacity, void *stream);

/* Device self-test of the float math as the trace kernel evaluates it (parity of the
 * kernels' arithmetic with the host).  fn: 0 u01 (input bits as uint32), 1/2 sin/cos(2 pi x),
 * 3/4 sin/cos of a small angle, 5 log, 6 exp, 7 erfinv, 8 Gaussian deviate of a uniform,
 * 9 sqrt, 10 division of pairs (out[i] = in[2i] / in[2i+1], i < n/2), 11 atan2 of pairs
 * (out[i] = atan2(in[2i], in[2i+1]), the cylinder receiver's azimuth).  9 and 10 use the
 * kernel's short exact sequences, which hold for x in [2^-90, 2^90] and for numerators 0 or
 * in [2^-90, 2^60] over denominators in [2^-60, 2^60].  12..14 are VALU issue probes, not
 * math: 256 iterations over 8 independent chains per element of 8 v_fma_f32 (12), 8
 * v_fma_f32 + 8 v_add_u32 (13), 8 v_fma_f32 + 2 v_exp_f32 (14) or 8 v_fma_f64 (15) per
 * lane, for measuring the VALU issue rate.  in/out device arrays of n floats.  16: the
 * FP64 geometry's slope-error erfinv(2u - 1) of the draw words in[i] (bits as uint32, u as
 * in 0); out holds n doubles (2n floats).  17..19 are atomic throughput probes: n threads
 * each add 64 times with 64-bit atomics at hashed addresses, into out as n / 2 words of
 * device memory without (17) and with (18) the old value returned, or into an LDS array of
 * the window kernels' size in each 1024-thread workgroup (19, old value returned; out[i]
 * gets a digest).  out is overwritten. */
om_status om_math_eval(om_ctx *ctx, int fn, const float *d_in, float *d_out, int n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
