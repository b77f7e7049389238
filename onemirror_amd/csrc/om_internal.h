// om_internal.h -- scene layout in HBM and the host/device glue of the tracer library.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/onemirror.h"
#include "om_common.h"

namespace om {

// Per-heliostat record, 32 words.  One workgroup traces rays of one heliostat, so the
// record is read with uniform (scalar) loads.
struct alignas(16) HelioRec {
    float pos[3];
    float n[3];
    float fu[3];            // frame of local2world around n (global_function.cuh:41-44)
    float fv[3];
    int ident;              // frame is the identity (|n.x|, |n.z| < Epsilon)
    float sx2, sy2, sz2;    // size / 2
    float gx, gy;           // facet gaps
    float subrow_len;       // subhelio_rowlength_collength.x (steps_for_raytracing.cu:75)
    float subcol_len;       // subhelio_rowlength_collength.y (:76)
    int sub_rows, sub_cols; // microsamples per facet row / column (:79-80)
    int cols_all;           // facet columns * sub_cols
    int mcount;             // microsamples of this heliostat
    int ray_off_lo, ray_off_hi; // global ray id of (m=0, l=0), for records
    int rows, cols;         // facet rows, facet columns
    int clear;              // 1: no ray toward the sun from this mirror can meet another heliostat
    int pad[3];
};
static_assert(sizeof(HelioRec) == 128, "HelioRec is 32 words");

// Kernel arguments of the trace kernels (passed by value).
struct TraceArgs {
    // sun (scene_instance_process.cu:6-226)
    v3 sun_u, sun_n, sun_v;
    int sun_ident;
    float pillbox;
    float buie_int465, buie_A, buie_B, buie_C, buie_inv_gp1;
    int nbins;
    const float4 *buie_bins;    // [nbins]: (cdf_y[i], cdf_y[i+1], slope[i], cdf_x[i])
    const int *buie_guide;      // [nbins]: a bin index at or below the bin of j / gscale
    float buie_gscale;          // nbins / int465
    const int2 *buie_split;     // [nbins]: (guide[j], bits of the upper edge cdf_y[guide[j] + 1])
    float slope_sigma;
    uint32_t key0, key1, w3;
    int L;
    float pitch, pitch2;
    // grid (grid.h:25-49) + occupied index range
    float gpx, gpy, gpz, gix, giy, giz;
    float inv_gix, inv_giy, inv_giz; // exact reciprocals when grid_pow2
    int grid_pow2;              // all cell edges are powers of two
    int gnx, gny, gnz;
    int occ_lo[3], occ_hi[3];
    const int *cell_start;
    const int *cell_items;
    const float4 *cell_geo; // [3 nnz]: (A, bits of the id), (E1, 0), (E2, 0) of cell_items[i]
    const float4 *geo;      // [3H]: (v0, 0), (v1 - v0, 0), (v3 - v0, 0)
    // receiver (receiver.cuh:79-82); cylinder: base of the axis, radius^2, height
    int recv_type;
    v3 rA, rE1, rE2;
    v3 cyl_base;                // cylinder and cone: base of the axis
    float cyl_r2, cyl_h;        // cylinder radius^2; cylinder and cone height
    float cone_r0, cone_k, cone_inv_area; // cone: bottom radius, dr/dy, 1 / (h (r0 + r1))
    int res_x, res_y;
    // heliostats and work
    const HelioRec *helio;
    const float2 *poisson;
    int poisson_cap;
    const int2 *tiles;      // (heliostat, first ray of the tile)
    int tile_base;          // first tile of this launch
    int ntiles;             // tiles of this launch
    int tile_rays;          // rays of a tile (the last tile of a heliostat may hold fewer)
    int rays_per_thread;
    // outputs
    uint32_t *counts;
    unsigned long long *energy;
    int32_t *rec_shade, *rec_block, *rec_pixel;
    long long rec_base;
    // packed accumulation (OM_ACC_PACKED)
    unsigned long long *part; // [8 << copy_shift][npx] packed (count << pack_shift) + energy
    uint32_t *carry;          // [npx] carries out of the energy field
    uint32_t *wrap;           // [npx] wraps of the packed word
    int npx;
    int copy_shift;           // 2^copy_shift map copies per XCD
    int copy_stride;          // words between copies of a pixel
    int px_stride;            // words between pixels of a copy
    int acc_packed;
    int pack_shift;
    // OM_ACC_WINDOW: rows [win_y0, win_y0 + win_h) x columns [win_x0, win_x0 + win_w) of the
    // map are accumulated in LDS by persistent workgroups, which store their windows to
    // win_part[workgroup][win_h * win_w] at the end; chunks of tiles come from 8 queues
    unsigned long long *win_part;
    uint32_t *win_queue;      // queue block (QUEUE_WORDS): [8] chunks taken from each queue
    const uint32_t *abort_flag; // nonzero stops every persistent launch (om_ctx_abort)
    int win_x0, win_y0, win_w, win_h;
    // Window of each queue's workgroups (azimuthal maps, whole-grid traces): the queue of XCD
    // x takes the tiles win_order[first..] (a permutation of the launch's tiles, null: their
    // order) and its workgroups hold rows [win_qy0[x], + win_h) and columns [win_qx0[x],
    // + win_w) modulo res_x; each workgroup records its queue in win_slice_q[workgroup] for
    // the decode (null: every workgroup holds win_x0, win_y0)
    const int *win_order;
    int *win_slice_q;
    int win_qx0[8], win_qy0[8];
    int literal_self;         // walks test the ray's own heliostat (om_params.literal_self)
};

// ---- FP64 geometry (om_params.precision = OM_PREC_FP64) ---------------------------------
// The scene in double as the FP64 model builds it: heliostat centres, aimed normals and
// their local2world frames, sizes and facet layout, the vertex parallelograms of the walk's
// candidates, the receiver and the Buie table, all from the FP32 inputs converted exactly.
struct alignas(16) Helio64 {
    double pos[3], n[3], u[3], v[3]; // centre, normal, frame of local2world around n
    double sz[3];                    // size (x width, y thickness, z height)
    double subrow, subcol;           // facet row and column lengths
    double gx, gy;                   // facet gaps
    int ident;                       // frame is the identity (|n.x|, |n.z| < Epsilon)
    int pad;
};
struct alignas(16) Geo64 {
    double A[3], E1[3], E2[3]; // v0, v1 - v0, v3 - v0
    int id;                    // the heliostat (entries of a grid's CSR list)
    int pad;
};

// Kernel arguments of the FP64 trace beside TraceArgs (which supplies the tiles, the grid's
// CSR and occupied range, the Philox key and stream, the maps and the records).
struct Trace64Args {
    const Helio64 *helio;
    const Geo64 *cell_geo;                  // the launch grid's CSR entries in CSR order
    const double *cdf_x, *cdf_y, *slope;    // Buie table, nbins bins
    const int *guide;                       // [nbins]: the bin of value j / gscale
    // Buie sunshape beyond 4.65 mrad: the polar angle (before the division by 1000) of the
    // sunshape draw u = (2k + 1) 2^-24 for k = tail_k0 .. 2^23 - 1, the draws whose value u A
    // reaches int465, each the same pow as the trace would evaluate (null: no such draw)
    const double *tail;
    uint32_t tail_k0;
    double gscale;                          // nbins / int465
    int nbins;
    double sun_n[3], sun_u[3], sun_v[3];
    int sun_ident;
    double pillbox, sigma, pitch;
    double int465, A, B, C, inv_gp1;
    double gpos[3], gint[3];                // the launch's grid
    double rA[3], rE1[3], rE2[3];           // rectangle receiver: v0, v1 - v0, v3 - v0
    double cyl_base[3], cyl_r, cyl_h, cone_r0, cone_r1;
    // window trace of a mixed launch: runs only when *gate > gate_cap (its redo list filled
    // up); null: always
    const uint32_t *gate;
    uint32_t gate_cap;
};

// ---- Certified FP32 geometry (om_params.precision = OM_PREC_MIXED) ----------------------
// A ray is sampled in FP32; its direction is the FP64 model's nominal reflection of its
// heliostat (HelioMix, from the model's doubles) plus an FP32 deviation, its receiver hit is
// found in FP32 relative to the receiver's reference point, and both walks run in FP32 over
// float images of the model's vertices.  Every decision of the FP32 arithmetic (a walk step's axis, a candidate's hit
// test, the receiver pixel) is taken only when it holds with the error bounds below: a ray
// with a decision inside its bound is traced again by the FP64 trace.  The bounds, for every
// ray of the scene:
//   e_dir   |d - d_model| (2-norm) of a direction formed from the FP32 draws, plus the FP32
//           rounding of its components in the walks: e_dr + 2^-21
//   e_pos   |float(origin) - origin| + |float(vertex) - vertex|: 2^-22 of the largest
//           coordinate of the scene
// (the walks' crossings start from the origin's offset inside its cell, formed in double:
// their error is that of the microsample's offset w from its heliostat, 2^-21 |w|_1)
// Per heliostat, the FP64 model's values rounded to float: the frame of the aimed normal n0
// (the identity axes for an identity frame), the reflection a = i0 - 2 (n0.i0) n0 of the
// nominal light direction i0 (the sun frame's axis), c0 = n0.i0, and the nominal ray's
// point nearest the receiver's reference point F: D = (F - pos).a, qf = pos + D a - F.  A ray
// is then a + dr with dr formed from small quantities only, and the receiver is met at
// F + c + tau d with c = qf + (o - pos) + D dr: no FP32 operand carries a coordinate of the
// field, so the rounding stays at the size of the deviations.
struct alignas(16) HelioMix {
    float u[3], n[3], v[3];
    float a[3];
    float qf[3];
    float c0, D;
    float e0;   // n0.sun: the energy |sun.n| = |e0 + sun.dn|
    float pad[2];
};
static_assert(sizeof(HelioMix) == 80, "HelioMix is 20 words");

struct MixArgs {
    float e_dir, e_pos;
    float gi_max;      // the launch grid's largest cell edge
    float e_dr;        // |dr - dr_model| (2-norm): 2^-19 (theta_max + 4 theta_s,max) + 2^-26
    float buie_vmax;   // the largest float value below the top of the Buie table's image
    // Buie tail draws k >= tail_k0 of Trace64Args: the value's excess over int465,
    // (k - tail_k0) tail_du + tail_dv0 (tail_du = 2^-23 A, tail_dv0 = u(tail_k0) A - int465 in
    // double), free of the cancellation of u A - int465 in float
    float tail_du, tail_dv0;
    double rgi[3];     // 1 / the launch grid's cell edges, rounded (the start cell's quotients)
    const HelioMix *hmix;
    v3 s_u, s_n, s_v;  // the model's sun frame rounded (identity axes for an identity frame)
    v3 sun;            // the sun direction rounded (the energy's)
    // rectangle receiver: unit normal, E2 x E1 (det = d.nr), dual vectors of E1 and E2 in the
    // plane (u = uF + (P - F).g1, v = vF + (P - F).g2) with their 1-norms
    v3 nn, nr, g1, g2;
    float uF, vF, g1n, g2n, nrn;
    // cylinder: radius, height, half height (F is the axis point at mid-height)
    float cyl_r, cyl_h;
    uint2 *redo;       // [redo_cap] (heliostat, ray of that heliostat) to trace again in FP64
    uint32_t *redo_n;  // [0]: rays put on the list this launch (may exceed redo_cap);
                       // [1]: rays put on the lists of the scene's last mixed trace;
                       // [2]: launches whose list filled up (traced again wholly in FP64)
    int redo_cap;
};

// One map of a batch for the decode (reduce_batch_kernel): packed copies and side maps,
// outputs, and the flux scale.
struct ReduceSlot {
    unsigned long long *part;
    uint32_t *carry, *wrap;
    int E, npx, ncopy, copy_stride, px_stride;
    uint32_t *counts;
    unsigned long long *energy;
    float *raw;   // float(energy * 2^-31), the smoothing input
    float *flux;  // may be null
    float scale;
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Device memory a trace uses beyond its inputs and outputs: the packed map copies with their
// carry and wrap maps, the window queue counters and window slices.  Either a library buffer
// that grows on demand (grow; the context's, for calls that may allocate) or fixed memory
// (p, bytes) that a trace never outgrows: a batch's own buffer, sized when the batch is
// built, or the static scratch of the allocation-free entry point.  The decode leaves every
// word it reads cleared, so fixed memory private to one user is cleared once; memory shared
// by callers (clear_each) is cleared at the start of every trace.
struct Workspace {
    DevBuf *grow = nullptr;
    void *p = nullptr;
    size_t bytes = 0;
    bool clear_each = false;
};

// One uniform grid of a field (RectGrid, grid.h:25-49) and the heliostats [h0, h1) it holds:
// per-cell CSR lists of global heliostat ids (CGridHelioMatch, grid.cpp:47-82), the range of
// occupied cell indices, and the work tiles [t0, t1) of its heliostats.  The rays of a
// heliostat walk the grid that holds it (recthelio_tracing.cpp:46-79 passes "the grid
// heliostat belongs to"); ids are global so that the vertex array is the scene's.
struct GridState {
    float pos[3] = {0, 0, 0}, size[3] = {0, 0, 0}, interval[3] = {1, 1, 1};
    int h0 = 0, h1 = 0;
    int gn[3] = {0, 0, 0};
    long long cells = 0, nnz = 0;
    int occ_lo[3] = {0, 0, 0}, occ_hi[3] = {-1, -1, -1};
    std::vector<int> cell_start, cell_items; // host CSR (root scenes only)
    int *d_cell_start = nullptr, *d_cell_items = nullptr;
    int t0 = 0, t1 = 0;
    // azimuthal maps (cylinder, cone): the grid's tiles in the order of 8 azimuth sectors of
    // its heliostats about the receiver axis (d_win_order, relative to t0), and each window
    // queue's window origin (win_sectors false: the centred window of every queue)
    bool win_sectors = false;
    int *d_win_order = nullptr;
    int win_qx0[8] = {0, 0, 0, 0, 0, 0, 0, 0}, win_qy0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

} // namespace om

struct om_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // om_ctx_abort / om_ctx_progress: the abort word (device memory, read by the kernels;
    // h_abort is its pinned staging copy), the queue block of the context's window traces,
    // the queue block of the launch last enqueued (this one or a batch's), and a stream and
    // pinned buffer that copy them while launches run
    uint32_t *d_abort = nullptr;
    uint32_t *h_abort = nullptr;
    uint32_t *d_queue = nullptr;
    uint32_t *cur_queue = nullptr;
    std::mutex mon_mu;
    hipStream_t mon_stream = nullptr;
    uint32_t *h_mon = nullptr;
    om::DevBuf scratch;     // raw maps for finalize/smooth
    om::DevBuf scratch2;
    om::DevBuf parts;       // per-XCD partial maps
    int win_wgs = -1;       // persistent workgroups of window accumulation (-1: not looked up)
    void *static_scratch = nullptr; // this device's address of the allocation-free workspace
};

struct om_scene {
    om_ctx *ctx = nullptr;
    om_params prm{};
    // host copy of the description
    int H = 0;
    std::vector<float> pos, size, gap;
    std::vector<int> matrix;
    std::vector<float> normal_in; // given heliostat normals [H][3]; empty: aim at the focus
    int recv_type = 0;
    float recv_pos[3], recv_size[3], recv_normal_in[3];
    int recv_face = 0;
    float recv_pixel_length = 0.01f;
    int recv_res_in[2] = {0, 0};
    std::vector<om::GridState> grids; // in heliostat order; each traced over its own CSR
    float sun_in[3];
    float dni = 1000.0f;
    // derived host state
    om::v3 sun;
    std::vector<om::v3> normal, v0, v1, v2, v3;
    om::v3 rv[4], recv_normal, focus;
    int geo_px = 0, res_x = 0, res_y = 0;
    float recv_pixel = 0.0f;
    float srec = 0.0f;            // pixel area
    om::v3 cyl_base{0, 0, 0};
    float cyl_r2 = 0.0f, cyl_h = 0.0f;
    float cone_r0 = 0.0f, cone_k = 0.0f, cone_inv_area = 0.0f;
    int nbins = 0;
    std::vector<float> cdf_x, cdf_y, slope;
    float buie_int465 = 0, buie_A = 0, buie_B = 0, buie_C = 0, buie_inv_gp1 = 0, buie_gamma = 0;
    std::vector<float> subrow_len, subcol_len;
    std::vector<int> sub_rows, sub_cols, mcount;
    std::vector<long long> ray_offset;
    long long total_micro = 0, total_rays = 0;
    int max_micro = 0;
    int poisson_cap = 0;
    float scale = 0.0f, total_energy = 0.0f;
    std::vector<om::HelioRec> recs;
    // device state
    om::HelioRec *d_helio = nullptr;
    float4 *d_geo = nullptr;
    std::vector<float4 *> d_cell_geo; // per grid: this scene's vertices of its CSR entries
    float4 *d_buie_bins = nullptr;
    int *d_buie_guide = nullptr;
    int2 *d_buie_split = nullptr;
    float2 *d_poisson = nullptr;
    int *d_mcount = nullptr;
    std::vector<int2> tiles;
    int2 *d_tiles = nullptr;
    // differentiable mode (om_diff.hip): copies of the splat map, energy sums, per-heliostat
    // normals and tile ranges, per-tile adjoint partial sums
    om::DevBuf diff_buf;
    int rays_per_thread = 8;
    // FP64 geometry: device block of Helio64[H], Geo64[H] and the double Buie table, and the
    // host-side values of the kernel arguments (om_params.precision = OM_PREC_FP64)
    char *d_f64 = nullptr;
    double *d_tail64 = nullptr;            // Trace64Args::tail
    std::vector<om::Geo64 *> d_cell_geo64; // per grid: this scene's double vertices of its CSR entries
    om::Trace64Args a64{};
    // OM_PREC_MIXED: per grid, float images of the double CSR vertices with their edge
    // lengths ((A, id bits), (E1, |E1|), (E2, |E2|)); the double Buie table rounded to float
    // with its split guide; the error bounds; the redo counters (MixArgs::redo_n)
    std::vector<float4 *> d_cell_geo_m;
    om::HelioMix *d_hmix = nullptr;
    om::MixArgs mix{};       // the scene's part of the mixed kernels' arguments
    float4 *d_buie_bins_m = nullptr;
    int2 *d_buie_split_m = nullptr;
    float buie_A_m = 0.0f, buie_vmax_m = 0.0f, buie_gscale_m = 0.0f;
    float mix_e_dir = 0.0f, mix_e_pos = 0.0f;
    uint32_t *d_mix_n = nullptr;
    // forks (om_scene_fork): the root scene whose grid, microsamples, tiles and Buie tables
    // this scene's device pointers share, and whose host CSR it reads; null for a root
    const om_scene *grid_src = nullptr;
};

namespace om {
void set_error(const std::string &msg);
om_status hip_fail(hipError_t e, const char *what);
// kernel arguments of a scene with the grid g (and its CSR) of the scene's grids
TraceArgs make_trace_args(const om_scene *s, int g = 0);
// the grid fields of a (box, cells, CSR, occupied range) set to grid g of s
void set_trace_grid(TraceArgs &a, const om_scene *s, int g);
// tiles [t0, t1) of heliostats [h0, h1) (tiles are in heliostat order)
void tile_range(const om_scene *s, int h0, int h1, int &t0, int &t1);
// kernels (om_kernels.hip)
om_status launch_poisson(om_scene *s, hipStream_t st);
om_status launch_trace(om_scene *s, const TraceArgs &a, bool records, hipStream_t st);
// the FP64 geometry trace (om_trace64.hip): the kernel arguments of s for grid g, the launch
Trace64Args trace64_args(const om_scene *s, int g);
// om_math_eval 16: out[i] = the FP64 trace's erfinv(2u - 1) of draw word in[i]
om_status launch_erfinv_ray(const uint32_t *in, double *out, int n, hipStream_t st);
// fills out[i] with the tail angle of draw k0 + i (Trace64Args::tail), i < n
om_status build_buie_tail64(double *out, uint32_t k0, uint32_t n, double A, double int465, double B, double C,
                            double inv_gp1);
om_status launch_trace64(om_scene *s, const TraceArgs &a, const Trace64Args &b, bool records, hipStream_t st);
// OM_PREC_MIXED: the kernel arguments of s for grid g (the FP32 fields of a replaced by the
// float images of the FP64 scene); whether s can take the mixed path (else it traces in FP64)
void mixed_args(const om_scene *s, int g, TraceArgs &a, MixArgs &x);
bool mixed_supported(const om_scene *s);
// the window launch (nwg persistent workgroups) and the FP64 trace of its redo list; the
// per-ray records (rays on the redo list traced in FP64 in place)
// acc: the packed copies and side maps of the launch (acc_words words), cleared when its redo
// list fills up
om_status launch_trace_mixed_window(om_scene *s, const TraceArgs &a, const Trace64Args &b, const MixArgs &x, int nwg,
                                    uint32_t *acc, long long acc_words, hipStream_t st);
om_status launch_trace_mixed_records(om_scene *s, const TraceArgs &a, const Trace64Args &b, const MixArgs &x,
                                     hipStream_t st);
// the FP64 trace with the LDS-window accumulation (nwg persistent workgroups)
om_status launch_trace64_window(om_scene *s, const TraceArgs &a, const Trace64Args &b, int nwg, hipStream_t st);
om_status launch_finalize(const unsigned long long *energy, float *raw, float *flux, float scale,
                          int n, hipStream_t st);
om_status launch_smooth(const float *in, float *out, int w, int h, int radius, float ratio,
                        float scale, hipStream_t st);
om_status launch_math(int fn, const float *in, float *out, int n, hipStream_t st);
om_status ensure_buf(DevBuf &b, size_t bytes);
// batched mode (om_kernels.hip)
om_status launch_trace_batch(int sunshape, bool poisson, int wgs, const TraceArgs *table, const int4 *items,
                             int n_items, unsigned int *queue, int nblocks, hipStream_t st);
int trace_batch_blocks_per_cu(int sunshape, bool poisson, int wgs);
om_status launch_reduce_batch(const ReduceSlot *slots, int nslot, int max_npx, hipStream_t st);
// batched mode with LDS windows: chunks (entry, heliostat, first ray, rays) in entry order;
// nwg persistent workgroups; queue: 8 counters, cleared
om_status launch_trace_batch_window(int sunshape, bool poisson, const TraceArgs *table, const int4 *chunks,
                                    int n_chunks, unsigned int *queue, int nwg, hipStream_t st);
int trace_batch_window_blocks_per_cu(int sunshape, bool poisson);
// batched mode with LDS windows over OM_PREC_MIXED scenes (om_trace64.hip): table holds
// n_entries entry arguments of mix_launch_bytes() each, made by mix_batch_entry from the
// entry's batch TraceArgs a (its copies and window), its redo list of redo_cap entries and its
// 4 counter words; counters holds 4 words per entry, then one flag per scene.  One launch
// runs: counter reset, the persistent mixed kernel, the FP64 trace of every redo list, and
// for each scene whose list filled up the clear of its maps and the FP64 trace of the whole
// scene (slots: the scenes' maps, entry_chunks: each entry's range of chunks).
struct MixBatch {
    int sunshape = 0;
    bool poisson = false;
    const void *table = nullptr;
    const int4 *chunks = nullptr;
    int n_chunks = 0;
    unsigned int *queue = nullptr;
    int nwg = 0;
    const int *entry_scene = nullptr;
    const int2 *entry_chunks = nullptr;
    int n_entries = 0;
    uint32_t *counters = nullptr;
    const ReduceSlot *slots = nullptr;
    int n_scenes = 0;
};
size_t mix_launch_bytes();
om_status mix_batch_entry(const om_scene *s, int g, const TraceArgs &a, uint2 *redo, int redo_cap,
                          uint32_t *redo_n, void *dst);
om_status launch_trace_batch_mixed(const MixBatch &m, hipStream_t st);
int trace_batch_window_mixed_blocks_per_cu(int sunshape, bool poisson);
// the window of a res_x x res_y map the LDS holds: the whole map when it fits, else the
// largest near-square block about the centre; the rays of one chunk of the window kernels
void window_rect(int rx, int ry, int &x0, int &y0, int &w, int &h);
int window_chunk_rays();
om_status launch_fill_u32(uint32_t *p, long long n, hipStream_t st);
// A queue block of a persistent launch: QUEUE_COUNTERS counters of items taken, the item
// count and the layout (QUEUE_EIGHTHS: 8 queues over contiguous eighths of the items;
// QUEUE_ONE: counter 0 over all of them), written by launch_queue_reset on the launch's
// stream, so that om_ctx_progress reads a consistent block of whichever launch runs.
constexpr int QUEUE_COUNTERS = 8, QUEUE_WORDS = 16;
enum { QUEUE_EIGHTHS = 0, QUEUE_ONE = 1 };
om_status launch_queue_reset(uint32_t *q, long long n_items, int layout, hipStream_t st);
// the queue block that om_ctx_progress reads from now on
void set_current_queue(om_ctx *ctx, uint32_t *q);
constexpr int RPT_LIMIT = 8; // rays per thread a trace workgroup holds hit slots for
// trace heliostats [h0, h1) into out (cleared first when clear is set), with the workspace ws
om_status trace_into(om_scene *s, int h0, int h1, const om_flux *out, bool clear, hipStream_t st,
                     const Workspace &ws);
// device address of the static workspace of om_trace_plain_noalloc on device (current),
// looked up once per device
om_status static_scratch(int device, void **addr);
// the context's growing workspace (calls that may allocate)
Workspace ctx_workspace(om_ctx *ctx);
// bytes of workspace a trace of s needs (0 when s cannot be traced with its accumulation)
size_t trace_workspace_bytes(om_scene *s);
// flux (and, with a smoothing output, the smoothing input raw[res_x * res_y]) from the
// energy map of out: the scaling loop of Gtest.cu:37-46 and image_smooth
om_status flux_finalize(om_scene *s, const om_flux *out, float *raw, hipStream_t st);
} // namespace om

#define OM_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess) return om::hip_fail(_e, #call);                              \
    } while (0)
