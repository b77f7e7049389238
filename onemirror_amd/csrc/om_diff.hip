// om_diff.hip -- differentiable mode: a Gaussian-kernel flux splat of the field's rays as a
// function of per-heliostat aim-point offsets and tracking-angle deltas, and its hand-written
// adjoint.  No reference counterpart (the reference only traces fixed aims).
//
// Model, per heliostat h with parameters theta = (a_u, a_v, t_u, t_v):
//   aim   = focus + a_u e1 + a_v e2         (e1, e2: unit receiver edges; metres)
//   n_a   = normalize(normalize(aim - pos) - sun)          (the aiming of heliostat.cpp:51-57)
//   n     = normalize(n_a + 1e-3 (t_u fu + t_v fv))        (fu, fv: frame of the nominal
//                                                            normal; t in milliradians)
// per ray (the plain path's Philox draws, microsample centre and sun ray):
//   nrm   = normalize(n + (nrm0 - n0))      (nrm0 = slope-error normal about the nominal n0:
//                                            the deviation is held fixed in world space)
//   r     = normalize(reflect(-to_sun, nrm)),  e = |sun . nrm|
//   p     = o + t r on the receiver plane, (Y, X) = (u res_y - 1/2, v res_x - 1/2) pixels
//   S[y][x] += e K(x - X) K(y - Y),  K(d) = (exp(-d^2/2) - eR (1 + (R^2 - d^2)/2)) / Z on
//             |d| < R = 3 (a Gaussian of sigma = 1 px tapered to zero value and slope at R)
//   E      = sum of e over the visible rays
// Visibility (not shaded, not blocked at theta, plane hit in front) is decided in the
// forward pass, stored as one ballot bit per ray and held fixed in the backward pass, which
// recomputes every ray from its Philox counter instead of storing rays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "om_device.h"

namespace om {

struct DiffArgs {
    const float4 *theta;        // [H] (a_u, a_v, t_u, t_v)
    float4 *n_theta;            // [H] parametrised normals
    unsigned long long *mask;   // [tiles][rays per thread][WG/64] visibility ballots
    unsigned long long *slices; // [G][Q][SPLAT_T^2] splat tiles of the workgroups, fixed point 2^-32
    unsigned long long *E;      // [8] energy, fixed point 2^-32
    const double *g;            // [npx] dL/dS
    float *gf;                  // [npx + 16] dL/dS in float (read 8 pixels at a time)
    const double *gE;           // dL/dE
    double *tile_grad;          // [tiles][4] per-tile sums of dL/dn (3 used)
    const int *tile_first;      // [H+1] first tile of each heliostat
    v3 A, E1, E2, N;            // receiver plane: corner, edges, normal
    float iE1, iE2;             // 1 / |E1|^2, 1 / |E2|^2
    v3 focus, e1, e2, sun;
    float kZ, kR, kEr;          // splat kernel 1/Z, radius, exp(-R^2/2)
    int res_x, res_y, npx, H;
    // tiled splat: records (X, Y, e) of the visible rays whose kernel window meets the map;
    // the map in ntx x nty tiles of SPLAT_T pixels, G workgroups per tile
    float *recX, *recY, *recE;  // [rec_cap]
    unsigned int *n_rec;        // records written
    int ntx, nty, Q, G;         // Q = ntx * nty
};

// The kernel-argument segment of diff_forward_kernel / diff_backward_kernel (TraceArgs,
// DiffArgs): explicit arguments are laid out in order, each at its type's alignment, as the
// members of this struct.
struct DiffKernArgs {
    TraceArgs a;
    DiffArgs p;
};
static_assert(sizeof(TraceArgs) % alignof(DiffArgs) == 0, "DiffArgs follows TraceArgs without padding");

constexpr int DIFF_COPIES = 8;
constexpr float DIFF_R = 3.0f;
constexpr int SPLAT_T = 128;     // splat tile edge: SPLAT_T^2 u64 words are 128 KiB of LDS
constexpr int SPLAT_WGS = 1024;  // threads of a splat workgroup (one workgroup per CU)

// pixels of the kernel window of (X, Y) inside the map; false when it misses the map, and
// then the window is empty on both axes (x1 < x0 and y1 < y0), so that a loop over its rows
// reads nothing.  A point 10^6 pixels or more off (maps are at most 1024 pixels a side) or
// not finite misses it, which also keeps the integer conversions in range.
__device__ __forceinline__ bool rec_window(const DiffArgs &p, float X, float Y, int &x0, int &x1, int &y0, int &y1)
{
    bool in = fabs_(X) < 1.0e6f && fabs_(Y) < 1.0e6f;
    if (in) {
        x0 = max(0, trunc_i(__builtin_ceilf(X - DIFF_R)));
        x1 = min(p.res_x - 1, trunc_i(__builtin_floorf(X + DIFF_R)));
        y0 = max(0, trunc_i(__builtin_ceilf(Y - DIFF_R)));
        y1 = min(p.res_y - 1, trunc_i(__builtin_floorf(Y + DIFF_R)));
        in = x0 <= x1 && y0 <= y1;
    }
    if (!in) {
        x0 = y0 = 0;
        x1 = y1 = -1;
    }
    return in;
}

// K(d) = (exp(-d^2/2) - eR (1 + (R^2 - d^2)/2)) / Z for the forward splat, from the
// hardware base-2 exponential (exp(-d^2/2) = 2^(-d^2 / (2 ln 2)), within a few ulp; the
// splat is compared with the FP64 model to a tolerance, and its integer sums keep it
// deterministic)
__device__ __forceinline__ float kern_fast(float d, const DiffArgs &p)
{
    return (__builtin_amdgcn_exp2f(-0.72134752f * d * d) - p.kEr * (1.0f + 0.5f * (p.kR * p.kR - d * d))) * p.kZ;
}

// K and K' at d from one exponential (the hardware base-2 exponential: the adjoint is
// compared with FP64 central differences to a tolerance)
__device__ __forceinline__ void kern2(float d, const DiffArgs &p, float &k, float &dk)
{
    const float ex = __builtin_amdgcn_exp2f(-0.72134752f * d * d);
    k = (ex - p.kEr * (1.0f + 0.5f * (p.kR * p.kR - d * d))) * p.kZ;
    dk = -d * (ex - p.kEr) * p.kZ;
}

// n(theta) of every heliostat
__global__ void diff_normals_kernel(const HelioRec *helio, DiffArgs p)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= p.H) return;
    const HelioRec &hr = helio[h];
    const float4 th = p.theta[h];
    const v3 pos = mk(hr.pos[0], hr.pos[1], hr.pos[2]);
    const v3 fu = mk(hr.fu[0], hr.fu[1], hr.fu[2]), fv = mk(hr.fv[0], hr.fv[1], hr.fv[2]);
    const v3 aim = p.focus + p.e1 * th.x + p.e2 * th.y;
    const v3 na = normalize(normalize(aim - pos) - p.sun);
    const v3 n = normalize(na + (fu * th.z + fv * th.w) * 1e-3f);
    p.n_theta[h] = make_float4(n.x, n.y, n.z, 0.0f);
}

// One ray of the model up to the splat position; false when the plane is not hit in front.
struct DiffRay {
    v3 i, nrm, r, w, q;
    float t, den, X, Y, e;
};
__device__ __forceinline__ bool diff_ray(const DiffArgs &p, v3 orig, v3 to_sun, v3 nrm0, v3 n0, v3 nth, DiffRay &o)
{
    o.q = nth + (nrm0 - n0);
    o.nrm = normalize<gfx>(o.q);
    o.i = -to_sun;
    o.w = reflect(o.i, o.nrm);
    o.r = normalize<gfx>(o.w);
    o.den = dot(o.r, p.N);
    if (!(fabs_(o.den) >= EPS_LT)) return false;
    o.t = dot(p.A - orig, p.N) / o.den;
    if (!(o.t >= EPS_LT)) return false;
    const v3 pp = orig + o.r * o.t;
    const float u = dot(pp - p.A, p.E1) * p.iE1, v = dot(pp - p.A, p.E2) * p.iE2;
    o.Y = u * (float)p.res_y - 0.5f;
    o.X = v * (float)p.res_x - 0.5f;
    o.e = fabs_(dot(p.sun, o.nrm));
    return true;
}

// Per-tile ray set-up shared by the forward and backward kernels (the plain path's sampling).
struct DiffTile {
    const HelioRec *hr;
    frame hf, sf;
    v3 hpos, nth;
    int h, L, nr;
    bool one_row, one_col;
    bool clear; // the shading walk finds nothing for any ray of this heliostat
};

// The sun ray of ray l of microsample m (the plain path's Philox draws and sunshape): the
// direction toward the sun, and the slope-error normal nrm0 about the nominal normal (the
// plain path's perturbed_normal of the same draws).  With the Buie sunshape the two loads of
// the bin lookup (buie_bin_split) are issued ahead of the normal and the sunshape azimuth,
// which do not depend on them and run while they are in flight, as in the plain trace.  The
// microsample's origin comes from the wave's start table (StartTable, om_device.h), the same
// bits as ray_start.
template <int SUN>
__device__ __forceinline__ void diff_dir(const TraceArgs &a, const DiffTile &T, int m, int l, v3 &to_sun, v3 &nrm0)
{
    const u4 x = philox_uniform_key(u4{(uint32_t)l, (uint32_t)m, (uint32_t)T.h, a.w3}, a.key0, a.key1);
    float th, sp, cp;
    if (SUN == OM_SUN_PILLBOX) {
        th = a.pillbox * sqrt_rn(u01(x.x));
        nrm0 = perturbed_normal(x, a.slope_sigma, T.hf);
        sincos2pi(u01(x.y), sp, cp);
    } else {
        const float v = u01(x.x) * a.buie_A;
        const int j = min(max(trunc_i(v * a.buie_gscale), 0), a.nbins - 1);
        const int2 g = a.buie_split[j];
        __builtin_amdgcn_sched_barrier(0); // the load above is issued before this work
        nrm0 = perturbed_normal(x, a.slope_sigma, T.hf);
        asm volatile("" ::"v"(nrm0.x), "v"(nrm0.y), "v"(nrm0.z)); // computed here
        __builtin_amdgcn_sched_barrier(0);
        int i = min(g.x + (v >= u2f((uint32_t)g.y) ? 1 : 0), a.nbins - 1);
        float4 bb = a.buie_bins[i];
        __builtin_amdgcn_sched_barrier(0);
        sincos2pi(u01(x.y), sp, cp);
        asm volatile("" ::"v"(sp), "v"(cp));
        __builtin_amdgcn_sched_barrier(0);
        if (v < a.buie_int465) {
            while (bb.x > v) bb = a.buie_bins[--i];
            while (bb.y <= v) bb = a.buie_bins[++i];
            th = (v - bb.x) * bb.z + bb.w;
        } else {
            th = pow_<gfx>((v - a.buie_int465) * a.buie_B + a.buie_C, a.buie_inv_gp1);
        }
        th = div_rn(th, 1000.0f);
    }
    float st, ct;
    sincos_small(th, st, ct);
    to_sun = -to_world(angle2xyz(st, ct, sp, cp), T.sf);
}

__device__ __forceinline__ DiffTile diff_tile(const TraceArgs &a, const DiffArgs &p, int h)
{
    DiffTile T;
    T.h = h;
    T.hr = a.helio + h;
    const HelioRec *hr = T.hr;
    T.hf.n = mk(hr->n[0], hr->n[1], hr->n[2]);
    T.hf.u = mk(hr->fu[0], hr->fu[1], hr->fu[2]);
    T.hf.v = mk(hr->fv[0], hr->fv[1], hr->fv[2]);
    T.hf.identity = hr->ident;
    T.sf.n = a.sun_n;
    T.sf.u = a.sun_u;
    T.sf.v = a.sun_v;
    T.sf.identity = a.sun_ident;
    T.hpos = mk(hr->pos[0], hr->pos[1], hr->pos[2]);
    const float4 n4 = p.n_theta[h];
    T.nth = mk(n4.x, n4.y, n4.z);
    T.L = a.L;
    T.nr = hr->mcount * a.L;
    T.one_row = hr->rows == 1;
    T.one_col = hr->cols == 1;
    T.clear = hr->clear != 0;
    return T;
}

template <int SUN, bool POISSON>
__global__ __launch_bounds__(WG) void diff_forward_kernel(TraceArgs, DiffArgs)
{
    // the arguments are read from the kernel-argument segment per ray pass, as in the plain
    // trace kernels (held in SGPRs they spill to VGPR lanes)
    const DiffKernArgs &k_args = *(const DiffKernArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    const TraceArgs &a = k_args.a;
    const DiffArgs &p = k_args.p;
    __shared__ float s_rx[RPT_MAX * WG], s_ry[RPT_MAX * WG], s_re[RPT_MAX * WG];
    __shared__ unsigned int s_wt[WG / 64], s_base;
    int nrec = 0; // records of this thread, staged in LDS slots [nrec][thread]
    const int tile = a.tile_base + xcd_chunk(blockIdx.x, gridDim.x);
    const int2 tl = a.tiles[tile];
    const DiffTile T = diff_tile(a, p, tl.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long esum = 0;
    int r = tl.y + (int)threadIdx.x;
    int m = r / T.L;
    int l = r - m * T.L;
    const int dm = WG / T.L, dl = WG - (WG / T.L) * T.L;
    StartTable tab; // the wave's rays of a pass are 64 consecutive rays
    for (int k = 0; k < a.rays_per_thread; ++k, r += WG) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= T.L) {
                l -= T.L;
                ++m;
            }
        }
        bool vis = false, rec = false;
        DiffRay o;
        int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
        asm volatile("" ::: "memory");
        tab.update<POISSON>(a, T, r < T.nr, m, lane); // the whole wave
        const RayStart rs = tab.get(a, m);
        if (r < T.nr) {
            v3 to_sun, nrm0;
            diff_dir<SUN>(a, T, m, l, to_sun, nrm0);
            const v3 orig = rs.orig;
            const cell3 c0 = rs.c0;
            if (T.clear || dda(a, orig, c0, to_sun, T.h) < 0) {
                if (diff_ray(p, orig, to_sun, nrm0, T.hf.n, T.nth, o) && dda(a, orig, c0, o.r, T.h) < 0) {
                    vis = true;
                    esum += (unsigned long long)((double)o.e * 4294967296.0);
                    rec = rec_window(p, o.X, o.Y, x0, x1, y0, y1);
                }
            }
        }
        const unsigned long long bits = __ballot(vis);
        if (lane == 0) p.mask[((size_t)tile * a.rays_per_thread + k) * (WG / 64) + wave] = bits;
        // a record for each ray whose window meets the map
        if (rec) {
            s_rx[nrec * WG + threadIdx.x] = o.X;
            s_ry[nrec * WG + threadIdx.x] = o.Y;
            s_re[nrec * WG + threadIdx.x] = o.e;
            ++nrec;
        }
    }
    if (esum) atomicAdd(p.E + (xcc_id() & 7u), esum);
    // workgroup prefix of the record counts, one counter atomic per workgroup
    unsigned int incl = (unsigned int)nrec;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_wt[wave] = incl;
    __syncthreads();
    unsigned int woff = 0, total = 0;
    for (int w = 0; w < WG / 64; ++w) {
        if (w < wave) woff += s_wt[w];
        total += s_wt[w];
    }
    if (threadIdx.x == 0) s_base = total ? atomicAdd(p.n_rec, total) : 0u;
    __syncthreads();
    const unsigned int pos = s_base + woff + incl - (unsigned int)nrec;
    for (int i = 0; i < nrec; ++i) {
        p.recX[pos + i] = s_rx[i * WG + threadIdx.x];
        p.recY[pos + i] = s_ry[i * WG + threadIdx.x];
        p.recE[pos + i] = s_re[i * WG + threadIdx.x];
    }
}

// e K(x - X) K(y - Y) of one record over the pixels [x0, x1] x [y0, y1] of its kernel
// window inside the tile at (tx0, ty0), as 2^-32 fixed point into the LDS tile.  w = e K K
// < 1 (K(0)^2 < 1, checked on the host), so w * 2^32 fits in 32 bits and its truncation is
// the fixed-point value (the scaling by 2^32 is exact).
__device__ __forceinline__ void splat_record(const DiffArgs &p, unsigned long long *s_tile, int tx0, int ty0,
                                             float X, float Y, float e, int x0, int x1, int y0, int y1)
{
    float kx[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) kx[i] = kern_fast((float)(x0 + i) - X, p) * 4294967296.0f;
    for (int yy = y0; yy <= y1; ++yy) {
        const float ky = e * kern_fast((float)yy - Y, p);
        unsigned long long *row = s_tile + (yy - ty0) * SPLAT_T + (x0 - tx0);
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            const float w = ky * kx[i];
            if (x0 + i <= x1 && w >= 1.0f) atomicAdd(row + i, (unsigned long long)(uint32_t)w);
        }
    }
}

// Splat of the records: workgroup b takes tile b % Q of the map and every G-th block of
// SPLAT_WGS records from block b / Q on (records come in heliostat order, and neighbouring
// heliostats light the same part of the map: strided blocks give every workgroup of a tile
// a like share of its work).  A wave queues the records of its lanes whose kernel window
// meets the tile in LDS and splats them 64 at a time, one per lane, so that lanes do not idle
// on the records that miss the tile.  Sums are 2^-32 fixed point in an LDS copy of the tile
// (integer sums: any order gives the same map), stored to this workgroup's slice at the end.
// The Q workgroups of one block run side by side, so a block is read from HBM about once.
constexpr int SPLAT_QUEUE = 128; // queued records per wave (a power of two, >= 2 x 64)

__global__ __launch_bounds__(SPLAT_WGS) void diff_splat_kernel(DiffArgs p)
{
    __shared__ unsigned long long s_tile[SPLAT_T * SPLAT_T];
    __shared__ float s_q[SPLAT_WGS / 64][3][SPLAT_QUEUE];
    const int t = (int)blockIdx.x % p.Q, g = (int)blockIdx.x / p.Q;
    const int tx0 = (t % p.ntx) * SPLAT_T, ty0 = (t / p.ntx) * SPLAT_T;
    const int tx1 = min(tx0 + SPLAT_T, p.res_x) - 1, ty1 = min(ty0 + SPLAT_T, p.res_y) - 1;
    const int lane = threadIdx.x & 63;
    float *qx = s_q[threadIdx.x >> 6][0], *qy = s_q[threadIdx.x >> 6][1], *qe = s_q[threadIdx.x >> 6][2];
    for (int i = threadIdx.x; i < SPLAT_T * SPLAT_T; i += SPLAT_WGS) s_tile[i] = 0ull;
    __syncthreads();
    const unsigned int n = *p.n_rec;
    unsigned int head = 0, tail = 0; // this wave's queue: entries [head, tail)
    const unsigned int step = (unsigned)p.G * SPLAT_WGS;
    for (unsigned int j0 = (unsigned)g * SPLAT_WGS; j0 < n; j0 += step) {
        // the whole wave runs this loop: j0 is uniform
        const unsigned int j = j0 + threadIdx.x;
        bool in = false;
        float X = 0.0f, Y = 0.0f, e = 0.0f;
        if (j < n) {
            X = p.recX[j];
            Y = p.recY[j];
            e = p.recE[j];
            int x0, x1, y0, y1;
            rec_window(p, X, Y, x0, x1, y0, y1);
            in = max(x0, tx0) <= min(x1, tx1) && max(y0, ty0) <= min(y1, ty1);
        }
        const unsigned long long m = __ballot(in);
        if (in) {
            const unsigned int k = (tail + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))) &
                                   (SPLAT_QUEUE - 1);
            qx[k] = X;
            qy[k] = Y;
            qe[k] = e;
        }
        tail += (unsigned)__popcll(m);
        if (tail - head >= 64u) {
            const unsigned int k = (head + lane) & (SPLAT_QUEUE - 1);
            const float Xq = qx[k], Yq = qy[k], eq = qe[k];
            int x0, x1, y0, y1;
            rec_window(p, Xq, Yq, x0, x1, y0, y1);
            splat_record(p, s_tile, tx0, ty0, Xq, Yq, eq, max(x0, tx0), min(x1, tx1), max(y0, ty0), min(y1, ty1));
            head += 64u;
        }
    }
    if (lane < (int)(tail - head)) {
        const unsigned int k = (head + lane) & (SPLAT_QUEUE - 1);
        const float Xq = qx[k], Yq = qy[k], eq = qe[k];
        int x0, x1, y0, y1;
        rec_window(p, Xq, Yq, x0, x1, y0, y1);
        splat_record(p, s_tile, tx0, ty0, Xq, Yq, eq, max(x0, tx0), min(x1, tx1), max(y0, ty0), min(y1, ty1));
    }
    __syncthreads();
    unsigned long long *out = p.slices + (size_t)blockIdx.x * (SPLAT_T * SPLAT_T);
    for (int i = threadIdx.x; i < SPLAT_T * SPLAT_T; i += SPLAT_WGS) out[i] = s_tile[i];
}

__global__ void diff_gconv_kernel(DiffArgs p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p.npx) p.gf[i] = (float)p.g[i];
    else if (i < p.npx + 16) p.gf[i] = 0.0f;
}

// slices -> double map; energy copies -> double (and cleared)
__global__ void diff_decode_kernel(DiffArgs p, double *S_out, double *E_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        unsigned long long e = 0;
        for (int c = 0; c < DIFF_COPIES; ++c) {
            e += p.E[c];
            p.E[c] = 0ull;
        }
        *E_out = (double)e * 2.3283064365386963e-10;
    }
    if (i >= p.npx) return;
    const int y = i / p.res_x, x = i - y * p.res_x;
    const int t = (y / SPLAT_T) * p.ntx + x / SPLAT_T;
    const unsigned long long *w = p.slices + (size_t)t * SPLAT_T * SPLAT_T + (y % SPLAT_T) * SPLAT_T + (x % SPLAT_T);
    unsigned long long s = 0;
    for (int g = 0; g < p.G; ++g) s += w[(size_t)g * p.Q * SPLAT_T * SPLAT_T];
    S_out[i] = (double)s * 2.3283064365386963e-10;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adjoint: every visible ray (per the forward ballots) is recomputed from its Philox counter
// without the walks, dL/dX, dL/dY, dL/de are gathered from dL/dS over its kernel window and
// carried back through the plane intersection, normalisation, reflection, energy and the
// normal's normalisation to dL/dn(theta); a tile's sum goes to tile_grad.
template <int SUN, bool POISSON>
__global__ __launch_bounds__(WG) void diff_backward_kernel(TraceArgs, DiffArgs)
{
    const DiffKernArgs &k_args = *(const DiffKernArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    const TraceArgs &a = k_args.a;
    const DiffArgs &p = k_args.p;
    __shared__ double s_red[WG / 64][3];
    const int tile = a.tile_base + xcd_chunk(blockIdx.x, gridDim.x);
    const int2 tl = a.tiles[tile];
    const DiffTile T = diff_tile(a, p, tl.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double gE = *p.gE;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    int r = tl.y + (int)threadIdx.x;
    int m = r / T.L;
    int l = r - m * T.L;
    const int dm = WG / T.L, dl = WG - (WG / T.L) * T.L;
    StartTable tab; // the wave's rays of a pass are 64 consecutive rays
    for (int k = 0; k < a.rays_per_thread; ++k, r += WG) {
        if (k > 0) {
            m += dm;
            l += dl;
            if (l >= T.L) {
                l -= T.L;
                ++m;
            }
        }
        // lane 0 holds the pass's first ray: past the heliostat's rays, every lane is
        if (__builtin_amdgcn_readfirstlane(r) >= T.nr) break;
        asm volatile("" ::: "memory");
        const unsigned long long bits = p.mask[((size_t)tile * a.rays_per_thread + k) * (WG / 64) + wave];
        tab.update<POISSON>(a, T, r < T.nr, m, lane); // the whole wave
        const RayStart rs = tab.get(a, m);
        if (r >= T.nr || !((bits >> lane) & 1ull)) continue;
        v3 to_sun, nrm0;
        diff_dir<SUN>(a, T, m, l, to_sun, nrm0);
        const v3 orig = rs.orig;
        DiffRay o;
        if (!diff_ray(p, orig, to_sun, nrm0, T.hf.n, T.nth, o)) continue; // (not reached)
        // gather over the kernel window (none when the window misses the map: then only the
        // energy term remains)
        int x0, x1, y0, y1;
        const bool win = rec_window(p, o.X, o.Y, x0, x1, y0, y1);
        float kx[7], dkx[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) kern2((float)(x0 + i) - o.X, p, kx[i], dkx[i]);
        // window pixels outside [x0, x1] carry zero weight
#pragma unroll
        for (int i = 0; i < 7; ++i)
            if (!win || x0 + i > x1) kx[i] = dkx[i] = 0.0f;
        float sA = 0.0f, sX = 0.0f, sY = 0.0f;
        for (int yy = y0; yy <= y1; ++yy) {
            float ky, kdy;
            kern2((float)yy - o.Y, p, ky, kdy);
            // eight consecutive pixels from x0 (the buffer is padded past the map's end)
            const float *grow = p.gf + (size_t)yy * p.res_x + x0;
            float gv[8];
            __builtin_memcpy(gv, grow, sizeof(gv));
            float rA = 0.0f, rX = 0.0f;
#pragma unroll
            for (int i = 0; i < 7; ++i) {
                rA += gv[i] * kx[i];
                rX += gv[i] * dkx[i];
            }
            sA += ky * rA;
            sX += ky * rX;
            sY += kdy * rA;
        }
        const float LX = -o.e * sX, LY = -o.e * sY, Le = sA + (float)gE;
        // (Y, X) = (u res_y, v res_x) - 1/2;  u = (p - A).E1 / |E1|^2, v likewise
        const float Lu = LY * (float)p.res_y, Lv = LX * (float)p.res_x;
        const v3 gp = p.E1 * (Lu * p.iE1) + p.E2 * (Lv * p.iE2);
        // p = o + t r, t = (A - o).N / (r.N)
        const v3 gr = (gp - p.N * (dot(gp, o.r) / o.den)) * o.t;
        // r = w / |w|
        const float wl = sqrt_rn(dot(o.w, o.w));
        const v3 gw = (gr - o.r * dot(gr, o.r)) * (1.0f / wl);
        // w = i - 2 (nrm.i) nrm
        const float di = dot(o.nrm, o.i);
        v3 gn = (gw * di + o.i * dot(gw, o.nrm)) * -2.0f;
        // e = |sun.nrm|
        gn = gn + p.sun * (dot(p.sun, o.nrm) >= 0.0f ? Le : -Le);
        // nrm = q / |q|, q = n(theta) + deviation
        const float ql = sqrt_rn(dot(o.q, o.q));
        const v3 gq = (gn - o.nrm * dot(gn, o.nrm)) * (1.0f / ql);
        gx += (double)gq.x;
        gy += (double)gq.y;
        gz += (double)gq.z;
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    gz = wave_sum(gz);
    if (lane == 0) {
        s_red[wave][0] = gx;
        s_red[wave][1] = gy;
        s_red[wave][2] = gz;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < WG / 64; ++w) t += s_red[w][threadIdx.x];
        p.tile_grad[(size_t)tile * 4 + threadIdx.x] = t;
    }
}

struct d3v {
    double x, y, z;
};
__device__ __forceinline__ d3v D(v3 a) { return d3v{a.x, a.y, a.z}; }
__device__ __forceinline__ double ddot(d3v a, d3v b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3v dadd(d3v a, d3v b) { return d3v{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3v dsub(d3v a, d3v b) { return d3v{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3v dscl(d3v a, double s) { return d3v{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dlen(d3v a) { return sqrt(ddot(a, a)); }
// adjoint of v -> v / |v| at v: g -> (g - u (g.u)) / |v|
__device__ __forceinline__ d3v dnorm_adj(d3v v, d3v g)
{
    const double l = dlen(v);
    const d3v u = dscl(v, 1.0 / l);
    return dscl(dsub(g, dscl(u, ddot(g, u))), 1.0 / l);
}

// dL/dtheta of every heliostat: its tiles' dL/dn summed in order, then through
// n = normalize(n_a + 1e-3 (t_u fu + t_v fv)), n_a = normalize(normalize(aim - pos) - sun),
// aim = focus + a_u e1 + a_v e2, in double.
__global__ void diff_params_kernel(const HelioRec *helio, DiffArgs p, float4 *grad)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= p.H) return;
    d3v gn{0.0, 0.0, 0.0};
    for (int t = p.tile_first[h]; t < p.tile_first[h + 1]; ++t)
        gn = dadd(gn, d3v{p.tile_grad[(size_t)t * 4], p.tile_grad[(size_t)t * 4 + 1], p.tile_grad[(size_t)t * 4 + 2]});
    const HelioRec &hr = helio[h];
    const float4 th = p.theta[h];
    const d3v pos{hr.pos[0], hr.pos[1], hr.pos[2]};
    const d3v fu{hr.fu[0], hr.fu[1], hr.fu[2]}, fv{hr.fv[0], hr.fv[1], hr.fv[2]};
    const d3v aim = dadd(dadd(D(p.focus), dscl(D(p.e1), th.x)), dscl(D(p.e2), th.y));
    const d3v wv = dsub(aim, pos);
    const d3v rd = dscl(wv, 1.0 / dlen(wv));
    const d3v ya = dsub(rd, D(p.sun));
    const d3v na = dscl(ya, 1.0 / dlen(ya));
    const d3v z = dadd(na, dscl(dadd(dscl(fu, th.z), dscl(fv, th.w)), 1e-3));
    const d3v gz = dnorm_adj(z, gn);
    const d3v gya = dnorm_adj(ya, gz);
    const d3v gw = dnorm_adj(wv, gya);
    grad[h] = make_float4((float)ddot(gw, D(p.e1)), (float)ddot(gw, D(p.e2)), (float)(1e-3 * ddot(gz, fu)),
                          (float)(1e-3 * ddot(gz, fv)));
}

// visibility of the rays of heliostats [h0, h1), one byte per ray in ray order
__global__ void diff_vis_kernel(TraceArgs a, const unsigned long long *mask, int t0, int t1, long long base,
                                uint8_t *out)
{
    const int tile = t0 + blockIdx.x;
    if (tile >= t1) return;
    const int2 tl = a.tiles[tile];
    const HelioRec *hr = a.helio + tl.x;
    const long long off = (((long long)hr->ray_off_hi << 32) | (unsigned)hr->ray_off_lo) - base;
    const int nr = hr->mcount * a.L;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int r = tl.y + (int)threadIdx.x;
    for (int k = 0; k < a.rays_per_thread && r < nr; ++k, r += WG) {
        const unsigned long long bits = mask[((size_t)tile * a.rays_per_thread + k) * (WG / 64) + wave];
        out[off + r] = (uint8_t)((bits >> lane) & 1ull);
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
struct DiffLayout {
    size_t E, n_rec, n_theta, tile_first, tile_grad, gf, slices, rec, bytes;
    size_t rec_cap;
    int ntx, nty, Q, G;
};

// one allocation per scene: the cleared part (energy copies, record counter) first, then the
// per-heliostat and per-tile arrays, the splat slices and the records
static DiffLayout diff_layout(const om_scene *s, int cus)
{
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    DiffLayout L;
    const size_t npx = (size_t)s->res_x * s->res_y;
    L.rec_cap = (size_t)s->total_rays;
    L.ntx = (s->res_x + SPLAT_T - 1) / SPLAT_T;
    L.nty = (s->res_y + SPLAT_T - 1) / SPLAT_T;
    L.Q = L.ntx * L.nty;
    L.G = std::max(1, cus / L.Q); // one splat workgroup per CU
    L.E = 0;
    L.n_rec = al(L.E + sizeof(unsigned long long) * DIFF_COPIES);
    L.n_theta = al(L.n_rec + sizeof(unsigned int));
    L.tile_first = al(L.n_theta + sizeof(float4) * (size_t)(s->H + 1));
    L.tile_grad = al(L.tile_first + sizeof(int) * (size_t)(s->H + 2));
    L.gf = al(L.tile_grad + sizeof(double) * 4 * (s->tiles.size() + 1));
    L.slices = al(L.gf + sizeof(float) * (npx + 16));
    L.rec = al(L.slices + sizeof(unsigned long long) * SPLAT_T * SPLAT_T * (size_t)L.Q * L.G);
    L.bytes = al(L.rec + sizeof(float) * 3 * L.rec_cap);
    return L;
}

static om_status diff_prepare(om_scene *s, DiffArgs &p, hipStream_t st)
{
    if (!s || !s->ctx) {
        set_error("differentiable mode needs a scene built on a context");
        return OM_ERR_INVALID;
    }
    if (s->recv_type != 0) {
        set_error("differentiable mode supports rectangle receivers");
        return OM_ERR_UNSUPPORTED;
    }
    if (s->prm.literal_self) {
        set_error("differentiable mode skips the ray's own heliostat (literal_self is not supported)");
        return OM_ERR_UNSUPPORTED;
    }
    if (s->prm.precision != OM_PREC_FP32) {
        set_error("differentiable mode traces in FP32 geometry (precisions OM_PREC_FP64 and OM_PREC_MIXED are "
                  "not supported)");
        return OM_ERR_UNSUPPORTED;
    }
    if (!s->normal_in.empty()) {
        set_error("differentiable mode aims the heliostats itself (scene has given normals)");
        return OM_ERR_UNSUPPORTED;
    }
    if (s->rays_per_thread > RPT_MAX) return OM_ERR_UNSUPPORTED;
    if (s->total_rays >= (1LL << 30) || s->res_x > 1024 || s->res_y > 1024) {
        set_error("differentiable mode supports up to 2^30 rays and maps up to 1024 pixels a side");
        return OM_ERR_UNSUPPORTED;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    int cus = 0;
    OM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->ctx->device));
    const DiffLayout L = diff_layout(s, cus);
    if (s->diff_buf.bytes < L.bytes) {
        om_status rc = ensure_buf(s->diff_buf, L.bytes);
        if (rc != OM_OK) return rc;
        char *b = (char *)s->diff_buf.p;
        // the decode leaves the energy copies cleared; clear once here
        if ((rc = launch_fill_u32((uint32_t *)b, (long long)(L.n_theta / 4), st)) != OM_OK) return rc;
        std::vector<int> first(s->H + 1, 0);
        size_t t = 0;
        for (int h = 0; h <= s->H; ++h) {
            while (t < s->tiles.size() && s->tiles[t].x < h) ++t;
            first[h] = (int)t;
        }
        // the host vector ends with this block: wait for the copy to have read it
        OM_HIP(hipMemcpyAsync(b + L.tile_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, st));
        OM_HIP(hipStreamSynchronize(st));
    }
    char *b = (char *)s->diff_buf.p;
    std::memset(&p, 0, sizeof(p));
    p.E = (unsigned long long *)(b + L.E);
    p.slices = (unsigned long long *)(b + L.slices);
    p.n_theta = (float4 *)(b + L.n_theta);
    p.tile_first = (const int *)(b + L.tile_first);
    p.tile_grad = (double *)(b + L.tile_grad);
    p.n_rec = (unsigned int *)(b + L.n_rec);
    p.gf = (float *)(b + L.gf);
    float *rec = (float *)(b + L.rec);
    p.recX = rec;
    p.recY = rec + L.rec_cap;
    p.recE = rec + 2 * L.rec_cap;
    p.ntx = L.ntx;
    p.nty = L.nty;
    p.Q = L.Q;
    p.G = L.G;
    p.A = s->rv[0];
    p.E1 = s->rv[1] - s->rv[0];
    p.E2 = s->rv[3] - s->rv[0];
    p.N = normalize(cross(p.E1, p.E2));
    p.iE1 = 1.0f / dot(p.E1, p.E1);
    p.iE2 = 1.0f / dot(p.E2, p.E2);
    p.focus = s->focus;
    p.e1 = normalize(p.E1);
    p.e2 = normalize(p.E2);
    p.sun = s->sun;
    const double R = DIFF_R, eR = std::exp(-0.5 * R * R);
    const double Z = std::sqrt(2.0 * 3.14159265358979323846) * std::erf(R / std::sqrt(2.0)) - eR * (2.0 * R + 2.0 * R * R * R / 3.0);
    p.kZ = (float)(1.0 / Z);
    // the splat's 32-bit fixed-point conversion needs e K(0)^2 < 1 (e <= 1)
    const double K0 = (1.0 - eR * (1.0 + 0.5 * R * R)) / Z;
    if (!(K0 * K0 < 0.5)) {
        set_error("splat kernel peak too large for the fixed-point splat");
        return OM_ERR_UNSUPPORTED;
    }
    p.kR = (float)R;
    p.kEr = (float)eR;
    p.res_x = s->res_x;
    p.res_y = s->res_y;
    p.npx = s->res_x * s->res_y;
    p.H = s->H;
    return OM_OK;
}

template <bool FWD>
static om_status launch_diff(om_scene *s, const TraceArgs &a, const DiffArgs &p, hipStream_t st);

// the diff trace of every grid's heliostats over that grid
template <bool FWD>
static om_status launch_diff_grids(om_scene *s, const DiffArgs &p, hipStream_t st)
{
    for (size_t g = 0; g < s->grids.size(); ++g) {
        TraceArgs a = make_trace_args(s, (int)g);
        a.tile_base = s->grids[g].t0;
        a.ntiles = s->grids[g].t1 - s->grids[g].t0;
        if (om_status rc = launch_diff<FWD>(s, a, p, st); rc != OM_OK) return rc;
    }
    return OM_OK;
}

template <bool FWD>
static om_status launch_diff(om_scene *s, const TraceArgs &a, const DiffArgs &p, hipStream_t st)
{
    const dim3 grid((unsigned)a.ntiles), block(WG);
    if (a.ntiles <= 0) return OM_OK;
    const bool po = s->prm.micro_mode == OM_MICRO_POISSON;
    if (s->prm.sunshape == OM_SUN_BUIE) {
        if (po) hipLaunchKernelGGL((FWD ? diff_forward_kernel<OM_SUN_BUIE, true> : diff_backward_kernel<OM_SUN_BUIE, true>), grid, block, 0, st, a, p);
        else hipLaunchKernelGGL((FWD ? diff_forward_kernel<OM_SUN_BUIE, false> : diff_backward_kernel<OM_SUN_BUIE, false>), grid, block, 0, st, a, p);
    } else {
        if (po) hipLaunchKernelGGL((FWD ? diff_forward_kernel<OM_SUN_PILLBOX, true> : diff_backward_kernel<OM_SUN_PILLBOX, true>), grid, block, 0, st, a, p);
        else hipLaunchKernelGGL((FWD ? diff_forward_kernel<OM_SUN_PILLBOX, false> : diff_backward_kernel<OM_SUN_PILLBOX, false>), grid, block, 0, st, a, p);
    }
    OM_HIP(hipGetLastError());
    return OM_OK;
}

} // namespace om

using namespace om;

extern "C" om_status om_diff_mask_words(const om_scene *s, int64_t *words)
{
    if (!s || !words) return OM_ERR_INVALID;
    *words = (int64_t)s->tiles.size() * s->rays_per_thread * (WG / 64);
    return OM_OK;
}

extern "C" om_status om_diff_forward(om_scene *s, const float *d_params, uint64_t *d_mask, double *d_splat,
                                     double *d_energy, void *stream)
{
    if (!s || !d_params || !d_mask || !d_splat || !d_energy) return OM_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    DiffArgs p;
    om_status rc = diff_prepare(s, p, st);
    if (rc != OM_OK) return rc;
    p.theta = (const float4 *)d_params;
    p.mask = (unsigned long long *)d_mask;
    hipLaunchKernelGGL(diff_normals_kernel, dim3((s->H + WG - 1) / WG), dim3(WG), 0, st, s->d_helio, p);
    OM_HIP(hipGetLastError());
    if ((rc = launch_fill_u32(p.n_rec, 1, st)) != OM_OK) return rc;
    if ((rc = launch_diff_grids<true>(s, p, st)) != OM_OK) return rc;
    hipLaunchKernelGGL(diff_splat_kernel, dim3((unsigned)(p.Q * p.G)), dim3(SPLAT_WGS), 0, st, p);
    OM_HIP(hipGetLastError());
    hipLaunchKernelGGL(diff_decode_kernel, dim3((p.npx + WG - 1) / WG), dim3(WG), 0, st, p, d_splat, d_energy);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

extern "C" om_status om_diff_backward(om_scene *s, const float *d_params, const uint64_t *d_mask,
                                      const double *d_dsplat, const double *d_denergy, float *d_grad, void *stream)
{
    if (!s || !d_params || !d_mask || !d_dsplat || !d_denergy || !d_grad) return OM_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    DiffArgs p;
    om_status rc = diff_prepare(s, p, st);
    if (rc != OM_OK) return rc;
    p.theta = (const float4 *)d_params;
    p.mask = (unsigned long long *)d_mask;
    p.g = d_dsplat;
    p.gE = d_denergy;
    hipLaunchKernelGGL(diff_normals_kernel, dim3((s->H + WG - 1) / WG), dim3(WG), 0, st, s->d_helio, p);
    OM_HIP(hipGetLastError());
    hipLaunchKernelGGL(diff_gconv_kernel, dim3((p.npx + 16 + WG - 1) / WG), dim3(WG), 0, st, p);
    OM_HIP(hipGetLastError());
    if ((rc = launch_diff_grids<false>(s, p, st)) != OM_OK) return rc;
    hipLaunchKernelGGL(diff_params_kernel, dim3((s->H + WG - 1) / WG), dim3(WG), 0, st, s->d_helio, p,
                       (float4 *)d_grad);
    OM_HIP(hipGetLastError());
    return OM_OK;
}

extern "C" om_status om_diff_visibility(om_scene *s, const uint64_t *d_mask, int h0, int h1, uint8_t *d_vis,
                                        int64_t capacity, void *stream)
{
    if (!s || !d_mask || !d_vis || h0 < 0 || h1 > s->H || h0 > h1) return OM_ERR_INVALID;
    if (!s->ctx) return OM_ERR_INVALID;
    const long long n = s->ray_offset[h1] - s->ray_offset[h0];
    if (capacity < n) {
        set_error("visibility buffer smaller than the rays of the heliostat range");
        return OM_ERR_SIZE;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    auto lo = std::lower_bound(s->tiles.begin(), s->tiles.end(), h0, [](const int2 &t, int h) { return t.x < h; });
    auto hi = std::lower_bound(s->tiles.begin(), s->tiles.end(), h1, [](const int2 &t, int h) { return t.x < h; });
    const int t0 = (int)(lo - s->tiles.begin()), t1 = (int)(hi - s->tiles.begin());
    if (t1 <= t0) return OM_OK;
    TraceArgs a = make_trace_args(s);
    hipLaunchKernelGGL(diff_vis_kernel, dim3(t1 - t0), dim3(WG), 0, (hipStream_t)stream, a,
                       (const unsigned long long *)d_mask, t0, t1, (long long)s->ray_offset[h0], d_vis);
    OM_HIP(hipGetLastError());
    return OM_OK;
}
