// om_scene.cpp -- host side of the tracer library: .scn parsing, scene construction
// (SolarScene::InitContent), upload of the scene to HBM, and the C ABI.
//
// Reference: Hu_smooth/ClassedSunRay (file:line cited per function).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <thread>

#include "om_internal.h"

using om::v3;

namespace {
thread_local std::string g_last_error;

v3 ld3(const float *p) { return om::mk(p[0], p[1], p[2]); }

template <class T>
om_status upload(T *&dptr, const T *host, size_t n)
{
    if (dptr) {
        (void)hipFree(dptr);
        dptr = nullptr;
    }
    size_t bytes = sizeof(T) * (n > 0 ? n : 1);
    OM_HIP(hipMalloc((void **)&dptr, bytes));
    if (n > 0) OM_HIP(hipMemcpy(dptr, host, sizeof(T) * n, hipMemcpyHostToDevice));
    return OM_OK;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}
} // namespace

namespace om {
void set_error(const std::string &msg) { g_last_error = msg; }

om_status hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return (e == hipErrorOutOfMemory) ? OM_ERR_NOMEM : OM_ERR_HIP;
}

om_status ensure_buf(DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return OM_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    OM_HIP(hipMalloc(&b.p, bytes));
    b.bytes = bytes;
    return OM_OK;
}
} // namespace om

extern "C" const char *om_status_string(om_status s)
{
    switch (s) {
    case OM_OK: return "ok";
    case OM_ERR_INVALID: return "invalid argument";
    case OM_ERR_IO: return "i/o error";
    case OM_ERR_PARSE: return "scene file parse error";
    case OM_ERR_HIP: return "HIP runtime error";
    case OM_ERR_NOMEM: return "out of memory";
    case OM_ERR_UNSUPPORTED: return "unsupported configuration";
    case OM_ERR_SIZE: return "buffer too small";
    }
    return "unknown status";
}

extern "C" const char *om_last_error(void) { return g_last_error.c_str(); }

// Defaults of common_var.cpp:9-27 with the reference's lattice microsamples and the
// Buie sun (Nc = 2048 lights per group, CSR 0.1, slope error 1 mrad, pitch 0.01 m).
extern "C" void om_default_params(om_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->seed = 0x5EEDULL;
    p->scene_index = 0;
    p->sunshape = OM_SUN_BUIE;
    p->pillbox_half_angle = 4.65e-3f;
    p->csr = 0.1f;
    p->buie_bins = 1024;
    p->slope_sigma = 0.001f;
    p->micro_mode = OM_MICRO_LATTICE;
    p->helio_pixel_length = 0.01f;
    p->poisson_k = 8;
    p->rays_per_sample = 2048;
    p->reflectivity = 0.88f;
    p->smooth_radius = 5;
    p->smooth_ratio = 0.02f;
    p->accumulate = OM_ACC_PACKED;
    p->pack_shift = 0;
}

extern "C" om_status om_ctx_create(int device, om_ctx **out)
{
    if (!out) return OM_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    OM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        om::set_error("no such device");
        return OM_ERR_INVALID;
    }
    OM_HIP(hipSetDevice(device));
    void *sp = nullptr;
    const om_status rc = om::static_scratch(device, &sp);
    if (rc != OM_OK) return rc;
    om_ctx *c = new om_ctx();
    c->device = device;
    c->static_scratch = sp;
    // the abort word and the window traces' queue block in device memory, their pinned
    // staging buffers, and the stream that copies them while launches run.  The word is in
    // device memory because the kernels read it once per chunk: read from host memory, those
    // small reads over PCIe (about 1e5 per trace of 2e8 rays) cost more than the trace.
    hipError_t e = hipHostMalloc((void **)&c->h_abort, 64, hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_mon, om::QUEUE_WORDS * sizeof(uint32_t), hipHostMallocCoherent);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_abort, 64);
    if (e == hipSuccess) e = hipMemset(c->d_abort, 0, 64);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_queue, om::QUEUE_WORDS * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(c->d_queue, 0, om::QUEUE_WORDS * sizeof(uint32_t));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->mon_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        om_ctx_destroy(c);
        return om::hip_fail(e, "context set-up");
    }
    c->h_abort[0] = 0u;
    c->cur_queue = c->d_queue;
    *out = c;
    return OM_OK;
}

extern "C" om_status om_ctx_abort(om_ctx *ctx, int on)
{
    if (!ctx || !ctx->d_abort || !ctx->mon_stream) return OM_ERR_INVALID;
    OM_HIP(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(ctx->mon_mu);
    ctx->h_abort[0] = on ? 1u : 0u;
    // a copy on the context's stream: it lands while launches on other streams run, and the
    // kernels' system-scope loads of the word read memory, not a cache
    OM_HIP(hipMemcpyAsync(ctx->d_abort, ctx->h_abort, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->mon_stream));
    OM_HIP(hipStreamSynchronize(ctx->mon_stream));
    return OM_OK;
}

extern "C" om_status om_ctx_progress(om_ctx *ctx, int64_t *taken, int64_t *total)
{
    if (!ctx || !ctx->mon_stream) return OM_ERR_INVALID;
    OM_HIP(hipSetDevice(ctx->device));
    uint32_t q[om::QUEUE_WORDS];
    {
        std::lock_guard<std::mutex> lk(ctx->mon_mu);
        OM_HIP(hipMemcpyAsync(ctx->h_mon, ctx->cur_queue, om::QUEUE_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              ctx->mon_stream));
        OM_HIP(hipStreamSynchronize(ctx->mon_stream));
        std::memcpy(q, ctx->h_mon, sizeof(q));
    }
    const long long n = q[om::QUEUE_COUNTERS];
    long long done = 0;
    if (q[om::QUEUE_COUNTERS + 1] == om::QUEUE_ONE) {
        done = std::min<long long>(q[0], n);
    } else {
        for (int x = 0; x < 8; ++x) done += std::min<long long>(q[x], n / 8 + (x < n % 8 ? 1 : 0));
    }
    if (taken) *taken = done;
    if (total) *total = n;
    return OM_OK;
}

extern "C" om_status om_ctx_destroy(om_ctx *ctx)
{
    if (!ctx) return OM_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->mon_stream) (void)hipStreamDestroy(ctx->mon_stream);
    if (ctx->d_queue) (void)hipFree(ctx->d_queue);
    if (ctx->d_abort) (void)hipFree(ctx->d_abort);
    if (ctx->h_mon) (void)hipHostFree(ctx->h_mon);
    if (ctx->h_abort) (void)hipHostFree(ctx->h_abort);
    if (ctx->scratch.p) (void)hipFree(ctx->scratch.p);
    if (ctx->scratch2.p) (void)hipFree(ctx->scratch2.p);
    if (ctx->parts.p) (void)hipFree(ctx->parts.p);
    delete ctx;
    return OM_OK;
}

// ---------------------------------------------------------------------------------------
// .scn parser: SceneFileProc::SceneFileRead(SolarScene*, std::string)
// (scene_file_proc.cpp:41-286) with its grammar and its quirks:
//  * lines whose first character is '#' and empty lines are skipped; a line holding only
//    blanks is not empty and reads as the key "" ;
//  * keys unknown to the block map to "pos" (std::map::operator[] default, :155,187,232);
//  * a failed numeric read stores 0 (std::istream semantics);
//  * the size of a heliostat is read from the line after its "helio" line (:3-11);
//  * "gap" and "matrix" hold until the next Grid block, which forgets both (:136-137), and a
//    heliostat line before both are set is skipped (:246-249);
//  * a grid's heliostats start at the sum of the "n" of the grids before it (:133, :199-201).
// The first receiver is kept (the reference traces receivers[0]); every grid is kept.
// Where the reference prints a warning and goes on into undefined behaviour, this parser
// stops with an error (see om_scn_parse).
// ---------------------------------------------------------------------------------------
namespace {
enum Key { K_POS = 0, K_SIZE, K_NORM, K_FACE, K_END, K_GAP, K_MATRIX, K_HELIO, K_INTER, K_N, K_TYPE };
enum Mode { M_NONE, M_GROUND, M_RECV, M_GRID, M_HELIO };

struct ParsedGrid {
    int type = 0;                 // Grid <type> (only 0, RectGrid, is defined)
    float pos[3] = {0, 0, 0}, size[3] = {0, 0, 0}, inter[3] = {1, 1, 1};
    int n = 0;                    // declared heliostats (num_helios_)
    int helio_type = 0;           // heliostat type of the grid (helio_type_)
    int first = 0;                // start_helio_pos_
    int listed = 0;               // heliostats read under it
    int skipped = 0;              // heliostat lines skipped (type != 0, no gap / matrix)
};

struct Parsed {
    std::vector<float> pos, size, gap;
    std::vector<int> matrix;
    int recv_type = 0;
    float recv_pos[3] = {0, 0, 0}, recv_size[3] = {0, 0, 0}, recv_norm[3] = {0, 0, 1};
    int recv_face = 0;
    int n_recv = 0;
    float ground[2] = {0, 0};
    int ngrid = 0;
    std::vector<ParsedGrid> grids;
};

om_status parse_fail(om_status st, const std::string &msg)
{
    om::set_error(msg);
    return st;
}

om_status parse_scn(const char *path, Parsed &P)
{
    std::ifstream f(path);
    if (!f.is_open()) return parse_fail(OM_ERR_IO, std::string("cannot open scene file ") + path);
    std::stringstream scene;
    scene << f.rdbuf();
    std::map<std::string, int> keys = {{"pos", K_POS}, {"size", K_SIZE}, {"norm", K_NORM},
                                       {"face", K_FACE}, {"end", K_END}, {"gap", K_GAP},
                                       {"matrix", K_MATRIX}, {"helio", K_HELIO},
                                       {"inter", K_INTER}, {"n", K_N}, {"type", K_TYPE}};
    Mode mode = M_NONE;
    bool gap_set = false, matrix_set = false, in_recv = false, in_grid = false;
    float gap[2] = {0, 0};
    int mat[2] = {1, 1};
    int next_first = 0;
    std::string line;
    while (std::getline(scene, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::stringstream ls;
        ls << line;
        std::string head;
        ls >> head;
        if (head == "ground") {
            mode = M_GROUND;
            ls >> P.ground[0] >> P.ground[1];
            continue;
        } else if (head == "Recv") {
            mode = M_RECV;
            int t = 0;
            ls >> t;
            if (t < 0 || t > 2)
                return parse_fail(OM_ERR_PARSE, "receiver type " + std::to_string(t) + " is not defined (0..2)");
            // a new receiver object; the first one that reaches "end" is receivers[0]
            in_recv = P.n_recv == 0;
            if (in_recv) {
                P.recv_type = t;
                for (int a = 0; a < 3; ++a) P.recv_pos[a] = P.recv_size[a] = 0.0f;
                P.recv_norm[0] = P.recv_norm[1] = 0.0f;
                P.recv_norm[2] = 1.0f;
                P.recv_face = 0;
            }
            continue;
        } else if (head == "Grid") {
            mode = M_GRID;
            ParsedGrid g;
            ls >> g.type;
            g.first = next_first;
            P.grids.push_back(g);
            in_grid = true;
            gap_set = matrix_set = false;
            continue;
        }
        int key = keys[head]; // unknown keys read as "pos"
        switch (mode) {
        case M_NONE:
            break;
        case M_GROUND:
            if (head == "ngrid") ls >> P.ngrid;
            break;
        case M_RECV:
            if (!in_recv) break;
            if (key == K_POS) ls >> P.recv_pos[0] >> P.recv_pos[1] >> P.recv_pos[2];
            else if (key == K_SIZE) ls >> P.recv_size[0] >> P.recv_size[1] >> P.recv_size[2];
            else if (key == K_NORM) ls >> P.recv_norm[0] >> P.recv_norm[1] >> P.recv_norm[2];
            else if (key == K_FACE) ls >> P.recv_face;
            else if (key == K_END) {
                ++P.n_recv;
                in_recv = false;
            }
            break;
        case M_GRID: {
            ParsedGrid &g = P.grids.back();
            if (g.type != 0)
                return parse_fail(OM_ERR_UNSUPPORTED, "grid type " + std::to_string(g.type) +
                                                          " is not supported (only 0, the rectangular grid)");
            if (!in_grid) break;
            if (key == K_POS) ls >> g.pos[0] >> g.pos[1] >> g.pos[2];
            else if (key == K_SIZE) ls >> g.size[0] >> g.size[1] >> g.size[2];
            else if (key == K_INTER) ls >> g.inter[0] >> g.inter[1] >> g.inter[2];
            else if (key == K_N) {
                int n = 0;
                ls >> n;
                g.n = n;
                next_first += n;
            } else if (key == K_TYPE) ls >> g.helio_type;
            else if (key == K_END) {
                in_grid = false;
                mode = M_HELIO;
            }
            break;
        }
        case M_HELIO: {
            ParsedGrid &g = P.grids.back();
            if (key == K_GAP) {
                ls >> gap[0] >> gap[1];
                gap_set = true;
            } else if (key == K_MATRIX) {
                ls >> mat[0] >> mat[1];
                matrix_set = true;
            } else if (key == K_HELIO) {
                if (!gap_set || !matrix_set || g.helio_type != 0) {
                    ++g.skipped; // "did not set the gap and matrix", or not a rectangle
                    break;
                }
                float p[3] = {0, 0, 0};
                ls >> p[0] >> p[1] >> p[2];
                std::string nxt;
                std::getline(scene, nxt);
                std::stringstream ns;
                ns << nxt;
                float sz[3] = {0, 0, 0};
                ns >> sz[0] >> sz[1] >> sz[2];
                P.pos.insert(P.pos.end(), p, p + 3);
                P.size.insert(P.size.end(), sz, sz + 3);
                P.gap.insert(P.gap.end(), gap, gap + 2);
                P.matrix.insert(P.matrix.end(), mat, mat + 2);
                ++g.listed;
            }
            break;
        }
        }
    }
    if (P.grids.empty()) return parse_fail(OM_ERR_PARSE, "scene file has no grid");
    if (P.n_recv == 0) return parse_fail(OM_ERR_PARSE, "scene file has no receiver");
    for (size_t k = 0; k < P.grids.size(); ++k) {
        const ParsedGrid &g = P.grids[k];
        const std::string gk = "grid " + std::to_string(k);
        if (g.helio_type != 0 && (g.n > 0 || g.skipped > 0))
            return parse_fail(OM_ERR_UNSUPPORTED, gk + ": heliostat type " + std::to_string(g.helio_type) +
                                                      " is not supported (only 0, the rectangular heliostat)");
        if (g.n < 0 || g.listed != g.n)
            return parse_fail(OM_ERR_PARSE, gk + " declares " + std::to_string(g.n) + " heliostats and lists " +
                                                std::to_string(g.listed) +
                                                (g.skipped ? " (heliostat lines before its gap and matrix are skipped)" : ""));
    }
    return OM_OK;
}

void put_grid(om_grid_desc &o, const float pos[3], const float size[3], const float inter[3], int first, int n)
{
    std::memcpy(o.pos, pos, 12);
    std::memcpy(o.size, size, 12);
    std::memcpy(o.interval, inter, 12);
    o.first_helio = first;
    o.n_helio = n;
}
} // namespace

extern "C" om_status om_scn_parse(const char *path, om_scene_desc *d)
{
    if (!path || !d) return OM_ERR_INVALID;
    std::memset(d, 0, sizeof(*d));
    Parsed P;
    om_status st = parse_scn(path, P);
    if (st != OM_OK) return st;
    int H = (int)(P.pos.size() / 3);
    float *pos = new float[3 * H + 3], *size = new float[3 * H + 3], *gap = new float[2 * H + 2];
    int *mat = new int[2 * H + 2];
    std::copy(P.pos.begin(), P.pos.end(), pos);
    std::copy(P.size.begin(), P.size.end(), size);
    std::copy(P.gap.begin(), P.gap.end(), gap);
    std::copy(P.matrix.begin(), P.matrix.end(), mat);
    om_grid_desc *grids = new om_grid_desc[P.grids.size()];
    for (size_t k = 0; k < P.grids.size(); ++k) {
        const ParsedGrid &g = P.grids[k];
        put_grid(grids[k], g.pos, g.size, g.inter, g.first, g.n);
    }
    d->n_helio = H;
    d->helio_pos = pos;
    d->helio_size = size;
    d->helio_gap = gap;
    d->helio_matrix = mat;
    d->recv_type = P.recv_type;
    std::memcpy(d->recv_pos, P.recv_pos, 12);
    std::memcpy(d->recv_size, P.recv_size, 12);
    std::memcpy(d->recv_normal, P.recv_norm, 12);
    d->recv_face = P.recv_face;
    d->recv_pixel_length = 0.01f; // solarenergy::receiver_pixel_length (common_var.cpp:12)
    std::memcpy(d->grid_pos, grids[0].pos, 12);
    std::memcpy(d->grid_size, grids[0].size, 12);
    std::memcpy(d->grid_interval, grids[0].interval, 12);
    d->n_grids = (int)P.grids.size();
    d->grids = grids;
    d->ground_length = P.ground[0];
    d->ground_width = P.ground[1];
    d->dni = 1000.0f;
    return OM_OK;
}

extern "C" void om_scene_desc_free(om_scene_desc *d)
{
    if (!d) return;
    delete[] d->helio_pos;
    delete[] d->helio_size;
    delete[] d->helio_gap;
    delete[] d->helio_matrix;
    delete[] d->grids;
    std::memset(d, 0, sizeof(*d));
}

// .scn writer: the grammar SceneFileRead reads (scene_file_proc.cpp:41-286).  "%.9g" prints
// every float with enough digits that the stream extraction gives it back exactly.
extern "C" om_status om_scn_write(const char *path, const om_scene_desc *d)
{
    if (!path || !d || d->n_helio < 0 || d->n_grids < 0 || (d->n_grids > 0 && !d->grids) ||
        (d->n_helio > 0 && (!d->helio_pos || !d->helio_size || !d->helio_gap || !d->helio_matrix)))
        return OM_ERR_INVALID;
    std::vector<om_grid_desc> grids;
    if (d->n_grids == 0) {
        om_grid_desc g;
        put_grid(g, d->grid_pos, d->grid_size, d->grid_interval, 0, d->n_helio);
        grids.push_back(g);
    } else {
        grids.assign(d->grids, d->grids + d->n_grids);
    }
    int next = 0;
    for (const om_grid_desc &g : grids) {
        if (g.first_helio != next || g.n_helio < 0) {
            om::set_error("grids must hold consecutive heliostat ranges from 0");
            return OM_ERR_INVALID;
        }
        next += g.n_helio;
    }
    if (next != d->n_helio) {
        om::set_error("the grids' heliostats must add up to n_helio");
        return OM_ERR_INVALID;
    }
    if (d->recv_type < 0 || d->recv_type > 2) {
        om::set_error("receiver type must be 0..2");
        return OM_ERR_INVALID;
    }
    // every number must read back: finite floats only
    auto finite = [](const float *x, size_t n) {
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(x[i])) return false;
        return true;
    };
    bool ok = finite(d->recv_pos, 3) && finite(d->recv_size, 3) && finite(d->recv_normal, 3) &&
              finite(&d->ground_length, 1) && finite(&d->ground_width, 1) &&
              finite(d->helio_pos, 3 * (size_t)d->n_helio) && finite(d->helio_size, 3 * (size_t)d->n_helio) &&
              finite(d->helio_gap, 2 * (size_t)d->n_helio);
    for (const om_grid_desc &g : grids) ok = ok && finite(g.pos, 3) && finite(g.size, 3) && finite(g.interval, 3);
    if (!ok) {
        om::set_error("scene values must be finite to be written");
        return OM_ERR_INVALID;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) {
        om::set_error(std::string("cannot write scene file ") + path);
        return OM_ERR_IO;
    }
    auto v3 = [&](const char *key, const float *x) { std::fprintf(f, "%s\t%.9g\t%.9g\t%.9g\n", key, x[0], x[1], x[2]); };
    std::fprintf(f, "# Ground Boundary\nground\t%.9g\t%.9g\nngrid\t%d\n\n", d->ground_length, d->ground_width,
                 (int)grids.size());
    std::fprintf(f, "# Receiver attributes\nRecv\t%d\n", d->recv_type);
    v3("pos", d->recv_pos);
    v3("size", d->recv_size);
    v3("norm", d->recv_normal);
    std::fprintf(f, "face\t%d\nend\n", d->recv_face);
    for (size_t k = 0; k < grids.size(); ++k) {
        const om_grid_desc &g = grids[k];
        std::fprintf(f, "\n# Grid%zu attributes\nGrid\t0\n", k);
        v3("pos", g.pos);
        v3("size", g.size);
        v3("inter", g.interval);
        std::fprintf(f, "n\t%d\ntype\t0\nend\n\n# Heliostats\n", g.n_helio);
        bool first = true;
        for (int i = g.first_helio; i < g.first_helio + g.n_helio; ++i) {
            const float *gp = d->helio_gap + 2 * (size_t)i;
            const int *mp = d->helio_matrix + 2 * (size_t)i;
            if (first || gp[0] != gp[-2] || gp[1] != gp[-1] ||
                std::signbit(gp[0]) != std::signbit(gp[-2]) || std::signbit(gp[1]) != std::signbit(gp[-1]))
                std::fprintf(f, "gap\t%.9g\t%.9g\n", gp[0], gp[1]);
            if (first || mp[0] != mp[-2] || mp[1] != mp[-1]) std::fprintf(f, "matrix\t%d\t%d\n", mp[0], mp[1]);
            first = false;
            v3("helio", d->helio_pos + 3 * (size_t)i);
            const float *sz = d->helio_size + 3 * (size_t)i;
            std::fprintf(f, "\t%.9g\t%.9g\t%.9g\n", sz[0], sz[1], sz[2]);
        }
    }
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) {
        om::set_error(std::string("error writing scene file ") + path);
        return OM_ERR_IO;
    }
    return OM_OK;
}

// ---------------------------------------------------------------------------------------
// scene construction
// ---------------------------------------------------------------------------------------
namespace {

// set_samplelights host part (scene_instance_process.cu:116-156, :223-260)
void build_buie(om_scene *s)
{
    int num_group = s->prm.buie_bins > 0 ? s->prm.buie_bins : 1024;
    float csr = s->prm.csr;
    float upper_lm = 4.65f;
    float k = 0.9f * logf(13.5f * csr) * powf(csr, -0.3f);
    float gamma = 2.2f * logf(0.52f * csr) * powf(csr, 0.43f) - 0.1f;
    float i93 = expf(k) / (gamma + 1) * powf(9.3f, gamma + 1);
    float i465 = expf(k) / (gamma + 1) * powf(upper_lm, gamma + 1);
    s->nbins = num_group;
    s->cdf_x.assign(num_group + 1, 0.0f);
    s->cdf_y.assign(num_group + 1, 0.0f);
    s->slope.assign(num_group, 0.0f);
    float li = upper_lm / (float)num_group;
    float x = 0.0f;
    float hp = cosf((float)(0.326 * x)) / cosf((float)(0.308 * x));
    for (int i = 1; i <= num_group; ++i) {
        x += li;
        float hc = cosf((float)(0.326 * x)) / cosf((float)(0.308 * x));
        float sub = (hp + hc) * li / 2.0f;
        s->cdf_x[i] = x;
        s->cdf_y[i] = sub + s->cdf_y[i - 1];
        s->slope[i - 1] = li / sub;
        hp = hc;
    }
    s->buie_gamma = gamma;
    s->buie_int465 = s->cdf_y[num_group];
    s->buie_A = s->buie_int465 + (i93 - i465);
    s->buie_B = (gamma + 1) / expf(k);
    s->buie_C = powf(4.65f, gamma + 1);
    s->buie_inv_gp1 = 1 / (gamma + 1);
}

// RectGrid::Cinit + CGridHelioMatch (grid.cpp:3-82) of grid G over its heliostats
// [G.h0, G.h1), with global heliostat ids (grid.cpp:53); binning clamped to the grid.
om_status build_grid(const om_scene *s, om::GridState &G)
{
    v3 gp = ld3(G.pos), gs = ld3(G.size), gi = ld3(G.interval);
    G.gn[0] = (int)ceilf(gs.x / gi.x);
    G.gn[1] = (int)ceilf(gs.y / gi.y);
    G.gn[2] = (int)ceilf(gs.z / gi.z);
    if (G.gn[0] <= 0 || G.gn[1] <= 0 || G.gn[2] <= 0) {
        om::set_error("grid has no cells");
        return OM_ERR_INVALID;
    }
    G.cells = (long long)G.gn[0] * G.gn[1] * G.gn[2];
    if (G.cells >= (1LL << 31) - 1) {
        om::set_error("grid too large");
        return OM_ERR_UNSUPPORTED;
    }
    std::vector<int> cnt((size_t)G.cells, 0);
    const int nh = G.h1 - G.h0;
    std::vector<int> rng((size_t)nh * 6);
    for (int a = 0; a < 3; ++a) {
        G.occ_lo[a] = 1 << 30;
        G.occ_hi[a] = -1;
    }
    long long nnz = 0;
    for (int i = G.h0; i < G.h1; ++i) {
        v3 sz = ld3(&s->size[3 * i]), p = ld3(&s->pos[3 * i]);
        float diag = om::sqrt_(om::dot(sz, sz));
        float radius = om::sqrt_(diag) / 2;
        v3 mn = om::mk(p.x - radius, p.y - radius, p.z - radius);
        v3 mx = om::mk(p.x + radius, p.y + radius, p.z + radius);
        v3 a0 = mn - gp, a1 = mx - gp;
        int lo[3] = {om::trunc_i(a0.x / gi.x), om::trunc_i(a0.y / gi.y), om::trunc_i(a0.z / gi.z)};
        int hi[3] = {om::trunc_i(a1.x / gi.x), om::trunc_i(a1.y / gi.y), om::trunc_i(a1.z / gi.z)};
        int *r = &rng[6 * (size_t)(i - G.h0)];
        for (int a = 0; a < 3; ++a) {
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > G.gn[a] - 1) hi[a] = G.gn[a] - 1;
            r[a] = lo[a];
            r[3 + a] = hi[a];
        }
        bool any = lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2];
        if (!any) continue;
        for (int a = 0; a < 3; ++a) {
            G.occ_lo[a] = std::min(G.occ_lo[a], lo[a]);
            G.occ_hi[a] = std::max(G.occ_hi[a], hi[a]);
        }
        for (int x = lo[0]; x <= hi[0]; ++x)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    cnt[(size_t)x * G.gn[1] * G.gn[2] + (size_t)y * G.gn[2] + z]++;
                    ++nnz;
                }
    }
    if (nnz >= (1LL << 31) - 1) {
        om::set_error("grid holds too many entries");
        return OM_ERR_UNSUPPORTED;
    }
    G.nnz = nnz;
    G.cell_start.assign((size_t)G.cells + 1, 0);
    for (long long c = 0; c < G.cells; ++c) G.cell_start[c + 1] = G.cell_start[c] + cnt[c];
    G.cell_items.assign((size_t)(nnz > 0 ? nnz : 1), 0);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = G.h0; i < G.h1; ++i) {
        const int *r = &rng[6 * (size_t)(i - G.h0)];
        for (int x = r[0]; x <= r[3]; ++x)
            for (int y = r[1]; y <= r[4]; ++y)
                for (int z = r[2]; z <= r[5]; ++z) {
                    size_t c = (size_t)x * G.gn[1] * G.gn[2] + (size_t)y * G.gn[2] + z;
                    G.cell_items[G.cell_start[c] + cnt[c]++] = i;
                }
    }
    return OM_OK;
}

// Cylinder receiver (new; CylinderReceiver is an empty stub in receiver.cuh:103-122): axis
// through recv_pos (mid-height), radius size.x, height size.y.  The map unrolls the outer
// surface, res_x columns over the azimuth and res_y rows over the height, given or from
// geo = int(1 / pixel length) as for rectangles.  Heliostats aim at the axis point.
om_status build_cylinder(om_scene *s, int geo)
{
    const float r = s->recv_size[0], h = s->recv_size[1];
    if (!(r > 0.0f && h > 0.0f)) {
        om::set_error("cylinder receiver needs a positive radius and height");
        return OM_ERR_INVALID;
    }
    int rx = s->recv_res_in[0], ry = s->recv_res_in[1];
    if (rx <= 0 || ry <= 0) {
        rx = (int)(6.28318531f * r * (float)geo);
        ry = (int)(h * (float)geo);
    }
    if (rx <= 0 || ry <= 0 || (long long)rx * ry >= (1LL << 28)) {
        om::set_error("cylinder map must have 1 .. 2^28 pixels");
        return OM_ERR_INVALID;
    }
    s->res_x = rx;
    s->res_y = ry;
    s->cyl_base = om::mk(s->recv_pos[0], s->recv_pos[1] - h / 2, s->recv_pos[2]);
    s->cyl_r2 = r * r;
    s->cyl_h = h;
    s->recv_pixel = h / (float)ry;
    s->srec = (6.28318531f * r / (float)rx) * (h / (float)ry);
    s->focus = ld3(s->recv_pos);
    s->recv_normal = om::mk(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < 4; ++i) s->rv[i] = om::mk(0.0f, 0.0f, 0.0f);
    return OM_OK;
}

// Truncated-cone receiver (new; CircularTruncatedConeReceiver is an empty stub in
// receiver.cuh:125-140, whose topradius_bottomradius_hight_ gives the order of size): axis
// through recv_pos (mid-height), top radius size.x, bottom radius size.y, height size.z.  The
// map unrolls the lateral surface, res_x columns over the azimuth and res_y rows of equal
// area over the height, given or from geo = int(1 / pixel length) (the mean circumference and
// the slant height in pixels).  Heliostats aim at the axis point.
om_status build_cone(om_scene *s, int geo)
{
    const float r1 = s->recv_size[0], r0 = s->recv_size[1], h = s->recv_size[2];
    if (!(r0 >= 0.0f && r1 >= 0.0f && r0 + r1 > 0.0f && h > 0.0f)) {
        om::set_error("cone receiver needs radii >= 0 (not both 0) and a positive height");
        return OM_ERR_INVALID;
    }
    const float slant = om::sqrt_(h * h + (r1 - r0) * (r1 - r0));
    int rx = s->recv_res_in[0], ry = s->recv_res_in[1];
    if (rx <= 0 || ry <= 0) {
        rx = (int)(3.14159265f * (r0 + r1) * (float)geo);
        ry = (int)(slant * (float)geo);
    }
    if (rx <= 0 || ry <= 0 || (long long)rx * ry >= (1LL << 28)) {
        om::set_error("cone map must have 1 .. 2^28 pixels");
        return OM_ERR_INVALID;
    }
    s->res_x = rx;
    s->res_y = ry;
    s->cyl_base = om::mk(s->recv_pos[0], s->recv_pos[1] - h / 2, s->recv_pos[2]);
    s->cyl_h = h;
    s->cone_r0 = r0;
    s->cone_k = (r1 - r0) / h;
    s->cone_inv_area = 1.0f / (h * (r0 + r1));
    s->recv_pixel = slant / (float)ry;
    s->srec = 3.14159265f * (r0 + r1) * slant / ((float)rx * (float)ry);
    s->focus = ld3(s->recv_pos);
    s->recv_normal = om::mk(0.0f, 0.0f, 0.0f);
    for (int i = 0; i < 4; ++i) s->rv[i] = om::mk(0.0f, 0.0f, 0.0f);
    return OM_OK;
}

// RectangleReceiver::CInit (receiver.cpp:24-117)
om_status build_receiver(om_scene *s)
{
    if (s->recv_type < 0 || s->recv_type > 2) {
        om::set_error("receiver type must be 0 (rectangle), 1 (cylinder) or 2 (truncated cone)");
        return OM_ERR_UNSUPPORTED;
    }
    if (!(s->recv_pixel_length > 0.0f)) {
        om::set_error("receiver pixel length must be positive");
        return OM_ERR_INVALID;
    }
    int geo = (int)(1.0f / s->recv_pixel_length);
    if (geo <= 0 && !(s->recv_type != 0 && s->recv_res_in[0] > 0 && s->recv_res_in[1] > 0)) {
        om::set_error("receiver pixel length above 1 m");
        return OM_ERR_UNSUPPORTED;
    }
    s->geo_px = geo;
    if (s->recv_type == 1) return build_cylinder(s, geo);
    if (s->recv_type == 2) return build_cone(s, geo);
    s->recv_pixel = 1.0f / (float)geo;
    s->srec = s->recv_pixel * s->recv_pixel;
    v3 sz = ld3(s->recv_size), p = ld3(s->recv_pos);
    v3 ln, v[4];
    switch (s->recv_face) {
    case 0:
        ln = om::mk(0.0f, 0.0f, 1.0f);
        v[0] = om::mk(-sz.x / 2, -sz.y / 2, sz.z / 2);
        v[1] = om::mk(-sz.x / 2, sz.y / 2, sz.z / 2);
        v[2] = om::mk(sz.x / 2, sz.y / 2, sz.z / 2);
        v[3] = om::mk(sz.x / 2, -sz.y / 2, sz.z / 2);
        break;
    case 1:
        ln = om::mk(1.0f, 0.0f, 0.0f);
        v[0] = om::mk(sz.x / 2, -sz.y / 2, sz.z / 2);
        v[1] = om::mk(sz.x / 2, sz.y / 2, sz.z / 2);
        v[2] = om::mk(sz.x / 2, sz.y / 2, -sz.z / 2);
        v[3] = om::mk(sz.x / 2, -sz.y / 2, -sz.z / 2);
        break;
    case 2:
        ln = om::mk(0.0f, 0.0f, -1.0f);
        v[0] = om::mk(sz.x / 2, -sz.y / 2, -sz.z / 2);
        v[1] = om::mk(sz.x / 2, sz.y / 2, -sz.z / 2);
        v[2] = om::mk(-sz.x / 2, sz.y / 2, -sz.z / 2);
        v[3] = om::mk(-sz.x / 2, -sz.y / 2, -sz.z / 2);
        break;
    case 3:
        ln = om::mk(-1.0f, 0.0f, 0.0f);
        v[0] = om::mk(-sz.x / 2, -sz.y / 2, -sz.z / 2);
        v[1] = om::mk(-sz.x / 2, sz.y / 2, -sz.z / 2);
        v[2] = om::mk(-sz.x / 2, sz.y / 2, sz.z / 2);
        v[3] = om::mk(-sz.x / 2, -sz.y / 2, sz.z / 2);
        break;
    default:
        om::set_error("receiver face must be 0..3");
        return OM_ERR_INVALID;
    }
    v3 n = om::normalize(ld3(s->recv_normal_in));
    s->recv_normal = n;
    for (int i = 0; i < 4; ++i) s->rv[i] = om::rotate_y(v[i], ln, n) + p;
    s->focus = (s->rv[0] + s->rv[2]) * (1.0f / 2);
    s->res_x = (int)(sz.x * (float)geo);
    s->res_y = (int)(sz.y * (float)geo);
    if (s->res_x <= 0 || s->res_y <= 0) {
        om::set_error("receiver has no pixels");
        return OM_ERR_INVALID;
    }
    return OM_OK;
}

// RectangleHelio::CRotate(focus, sun) (heliostat.cpp:16-62), or CRotate(dir) with the given
// normals (heliostat.cpp:22-27)
void aim_heliostats(om_scene *s)
{
    s->normal.resize(s->H);
    s->v0.resize(s->H);
    s->v1.resize(s->H);
    s->v2.resize(s->H);
    s->v3.resize(s->H);
    for (int i = 0; i < s->H; ++i) {
        v3 p = ld3(&s->pos[3 * i]), sz = ld3(&s->size[3 * i]);
        v3 n;
        if (!s->normal_in.empty()) {
            n = om::normalize(ld3(&s->normal_in[3 * i])); // Cset_normal(dir) (heliostat.cpp:73-76)
        } else {
            v3 rd = om::normalize(s->focus - p);
            n = om::normalize(rd - s->sun);
        }
        s->normal[i] = n;
        om::frame f = om::make_frame(n);
        v3 l0 = om::mk(-sz.x / 2, sz.y / 2, -sz.z / 2);
        v3 l1 = om::mk(l0.x, l0.y, l0.z + sz.z);
        v3 l2 = om::mk(l1.x + sz.x, l1.y, l1.z);
        v3 l3 = om::mk(l0.x + sz.x, l0.y, l0.z);
        s->v0[i] = om::to_world(l0, f) + p;
        s->v1[i] = om::to_world(l1, f) + p;
        s->v2[i] = om::to_world(l2, f) + p;
        s->v3[i] = om::to_world(l3, f) + p;
    }
}

// microsample layout (steps_for_raytracing.cu:67-82, :105-125)
void micro_layout(om_scene *s)
{
    int H = s->H;
    float pitch = s->prm.helio_pixel_length;
    s->subrow_len.resize(H);
    s->subcol_len.resize(H);
    s->sub_rows.resize(H);
    s->sub_cols.resize(H);
    s->mcount.assign(H, 0);
    s->poisson_cap = 0;
    float r = pitch / om::sqrt_(1.5f);
    float cell = r / om::sqrt_(2.0f);
    for (int i = 0; i < H; ++i) {
        int rows = s->matrix[2 * i], cols = s->matrix[2 * i + 1];
        float gx = s->gap[2 * i], gy = s->gap[2 * i + 1];
        s->subrow_len[i] = (s->size[3 * i + 2] - gy * (float)(rows - 1)) / (float)rows;
        s->subcol_len[i] = (s->size[3 * i + 0] - gx * (float)(cols - 1)) / (float)cols;
        s->sub_rows[i] = om::trunc_i(s->subrow_len[i] / pitch);
        s->sub_cols[i] = om::trunc_i(s->subcol_len[i] / pitch);
        if (s->prm.micro_mode == OM_MICRO_POISSON) {
            float width = s->subcol_len[i] * (float)cols;
            float height = s->subrow_len[i] * (float)rows;
            int wi = (int)(ceilf(width / cell) + 2), hi = (int)(ceilf(height / cell) + 2);
            s->poisson_cap = std::max(s->poisson_cap, wi * hi);
        } else {
            s->mcount[i] = s->sub_rows[i] * s->sub_cols[i] * rows * cols;
        }
    }
}

void finish_micro(om_scene *s)
{
    int H = s->H;
    int L = s->prm.rays_per_sample;
    s->ray_offset.assign(H + 1, 0);
    s->total_micro = 0;
    s->max_micro = 0;
    for (int i = 0; i < H; ++i) {
        s->ray_offset[i + 1] = s->ray_offset[i] + (long long)s->mcount[i] * L;
        s->total_micro += s->mcount[i];
        s->max_micro = std::max(s->max_micro, s->mcount[i]);
    }
    s->total_rays = s->ray_offset[H];
}

// results_with_given_normal normalisation (Gtest.cu:14,29,37-46)
void flux_scale(om_scene *s)
{
    float Id = s->dni;
    float total = 0.0f;
    for (int i = 0; i < s->H; ++i)
        total += om::fabs_(om::dot(s->normal[i], s->sun) * Id * s->size[3 * i] * s->size[3 * i + 2]);
    float Ssub = s->prm.helio_pixel_length * s->prm.helio_pixel_length;
    float rou = s->prm.reflectivity;
    int Nc = s->prm.rays_per_sample;
    float Srec = s->srec;
    total *= rou / Srec;
    s->total_energy = total;
    s->scale = Id * Ssub * rou / (float)Nc / Srec;
}

// HelioRec array and blocker geometry (A, E1 = B - A, E2 = C - A of
// rayParallelogramIntersect with A = v0, B = v1, C = v3: rectgrid_raytracing.cu:29-31)
// Largest angle between a sun ray of the sunshape and the sun direction, rounded up: the
// pillbox's half angle, or the Buie inverse's top, 9.3 mrad (the tail's inverse at the top
// of the value range is (9.3^(g+1))^(1/(g+1)); the bins stop at 4.65 mrad).
double sunshape_max_angle(const om_scene *s)
{
    const double th = s->prm.sunshape == OM_SUN_BUIE ? 9.3e-3 : (double)s->prm.pillbox_half_angle;
    return th * 1.01 + 1e-5;
}

// Heliostats whose shading walk finds nothing whatever the ray: every other heliostat is
// kept out of the set of rays that leave this mirror toward the sun within the sunshape's
// largest angle alpha.  Looked along the sun direction d0, a ray from mirror point p that
// meets heliostat k at x (t = |x - p| <= T, the largest corner distance) projects within
// T sin(alpha) of p's projection; so k is out of reach when, on the plane normal to d0, the
// projections of the two parallelograms are farther apart than T alpha + 0.1 m (a
// separating edge normal of either quadrilateral), or when every point of k lies more than
// 0.1 m behind every mirror point along every such direction.  Candidates whose plane comes
// within 0.05 rad of the rays are never cleared.  The 0.1 m and 0.05 rad margins exceed the
// float error of the kernel's intersection (coordinates within 2^20 m give errors of about
// 1e-4 m, divided by the sine of the grazing angle), so a cleared heliostat's walk returns
// -1 for every ray, as the walk itself would.  Only heliostats the rays can reach before
// climbing above every mirror are candidates (horizontal reach from the sun's elevation
// less alpha); with the sun within 0.05 rad of that elevation nothing is cleared.
void mark_clear(om_scene *s)
{
    const int H = s->H;
    for (om::HelioRec &r : s->recs) r.clear = 0;
    if (H == 0) return;
    const double alpha = sunshape_max_angle(s);
    const double sl = std::sqrt((double)s->sun.x * s->sun.x + (double)s->sun.y * s->sun.y +
                                (double)s->sun.z * s->sun.z);
    const double d0[3] = {-s->sun.x / sl, -s->sun.y / sl, -s->sun.z / sl};
    const double emin = std::asin(std::min(1.0, d0[1])) - alpha;
    if (!(emin > 0.05)) return;
    const double cot = std::cos(emin) / std::sin(emin);
    // plane basis normal to d0
    double e1[3], e2[3];
    {
        const double ax[3] = {std::fabs(d0[0]) < 0.9 ? 1.0 : 0.0, std::fabs(d0[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        e1[0] = d0[1] * ax[2] - d0[2] * ax[1];
        e1[1] = d0[2] * ax[0] - d0[0] * ax[2];
        e1[2] = d0[0] * ax[1] - d0[1] * ax[0];
        const double l = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (double &c : e1) c /= l;
        e2[0] = d0[1] * e1[2] - d0[2] * e1[1];
        e2[1] = d0[2] * e1[0] - d0[0] * e1[2];
        e2[2] = d0[0] * e1[1] - d0[1] * e1[0];
    }
    // mirror quadrilaterals (ray origins: loc = (+-sx2, sy2, +-sz2) through the frame) and
    // blocker quadrilaterals (v0, v1, v2, v3), in double
    struct Quad {
        double p[4][3];
        double cx, cz, r; // horizontal centre and radius
        double ylo, yhi;
    };
    auto finish = [](Quad &q) {
        q.cx = q.cz = 0.0;
        q.ylo = 1e300;
        q.yhi = -1e300;
        for (auto &c : q.p) {
            q.cx += c[0] / 4;
            q.cz += c[2] / 4;
            q.ylo = std::min(q.ylo, c[1]);
            q.yhi = std::max(q.yhi, c[1]);
        }
        q.r = 0.0;
        for (auto &c : q.p) q.r = std::max(q.r, std::hypot(c[0] - q.cx, c[2] - q.cz));
    };
    std::vector<Quad> mir(H), blk(H);
    double ytop = -1e300;
    for (int i = 0; i < H; ++i) {
        const om::HelioRec &r = s->recs[i];
        om::frame f;
        f.n = om::mk(r.n[0], r.n[1], r.n[2]);
        f.u = om::mk(r.fu[0], r.fu[1], r.fu[2]);
        f.v = om::mk(r.fv[0], r.fv[1], r.fv[2]);
        f.identity = r.ident;
        // the mirror grown by one microsample pitch on each side holds every ray origin
        const float gx = r.sx2 + s->prm.helio_pixel_length, gz = r.sz2 + s->prm.helio_pixel_length;
        const float sx[4] = {-gx, gx, gx, -gx}, sz[4] = {-gz, -gz, gz, gz};
        for (int c = 0; c < 4; ++c) {
            const v3 w = om::to_world(om::mk(sx[c], r.sy2, sz[c]), f);
            mir[i].p[c][0] = (double)w.x + r.pos[0];
            mir[i].p[c][1] = (double)w.y + r.pos[1];
            mir[i].p[c][2] = (double)w.z + r.pos[2];
        }
        const v3 q[4] = {s->v0[i], s->v1[i], s->v1[i] + s->v3[i] - s->v0[i], s->v3[i]};
        for (int c = 0; c < 4; ++c) {
            blk[i].p[c][0] = q[c].x;
            blk[i].p[c][1] = q[c].y;
            blk[i].p[c][2] = q[c].z;
        }
        finish(mir[i]);
        finish(blk[i]);
        ytop = std::max(ytop, blk[i].yhi);
    }
    // horizontal buckets of the blockers
    double rmax = 0.0, xlo = 1e300, zlo = 1e300, xhi = -1e300, zhi = -1e300;
    for (const Quad &q : blk) {
        rmax = std::max(rmax, q.r);
        xlo = std::min(xlo, q.cx);
        xhi = std::max(xhi, q.cx);
        zlo = std::min(zlo, q.cz);
        zhi = std::max(zhi, q.cz);
    }
    const double bs = std::max(8.0, 4.0 * rmax);
    const int nbx = (int)std::min(4096.0, std::floor((xhi - xlo) / bs) + 1);
    const int nbz = (int)std::min(4096.0, std::floor((zhi - zlo) / bs) + 1);
    std::vector<std::vector<int>> bucket((size_t)nbx * nbz);
    auto bx = [&](double x) { return std::min(nbx - 1, std::max(0, (int)std::floor((x - xlo) / bs))); };
    auto bz = [&](double z) { return std::min(nbz - 1, std::max(0, (int)std::floor((z - zlo) / bs))); };
    for (int k = 0; k < H; ++k) bucket[(size_t)bx(blk[k].cx) * nbz + bz(blk[k].cz)].push_back(k);
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto mark = [&](int h) {
        const Quad &m = mir[h];
        const double reach = (ytop - m.ylo) * cot + m.r + rmax + 1.0;
        if (reach > 1e5) return;
        double P[4][2];
        for (int c = 0; c < 4; ++c) {
            P[c][0] = dot(m.p[c], e1);
            P[c][1] = dot(m.p[c], e2);
        }
        bool clear = true;
        for (int ix = bx(m.cx - reach); clear && ix <= bx(m.cx + reach); ++ix)
            for (int iz = bz(m.cz - reach); clear && iz <= bz(m.cz + reach); ++iz)
                for (int k : bucket[(size_t)ix * nbz + iz]) {
                    if (k == h) continue;
                    const Quad &b = blk[k];
                    if (std::hypot(b.cx - m.cx, b.cz - m.cz) > reach) continue;
                    // every point of k behind every mirror point along every ray direction
                    double T = 0.0, ahead = -1e300;
                    for (int c = 0; c < 4; ++c)
                        for (int e = 0; e < 4; ++e) {
                            const double w[3] = {b.p[c][0] - m.p[e][0], b.p[c][1] - m.p[e][1], b.p[c][2] - m.p[e][2]};
                            const double l = std::sqrt(dot(w, w));
                            T = std::max(T, l);
                            ahead = std::max(ahead, dot(w, d0) + l * alpha);
                        }
                    if (ahead < -0.1) continue;
                    // the rays cross k's plane at more than 0.05 rad
                    const double E1[3] = {b.p[1][0] - b.p[0][0], b.p[1][1] - b.p[0][1], b.p[1][2] - b.p[0][2]};
                    const double E2[3] = {b.p[3][0] - b.p[0][0], b.p[3][1] - b.p[0][1], b.p[3][2] - b.p[0][2]};
                    double nk[3] = {E1[1] * E2[2] - E1[2] * E2[1], E1[2] * E2[0] - E1[0] * E2[2],
                                    E1[0] * E2[1] - E1[1] * E2[0]};
                    const double nl = std::sqrt(dot(nk, nk));
                    if (!(nl > 0.0) || std::fabs(dot(nk, d0)) / nl < std::sin(alpha + 0.05)) {
                        clear = false;
                        break;
                    }
                    // separating edge normal of the projections, margin T alpha + 0.1
                    double Q[4][2];
                    for (int c = 0; c < 4; ++c) {
                        Q[c][0] = dot(b.p[c], e1);
                        Q[c][1] = dot(b.p[c], e2);
                    }
                    const double margin = T * alpha + 0.1;
                    bool sep = false;
                    for (int poly = 0; poly < 2 && !sep; ++poly) {
                        const double(*R)[2] = poly ? Q : P;
                        for (int e = 0; e < 4 && !sep; ++e) {
                            const double ex = R[(e + 1) & 3][0] - R[e][0], ey = R[(e + 1) & 3][1] - R[e][1];
                            const double l = std::hypot(ex, ey);
                            if (!(l > 1e-9)) continue;
                            const double nx = -ey / l, ny = ex / l;
                            double pa = 1e300, pb = -1e300, qa = 1e300, qb = -1e300;
                            for (int c = 0; c < 4; ++c) {
                                const double a1 = P[c][0] * nx + P[c][1] * ny, b1 = Q[c][0] * nx + Q[c][1] * ny;
                                pa = std::min(pa, a1);
                                pb = std::max(pb, a1);
                                qa = std::min(qa, b1);
                                qb = std::max(qb, b1);
                            }
                            sep = pb + margin < qa || qb + margin < pa;
                        }
                    }
                    if (!sep) {
                        clear = false;
                        break;
                    }
                }
        s->recs[h].clear = clear ? 1 : 0;
    };
    // heliostats are independent: a few host threads over interleaved heliostats
    const int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (nt <= 1 || H < 512) {
        for (int h = 0; h < H; ++h) mark(h);
        return;
    }
    std::vector<std::thread> pool;
    int started = 0;
    try {
        for (; started < nt; ++started)
            pool.emplace_back([&, started] {
                for (int h = started; h < H; h += nt) mark(h);
            });
    } catch (...) {
        // no more threads: the heliostats of the threads not started are marked here
        for (int t = started; t < nt; ++t)
            for (int h = t; h < H; h += nt) mark(h);
    }
    for (std::thread &th : pool) th.join();
}

void pack_heliostats(om_scene *s)
{
    int H = s->H;
    s->recs.assign(H, om::HelioRec{});
    for (int i = 0; i < H; ++i) {
        om::HelioRec &r = s->recs[i];
        std::memset(&r, 0, sizeof(r));
        om::frame f = om::make_frame(s->normal[i]);
        for (int a = 0; a < 3; ++a) r.pos[a] = s->pos[3 * i + a];
        r.n[0] = f.n.x; r.n[1] = f.n.y; r.n[2] = f.n.z;
        r.fu[0] = f.u.x; r.fu[1] = f.u.y; r.fu[2] = f.u.z;
        r.fv[0] = f.v.x; r.fv[1] = f.v.y; r.fv[2] = f.v.z;
        r.ident = f.identity;
        r.sx2 = s->size[3 * i] / 2;
        r.sy2 = s->size[3 * i + 1] / 2;
        r.sz2 = s->size[3 * i + 2] / 2;
        r.gx = s->gap[2 * i];
        r.gy = s->gap[2 * i + 1];
        r.subrow_len = s->subrow_len[i];
        r.subcol_len = s->subcol_len[i];
        r.sub_rows = s->sub_rows[i];
        r.sub_cols = s->sub_cols[i];
        r.cols_all = s->matrix[2 * i + 1] * s->sub_cols[i];
        r.rows = s->matrix[2 * i];
        r.cols = s->matrix[2 * i + 1];
        r.mcount = s->mcount[i];
        long long off = s->ray_offset.empty() ? 0 : s->ray_offset[i];
        r.ray_off_lo = (int)(off & 0xffffffffLL);
        r.ray_off_hi = (int)(off >> 32);
    }
    mark_clear(s);
}

std::vector<float4> pack_geo(const om_scene *s)
{
    std::vector<float4> g((size_t)s->H * 3 + 3);
    for (int i = 0; i < s->H; ++i) {
        v3 A = s->v0[i], E1 = s->v1[i] - s->v0[i], E2 = s->v3[i] - s->v0[i];
        g[3 * i + 0] = make_float4(A.x, A.y, A.z, 0.0f);
        g[3 * i + 1] = make_float4(E1.x, E1.y, E1.z, 0.0f);
        g[3 * i + 2] = make_float4(E2.x, E2.y, E2.z, 0.0f);
    }
    return g;
}

// Work tiles: (heliostat, first ray) covering rays_per_thread * 256 rays each.
void build_tiles(om_scene *s)
{
    long long per_tile = 256LL * s->rays_per_thread;
    s->tiles.clear();
    for (int i = 0; i < s->H; ++i) {
        long long nr = (long long)s->mcount[i] * s->prm.rays_per_sample;
        for (long long r = 0; r < nr; r += per_tile) s->tiles.push_back(make_int2(i, (int)r));
    }
    for (om::GridState &G : s->grids) om::tile_range(s, G.h0, G.h1, G.t0, G.t1);
}

// ---- FP64 geometry (om_params.precision = OM_PREC_FP64) ---------------------------------
// The scene in double as the FP64 model builds it (the same operations in the same order, so
// that the values are the same bits): sun, receiver corners by face and rotateY
// (receiver.cpp:24-117), the focus, each heliostat's aimed normal and local2world frame
// (heliostat.cpp:16-62, global_function.cuh:14-46), its vertices, and the Buie table
// (scene_instance_process.cu:116-226) with libm in double.
struct D3 {
    double x, y, z;
};
D3 d3(double x, double y, double z) { return D3{x, y, z}; }
D3 dadd(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
D3 dsub(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
D3 dscl(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dcross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
D3 dnorm(D3 v) { return dscl(v, 1.0 / std::sqrt(ddot(v, v))); }
D3 dload(const float *p) { return d3(p[0], p[1], p[2]); }
void dstore(double *o, D3 v)
{
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}
struct DFrame {
    D3 n, u, v;
    bool ident;
};
DFrame dframe(D3 n)
{
    DFrame f;
    f.n = n;
    f.ident = std::fabs(n.x) < 1e-6 && std::fabs(n.z) < 1e-6;
    f.u = f.ident ? d3(1, 0, 0) : dnorm(dcross(d3(0.0, 1.0, 0.0), n));
    f.v = f.ident ? d3(0, 0, 1) : dnorm(dcross(f.u, n));
    return f;
}
D3 dworld(D3 d, const DFrame &f)
{
    if (f.ident) return d;
    return d3(d.x * f.u.x + d.y * f.n.x + d.z * f.v.x, d.x * f.u.y + d.y * f.n.y + d.z * f.v.y,
              d.x * f.u.z + d.y * f.n.z + d.z * f.v.z);
}

// The float nearest x from above (a magnitude the error bounds may overstate, never understate)
float up_f(double x)
{
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

// OM_PREC_MIXED: float images of the FP64 scene for the FP32 walks and sampling, and the
// scene's error bounds (MixArgs).  geo: each heliostat's double vertices; cx, cy, sl: the
// double Buie table.
// OM_PREC_MIXED: per heliostat and for the receiver, the values the mixed kernels take from the
// FP64 model (om_internal.h, HelioMix and MixArgs): frames, the nominal reflected direction
// and the nominal ray's point nearest the receiver's reference point F (the rectangle's
// centre, the cylinder's axis point at mid-height), all formed in double and rounded once.
void mix_helio(om_scene *s, const std::vector<om::Helio64> &hel, const DFrame &sf, D3 F, std::vector<om::HelioMix> &out)
{
    const D3 i0 = sf.ident ? d3(0, 1, 0) : sf.n;
    out.assign((size_t)std::max(s->H, 1), om::HelioMix{});
    for (int i = 0; i < s->H; ++i) {
        const om::Helio64 &h = hel[(size_t)i];
        D3 u = d3(h.u[0], h.u[1], h.u[2]), n = d3(h.n[0], h.n[1], h.n[2]), v = d3(h.v[0], h.v[1], h.v[2]);
        if (h.ident) {
            u = d3(1, 0, 0);
            n = d3(0, 1, 0);
            v = d3(0, 0, 1);
        }
        const double c0 = ddot(n, i0);
        const D3 a = dsub(i0, dscl(n, 2.0 * c0));
        const D3 pos = d3(h.pos[0], h.pos[1], h.pos[2]);
        const double D = ddot(dsub(F, pos), a);
        const D3 qf = dsub(dadd(pos, dscl(a, D)), F);
        om::HelioMix &m = out[(size_t)i];
        const D3 vs[5] = {u, n, v, a, qf};
        float *dst[5] = {m.u, m.n, m.v, m.a, m.qf};
        for (int k = 0; k < 5; ++k) {
            dst[k][0] = (float)vs[k].x;
            dst[k][1] = (float)vs[k].y;
            dst[k][2] = (float)vs[k].z;
        }
        m.c0 = (float)c0;
        m.D = (float)D;
        m.e0 = (float)ddot(n, sf.n);
    }
}

om::v3 fv3(D3 a) { return om::mk((float)a.x, (float)a.y, (float)a.z); }
float norm1_up(D3 a) { return up_f((std::fabs(a.x) + std::fabs(a.y) + std::fabs(a.z)) * (1.0 + 1e-6)); }

om_status upload_mixed(om_scene *s, const std::vector<om::Helio64> &hel, const std::vector<om::Geo64> &geo,
                       const std::vector<double> &cx, const std::vector<double> &cy, const std::vector<double> &sl,
                       const DFrame &sf, D3 focus, const D3 *rv)
{
    om_status st;
    om::MixArgs &x = s->mix;
    x = om::MixArgs{};
    // sun frame (the identity axes for an identity frame: world() is then the identity)
    x.s_u = sf.ident ? om::mk(1, 0, 0) : fv3(sf.u);
    x.s_n = sf.ident ? om::mk(0, 1, 0) : fv3(sf.n);
    x.s_v = sf.ident ? om::mk(0, 0, 1) : fv3(sf.v);
    x.sun = fv3(sf.n);
    // receiver: the reference point F and the plane's constants (rectangle) or the axis
    // point at mid-height (cylinder)
    D3 F = focus;
    if (s->recv_type == 0) {
        const D3 A = rv[0], E1 = dsub(rv[1], rv[0]), E2 = dsub(rv[3], rv[0]);
        const D3 nr = dcross(E2, E1), nn = dnorm(nr);
        const D3 g1r = dcross(E2, nn), g2r = dcross(nn, E1);
        const D3 g1 = dscl(g1r, 1.0 / ddot(E1, g1r)), g2 = dscl(g2r, 1.0 / ddot(E2, g2r));
        x.nn = fv3(nn);
        x.nr = fv3(nr);
        x.g1 = fv3(g1);
        x.g2 = fv3(g2);
        x.uF = (float)ddot(dsub(F, A), g1);
        x.vF = (float)ddot(dsub(F, A), g2);
        x.g1n = norm1_up(g1);
        x.g2n = norm1_up(g2);
        x.nrn = norm1_up(nr);
    } else if (s->recv_type == 1) {
        F = dload(s->recv_pos);
        x.cyl_r = (float)s->recv_size[0];
        x.cyl_h = (float)s->recv_size[1];
    }
    std::vector<om::HelioMix> hm;
    mix_helio(s, hel, sf, F, hm);
    if ((st = upload(s->d_hmix, hm.data(), hm.size())) != OM_OK) return st;
    x.hmix = s->d_hmix;
    // Origins lie on the mirrors, so no coordinate of an origin or a vertex exceeds the
    // vertices' largest; the float of each is within 2^-24 of it per component, and the two
    // together within 2^-22 of it (2-norm, with a margin).
    double rmax = 1.0;
    for (int i = 0; i < s->H; ++i) {
        const om::Geo64 &G = geo[(size_t)i];
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 3; ++k)
                rmax = std::max(rmax, std::fabs(G.A[k] + ((c & 1) ? G.E1[k] : 0.0) + ((c & 2) ? G.E2[k] : 0.0)));
    }
    s->mix_e_pos = up_f(std::ldexp(rmax, -22));
    // Directions from the FP32 draws: the sunshape polar angle theta (at most 9.3 mrad, the
    // Buie tail's end, or the pillbox half angle) and the slope-error angle theta_s (at most
    // 5.3 sigma, the deviate of the extreme draw 1 - 2^-24) enter as sines and cosines of
    // relative error a few 2^-24 (the erfinv's about 2^-21), the deviations from the nominal
    // directions scale with them and the reflection doubles the normal's, so a direction's
    // deviation is within about 2^-21 (theta + 2 theta_s) of the model's, plus at most about
    // 3e-9 rad of the float Buie lookup; the bound takes 2^-19 (theta + 4 theta_s) + 2^-26 at
    // the largest angles (the kernels take it per ray at the ray's angles).
    const double th_max = s->prm.sunshape == OM_SUN_BUIE ? 9.3e-3 : (double)s->prm.pillbox_half_angle;
    const double ths_max = 5.5 * (double)s->prm.slope_sigma;
    s->mix_e_dir = up_f(std::ldexp(th_max + 4.0 * ths_max, -19) + std::ldexp(1.0, -26));
    // the deviations dr carry that error; a direction a + dr, rounded to float in the walks,
    // adds three roundings of its components (within 2^-21 in 2-norm)
    x.e_dr = s->mix_e_dir;
    x.e_dir = s->mix_e_dir + std::ldexp(1.0f, -21);
    x.e_pos = s->mix_e_pos;
    // the double Buie table rounded to float, with the split guide of the FP32 lookup
    dfree(s->d_buie_bins_m);
    dfree(s->d_buie_split_m);
    const int ng = s->nbins;
    if (s->prm.sunshape == OM_SUN_BUIE && ng > 0) {
        std::vector<float> fy((size_t)ng + 1);
        for (int i = 0; i <= ng; ++i) fy[i] = (float)cy[i];
        std::vector<float4> bins((size_t)ng);
        for (int i = 0; i < ng; ++i) bins[i] = make_float4(fy[i], fy[i + 1], (float)sl[i], (float)cx[i]);
        const float top = fy[ng];
        s->buie_gscale_m = (float)ng / top;
        std::vector<int2> split((size_t)ng);
        for (int j = 0, i = 0; j < ng; ++j) {
            const float vj = (float)j / s->buie_gscale_m;
            while (i < ng - 1 && fy[i + 1] <= vj) ++i;
            const float edge = fy[i + 1];
            int bits;
            std::memcpy(&bits, &edge, 4);
            split[j] = make_int2(i, bits);
        }
        s->buie_A_m = (float)s->a64.A;
        // the draws below the tail (k < tail_k0) have u A < int465 in double; their float
        // value is held below the table's top so that the lookup stays in the table
        s->buie_vmax_m = std::nextafter(top, 0.0f);
        if ((st = upload(s->d_buie_bins_m, bins.data(), bins.size())) != OM_OK) return st;
        if ((st = upload(s->d_buie_split_m, split.data(), split.size())) != OM_OK) return st;
    }
    // the candidates of every grid cell in CSR order: (A, id bits), (E1, |E1|), (E2, |E2|)
    const om_scene *root = s->grid_src ? s->grid_src : s;
    for (float4 *&p : s->d_cell_geo_m) dfree(p);
    s->d_cell_geo_m.assign(s->grids.size(), nullptr);
    for (size_t k = 0; k < s->grids.size(); ++k) {
        const std::vector<int> &items = root->grids[k].cell_items;
        std::vector<float4> cg(3 * items.size());
        for (size_t i = 0; i < items.size(); ++i) {
            const int h = items[i];
            const om::Geo64 &G = geo[(size_t)h];
            const double l1 = std::sqrt(G.E1[0] * G.E1[0] + G.E1[1] * G.E1[1] + G.E1[2] * G.E1[2]);
            const double l2 = std::sqrt(G.E2[0] * G.E2[0] + G.E2[1] * G.E2[1] + G.E2[2] * G.E2[2]);
            cg[3 * i] = make_float4((float)G.A[0], (float)G.A[1], (float)G.A[2], 0.0f);
            std::memcpy(&cg[3 * i].w, &h, 4);
            cg[3 * i + 1] = make_float4((float)G.E1[0], (float)G.E1[1], (float)G.E1[2], up_f(l1 * (1.0 + 1e-6)));
            cg[3 * i + 2] = make_float4((float)G.E2[0], (float)G.E2[1], (float)G.E2[2], up_f(l2 * (1.0 + 1e-6)));
        }
        if ((st = upload(s->d_cell_geo_m[k], cg.data(), cg.size())) != OM_OK) return st;
    }
    if (!s->d_mix_n) {
        const uint32_t zero[3] = {0u, 0u, 0u};
        if ((st = upload(s->d_mix_n, zero, 3)) != OM_OK) return st;
    }
    return OM_OK;
}

om_status upload_fp64(om_scene *s)
{
    om::Trace64Args &q = s->a64;
    q = om::Trace64Args{};
    const D3 sun = dnorm(dload(s->sun_in));
    const DFrame sf = dframe(sun);
    dstore(q.sun_n, sf.n);
    dstore(q.sun_u, sf.u);
    dstore(q.sun_v, sf.v);
    q.sun_ident = sf.ident ? 1 : 0;
    // receiver corners (the construction of receiver.cpp:24-117 in double)
    const D3 sz = dload(s->recv_size), pos = dload(s->recv_pos);
    D3 ln, v[4];
    switch (s->recv_face) {
    case 0:
        ln = d3(0, 0, 1);
        v[0] = d3(-sz.x / 2, -sz.y / 2, sz.z / 2); v[1] = d3(-sz.x / 2, sz.y / 2, sz.z / 2);
        v[2] = d3(sz.x / 2, sz.y / 2, sz.z / 2); v[3] = d3(sz.x / 2, -sz.y / 2, sz.z / 2);
        break;
    case 1:
        ln = d3(1, 0, 0);
        v[0] = d3(sz.x / 2, -sz.y / 2, sz.z / 2); v[1] = d3(sz.x / 2, sz.y / 2, sz.z / 2);
        v[2] = d3(sz.x / 2, sz.y / 2, -sz.z / 2); v[3] = d3(sz.x / 2, -sz.y / 2, -sz.z / 2);
        break;
    case 2:
        ln = d3(0, 0, -1);
        v[0] = d3(sz.x / 2, -sz.y / 2, -sz.z / 2); v[1] = d3(sz.x / 2, sz.y / 2, -sz.z / 2);
        v[2] = d3(-sz.x / 2, sz.y / 2, -sz.z / 2); v[3] = d3(-sz.x / 2, -sz.y / 2, -sz.z / 2);
        break;
    default:
        ln = d3(-1, 0, 0);
        v[0] = d3(-sz.x / 2, -sz.y / 2, -sz.z / 2); v[1] = d3(-sz.x / 2, sz.y / 2, -sz.z / 2);
        v[2] = d3(-sz.x / 2, sz.y / 2, sz.z / 2); v[3] = d3(-sz.x / 2, -sz.y / 2, sz.z / 2);
        break;
    }
    const D3 rn = dnorm(dload(s->recv_normal_in));
    D3 rv[4];
    for (int i = 0; i < 4; ++i) {
        const int dir = (dcross(ln, rn).y > 0) ? 1 : -1;
        const double c = ddot(ln, rn);
        const double sn = dir * std::sqrt(std::fmax(0.0, 1 - c * c));
        const D3 o = v[i];
        rv[i] = dadd(d3(c * o.x + sn * o.z, o.y, -sn * o.x + c * o.z), pos);
    }
    D3 focus = dscl(dadd(rv[0], rv[2]), 0.5);
    dstore(q.rA, rv[0]);
    dstore(q.rE1, dsub(rv[1], rv[0]));
    dstore(q.rE2, dsub(rv[3], rv[0]));
    if (s->recv_type == 1) {
        const double r = s->recv_size[0], h = s->recv_size[1];
        focus = pos;
        dstore(q.cyl_base, d3(pos.x, pos.y - h / 2, pos.z));
        q.cyl_r = r;
        q.cyl_h = h;
    } else if (s->recv_type == 2) {
        const double h = s->recv_size[2];
        focus = pos;
        dstore(q.cyl_base, d3(pos.x, pos.y - h / 2, pos.z));
        q.cone_r0 = s->recv_size[1];
        q.cone_r1 = s->recv_size[0];
        q.cyl_h = h;
    }
    // heliostats
    const int H = s->H;
    std::vector<om::Helio64> hel((size_t)std::max(H, 1));
    std::vector<om::Geo64> geo((size_t)std::max(H, 1), om::Geo64{});
    for (int i = 0; i < H; ++i) {
        const D3 p = dload(&s->pos[3 * (size_t)i]), z = dload(&s->size[3 * (size_t)i]);
        const D3 nn = !s->normal_in.empty() ? dnorm(dload(&s->normal_in[3 * (size_t)i]))
                                            : dnorm(dsub(dnorm(dsub(focus, p)), sun));
        const DFrame f = dframe(nn);
        om::Helio64 &hr = hel[i];
        dstore(hr.pos, p);
        dstore(hr.n, f.n);
        dstore(hr.u, f.u);
        dstore(hr.v, f.v);
        dstore(hr.sz, z);
        hr.ident = f.ident ? 1 : 0;
        const double gx = s->gap[2 * (size_t)i], gy = s->gap[2 * (size_t)i + 1];
        const int rows = s->matrix[2 * (size_t)i], cols = s->matrix[2 * (size_t)i + 1];
        hr.gx = gx;
        hr.gy = gy;
        hr.subrow = (z.z - gy * (rows - 1)) / rows;
        hr.subcol = (z.x - gx * (cols - 1)) / cols;
        const D3 l0 = d3(-z.x / 2, z.y / 2, -z.z / 2);
        const D3 l1 = d3(l0.x, l0.y, l0.z + z.z);
        const D3 l3 = d3(l0.x + z.x, l0.y, l0.z);
        const D3 v0 = dadd(dworld(l0, f), p), v1 = dadd(dworld(l1, f), p), v3 = dadd(dworld(l3, f), p);
        dstore(geo[i].A, v0);
        dstore(geo[i].E1, dsub(v1, v0));
        dstore(geo[i].E2, dsub(v3, v0));
        geo[i].id = i;
    }
    // Buie table in double (inter_larger_465_intensity, parameters_generate)
    const int ng = s->nbins;
    const double csr = s->prm.csr, upper = 4.65;
    const double k = 0.9 * std::log(13.5 * csr) * std::pow(csr, -0.3);
    const double g = 2.2 * std::log(0.52 * csr) * std::pow(csr, 0.43) - 0.1;
    const double i93 = std::exp(k) / (g + 1) * std::pow(9.3, g + 1), i465 = std::exp(k) / (g + 1) * std::pow(upper, g + 1);
    std::vector<double> cx((size_t)ng + 1), cy((size_t)ng + 1), sl((size_t)std::max(ng, 1));
    const double li = upper / ng;
    cx[0] = 0;
    cy[0] = 0;
    double hp = 1.0;
    for (int i = 1; i <= ng; ++i) {
        const double x = li * i;
        const double hc = std::cos(0.326 * x) / std::cos(0.308 * x);
        const double sa = (hp + hc) * li / 2;
        cx[i] = x;
        cy[i] = sa + cy[i - 1];
        sl[i - 1] = li / sa;
        hp = hc;
    }
    q.nbins = ng;
    q.int465 = cy[ng];
    q.A = q.int465 + (i93 - i465);
    q.B = (g + 1) / std::exp(k);
    q.C = std::pow(upper, g + 1);
    q.inv_gp1 = 1 / (g + 1);
    q.pillbox = s->prm.pillbox_half_angle;
    q.sigma = s->prm.slope_sigma;
    q.pitch = s->prm.helio_pixel_length;
    // guide of the Buie lookup: the bin of value j / gscale (the largest i with
    // cdf_y[i] <= j / gscale); the device walk from it ends on the bin a binary search finds
    q.gscale = ng / q.int465;
    std::vector<int> guide((size_t)std::max(ng, 1), 0);
    for (int j = 0, i = 0; j < ng; ++j) {
        const double vj = j / q.gscale;
        while (i < ng - 1 && cy[i + 1] <= vj) ++i;
        guide[j] = i;
    }
    // one device block: Helio64[H] | cdf_x | cdf_y | slope | guide
    const size_t nh = hel.size();
    const size_t b_hel = sizeof(om::Helio64) * nh;
    const size_t b_tab = sizeof(double) * ((size_t)ng + 1), b_sl = sizeof(double) * sl.size();
    std::vector<char> blk(b_hel + 2 * b_tab + b_sl + sizeof(int) * guide.size());
    char *o = blk.data();
    std::memcpy(o, hel.data(), b_hel);
    std::memcpy(o + b_hel, cx.data(), b_tab);
    std::memcpy(o + b_hel + b_tab, cy.data(), b_tab);
    std::memcpy(o + b_hel + 2 * b_tab, sl.data(), b_sl);
    std::memcpy(o + b_hel + 2 * b_tab + b_sl, guide.data(), sizeof(int) * guide.size());
    dfree(s->d_f64);
    om_status st = upload(s->d_f64, blk.data(), blk.size());
    if (st != OM_OK) return st;
    q.helio = (const om::Helio64 *)s->d_f64;
    q.cdf_x = (const double *)(s->d_f64 + b_hel);
    q.cdf_y = (const double *)(s->d_f64 + b_hel + b_tab);
    q.slope = (const double *)(s->d_f64 + b_hel + 2 * b_tab);
    q.guide = (const int *)(s->d_f64 + b_hel + 2 * b_tab + b_sl);
    // the Buie tail: the draws k whose value u01(k << 9) A reaches int465 form the range
    // [k0, 2^23) (u01 grows with k); their angles are evaluated once, on the device
    dfree(s->d_tail64);
    q.tail = nullptr;
    q.tail_k0 = 1u << 23;
    if (s->prm.sunshape == OM_SUN_BUIE) {
        uint32_t lo = 0, hi = 1u << 23; // the first k in [lo, hi] with u A >= int465 (hi: none)
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((double)om::u01(mid << 9) * q.A >= q.int465) hi = mid;
            else lo = mid + 1;
        }
        q.tail_k0 = lo;
        const uint32_t n = (1u << 23) - lo;
        if (n > 0) {
            if (hipMalloc((void **)&s->d_tail64, sizeof(double) * n) != hipSuccess)
                return om::hip_fail(hipGetLastError(), "hipMalloc (Buie tail table)");
            if ((st = om::build_buie_tail64(s->d_tail64, lo, n, q.A, q.int465, q.B, q.C, q.inv_gp1)) != OM_OK)
                return st;
            q.tail = s->d_tail64;
        }
    }
    // the candidates of every grid cell in CSR order with their ids (a walk reads a
    // candidate's vertices without first reading its id)
    const om_scene *root = s->grid_src ? s->grid_src : s;
    for (om::Geo64 *&p : s->d_cell_geo64) dfree(p);
    s->d_cell_geo64.assign(s->grids.size(), nullptr);
    for (size_t k = 0; k < s->grids.size(); ++k) {
        const std::vector<int> &items = root->grids[k].cell_items;
        std::vector<om::Geo64> cg(items.size());
        for (size_t i = 0; i < items.size(); ++i) cg[i] = geo[(size_t)items[i]];
        if ((st = upload(s->d_cell_geo64[k], cg.data(), cg.size())) != OM_OK) return st;
    }
    if (s->prm.precision == OM_PREC_MIXED) return upload_mixed(s, hel, geo, cx, cy, sl, sf, focus, rv);
    return OM_OK;
}

om_status upload_geometry(om_scene *s)
{
    if (s->prm.precision == OM_PREC_FP64 || s->prm.precision == OM_PREC_MIXED) {
        om_status st = upload_fp64(s);
        if (st != OM_OK) return st;
    }
    pack_heliostats(s);
    om_status st = upload(s->d_helio, s->recs.data(), s->recs.size());
    if (st != OM_OK) return st;
    std::vector<float4> g = pack_geo(s);
    if ((st = upload(s->d_geo, g.data(), g.size())) != OM_OK) return st;
    // the blockers of every grid cell in CSR order, each with its id, so that a walk reads
    // a candidate's vertices without first reading its id
    const om_scene *root = s->grid_src ? s->grid_src : s;
    s->d_cell_geo.resize(s->grids.size(), nullptr);
    for (size_t k = 0; k < s->grids.size(); ++k) {
        const std::vector<int> &items = root->grids[k].cell_items;
        std::vector<float4> cg(3 * items.size());
        for (size_t i = 0; i < items.size(); ++i) {
            const int h = items[i];
            cg[3 * i] = g[3 * (size_t)h];
            std::memcpy(&cg[3 * i].w, &h, 4);
            cg[3 * i + 1] = g[3 * (size_t)h + 1];
            cg[3 * i + 2] = g[3 * (size_t)h + 2];
        }
        if ((st = upload(s->d_cell_geo[k], cg.data(), cg.size())) != OM_OK) return st;
    }
    return OM_OK;
}

// Azimuthal maps (cylinder, cone): a heliostat's rays land on the side of the receiver that
// faces it, so the LDS-window traces give each of their 8 queues one azimuth sector of the
// grid's heliostats about the receiver axis (sectors of equal tile counts; heliostat order
// within a sector keeps neighbouring heliostats together), and the workgroups of a queue
// hold the window centred on its sector's mean azimuth.  On the 25,000-heliostat cylinder
// field the windows then take 68 % of the hits instead of the centred window's 22 %.  Maps
// that fit one window whole keep it.  The window kernels split a grid's n = tiles x parts
// chunks into queues x of q = n / 8 (+1 for x < n % 8), as here.
om_status build_win_sectors(om_scene *s)
{
    for (om::GridState &G : s->grids) {
        dfree(G.d_win_order);
        G.win_sectors = false;
    }
    if (s->recv_type != 1 && s->recv_type != 2) return OM_OK;
    int x0, y0, w, h;
    om::window_rect(s->res_x, s->res_y, x0, y0, w, h);
    if (w >= s->res_x && h >= s->res_y) return OM_OK;
    const int tile_rays = 256 * s->rays_per_thread;
    const int parts = tile_rays / std::min(tile_rays, om::window_chunk_rays());
    const double two_pi = 6.283185307179586;
    for (om::GridState &G : s->grids) {
        const int nt = G.t1 - G.t0;
        if (nt < 8 || G.h1 <= G.h0) continue;
        // each heliostat's first tile and tile count, and its azimuth about the axis
        const int nh = G.h1 - G.h0;
        std::vector<int> first(nh, 0), count(nh, 0);
        for (int t = G.t0; t < G.t1; ++t) {
            const int k = s->tiles[(size_t)t].x - G.h0;
            if (count[k]++ == 0) first[k] = t - G.t0;
        }
        std::vector<double> th(nh);
        for (int k = 0; k < nh; ++k) {
            const int hh = G.h0 + k;
            double a = std::atan2((double)s->pos[3 * (size_t)hh + 2] - s->cyl_base.z,
                                  (double)s->pos[3 * (size_t)hh] - s->cyl_base.x);
            th[k] = a < 0.0 ? a + two_pi : a;
        }
        std::vector<int> by_az(nh);
        for (int k = 0; k < nh; ++k) by_az[k] = k;
        std::stable_sort(by_az.begin(), by_az.end(), [&](int u, int v) { return th[u] < th[v]; });
        std::vector<int> sector(nh, 0);
        long long cum = 0;
        for (int k : by_az) {
            sector[k] = (int)std::min<long long>(7, (8 * (2 * cum + count[k])) / (2LL * nt));
            cum += count[k];
        }
        std::vector<int> order_h(nh);
        for (int k = 0; k < nh; ++k) order_h[k] = k;
        std::stable_sort(order_h.begin(), order_h.end(), [&](int u, int v) { return sector[u] < sector[v]; });
        std::vector<int> order;
        order.reserve((size_t)nt);
        for (int k : order_h)
            for (int j = 0; j < count[k]; ++j) order.push_back(first[k] + j);
        // each queue's window: centred on the circular mean azimuth of its tiles
        const long long n = (long long)nt * parts, q = n / 8, rmd = n % 8;
        for (int x = 0; x < 8; ++x) {
            const long long c0 = x * q + std::min<long long>(x, rmd), cnt = q + (x < rmd ? 1 : 0);
            double cs = 0.0, sn = 0.0;
            for (long long c = c0; c < c0 + cnt; c += parts) {
                const int k = s->tiles[(size_t)(G.t0 + order[(size_t)(c / parts)])].x - G.h0;
                cs += std::cos(th[k]);
                sn += std::sin(th[k]);
            }
            double a = std::atan2(sn, cs);
            if (a < 0.0) a += two_pi;
            const int col = (int)std::floor(a / two_pi * s->res_x - 0.5 * w);
            G.win_qx0[x] = ((col % s->res_x) + s->res_x) % s->res_x;
            G.win_qy0[x] = y0;
        }
        om_status st = upload(G.d_win_order, order.data(), order.size());
        if (st != OM_OK) return st;
        G.win_sectors = true;
    }
    return OM_OK;
}

om_status choose_tiles(om_scene *s)
{
    // enough workgroups to fill 256 CUs several times over, tiles of at least 256 rays
    long long total = s->total_rays;
    int rpt = 8; // at most the hit slots of a trace workgroup
    while (rpt > 1 && total / (256LL * rpt) < 8192) rpt /= 2;
    s->rays_per_thread = rpt;
    build_tiles(s);
    om_status st = upload(s->d_tiles, s->tiles.data(), s->tiles.size());
    if (st != OM_OK) return st;
    return build_win_sectors(s);
}

void free_scene_dev(om_scene *s)
{
    dfree(s->d_f64);
    dfree(s->d_tail64);
    for (om::Geo64 *&p : s->d_cell_geo64) dfree(p);
    for (float4 *&p : s->d_cell_geo_m) dfree(p);
    dfree(s->d_hmix);
    dfree(s->d_buie_bins_m);
    dfree(s->d_buie_split_m);
    dfree(s->d_mix_n);
    dfree(s->d_helio);
    dfree(s->d_geo);
    for (float4 *&p : s->d_cell_geo) dfree(p);
    if (s->diff_buf.p) (void)hipFree(s->diff_buf.p);
    s->diff_buf.p = nullptr;
    s->diff_buf.bytes = 0;
    if (s->grid_src) return; // the rest belongs to the root scene
    for (om::GridState &G : s->grids) {
        dfree(G.d_cell_start);
        dfree(G.d_cell_items);
        dfree(G.d_win_order);
    }
    dfree(s->d_buie_bins);
    dfree(s->d_buie_guide);
    dfree(s->d_buie_split);
    dfree(s->d_poisson);
    dfree(s->d_mcount);
    dfree(s->d_tiles);
}

// Coordinates and sizes the tracer accepts: finite and within 2^20 m; cell edges within
// [2^-20, 2^20] m.  Inside these bounds every division and square root of the trace kernel
// has operands in the range of its short exact sequences.
constexpr float COORD_MAX = 1048576.0f;
constexpr float EDGE_MIN = 1.0f / 1048576.0f;

bool bounded(const float *x, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!(std::fabs(x[i]) <= COORD_MAX)) return false;
    return true;
}

// Every ray origin must lie strictly inside the grid: a walk indexes the cell arrays from
// the origin's cell and checks only the axes it steps (rectgrid_raytracing.cu:43-48; an
// origin outside the grid reads outside the arrays there).  Origins lie in the mirror
// rectangle v0 v1 v2 v3; the slack covers the rounding of an origin and of its cell index.
om_status check_in_grid(const om_scene *s)
{
    for (size_t g = 0; g < s->grids.size(); ++g) {
        const om::GridState &G = s->grids[g];
        for (int i = G.h0; i < G.h1; ++i) {
            const v3 c[4] = {s->v0[i], s->v1[i], s->v2[i], s->v3[i]};
            const double sx = s->size[3 * i], sy = s->size[3 * i + 1], sz = s->size[3 * i + 2];
            const double R = 0.5 * std::sqrt(sx * sx + sy * sy + sz * sz);
            for (int a = 0; a < 3; ++a) {
                double lo = HUGE_VAL, hi = -HUGE_VAL;
                for (int k = 0; k < 4; ++k) {
                    const double x = a == 0 ? c[k].x : (a == 1 ? c[k].y : c[k].z);
                    lo = std::min(lo, x);
                    hi = std::max(hi, x);
                }
                const double gp = G.pos[a], ext = (double)G.gn[a] * G.interval[a];
                const double slack = 1e-5 * (std::fabs((double)s->pos[3 * i + a]) + R + std::fabs(gp) + ext);
                if (!(lo - slack > gp && hi + slack < gp + ext)) {
                    char msg[160];
                    std::snprintf(msg, sizeof msg, "heliostat %d reaches outside the grid on axis %c "
                                  "(mirror %.6g..%.6g, grid %d %.6g..%.6g)", i, "xyz"[a], lo, hi, (int)g, gp, gp + ext);
                    om::set_error(msg);
                    return OM_ERR_UNSUPPORTED;
                }
            }
        }
    }
    return OM_OK;
}

// the grids of a description: the grids array, or the one grid of the grid_* fields
std::vector<om_grid_desc> desc_grids(const om_scene_desc *d)
{
    if (d->n_grids > 0) return std::vector<om_grid_desc>(d->grids, d->grids + d->n_grids);
    om_grid_desc g;
    put_grid(g, d->grid_pos, d->grid_size, d->grid_interval, 0, d->n_helio);
    return {g};
}

om_status validate(const om_scene_desc *d, const om_params *p)
{
    if (!d || !p) return OM_ERR_INVALID;
    if (d->n_helio < 0 || (d->n_helio > 0 && (!d->helio_pos || !d->helio_size || !d->helio_gap ||
                                              !d->helio_matrix))) {
        om::set_error("heliostat arrays missing");
        return OM_ERR_INVALID;
    }
    if (p->rays_per_sample <= 0 || !(p->helio_pixel_length > 0.0f)) {
        om::set_error("rays_per_sample and helio_pixel_length must be positive");
        return OM_ERR_INVALID;
    }
    if (p->sunshape != OM_SUN_PILLBOX && p->sunshape != OM_SUN_BUIE) {
        om::set_error("unknown sunshape");
        return OM_ERR_INVALID;
    }
    if (p->micro_mode != OM_MICRO_LATTICE && p->micro_mode != OM_MICRO_POISSON) {
        om::set_error("unknown microsample mode");
        return OM_ERR_INVALID;
    }
    if (p->precision != OM_PREC_FP32 && p->precision != OM_PREC_FP64 && p->precision != OM_PREC_MIXED) {
        om::set_error("precision must be OM_PREC_FP32, OM_PREC_FP64 or OM_PREC_MIXED");
        return OM_ERR_INVALID;
    }
    if (p->redo_cap < 0) {
        om::set_error("redo_cap must be 0 or positive");
        return OM_ERR_INVALID;
    }
    if (p->ray_compact != 0 && p->ray_compact != 1) {
        om::set_error("ray_compact must be 0 or 1");
        return OM_ERR_INVALID;
    }
    if (p->sunshape == OM_SUN_BUIE && (p->buie_bins <= 0 || p->buie_bins > 8192 || !(p->csr > 0.0f))) {
        om::set_error("buie_bins must be in 1..8192 and csr positive");
        return OM_ERR_INVALID;
    }
    if (p->accumulate != OM_ACC_PACKED && p->accumulate != OM_ACC_DEVICE && p->accumulate != OM_ACC_WINDOW) {
        om::set_error("unknown accumulation mode");
        return OM_ERR_INVALID;
    }
    if (p->pack_shift != 0 && (p->pack_shift < 32 || p->pack_shift > 62)) {
        om::set_error("pack_shift must be 0 or in 32..62");
        return OM_ERR_INVALID;
    }
    if (p->micro_mode == OM_MICRO_POISSON && p->poisson_k <= 0) {
        om::set_error("poisson_k must be positive");
        return OM_ERR_INVALID;
    }
    for (int i = 0; i < d->n_helio; ++i)
        if (d->helio_matrix[2 * i] <= 0 || d->helio_matrix[2 * i + 1] <= 0) {
            om::set_error("heliostat facet matrix must be positive");
            return OM_ERR_INVALID;
        }
    if (d->n_grids < 0 || (d->n_grids > 0 && !d->grids)) {
        om::set_error("grid array missing");
        return OM_ERR_INVALID;
    }
    const std::vector<om_grid_desc> grids = desc_grids(d);
    if (p->grid_select != OM_GRIDS_ALL && (p->grid_select < 0 || p->grid_select >= (int)grids.size())) {
        om::set_error("grid_select must be OM_GRIDS_ALL or the index of a grid");
        return OM_ERR_INVALID;
    }
    int next = 0;
    for (const om_grid_desc &g : grids) {
        if (g.first_helio != next || g.n_helio < 0) {
            om::set_error("grids must hold consecutive heliostat ranges from 0");
            return OM_ERR_INVALID;
        }
        next += g.n_helio;
        for (int a = 0; a < 3; ++a)
            if (!(g.interval[a] > 0.0f)) {
                om::set_error("grid interval must be positive");
                return OM_ERR_INVALID;
            }
    }
    if (next != d->n_helio) {
        om::set_error("the grids' heliostats must add up to n_helio");
        return OM_ERR_INVALID;
    }
    const size_t H = (size_t)d->n_helio;
    if (d->helio_normal) {
        if (!bounded(d->helio_normal, 3 * H)) {
            om::set_error("coordinates and sizes must be finite and within 2^20 m");
            return OM_ERR_UNSUPPORTED;
        }
        for (size_t i = 0; i < H; ++i) {
            const float *n = d->helio_normal + 3 * i;
            if (!(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] > 1e-12f)) {
                om::set_error("given heliostat normals must be nonzero");
                return OM_ERR_INVALID;
            }
        }
    }
    if (!bounded(d->helio_pos, 3 * H) || !bounded(d->helio_size, 3 * H) || !bounded(d->helio_gap, 2 * H) ||
        !bounded(d->recv_pos, 3) || !bounded(d->recv_size, 3)) {
        om::set_error("coordinates and sizes must be finite and within 2^20 m");
        return OM_ERR_UNSUPPORTED;
    }
    for (const om_grid_desc &g : grids) {
        if (!bounded(g.pos, 3) || !bounded(g.size, 3)) {
            om::set_error("coordinates and sizes must be finite and within 2^20 m");
            return OM_ERR_UNSUPPORTED;
        }
        for (int a = 0; a < 3; ++a)
            if (!(g.interval[a] >= EDGE_MIN && g.interval[a] <= COORD_MAX)) {
                om::set_error("grid interval must be within [2^-20, 2^20] m");
                return OM_ERR_UNSUPPORTED;
            }
    }
    if (!bounded(d->sun_dir, 3) || (d->sun_dir[0] == 0.0f && d->sun_dir[1] == 0.0f && d->sun_dir[2] == 0.0f)) {
        om::set_error("sun direction must be finite and nonzero");
        return OM_ERR_INVALID;
    }
    return OM_OK;
}

om_status build_all(om_scene *s)
{
    om_status st;
    // InitContent order (solar_scene.cpp:83-99): sun, grid, receiver, heliostats
    s->sun = om::normalize(ld3(s->sun_in));
    build_buie(s);
    for (om::GridState &G : s->grids)
        if ((st = build_grid(s, G)) != OM_OK) return st;
    if ((st = build_receiver(s)) != OM_OK) return st;
    aim_heliostats(s);
    if ((st = check_in_grid(s)) != OM_OK) return st;
    micro_layout(s);
    if (s->H > 0 && s->H >= (1 << 28)) {
        om::set_error("too many heliostats");
        return OM_ERR_UNSUPPORTED;
    }
    if (!s->ctx) {
        // host-only scene: geometry, grid, receiver and sun, no device state; Poisson
        // microsamples are drawn on the GPU, so their counts stay 0 here
        finish_micro(s);
        flux_scale(s);
        pack_heliostats(s);
        return OM_OK;
    }
    // device copies
    OM_HIP(hipSetDevice(s->ctx->device));
    for (om::GridState &G : s->grids) {
        if ((st = upload(G.d_cell_start, G.cell_start.data(), G.cell_start.size())) != OM_OK) return st;
        if ((st = upload(G.d_cell_items, G.cell_items.data(), G.cell_items.size())) != OM_OK) return st;
    }
    {
        // Buie lookup tables: packed bins and a guide from the value scale to a bin
        int n = s->nbins;
        std::vector<float4> bins(n);
        for (int i = 0; i < n; ++i)
            bins[i] = make_float4(s->cdf_y[i], s->cdf_y[i + 1], s->slope[i], s->cdf_x[i]);
        std::vector<int> guide(n);
        float gscale = (float)n / s->buie_int465;
        int i = 0;
        for (int j = 0; j < n; ++j) {
            float vj = (float)j / gscale;
            while (i < n - 1 && s->cdf_y[i + 1] <= vj) ++i;
            guide[j] = i;
        }
        // the guide with the upper edge of each entry's bin (the last bin's edge is int465,
        // above every value looked up)
        std::vector<int2> split(n);
        for (int j = 0; j < n; ++j) {
            float edge = s->cdf_y[guide[j] + 1];
            int bits;
            std::memcpy(&bits, &edge, 4);
            split[j] = make_int2(guide[j], bits);
        }
        if ((st = upload(s->d_buie_bins, bins.data(), bins.size())) != OM_OK) return st;
        if ((st = upload(s->d_buie_guide, guide.data(), guide.size())) != OM_OK) return st;
        if ((st = upload(s->d_buie_split, split.data(), split.size())) != OM_OK) return st;
    }
    if (s->prm.micro_mode == OM_MICRO_POISSON) {
        // Poisson points are drawn on the GPU (one workgroup per heliostat); their
        // counts come back to size the ray index space.
        s->ray_offset.assign(s->H + 1, 0);
        if ((st = upload_geometry(s)) != OM_OK) return st;
        dfree(s->d_poisson);
        dfree(s->d_mcount);
        if (s->H > 0) {
            OM_HIP(hipMalloc((void **)&s->d_poisson, sizeof(float2) * (size_t)s->poisson_cap * s->H));
            OM_HIP(hipMalloc((void **)&s->d_mcount, sizeof(int) * s->H));
            if ((st = om::launch_poisson(s, nullptr)) != OM_OK) return st;
            OM_HIP(hipMemcpy(s->mcount.data(), s->d_mcount, sizeof(int) * s->H, hipMemcpyDeviceToHost));
        }
    }
    finish_micro(s);
    if (s->total_rays >= (1LL << 40)) {
        om::set_error("more than 2^40 rays");
        return OM_ERR_UNSUPPORTED;
    }
    flux_scale(s);
    if ((st = upload_geometry(s)) != OM_OK) return st;
    return choose_tiles(s);
}

} // namespace

namespace om {
void set_trace_grid(TraceArgs &a, const om_scene *s, int g)
{
    const GridState &G = s->grids[g];
    const GridState &D = (s->grid_src ? s->grid_src : s)->grids[g]; // device CSR owner
    a.gpx = G.pos[0];
    a.gpy = G.pos[1];
    a.gpz = G.pos[2];
    a.gix = G.interval[0];
    a.giy = G.interval[1];
    a.giz = G.interval[2];
    auto pow2 = [](float x) {
        int e;
        return x > 0.0f && std::frexp(x, &e) == 0.5f && e > -100 && e < 100;
    };
    a.grid_pow2 = pow2(a.gix) && pow2(a.giy) && pow2(a.giz);
    a.inv_gix = 1.0f / a.gix;
    a.inv_giy = 1.0f / a.giy;
    a.inv_giz = 1.0f / a.giz;
    a.gnx = G.gn[0];
    a.gny = G.gn[1];
    a.gnz = G.gn[2];
    for (int i = 0; i < 3; ++i) {
        a.occ_lo[i] = G.occ_lo[i];
        a.occ_hi[i] = G.occ_hi[i];
    }
    a.cell_start = D.d_cell_start;
    a.cell_items = D.d_cell_items;
    a.cell_geo = (size_t)g < s->d_cell_geo.size() ? s->d_cell_geo[g] : nullptr;
}

TraceArgs make_trace_args(const om_scene *s, int g)
{
    TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.literal_self = s->prm.literal_self ? 1 : 0;
    a.abort_flag = s->ctx ? s->ctx->d_abort : nullptr;
    frame sf = make_frame(s->sun);
    a.sun_u = sf.u;
    a.sun_n = sf.n;
    a.sun_v = sf.v;
    a.sun_ident = sf.identity;
    a.pillbox = s->prm.pillbox_half_angle;
    a.buie_int465 = s->buie_int465;
    a.buie_A = s->buie_A;
    a.buie_B = s->buie_B;
    a.buie_C = s->buie_C;
    a.buie_inv_gp1 = s->buie_inv_gp1;
    a.nbins = (s->prm.sunshape == OM_SUN_BUIE) ? s->nbins : 0;
    a.buie_bins = s->d_buie_bins;
    a.buie_guide = s->d_buie_guide;
    a.buie_gscale = (float)s->nbins / s->buie_int465;
    a.buie_split = s->d_buie_split;
    a.slope_sigma = s->prm.slope_sigma;
    a.key0 = (uint32_t)s->prm.seed;
    a.key1 = (uint32_t)(s->prm.seed >> 32);
    a.w3 = (s->prm.scene_index << 4) | PURPOSE_RAY;
    a.L = s->prm.rays_per_sample;
    a.pitch = s->prm.helio_pixel_length;
    a.pitch2 = s->prm.helio_pixel_length / 2;
    set_trace_grid(a, s, g);
    a.geo = s->d_geo;
    a.recv_type = s->recv_type;
    a.cyl_base = s->cyl_base;
    a.cyl_r2 = s->cyl_r2;
    a.cyl_h = s->cyl_h;
    a.cone_r0 = s->cone_r0;
    a.cone_k = s->cone_k;
    a.cone_inv_area = s->cone_inv_area;
    a.rA = s->rv[0];
    a.rE1 = s->rv[1] - s->rv[0];
    a.rE2 = s->rv[3] - s->rv[0];
    a.res_x = s->res_x;
    a.res_y = s->res_y;
    a.helio = s->d_helio;
    a.poisson = s->d_poisson;
    a.poisson_cap = s->poisson_cap;
    a.tiles = s->d_tiles;
    a.ntiles = (int)s->tiles.size();
    a.rays_per_thread = s->rays_per_thread;
    a.tile_rays = 256 * s->rays_per_thread;
    return a;
}
} // namespace om

static om_status scene_from_desc(om_ctx *ctx, const om_scene_desc *d, const om_params *p,
                                 om_scene **out)
{
    om_status st = validate(d, p);
    if (st != OM_OK) return st;
    om_scene *s = new om_scene();
    s->ctx = ctx;
    s->prm = *p;
    s->H = d->n_helio;
    s->pos.assign(d->helio_pos, d->helio_pos + 3 * (size_t)s->H);
    s->size.assign(d->helio_size, d->helio_size + 3 * (size_t)s->H);
    s->gap.assign(d->helio_gap, d->helio_gap + 2 * (size_t)s->H);
    s->matrix.assign(d->helio_matrix, d->helio_matrix + 2 * (size_t)s->H);
    if (d->helio_normal) s->normal_in.assign(d->helio_normal, d->helio_normal + 3 * (size_t)s->H);
    s->recv_type = d->recv_type;
    std::memcpy(s->recv_pos, d->recv_pos, 12);
    std::memcpy(s->recv_size, d->recv_size, 12);
    std::memcpy(s->recv_normal_in, d->recv_normal, 12);
    s->recv_face = d->recv_face;
    s->recv_pixel_length = d->recv_pixel_length;
    s->recv_res_in[0] = d->recv_res_x;
    s->recv_res_in[1] = d->recv_res_y;
    // grids: every grid over its heliostats, or grid k alone with its heliostats as the
    // scene's (the reference traces grid 0's heliostats over grid 0, Gtest.cu:12,21-23)
    std::vector<om_grid_desc> gl = desc_grids(d);
    if (p->grid_select != OM_GRIDS_ALL) {
        const om_grid_desc k = gl[p->grid_select];
        const size_t f = (size_t)k.first_helio;
        s->H = k.n_helio;
        s->pos.assign(d->helio_pos + 3 * f, d->helio_pos + 3 * (f + s->H));
        s->size.assign(d->helio_size + 3 * f, d->helio_size + 3 * (f + s->H));
        s->gap.assign(d->helio_gap + 2 * f, d->helio_gap + 2 * (f + s->H));
        s->matrix.assign(d->helio_matrix + 2 * f, d->helio_matrix + 2 * (f + s->H));
        if (d->helio_normal) s->normal_in.assign(d->helio_normal + 3 * f, d->helio_normal + 3 * (f + s->H));
        gl = {k};
        gl[0].first_helio = 0;
    }
    for (const om_grid_desc &g : gl) {
        om::GridState G;
        std::memcpy(G.pos, g.pos, 12);
        std::memcpy(G.size, g.size, 12);
        std::memcpy(G.interval, g.interval, 12);
        G.h0 = g.first_helio;
        G.h1 = g.first_helio + g.n_helio;
        s->grids.push_back(std::move(G));
    }
    std::memcpy(s->sun_in, d->sun_dir, 12);
    s->dni = d->dni;
    st = build_all(s);
    if (st != OM_OK) {
        free_scene_dev(s);
        delete s;
        return st;
    }
    *out = s;
    return OM_OK;
}

extern "C" om_status om_scene_create(om_ctx *ctx, const om_scene_desc *desc, const om_params *params,
                                     om_scene **out)
{
    if (!desc || !params || !out) return OM_ERR_INVALID;
    *out = nullptr;
    return scene_from_desc(ctx, desc, params, out);
}

extern "C" om_status om_scene_from_scn(om_ctx *ctx, const char *path, const float sun_dir[3], float dni,
                                       const om_params *params, om_scene **out)
{
    if (!ctx || !path || !sun_dir || !params || !out) return OM_ERR_INVALID;
    *out = nullptr;
    om_scene_desc d;
    om_status st = om_scn_parse(path, &d);
    if (st != OM_OK) return st;
    std::memcpy(d.sun_dir, sun_dir, 12);
    d.dni = dni;
    st = scene_from_desc(ctx, &d, params, out);
    om_scene_desc_free(&d);
    return st;
}

extern "C" om_status om_scene_destroy(om_scene *s)
{
    if (!s) return OM_ERR_INVALID;
    if (s->ctx) {
        (void)hipSetDevice(s->ctx->device);
        free_scene_dev(s);
    }
    delete s;
    return OM_OK;
}

extern "C" om_status om_scene_get_info(const om_scene *s, om_scene_info *info)
{
    if (!s || !info) return OM_ERR_INVALID;
    info->n_helio = s->H;
    info->res_x = s->res_x;
    info->res_y = s->res_y;
    for (int a = 0; a < 3; ++a) info->grid_num[a] = s->grids[0].gn[a];
    info->grid_cells = 0;
    info->grid_nnz = 0;
    for (const om::GridState &G : s->grids) {
        info->grid_cells += G.cells;
        info->grid_nnz += G.nnz;
    }
    info->n_grids = (int)s->grids.size();
    info->pixel_area = s->srec;
    info->total_micro = s->total_micro;
    info->total_rays = s->total_rays;
    info->max_micro = s->max_micro;
    info->pixel_length = s->recv_pixel;
    info->flux_scale = s->scale;
    info->total_energy = s->total_energy;
    return OM_OK;
}

extern "C" om_status om_scene_set_sun(om_scene *s, const float sun_dir[3], float dni)
{
    if (!s || !sun_dir) return OM_ERR_INVALID;
    if (!bounded(sun_dir, 3) || (sun_dir[0] == 0.0f && sun_dir[1] == 0.0f && sun_dir[2] == 0.0f)) {
        om::set_error("sun direction must be finite and nonzero");
        return OM_ERR_INVALID;
    }
    float old_sun[3];
    std::memcpy(old_sun, s->sun_in, 12);
    std::memcpy(s->sun_in, sun_dir, 12);
    s->sun = om::normalize(ld3(s->sun_in));
    aim_heliostats(s);
    om_status st = check_in_grid(s);
    if (st != OM_OK) {
        // keep the scene as it was
        std::memcpy(s->sun_in, old_sun, 12);
        s->sun = om::normalize(ld3(s->sun_in));
        aim_heliostats(s);
        return st;
    }
    s->dni = dni;
    flux_scale(s);
    if (!s->ctx) {
        pack_heliostats(s);
        return OM_OK;
    }
    OM_HIP(hipSetDevice(s->ctx->device));
    // the Poisson points are drawn in each heliostat's local frame, so they survive re-aiming
    return upload_geometry(s);
}

// One time step of an annual run on the same field (main.cpp:56-116 builds a new scene per
// step): the source's device grid, microsamples, tiles and Buie tables are shared, the
// heliostats are re-aimed for the new sun and get their own records and vertices.
extern "C" om_status om_scene_fork(const om_scene *src, const float sun_dir[3], float dni, uint32_t scene_index,
                                   om_scene **out)
{
    if (!src || !sun_dir || !out) return OM_ERR_INVALID;
    *out = nullptr;
    if (!src->ctx) {
        om::set_error("host-only scene (created without a context) cannot be forked");
        return OM_ERR_INVALID;
    }
    if (!bounded(sun_dir, 3) || (sun_dir[0] == 0.0f && sun_dir[1] == 0.0f && sun_dir[2] == 0.0f)) {
        om::set_error("sun direction must be finite and nonzero");
        return OM_ERR_INVALID;
    }
    // copy everything but the host CSR, which stays with the root
    om_scene &m = const_cast<om_scene &>(*src);
    const size_t ng = m.grids.size();
    std::vector<std::vector<int>> cs(ng), ci(ng);
    for (size_t g = 0; g < ng; ++g) {
        cs[g].swap(m.grids[g].cell_start);
        ci[g].swap(m.grids[g].cell_items);
    }
    om_scene *s = new om_scene(*src);
    for (size_t g = 0; g < ng; ++g) {
        m.grids[g].cell_start.swap(cs[g]);
        m.grids[g].cell_items.swap(ci[g]);
    }
    s->grid_src = src->grid_src ? src->grid_src : src;
    s->d_helio = nullptr;
    s->d_geo = nullptr;
    s->d_f64 = nullptr;
    s->d_tail64 = nullptr;
    s->d_cell_geo64.assign(ng, nullptr);
    s->d_cell_geo_m.assign(ng, nullptr);
    s->d_hmix = nullptr;
    s->d_buie_bins_m = nullptr;
    s->d_buie_split_m = nullptr;
    s->d_mix_n = nullptr;
    s->d_cell_geo.assign(ng, nullptr);
    s->diff_buf = om::DevBuf{};
    s->prm.scene_index = scene_index;
    std::memcpy(s->sun_in, sun_dir, 12);
    s->sun = om::normalize(ld3(s->sun_in));
    aim_heliostats(s);
    om_status st = check_in_grid(s);
    if (st == OM_OK) {
        s->dni = dni;
        flux_scale(s);
        if (hipSetDevice(s->ctx->device) != hipSuccess) st = om::hip_fail(hipGetLastError(), "hipSetDevice");
        else st = upload_geometry(s);
    }
    if (st != OM_OK) {
        free_scene_dev(s);
        delete s;
        return st;
    }
    *out = s;
    return OM_OK;
}

extern "C" om_status om_scene_export(const om_scene *s, const char *what, void *host, size_t bytes)
{
    if (!s || !what || !host) return OM_ERR_INVALID;
    std::string w(what);
    std::vector<char> buf;
    auto put = [&](const void *p, size_t n) {
        const char *c = (const char *)p;
        buf.insert(buf.end(), c, c + n);
    };
    if (w == "helio_normal") put(s->normal.data(), 12 * (size_t)s->H);
    else if (w == "helio_clear") {
        for (const om::HelioRec &r : s->recs) put(&r.clear, 4);
    }
    else if (w == "helio_v0") put(s->v0.data(), 12 * (size_t)s->H);
    else if (w == "helio_v1") put(s->v1.data(), 12 * (size_t)s->H);
    else if (w == "helio_v3") put(s->v3.data(), 12 * (size_t)s->H);
    else if (w.compare(0, 5, "grid_") == 0) {
        // grid_start / grid_items / grid_num / grid_helio, of grid 0 or of grid K (".K")
        std::string base = w;
        int k = 0;
        const size_t dot = w.find('.');
        if (dot != std::string::npos) {
            base = w.substr(0, dot);
            const std::string num = w.substr(dot + 1);
            if (num.empty() || num.size() > 6 || num.find_first_not_of("0123456789") != std::string::npos) {
                om::set_error("unknown export " + w);
                return OM_ERR_INVALID;
            }
            k = std::stoi(num);
        }
        if (k >= (int)s->grids.size()) {
            om::set_error("export " + w + ": no such grid");
            return OM_ERR_INVALID;
        }
        const om::GridState &G = (s->grid_src ? s->grid_src : s)->grids[k];
        if (base == "grid_start") put(G.cell_start.data(), 4 * G.cell_start.size());
        else if (base == "grid_items") put(G.cell_items.data(), 4 * (size_t)G.nnz);
        else if (base == "grid_num") put(G.gn, 12);
        else if (base == "grid_helio") {
            const int hn[2] = {G.h0, G.h1 - G.h0};
            put(hn, 8);
        } else {
            om::set_error("unknown export " + w);
            return OM_ERR_INVALID;
        }
    } else if (w == "receiver") {
        put(s->rv, 48);
        put(&s->focus, 12);
    } else if (w == "buie") {
        put(s->cdf_x.data(), 4 * s->cdf_x.size());
        put(s->cdf_y.data(), 4 * s->cdf_y.size());
        put(s->slope.data(), 4 * s->slope.size());
    } else if (w == "micro_count") put(s->mcount.data(), 4 * (size_t)s->H);
    else if (w == "ray_offset") put(s->ray_offset.data(), 8 * s->ray_offset.size());
    else if (w == "sun") {
        put(&s->sun, 12);
        float c[5] = {s->buie_int465, s->buie_A, s->buie_B, s->buie_C, s->buie_inv_gp1};
        put(c, 20);
    } else if (w == "poisson") {
        if (s->prm.micro_mode != OM_MICRO_POISSON) {
            om::set_error("scene has no Poisson microsamples");
            return OM_ERR_INVALID;
        }
        std::vector<float2> all((size_t)s->poisson_cap * s->H);
        if (!all.empty() && s->d_poisson)
            OM_HIP(hipMemcpy(all.data(), s->d_poisson, sizeof(float2) * all.size(), hipMemcpyDeviceToHost));
        std::vector<float2> out((size_t)s->max_micro * s->H, make_float2(0.0f, 0.0f));
        for (int h = 0; h < s->H; ++h)
            for (int m = 0; m < s->mcount[h]; ++m)
                out[(size_t)h * s->max_micro + m] = all[(size_t)h * s->poisson_cap + m];
        put(out.data(), sizeof(float2) * out.size());
    } else {
        om::set_error("unknown export " + w);
        return OM_ERR_INVALID;
    }
    if (bytes < buf.size()) {
        om::set_error("export buffer too small");
        return OM_ERR_SIZE;
    }
    std::memcpy(host, buf.data(), buf.size());
    return OM_OK;
}

// ImageSaver::savetxt(..., float totalEnergy) (image_save.cpp:26-47): rows are summed from
// the last to the first, pixels below Epsilon skipped.
extern "C" float om_intercept_ratio(const float *flux, int width, int height, float total_energy)
{
    float in = 0.0f;
    for (int r = 0; r < height; ++r)
        for (int c = 0; c < width; ++c) {
            float v = flux[(size_t)(height - 1 - r) * width + c];
            if (v < om::EPS_LT) continue;
            in += v;
        }
    return in / total_energy;
}
