// A C++ consumer of include/onemirror.h, built by tests/test_abi.py with g++ (no HIP
// compiler) and linked against libonemirror_amd.so.  It prints, as one JSON object:
//   "layout": sizeof, alignof and the offset of every field of every public struct, which
//             the test compares with the Python (ctypes) mirrors of the same structs;
//   "calls":  results of the entry points that need no GPU (defaults, status strings, sun
//             direction, clear-sky DNI, a .scn file parsed and built as a host-only scene),
//             which the test compares with the same calls made through ctypes.
// Usage: abi_layout <path of a .scn file>
#include "onemirror.h"

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <string>

namespace {

std::string g_out;
bool g_first_struct = true, g_first_field = true;

void begin_struct(const char *name, size_t size, size_t align)
{
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s\"%s\": {\"size\": %zu, \"align\": %zu, \"fields\": {",
                  g_first_struct ? "" : ", ", name, size, align);
    g_out += buf;
    g_first_struct = false;
    g_first_field = true;
}

void field(const char *name, size_t off, size_t size)
{
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s\"%s\": [%zu, %zu]", g_first_field ? "" : ", ", name, off, size);
    g_out += buf;
    g_first_field = false;
}

void end_struct() { g_out += "}}"; }

#define OM_STRUCT(T) begin_struct(#T, sizeof(T), alignof(T))
#define OM_FIELD(T, f) field(#f, offsetof(T, f), sizeof(((T *)nullptr)->f))

void layout()
{
    OM_STRUCT(om_grid_desc);
    OM_FIELD(om_grid_desc, pos); OM_FIELD(om_grid_desc, size); OM_FIELD(om_grid_desc, interval);
    OM_FIELD(om_grid_desc, first_helio); OM_FIELD(om_grid_desc, n_helio);
    end_struct();

    OM_STRUCT(om_scene_desc);
    OM_FIELD(om_scene_desc, n_helio); OM_FIELD(om_scene_desc, helio_pos); OM_FIELD(om_scene_desc, helio_size);
    OM_FIELD(om_scene_desc, helio_gap); OM_FIELD(om_scene_desc, helio_matrix); OM_FIELD(om_scene_desc, recv_type);
    OM_FIELD(om_scene_desc, recv_pos); OM_FIELD(om_scene_desc, recv_size); OM_FIELD(om_scene_desc, recv_normal);
    OM_FIELD(om_scene_desc, recv_face); OM_FIELD(om_scene_desc, recv_pixel_length);
    OM_FIELD(om_scene_desc, grid_pos); OM_FIELD(om_scene_desc, grid_size); OM_FIELD(om_scene_desc, grid_interval);
    OM_FIELD(om_scene_desc, sun_dir); OM_FIELD(om_scene_desc, dni); OM_FIELD(om_scene_desc, recv_res_x);
    OM_FIELD(om_scene_desc, recv_res_y); OM_FIELD(om_scene_desc, n_grids); OM_FIELD(om_scene_desc, grids);
    OM_FIELD(om_scene_desc, ground_length); OM_FIELD(om_scene_desc, ground_width);
    OM_FIELD(om_scene_desc, helio_normal);
    end_struct();

    OM_STRUCT(om_params);
    OM_FIELD(om_params, seed); OM_FIELD(om_params, scene_index); OM_FIELD(om_params, sunshape);
    OM_FIELD(om_params, pillbox_half_angle); OM_FIELD(om_params, csr); OM_FIELD(om_params, buie_bins);
    OM_FIELD(om_params, slope_sigma); OM_FIELD(om_params, micro_mode); OM_FIELD(om_params, helio_pixel_length);
    OM_FIELD(om_params, poisson_k); OM_FIELD(om_params, rays_per_sample); OM_FIELD(om_params, reflectivity);
    OM_FIELD(om_params, smooth_radius); OM_FIELD(om_params, smooth_ratio); OM_FIELD(om_params, accumulate);
    OM_FIELD(om_params, pack_shift); OM_FIELD(om_params, grid_select); OM_FIELD(om_params, literal_self);
    OM_FIELD(om_params, precision); OM_FIELD(om_params, ray_compact); OM_FIELD(om_params, redo_cap);
    end_struct();

    OM_STRUCT(om_flux);
    OM_FIELD(om_flux, counts); OM_FIELD(om_flux, energy); OM_FIELD(om_flux, flux);
    OM_FIELD(om_flux, flux_smooth); OM_FIELD(om_flux, capacity);
    end_struct();

    OM_STRUCT(om_scene_info);
    OM_FIELD(om_scene_info, n_helio); OM_FIELD(om_scene_info, res_x); OM_FIELD(om_scene_info, res_y);
    OM_FIELD(om_scene_info, grid_num); OM_FIELD(om_scene_info, grid_cells); OM_FIELD(om_scene_info, grid_nnz);
    OM_FIELD(om_scene_info, total_micro); OM_FIELD(om_scene_info, total_rays); OM_FIELD(om_scene_info, max_micro);
    OM_FIELD(om_scene_info, pixel_length); OM_FIELD(om_scene_info, flux_scale);
    OM_FIELD(om_scene_info, total_energy); OM_FIELD(om_scene_info, n_grids); OM_FIELD(om_scene_info, pixel_area);
    end_struct();

    OM_STRUCT(om_result);
    OM_FIELD(om_result, flux); OM_FIELD(om_result, flux_smooth); OM_FIELD(om_result, ratio);
    OM_FIELD(om_result, ratio_smooth); OM_FIELD(om_result, total_energy); OM_FIELD(om_result, rays);
    end_struct();

    OM_STRUCT(om_launch);
    OM_FIELD(om_launch, workgroup); OM_FIELD(om_launch, rays_per_thread);
    end_struct();

    OM_STRUCT(om_tune);
    OM_FIELD(om_tune, best); OM_FIELD(om_tune, best_ms); OM_FIELD(om_tune, rays_per_s);
    OM_FIELD(om_tune, n_trials); OM_FIELD(om_tune, trials);
    end_struct();
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: abi_layout <scene.scn>\n");
        return 2;
    }
    g_out = "{\"layout\": {";
    layout();
    g_out += "}, \"calls\": {";
    char buf[1024];
    om_params p;
    om_default_params(&p);
    std::snprintf(buf, sizeof(buf),
                  "\"default_params\": {\"seed\": %llu, \"sunshape\": %d, \"csr\": %.9g, \"buie_bins\": %d, "
                  "\"rays_per_sample\": %d, \"reflectivity\": %.9g, \"smooth_radius\": %d, \"smooth_ratio\": %.9g, "
                  "\"accumulate\": %d, \"literal_self\": %d}",
                  (unsigned long long)p.seed, p.sunshape, p.csr, p.buie_bins, p.rays_per_sample, p.reflectivity,
                  p.smooth_radius, p.smooth_ratio, p.accumulate, p.literal_self);
    g_out += buf;
    std::snprintf(buf, sizeof(buf), ", \"status_size\": \"%s\"", om_status_string(OM_ERR_SIZE));
    g_out += buf;
    float sd[3];
    const om_status rs = om_sun_direction(30.0f, 135.0f, sd);
    std::snprintf(buf, sizeof(buf), ", \"sun_direction\": [%d, %.9g, %.9g, %.9g]", (int)rs, sd[0], sd[1], sd[2]);
    g_out += buf;
    std::snprintf(buf, sizeof(buf), ", \"dni_hottel\": %.9g", om_dni_hottel(40.0f, 6, 21, 1500.0f));
    g_out += buf;
    // a .scn file parsed and built as a host-only scene (no context: no GPU needed)
    om_scene_desc d;
    om_status st = om_scn_parse(argv[1], &d);
    int nh = -1, rx = -1, ry = -1, gn0 = -1;
    long long rays = -1;
    if (st == OM_OK) {
        d.sun_dir[0] = 0.0f;
        d.sun_dir[1] = -0.70710677f;
        d.sun_dir[2] = -0.70710677f;
        d.dni = 1000.0f;
        om_scene *s = nullptr;
        st = om_scene_create(nullptr, &d, &p, &s);
        if (st == OM_OK) {
            om_scene_info info;
            st = om_scene_get_info(s, &info);
            nh = info.n_helio;
            rx = info.res_x;
            ry = info.res_y;
            gn0 = info.grid_num[0];
            rays = (long long)info.total_rays;
            om_scene_destroy(s);
        }
        om_scene_desc_free(&d);
    }
    std::snprintf(buf, sizeof(buf),
                  ", \"scn_host_scene\": {\"status\": %d, \"n_helio\": %d, \"res_x\": %d, \"res_y\": %d, "
                  "\"grid_num0\": %d, \"total_rays\": %lld}",
                  (int)st, nh, rx, ry, gn0, rays);
    g_out += buf;
    // an invalid call returns a status instead of ending the process
    const om_status bad = om_scene_get_info(nullptr, nullptr);
    std::snprintf(buf, sizeof(buf), ", \"null_info_status\": %d", (int)bad);
    g_out += buf;
    g_out += "}}";
    std::printf("%s\n", g_out.c_str());
    return 0;
}
