// om_io.cpp -- the reference's text outputs and normals file: ImageSaver::savetxt
// (image_save.cpp:5-47) and load_noraml (scene_normals.cpp:3-13).
//
// iostream formatting is printf formatting of the value promoted to double: "fixed <<
// setprecision(p) << x" is "%.*f" and a default-formatted float is "%g" (precision 6).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "om_internal.h"

namespace {

FILE *open_append(const char *path)
{
    FILE *f = std::fopen(path, "a");
    if (!f) om::set_error(std::string("cannot open ") + path + " for appending");
    return f;
}

om_status close_checked(FILE *f, const char *path)
{
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) {
        om::set_error(std::string("error writing ") + path);
        return OM_ERR_IO;
    }
    return OM_OK;
}

} // namespace

// ImageSaver::savetxt(filename, w, h, h_data, precision) (image_save.cpp:5-24): appends the
// map, last row first, "0," for a pixel below 1e-6 and "%.<precision>f," otherwise, one
// line per row.
extern "C" om_status om_savetxt_map(const char *path, int width, int height, const float *map, int precision)
{
    if (!path || !map || width <= 0 || height <= 0 || precision < 0 || precision > 60) return OM_ERR_INVALID;
    FILE *f = open_append(path);
    if (!f) return OM_ERR_IO;
    for (int r = 0; r < height; ++r) {
        const float *row = map + (size_t)(height - 1 - r) * width;
        for (int c = 0; c < width; ++c) {
            if (row[c] < om::EPS_LT) std::fputs("0,", f);
            else std::fprintf(f, "%.*f,", precision, (double)row[c]);
        }
        std::fputc('\n', f);
    }
    return close_checked(f, path);
}

// ImageSaver::savetxt(filename, w, h, h_data, totalEnergy) (image_save.cpp:26-47): appends
// one line, the intercept ratio (om_intercept_ratio) in default float formatting.
extern "C" om_status om_savetxt_ratio(const char *path, int width, int height, const float *map, float total_energy)
{
    if (!path || !map || width <= 0 || height <= 0) return OM_ERR_INVALID;
    const float ratio = om_intercept_ratio(map, width, height, total_energy);
    FILE *f = open_append(path);
    if (!f) return OM_ERR_IO;
    std::fprintf(f, "%g\n", (double)ratio);
    return close_checked(f, path);
}

// load_noraml (scene_normals.cpp:3-13): a count N, then N normals as x y z, whitespace
// separated.  A file with fewer than N readable normals is an error (the reference leaves
// the rest of its array uninitialised).  out == NULL asks for N only.
extern "C" om_status om_load_normals(const char *path, float *out, int capacity, int *n_out)
{
    if (!path || !n_out) return OM_ERR_INVALID;
    std::ifstream f(path);
    if (!f.is_open()) {
        om::set_error(std::string("cannot open normals file ") + path);
        return OM_ERR_IO;
    }
    long long n = -1;
    if (!(f >> n) || n < 0 || n > (1LL << 28)) {
        om::set_error("normals file must start with a count of 0 .. 2^28");
        return OM_ERR_PARSE;
    }
    *n_out = (int)n;
    if (!out) return OM_OK;
    if (capacity < n) {
        om::set_error("normals buffer smaller than the file's count");
        return OM_ERR_SIZE;
    }
    for (long long i = 0; i < 3 * n; ++i) {
        if (!(f >> out[i])) {
            om::set_error("normals file holds fewer normals than its count");
            return OM_ERR_PARSE;
        }
    }
    return OM_OK;
}
