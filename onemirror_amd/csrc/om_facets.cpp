// om_facets.cpp -- the reference's faceted heliostat scene: nine flat facets canted along a
// paraboloid and turned to a master normal (SceneFileProc::SceneFileRead(SolarScene*, a, b,
// mirrorNormal, X, Y, Z), scene_file_proc.cpp:288-534), the master normal of the annual
// driver (main.cpp:84-91), and the scene the driver traces with it.
//
// The arithmetic keeps the reference's types: float facet coordinates and matrices, the
// double constants it writes as literals (sqrt(2.0)/3 + 0.1, 1.4142136/3, receiverH = 13.01)
// rounded where the reference assigns them to floats.
#include <cmath>
#include <cstring>

#include "om_internal.h"

namespace {

// receiverH of main.cpp:8, a double literal
constexpr double RECEIVER_H = 13.01;

// SceneFileProc::matrixMulti (scene_file_proc.cpp:536-543): c = A b, summed in j order
void matrix_multi(const float A[3][3], const float b[3], float c[3])
{
    for (int i = 0; i < 3; ++i) {
        c[i] = 0.0f;
        for (int j = 0; j < 3; ++j) c[i] += A[i][j] * b[j];
    }
}

} // namespace

extern "C" om_status om_master_normal(float X, float Y, float Z, const float sun_dir[3], float out[3])
{
    if (!sun_dir || !out) return OM_ERR_INVALID;
    const float a = sun_dir[0], b = sun_dir[1], c = sun_dir[2];
    // main.cpp:84-91; X*X + Z*Z is a float sum, the receiver height term a double
    const float distance = (float)std::sqrt((double)(X * X + Z * Z) + (RECEIVER_H - (double)Y) * (RECEIVER_H - (double)Y));
    if (!(distance > 0.0f)) {
        om::set_error("the heliostat sits at the receiver point");
        return OM_ERR_INVALID;
    }
    const float rox = -X / distance;
    const float roz = -Z / distance;
    const float roy = (float)((RECEIVER_H - (double)Y) / (double)distance);
    const float NR = std::sqrt((-a + rox) * (-a + rox) + (-b + roy) * (-b + roy) + (-c + roz) * (-c + roz));
    if (!(NR > 0.0f)) {
        om::set_error("sun and receiver directions cancel");
        return OM_ERR_INVALID;
    }
    out[0] = (-a + rox) / NR;
    out[1] = (-b + roy) / NR;
    out[2] = (-c + roz) / NR;
    return OM_OK;
}

extern "C" om_status om_facet_heliostat(float a, float b, const float mirror_normal[3], float X, float Y, float Z,
                                        float pos[27], float normal[27], float size[27])
{
    if (!mirror_normal || !pos || !normal || !size) return OM_ERR_INVALID;
    if (!(a != 0.0f) || !(b != 0.0f)) {
        om::set_error("paraboloid coefficients a and b must be nonzero");
        return OM_ERR_INVALID;
    }
    // facet centres on a 3 x 3 lattice of pitch sqrt(2)/3 + 0.1 (:289-290), lifted onto
    // y = x^2/a^2 + z^2/b^2 (:298-300)
    const float e = (float)(std::sqrt(2.0) / 3 + 0.1);
    const float cx[9] = {-e, 0.0f, e, -e, 0.0f, e, -e, 0.0f, e};
    const float cz[9] = {-e, -e, -e, 0.0f, 0.0f, 0.0f, e, e, e};
    float cy[9];
    for (int i = 0; i < 9; ++i) cy[i] = cx[i] * cx[i] / a / a + cz[i] * cz[i] / b / b;
    // the matrix from the paraboloid frame to the ground (:302-325): column 1 is the master
    // normal; column 0 is horizontal, scaled by sqrt(n0^2 + n1^2) as written (a unit vector
    // only when n1 = 0); column 2 = column 0 x column 1
    const float *n = mirror_normal;
    const float temp = std::sqrt(n[0] * n[0] + n[1] * n[1]);
    float M[3][3];
    if (n[2] > 0.0f) {
        M[0][0] = n[2] / temp;
        M[2][0] = -n[0] / temp;
        M[1][0] = 0.0f;
    } else if (n[2] < 0.0f) {
        M[0][0] = -n[2] / temp;
        M[2][0] = n[0] / temp;
        M[1][0] = 0.0f;
    } else {
        M[0][0] = 0.0f;
        M[1][0] = 0.0f;
        M[2][0] = 1.0f;
    }
    for (int i = 0; i < 3; ++i) M[i][1] = n[i];
    M[0][2] = M[1][0] * M[2][1] - M[2][0] * M[1][1];
    M[1][2] = M[2][0] * M[0][1] - M[0][0] * M[2][1];
    M[2][2] = M[0][0] * M[1][1] - M[1][0] * M[0][1];
    // facet centres and paraboloid normals (-2x/a^2, 1, -2z/b^2) to the ground (:327-339)
    const float sx = (float)(1.4142136 / 3), sy = 0.1f;
    for (int i = 0; i < 9; ++i) {
        const float c[3] = {cx[i], cy[i], cz[i]};
        const float nl[3] = {-c[0] * 2 / a / a, 1.0f, -c[2] * 2 / b / b};
        float ct[3], nt[3];
        matrix_multi(M, c, ct);
        matrix_multi(M, nl, nt);
        pos[3 * i + 0] = ct[0] + X;
        pos[3 * i + 1] = ct[1] + Y;
        pos[3 * i + 2] = ct[2] + Z;
        std::memcpy(normal + 3 * i, nt, 12);
        // facet size (:478-481)
        size[3 * i + 0] = sx;
        size[3 * i + 1] = sy;
        size[3 * i + 2] = sx;
    }
    auto finite = [](const float *x, int k) {
        for (int i = 0; i < k; ++i)
            if (!std::isfinite(x[i])) return false;
        return true;
    };
    if (!finite(pos, 27) || !finite(normal, 27)) {
        om::set_error("facet geometry is not finite (master normal with n0 = n1 = 0?)");
        return OM_ERR_INVALID;
    }
    return OM_OK;
}

// The rest of the faceted scene (scene_file_proc.cpp:360-462): ground 1024 x 1024, one
// rectangle receiver 1.4 x 1.4 x 3 m at (0, 13, -1.5) facing +z, one grid of 110 m cells
// over (-5.5, -5.5, 94.5) + (110, 110, 220), nine facets without gaps in one 1 x 1 matrix,
// their normals given (InitContent(normals), solar_scene.cpp:101-115).
extern "C" om_status om_facet_scene_desc(float a, float b, const float mirror_normal[3], float X, float Y, float Z,
                                         float pos[27], float normal[27], float size[27], float gap[18],
                                         int matrix[18], om_scene_desc *desc)
{
    if (!gap || !matrix || !desc) return OM_ERR_INVALID;
    om_status st = om_facet_heliostat(a, b, mirror_normal, X, Y, Z, pos, normal, size);
    if (st != OM_OK) return st;
    for (int i = 0; i < 18; ++i) {
        gap[i] = 0.0f;
        matrix[i] = 1;
    }
    std::memset(desc, 0, sizeof(*desc));
    desc->n_helio = 9;
    desc->helio_pos = pos;
    desc->helio_size = size;
    desc->helio_gap = gap;
    desc->helio_matrix = matrix;
    desc->helio_normal = normal;
    desc->recv_type = 0;
    const float rp[3] = {0.0f, 13.0f, -1.5f}, rs[3] = {1.4f, 1.4f, 3.0f}, rn[3] = {0.0f, 0.0f, 1.0f};
    std::memcpy(desc->recv_pos, rp, 12);
    std::memcpy(desc->recv_size, rs, 12);
    std::memcpy(desc->recv_normal, rn, 12);
    desc->recv_face = 0;
    desc->recv_pixel_length = 0.01f; // solarenergy::receiver_pixel_length (common_var.cpp:12)
    const float gp[3] = {-5.5f, -5.5f, 94.5f}, gs[3] = {110.0f, 110.0f, 220.0f}, gi[3] = {110.0f, 110.0f, 110.0f};
    std::memcpy(desc->grid_pos, gp, 12);
    std::memcpy(desc->grid_size, gs, 12);
    std::memcpy(desc->grid_interval, gi, 12);
    desc->ground_length = 1024.0f;
    desc->ground_width = 1024.0f;
    desc->dni = 1000.0f;
    const float down[3] = {0.0f, -1.0f, 0.0f};
    std::memcpy(desc->sun_dir, down, 12);
    return OM_OK;
}
