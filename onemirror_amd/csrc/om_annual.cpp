// om_annual.cpp -- host helpers of the annual time-step driver (main.cpp:56-116): the sun
// direction from zenith and azimuth and the clear-sky direct normal irradiance of getDNI
// (SPA.cpp:1174-1189, Hottel's beam transmittance).  Solar position itself (SPA) stays with
// the caller: these take the angles as input.
#include <cmath>

#include "om_internal.h"

namespace {
constexpr double PI_D = 3.14159265358979323846;
}

// main.cpp:71-73: light travel direction of the sun at zenith z and azimuth az (degrees,
// azimuth clockwise from north): (-sin z sin az, -cos z, sin z cos az).
extern "C" om_status om_sun_direction(float zenith_deg, float azimuth_deg, float out[3])
{
    if (!out) return OM_ERR_INVALID;
    if (!std::isfinite(zenith_deg) || !std::isfinite(azimuth_deg)) {
        om::set_error("sun angles must be finite");
        return OM_ERR_INVALID;
    }
    const double z = zenith_deg / 180.0 * PI_D, az = azimuth_deg / 180.0 * PI_D;
    out[0] = (float)(-std::sin(z) * std::sin(az));
    out[1] = (float)(-std::cos(z));
    out[2] = (float)(std::sin(z) * std::cos(az));
    return OM_OK;
}

// getDNI (SPA.cpp:1174-1189): extraterrestrial s = 1353 + 45 cos(2 pi (n + 10) / 365) with n
// the day of a 365-day year, times Hottel's clear-sky beam transmittance
// a0 + a1 exp(-k / cos z), a0 = 0.4237 - 0.00821 (6 - A)^2, a1 = 0.5055 + 0.00595 (6.5 - A)^2,
// k = 0.2711 + 0.01858 (2.5 - A)^2 for a site at A km.  W/m^2.
extern "C" float om_dni_hottel(float zenith_deg, int month, int day, float elevation_m)
{
    static const int month_days[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if (month < 1 || month > 12) return NAN;
    double n = 0.0;
    for (int i = 0; i < month - 1; ++i) n += month_days[i];
    n += day;
    const double A = elevation_m / 1000.0;
    const double a0 = 0.4237 - 0.00821 * (6.0 - A) * (6.0 - A);
    const double a1 = 0.5055 + 0.00595 * (6.5 - A) * (6.5 - A);
    const double k = 0.2711 + 0.01858 * (2.5 - A) * (2.5 - A);
    const double s = 1353.0 + 45.0 * std::cos(2.0 * PI_D * (n + 10.0) / 365.0);
    return (float)(s * (a0 + a1 * std::exp(-k / std::sin((90.0 - zenith_deg) * PI_D / 180.0))));
}
