// Input producers of the chemistry step on the device: solar zenith angle (zenith,
// F:5308-5406) and geometric altitude (altitude, src/altitude.f90 A:1-148) of the
// trajectory points, so that the model's driver passes only T, p, lon, lat and the
// day number.  Expressions keep the reference's evaluation order.
#include "chem_device.h"

#include <math.h>

namespace rb {

// SZA of each point; the day's declination terms are computed once on the host.
__global__ void k_zenith(const double *__restrict__ lon, const double *__restrict__ lat,
                         int64_t n, ZenithTerms z, double *__restrict__ sza)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rlt = lat[i] * z.pi / 180.;
    const double lbut = 12. - z.eqh - lon[i] * 24. / 360.;
    const double lzut = 15. * (z.ut - lbut);
    const double zpt = lzut * z.pi / 180.;
    const double csz = sin(rlt) * z.sin_decl + cos(rlt) * z.cos_decl * cos(zpt);
    const double zr = acos(csz);
    sza[i] = fmax(zr * 180. / z.pi, 0.);
}

// Regular grid, one thread per column: level pressures p = aaa + bbb psol, virtual
// temperature, and geopotential height (km) integrated up from the surface (A:39-66).
__global__ void k_georeg(const double *__restrict__ relief, const double *__restrict__ psol,
                         const double *__restrict__ aaa, const double *__restrict__ bbb,
                         const double *__restrict__ t, const double *__restrict__ h2o,
                         int64_t plane, int niv, double *__restrict__ p, double *__restrict__ tv,
                         double *__restrict__ geo)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    const double ctegas = 287.04 / 9.80665;
    for (int k = 0; k < niv; ++k) {
        const int64_t o = (int64_t)k * plane + c;
        p[o] = aaa[k] + bbb[k] * psol[c];
        tv[o] = t[o] * (1. + 0.608 * h2o[o] * 18. / 28.97);
    }
    int64_t o = (int64_t)(niv - 1) * plane + c;
    double g = relief[c] + ctegas * tv[o] * log(psol[c] / p[o]) * 1.e-3;
    geo[o] = g;
    for (int k = niv - 2; k >= 0; --k) {
        const int64_t u = (int64_t)k * plane + c;
        g = g + ctegas * ((tv[o] + tv[u]) / 2.) * log(p[o] / p[u]) * 1.e-3;
        geo[u] = g;
        o = u;
    }
}

// Trajectory points: nearest regular column (dlon x dlat spacing, indices clamped into
// the grid), pressure clamped to the column's bottom level, geopotential of the
// uppermost level (>= 2) at or below that pressure plus the hypsometric increment, then
// geometric altitude (A:68-146).
__global__ void k_altitude(const double *__restrict__ p, const double *__restrict__ tv,
                           const double *__restrict__ geo, const double *__restrict__ tlon,
                           const double *__restrict__ tlat, const double *__restrict__ tpre,
                           int mlon, int mlat, int niv, double dlon, double dlat,
                           double *__restrict__ alt)
{
    const int64_t plane = (int64_t)mlon * mlat;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= plane * niv) return;
    const double ctegas = 287.04 / 9.80665;
    const double pi = 3.141592653589793;   // 2 asin(1)
    int il = (int)round(tlon[b] / dlon) + 1;
    if (il > mlon) il = 1;
    int jl = (int)round((90. - tlat[b]) / dlat) + 1;
    il = max(il, 1);
    jl = min(max(jl, 1), mlat);
    const int64_t c = (int64_t)(jl - 1) * mlon + (il - 1);
    const double pirr = fmin(tpre[b], p[(int64_t)(niv - 1) * plane + c]);
    int kr = niv;
    for (int k = 2; k <= niv; ++k)
        if (p[(int64_t)(k - 1) * plane + c] >= pirr) { kr = k; break; }
    const int64_t o = (int64_t)(kr - 1) * plane + c, u = o - plane;
    const double dz = ctegas * (tv[o] + tv[u]) / 2. * log(p[o] / pirr) * 1.e-3;
    const double g = geo[o] + dz;
    const double xlat = tlat[b] * pi / 180.;
    alt[b] = (1. + 0.002644 * cos(2. * xlat)) * g + (1. + 0.0089 * cos(2. * xlat)) * (g * g) / 6245.;
}

ZenithTerms zenith_terms(double day)
{
    ZenithTerms z;
    z.pi = 2. * asin(1.0);
    z.ut = (day - trunc(day)) * 24.;
    const double tz = 2. * z.pi * day / 365.;
    const double sintz = sin(tz), costz = cos(tz);
    const double sin2tz = 2. * sintz * costz;
    const double cos2tz = costz * costz - sintz * sintz;
    const double sin3tz = sintz * cos2tz + costz * sin2tz;
    const double cos3tz = costz * cos2tz - sintz * sin2tz;
    // Spencer (1971) series for the declination and the equation of time (F:5362-5376)
    const double rdecl = 0.006918 - 0.399912 * costz + 0.070257 * sintz - 0.006758 * cos2tz +
                         0.000907 * sin2tz - 0.002697 * cos3tz + 0.001480 * sin3tz;
    const double eqr = 0.000075 + 0.001868 * costz - 0.032077 * sintz - 0.014615 * cos2tz -
                       0.040849 * sin2tz;
    z.eqh = eqr * 24. / (2. * z.pi);
    z.sin_decl = sin(rdecl);
    z.cos_decl = cos(rdecl);
    return z;
}

hipError_t launch_zenith(const double *lon, const double *lat, int64_t n, const ZenithTerms &z,
                         double *sza, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zenith, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lon, lat, n, z,
                       sza);
    return hipGetLastError();
}

hipError_t launch_altitude(const double *relief, const double *psol, const double *aaa,
                           const double *bbb, const double *t, const double *h2o,
                           const double *tlon, const double *tlat, const double *tpre, int mlon,
                           int mlat, int niv, double dlon, double dlat, double *work, double *alt,
                           hipStream_t s)
{
    const int64_t plane = (int64_t)mlon * mlat, field = plane * niv;
    double *p = work, *tv = work + field, *geo = work + 2 * field;
    hipLaunchKernelGGL(k_georeg, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, relief, psol,
                       aaa, bbb, t, h2o, plane, niv, p, tv, geo);
    hipLaunchKernelGGL(k_altitude, dim3((unsigned)((field + 255) / 256)), dim3(256), 0, s, p, tv, geo,
                       tlon, tlat, tpre, mlon, mlat, niv, dlon, dlat, alt);
    return hipGetLastError();
}

}  // namespace rb
