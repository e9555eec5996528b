// Device helpers shared by the transport kernels: the cell of a point and its Lagrange
// weights (posx/posy/posz, poidx/poidy/poidz, src/reprobus_l137_20230723.f:7371-7647),
// operation for operation, in FP64 (built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "chem_device.h"

namespace rb {

// Cell (1-based ig0, kg0 and, with want_z, lg0) and weights w[0..3] along x, w[4..7] along
// y and, with want_z, w[8..11] along z of the point (xx, yy, zz) in a column of surface
// pressure pp; aaa, bbb are the hybrid coefficients (hPa, 1).
__device__ __forceinline__ void cell_weights(const TransportGrid &t, const double *__restrict__ aaa,
                                             const double *__restrict__ bbb, double xx, double yy,
                                             double zz, double pp, bool want_z, int c[3],
                                             double w[12])
{
    // posx F:7402-7406, posy F:7442-7446
    const int ix = (int)(xx / t.ddlon);
    const double dx = xx - (double)ix * t.ddlon;
    const int iy = (int)((t.xpih - yy) / t.ddlat);
    const double dy = t.xpih - (double)iy * t.ddlat - yy;
    c[0] = ix + 1;
    c[1] = iy + 1;
    // poidx F:7529-7536
    {
        const double dxm1 = dx + t.ddlon, dxp0 = dx, dxp1 = dx - t.ddlon, dxp2 = dx - t.ddlon * 2.;
        w[0] = t.xcoef[0] * dxp0 * dxp1 * dxp2;
        w[1] = t.xcoef[1] * dxm1 * dxp1 * dxp2;
        w[2] = t.xcoef[2] * dxm1 * dxp0 * dxp2;
        w[3] = t.xcoef[3] * dxm1 * dxp0 * dxp1;
    }
    // poidy F:7571-7578
    {
        const double dym1 = dy + t.ddlat, dyp0 = dy, dyp1 = dy - t.ddlat, dyp2 = dy - t.ddlat * 2.;
        w[4] = t.ycoef[0] * dyp0 * dyp1 * dyp2;
        w[5] = t.ycoef[1] * dym1 * dyp1 * dyp2;
        w[6] = t.ycoef[2] * dym1 * dyp0 * dyp2;
        w[7] = t.ycoef[3] * dym1 * dyp0 * dyp1;
    }
    if (!want_z) return;
    // posz F:7485-7494: the last level l in 2..niv-1 whose pressure is <= zz
    int ll = 1;
    for (int l = 2; l <= t.niv - 1; ++l) {
        const double pres = (aaa[l - 1] + bbb[l - 1] * pp) * 100.;
        if (zz - pres >= 0.) ll = l;
    }
    c[2] = ll;
    // poidz F:7614-7644
    const double prem1 = ll > 1 ? (aaa[ll - 2] + bbb[ll - 2] * pp) * 100. : 5.e-3 * 100.;
    const double prep0 = (aaa[ll - 1] + bbb[ll - 1] * pp) * 100.;
    const double prep1 = (aaa[ll] + bbb[ll] * pp) * 100.;
    const double prep2 = ll < t.niv - 1 ? (aaa[ll + 1] + bbb[ll + 1] * pp) * 100. : 1200. * 100.;
    const double dzm1 = zz - prem1, dzp0 = zz - prep0, dzp1 = zz - prep1, dzp2 = zz - prep2;
    const double dzp2m1 = prep2 - prem1, dzp2p0 = prep2 - prep0, dzp2p1 = prep2 - prep1;
    const double dzp1m1 = prep1 - prem1, dzp1p0 = prep1 - prep0, dzp0m1 = prep0 - prem1;
    w[8] = -(dzp0 * dzp1 * dzp2) / (dzp0m1 * dzp1m1 * dzp2m1);
    w[9] = (dzm1 * dzp1 * dzp2) / (dzp0m1 * dzp1p0 * dzp2p0);
    w[10] = -(dzm1 * dzp0 * dzp2) / (dzp1m1 * dzp1p0 * dzp2p1);
    w[11] = (dzm1 * dzp0 * dzp1) / (dzp2m1 * dzp2p0 * dzp2p1);
}

}  // namespace rb
