// Back-trajectories of the semi-Lagrangian transport (backtra,
// src/reprobus_l137_20230723.f:6156-6470) on an nlon x nlat x niv grid (the model: 180 x 91 x
// 137): for every node, the point the air arriving there left dtdyn seconds earlier, in nrapp
// sub-steps of two-pass midpoint trajectories, with the positions and temperatures along the
// way and, at the end, the departure cell and Lagrange weights the interpolation reads.
//
// Per sub-step two launches, one lane per point each:
//   k_backtra_fields  the forcing interpolated in time (instant1/instant2, F:6473-6554) and
//                     the winds of inivit (F:6557-6722) -- scaled to radians and Pa per
//                     sub-step, and in the frame rotated by 90 degrees about the y axis that
//                     the trajectories use poleward of 45 degrees -- on the base grid;
//   k_backtra_step    the point's sub-step: first guess, midpoint, departure point, cell,
//                     weights and temperature (F:6285-6469).
// The reference's extended grids (x(0:nlon+2, 0:nlat+1, 0:niv+1) of inivit, interp and
// intert: longitudes periodic, polar rows mirrored through the pole, with the sign of the
// winds flipped there, levels clamped) are index maps onto the base fields, not copies.
//
// The arithmetic follows the reference operation for operation (FP64, -ffp-contract=off);
// sin, cos, asin, acos and exp are the correctly rounded ones of crmath.h, which agree with
// the host libm the reference links wherever that library rounds correctly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "chem_device.h"
#include "crmath.h"
#include "transport_device.h"

namespace rb {

namespace {

constexpr int F_U1S = 0, F_V1S = 1, F_W1S = 2, F_U1M = 3, F_V1M = 4, F_T = 5, F_PS = 6;

// value of the extended field at (i, k, l) (i in 0..nlon+2, k in 0..nlat+1, l in 0..niv+1):
// the base field f at the mapped node, times sg on the mirrored polar rows (-1 for the
// non-rotated winds, F:6626-6632).  Indices are clamped into the extended grid, so that a
// point whose cell fell outside it (counted by the caller) still reads inside the fields;
// a point on the south-pole row (kg0 = nlat) reaches one row past the reference's extended
// grid, with weight 0 when it is on the row, and reads row nlat + 1 there.
__device__ __forceinline__ double ext(const double *__restrict__ f, const TransportGrid &t, int i,
                                     int k, int l, double sg)
{
    i = i < 0 ? 0 : (i > t.nlon + 2 ? t.nlon + 2 : i);
    k = k < 0 ? 0 : (k > t.nlat + 1 ? t.nlat + 1 : k);
    int ii = i == 0 ? t.nlon : (i > t.nlon ? i - t.nlon : i);
    int kk = k;
    bool mirror = false;
    if (k == 0 || k == t.nlat + 1) {
        ii = (ii + t.nlon / 2 - 1) % t.nlon + 1;         // ip(i), F:6089-6091
        kk = k == 0 ? 2 : t.nlat - 1;
        mirror = true;
    }
    const int ll = l < 1 ? 1 : (l > t.niv ? t.niv : l);
    const double v = f[((int64_t)(ll - 1) * t.nlat + (kk - 1)) * t.nlon + (ii - 1)];
    return mirror ? sg * v : v;
}

// the 2-d extended surface pressure of interp (F:7029-7050)
__device__ __forceinline__ double ext2(const double *__restrict__ f, const TransportGrid &t, int i,
                                      int k)
{
    return ext(f, t, i, k, 1, 1.);
}

// tricubic Lagrange sum over the extended field (interv F:6784-6978 without the limiter,
// intert F:7181-7250 with it at every stage)
template <bool LIMIT>
__device__ double tricubic(const double *__restrict__ f, const TransportGrid &t, double sg,
                           const int c[3], const double w[12])
{
    double z[4];
#pragma unroll
    for (int ls = 0; ls < 4; ++ls) {
        double y[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int k = c[1] + ks - 1, l = c[2] + ls - 1;
            const double xm1 = ext(f, t, c[0] - 1, k, l, sg), xp0 = ext(f, t, c[0], k, l, sg);
            const double xp1 = ext(f, t, c[0] + 1, k, l, sg), xp2 = ext(f, t, c[0] + 2, k, l, sg);
            double v = w[0] * xm1 + w[1] * xp0 + w[2] * xp1 + w[3] * xp2;
            if (LIMIT) v = fmax(fmin(v, fmax(xp0, xp1)), fmin(xp0, xp1));
            y[ks] = v;
        }
        double v = w[4] * y[0] + w[5] * y[1] + w[6] * y[2] + w[7] * y[3];
        if (LIMIT) v = fmax(fmin(v, fmax(y[1], y[2])), fmin(y[1], y[2]));
        z[ls] = v;
    }
    double v = w[8] * z[0] + w[9] * z[1] + w[10] * z[2] + w[11] * z[3];
    if (LIMIT) v = fmax(fmin(v, fmax(z[1], z[2])), fmin(z[1], z[2]));
    return v;
}

// interp (F:6982-7108): the surface pressure at (xx, yy), bicubic with the limiter
__device__ double interp_ps(const double *__restrict__ ps, const TransportGrid &t, double xx,
                            double yy, bool &outside)
{
    int c[3];
    double w[12];
    cell_weights(t, nullptr, nullptr, xx, yy, 0., 0., false, c, w);
    outside |= !(c[0] >= 1 && c[0] <= t.nlon && c[1] >= 1 && c[1] <= t.nlat);
    double y[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int k = c[1] + ks - 1;
        const double xm1 = ext2(ps, t, c[0] - 1, k), xp0 = ext2(ps, t, c[0], k);
        const double xp1 = ext2(ps, t, c[0] + 1, k), xp2 = ext2(ps, t, c[0] + 2, k);
        const double v = w[0] * xm1 + w[1] * xp0 + w[2] * xp1 + w[3] * xp2;
        y[ks] = fmax(fmin(v, fmax(xp0, xp1)), fmin(xp0, xp1));
    }
    const double v = w[4] * y[0] + w[5] * y[1] + w[6] * y[2] + w[7] * y[3];
    return fmax(fmin(v, fmax(y[1], y[2])), fmin(y[1], y[2]));
}

// corlat (F:7283-7329): longitudes into [0, 2 pi), latitudes past a pole reflected
__device__ __forceinline__ void corlat(const BacktraArgs &a, double &x, double &y)
{
    x = x - a.xpi2 * (double)(int)(x / a.xpi2);
    if (x < 0.) x = x + a.xpi2;
    y = y - a.xpi2 * (double)(int)(y / a.xpi2);
    const double temp0 = a.t.xpih - fabs(y);
    const double temp1 = x - a.xpi < 0. ? x + a.xpi : x - a.xpi;
    const double temp2 = copysign(a.xpi, y) - y;
    if (temp0 < 0.) {
        x = temp1;
        y = temp2;
    }
}

// corniv (F:7333-7367): the pressure kept between the top and bottom levels of the column
__device__ __forceinline__ double corniv(const BacktraArgs &a, double z, double p)
{
    const double pmin = (a.aaa[0] + a.bbb[0] * p) * 100.;
    const double pmax = (a.aaa[a.t.niv - 1] + a.bbb[a.t.niv - 1] * p) * 100.;
    return fmin(pmax, fmax(pmin, z));
}

// the three winds at (xx, yy, zz) in a column of surface pressure pp (interv): the same
// sums as three tricubic<false> calls, with each stencil node's index worked out once
__device__ void winds_at(const BacktraArgs &a, const double *ue, const double *ve, double sg,
                         double xx, double yy, double zz, double pp, double &u2, double &v2,
                         double &w2, bool &outside)
{
    const TransportGrid &t = a.t;
    int c[3];
    double w[12];
    cell_weights(t, a.aaa, a.bbb, xx, yy, zz, pp, true, c, w);
    outside |= !(c[0] >= 1 && c[0] <= t.nlon && c[1] >= 1 && c[1] <= t.nlat);
    const double *we = a.f + F_W1S * a.n;
    // the stencil's 4 x 4 columns of a level: offsets in a level plane and the mirror flag
    int32_t po[4][4];
    bool mir[4];
    const int i0 = c[0] < 1 ? 1 : (c[0] > t.nlon ? t.nlon : c[0]);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int k = c[1] + b - 1;
        k = k < 0 ? 0 : (k > t.nlat + 1 ? t.nlat + 1 : k);
        mir[b] = k == 0 || k == t.nlat + 1;
        const int kk = k == 0 ? 2 : (k == t.nlat + 1 ? t.nlat - 1 : k);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = i0 + e - 1;
            int ii = i == 0 ? t.nlon : (i > t.nlon ? i - t.nlon : i);
            if (mir[b]) ii = (ii + t.nlon / 2 - 1) % t.nlon + 1;     // ip(i), F:6089-6091
            po[b][e] = (kk - 1) * t.nlon + (ii - 1);
        }
    }
    const int64_t plane = (int64_t)t.nlon * t.nlat;
    double zu[4], zv[4], zw[4];
#pragma unroll
    for (int ls = 0; ls < 4; ++ls) {
        const int l = c[2] + ls - 1;
        const int64_t lo = (int64_t)((l < 1 ? 1 : (l > t.niv ? t.niv : l)) - 1) * plane;
        double yu[4], yv[4], yw[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double xu[4], xv[4], xw[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t o = lo + po[b][e];
                xu[e] = mir[b] ? sg * ue[o] : ue[o];
                xv[e] = mir[b] ? sg * ve[o] : ve[o];
                xw[e] = we[o];
            }
            yu[b] = w[0] * xu[0] + w[1] * xu[1] + w[2] * xu[2] + w[3] * xu[3];
            yv[b] = w[0] * xv[0] + w[1] * xv[1] + w[2] * xv[2] + w[3] * xv[3];
            yw[b] = w[0] * xw[0] + w[1] * xw[1] + w[2] * xw[2] + w[3] * xw[3];
        }
        zu[ls] = w[4] * yu[0] + w[5] * yu[1] + w[6] * yu[2] + w[7] * yu[3];
        zv[ls] = w[4] * yv[0] + w[5] * yv[1] + w[6] * yv[2] + w[7] * yv[3];
        zw[ls] = w[4] * yw[0] + w[5] * yw[1] + w[6] * yw[2] + w[7] * yw[3];
    }
    u2 = w[8] * zu[0] + w[9] * zu[1] + w[10] * zu[2] + w[11] * zu[3];
    v2 = w[8] * zv[0] + w[9] * zv[1] + w[10] * zv[2] + w[11] * zv[3];
    w2 = w[8] * zw[0] + w[9] * zw[1] + w[10] * zw[2] + w[11] * zw[3];
}

__global__ void __launch_bounds__(256) k_backtra_fields(BacktraArgs a)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    const int nlon = a.t.nlon, nlat = a.t.nlat;
    const int i = (int)(q % nlon) + 1, k = (int)((q / nlon) % nlat) + 1;
    // instant1 (F:6500-6507)
    const double uj = a.cp1 * a.p1.u[q] + a.cm1 * a.m1.u[q];
    const double vj = a.cp1 * a.p1.v[q] + a.cm1 * a.m1.v[q];
    const double tj = a.cp1 * a.p1.t[q] + a.cm1 * a.m1.t[q];
    const double wj = a.cp1 * a.p1.w[q] + a.cm1 * a.m1.w[q];
    // inivit F:6598-6604
    const double u1s = uj * a.coefh, v1s = vj * a.coefh, w1s = wj * a.coefv;
    double u1m = 0., v1m = 0.;
    // F:6606-6625: the rotated frame, not set on the row where clat = 1
    if (k != (nlat - 1) / 2 + 1) {
        const int64_t r = (int64_t)(k - 1) * nlon + (i - 1);
        v1m = (u1s * a.tb.slon[i - 1] + v1s * a.tb.clon[i - 1] * a.tb.slat[k - 1]) / a.tb.clatr[r];
        u1m = (u1s * a.tb.clon[i - 1] - v1s * a.tb.slon[i - 1] * a.tb.slat[k - 1] +
               v1m * a.tb.slonr[r] * a.tb.slatr[r]) / a.tb.clonr[r];
    }
    a.f[F_U1S * a.n + q] = u1s;
    a.f[F_V1S * a.n + q] = v1s;
    a.f[F_W1S * a.n + q] = w1s;
    a.f[F_U1M * a.n + q] = u1m;
    a.f[F_V1M * a.n + q] = v1m;
    a.f[F_T * a.n + q] = tj;
    // instant2 (F:6542-6549): the surface pressure in hPa
    const int64_t plane = (int64_t)nlon * nlat;
    if (q < plane)
        a.f[F_PS * a.n + q] =
            .01 * (a.cp1 * crm::exp_cr(a.p1.lnps[q]) + a.cm1 * crm::exp_cr(a.m1.lnps[q]));
}

// Workgroups go round-robin to the 8 XCDs: workgroup j takes tile (j mod 8) * per + j / 8,
// so each XCD works on one contiguous eighth of the grid and the fields its stencils share
// stay in its L2.
__global__ void __launch_bounds__(256) k_backtra_step(BacktraArgs a, int ntiles)
{
    const int per = (ntiles + 7) / 8;
    const int64_t tile = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t q = tile * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    const int nlon = a.t.nlon, nlat = a.t.nlat;
    const int64_t plane = (int64_t)nlon * nlat;
    const int i = (int)(q % nlon) + 1, k = (int)((q / nlon) % nlat) + 1, l = (int)(q / plane) + 1;
    const double *ps = a.f + F_PS * a.n;
    const double *te = a.f + F_T * a.n;
    double x0, y0, z0, p0;
    if (a.iter == 1) {                                          // F:6262-6276
        const double pj = ps[(int64_t)(k - 1) * nlon + (i - 1)];
        x0 = a.tb.dlon[i];
        y0 = a.tb.dlat[k];
        z0 = (a.aaa[l - 1] + a.bbb[l - 1] * pj) * 100.;
        p0 = pj;
        if (a.traj[0]) {
            const int64_t o = (int64_t)(a.nrapp - 1) * a.n + q;
            a.traj[0][o] = x0 * 180. / a.xpi;
            a.traj[1][o] = y0 * 180. / a.xpi;
            a.traj[2][o] = z0 / 100.;
            a.traj[3][o] = te[q];
        }
    } else {
        x0 = a.x0[q];
        y0 = a.y0[q];
        z0 = a.z0[q];
        p0 = a.p0[q];
    }
    bool outside = false;
    // first guess of the midpoint (F:6285-6296) and, poleward of 45 degrees, the point in
    // the rotated frame (F:6298-6315)
    double xm = x0, ym = y0, zm = z0, pm = p0;
    const bool ind = fabs(ym) >= a.dlatlim;
    double x0r = 0., y0r = 0., sy0r = 0., cy0r = 0.;
    if (ind) {
        double sx, cx, sy, cy;
        crm::sincos_cr(x0, sx, cx);
        crm::sincos_cr(y0, sy, cy);
        sy0r = -cx * cy;
        y0r = crm::asin_cr(sy0r);
        cy0r = sqrt(1. - sy0r * sy0r);
        const double cx0r = cy0r == 0. ? 0. : fmax(-1., fmin(1., sy / cy0r));
        x0r = copysign(crm::acos_cr(cx0r), sx);
    }
    const double *ue = a.f + (ind ? F_U1M : F_U1S) * a.n;
    const double *ve = a.f + (ind ? F_V1M : F_V1S) * a.n;
    const double sg = ind ? 1. : -1.;
    double u2, v2, w2;
    if (a.iter == 1) {                                          // F:6329-6343
        u2 = ue[q];
        v2 = ve[q];
        w2 = a.f[F_W1S * a.n + q];
    } else {                                                    // interv, F:6349
        winds_at(a, ue, ve, sg, xm, ym, zm, pm, u2, v2, w2, outside);
    }
    // the midpoint, F:6356-6384
    if (!ind) {
        xm = x0 - 0.5 * u2 / crm::cos_cr(ym);
        ym = y0 - 0.5 * v2;
    } else {
        // symr = -cos(xm) cos(ym) at the first guess (xm, ym) = (x0, y0): sy0r, and its
        // cymr is cy0r
        const double xmr = x0r - 0.5 * u2 / cy0r;
        const double ymr = y0r - 0.5 * v2;
        double sxr, cxr, syr, cyr;
        crm::sincos_cr(xmr, sxr, cxr);
        crm::sincos_cr(ymr, syr, cyr);
        const double sym = cxr * cyr;
        ym = crm::asin_cr(sym);
        const double cym = sqrt(1. - sym * sym);
        const double cxm = cym == 0. ? 0. : fmax(-1., fmin(1., -syr / cym));
        xm = copysign(crm::acos_cr(cxm), sxr);
    }
    zm = z0 - 0.5 * w2;
    corlat(a, xm, ym);                                          // F:6388-6394
    pm = interp_ps(ps, a.t, xm, ym, outside);
    zm = corniv(a, zm, pm);
    // the wind at the midpoint (F:6402) and the departure point (F:6404-6430)
    winds_at(a, ue, ve, sg, xm, ym, zm, pm, u2, v2, w2, outside);
    double xx = x0, yy = y0;
    if (!ind) {
        xx = xx - u2 / crm::cos_cr(ym);
        yy = yy - v2;
    } else {
        const double symr = -crm::cos_cr(xm) * crm::cos_cr(ym);
        const double cymr = sqrt(1. - symr * symr);
        x0r = x0r - u2 / cymr;
        y0r = y0r - v2;
        double sxr, cxr, syr, cyr;
        crm::sincos_cr(x0r, sxr, cxr);
        crm::sincos_cr(y0r, syr, cyr);
        const double sy0 = cxr * cyr;
        yy = crm::asin_cr(sy0);
        const double cy0 = sqrt(1. - sy0 * sy0);
        const double cx0 = cy0 == 0. ? 0. : fmax(-1., fmin(1., -syr / cy0));
        xx = copysign(crm::acos_cr(cx0), sxr);
    }
    double zz = z0 - w2;
    corlat(a, xx, yy);                                          // F:6434-6440
    const double pp = interp_ps(ps, a.t, xx, yy, outside);
    zz = corniv(a, zz, pp);
    a.x0[q] = xx;
    a.y0[q] = yy;
    a.z0[q] = zz;
    a.p0[q] = pp;
    // the departure cell and weights (F:6444-6451), the temperature there (intert, F:6455)
    int c[3];
    double w[12];
    cell_weights(a.t, a.aaa, a.bbb, xx, yy, zz, pp, true, c, w);
    outside |= !(c[0] >= 1 && c[0] <= a.t.nlon && c[1] >= 1 && c[1] <= a.t.nlat);
#pragma unroll
    for (int m = 0; m < 3; ++m) a.idx[m * a.n + q] = c[m];
#pragma unroll
    for (int m = 0; m < 12; ++m) a.w[m * a.n + q] = w[m];
    if (a.traj[0] && a.iter < a.nrapp) {                        // F:6459-6469
        const double tt = tricubic<true>(te, a.t, 1., c, w);
        const int64_t o = (int64_t)(a.nrapp - a.iter - 1) * a.n + q;
        a.traj[0][o] = xx * 180. / a.xpi;
        a.traj[1][o] = yy * 180. / a.xpi;
        a.traj[2][o] = zz / 100.;
        a.traj[3][o] = tt;
    }
    if (outside) atomicAdd(a.bad, 1ull);
}

}  // namespace

size_t backtra_table_size(int nlon, int nlat)
{
    return (size_t)(nlon + 3) + (nlat + 2) + 2 * nlon + nlat + 4 * (size_t)nlon * nlat;
}

void backtra_tables(int nlon, int nlat, double *h)
{
    // initransp F:6036-6073 with the host libm, as the model evaluates it: the flang build
    // computes each node's sine and cosine with one sincos call
    double *dlon = h, *dlat = dlon + nlon + 3, *slon = dlat + nlat + 2, *clon = slon + nlon;
    double *slat = clon + nlon, *slonr = slat + nlat, *clonr = slonr + (size_t)nlon * nlat;
    double *slatr = clonr + (size_t)nlon * nlat, *clatr = slatr + (size_t)nlon * nlat;
    const double xpi = 2. * asin(1.), xpi2 = 2. * xpi, xpih = xpi / 2.;
    const double ddlat = xpi / (double)(nlat - 1), ddlon = xpi2 / (double)nlon;
    for (int k = 0; k <= nlat + 1; ++k) dlat[k] = xpih - ((double)k - 1.) * ddlat;
    for (int i = 0; i <= nlon + 2; ++i) dlon[i] = ddlon * ((double)i - 1.);
    for (size_t c = 0; c < 4 * (size_t)nlon * nlat; ++c) slonr[c] = 0.;
    for (int i = 1; i <= nlon; ++i) sincos(dlon[i], &slon[i - 1], &clon[i - 1]);
    for (int k = 1; k <= nlat; ++k) {
        double clat;
        sincos(dlat[k], &slat[k - 1], &clat);
        if (clat == 1.) continue;
        for (int i = 1; i <= nlon; ++i) {
            const size_t c = (size_t)(k - 1) * nlon + (i - 1);
            slatr[c] = -clon[i - 1] * clat;
            clatr[c] = sqrt(1. - slatr[c] * slatr[c]);
            clonr[c] = fmax(-1., fmin(slat[k - 1] / clatr[c], 1.));
            const double sl = sqrt(1. - clonr[c] * clonr[c]);
            slonr[c] = copysign(sl, slon[i - 1]);
        }
    }
}

BtTables backtra_table_ptrs(int nlon, int nlat, const double *d)
{
    BtTables t;
    t.dlon = d;
    t.dlat = t.dlon + nlon + 3;
    t.slon = t.dlat + nlat + 2;
    t.clon = t.slon + nlon;
    t.slat = t.clon + nlon;
    t.slonr = t.slat + nlat;
    t.clonr = t.slonr + (size_t)nlon * nlat;
    t.slatr = t.clonr + (size_t)nlon * nlat;
    t.clatr = t.slatr + (size_t)nlon * nlat;
    return t;
}

hipError_t launch_backtra_fields(const BacktraArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_backtra_fields, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_backtra_step(const BacktraArgs &a, hipStream_t s)
{
    const int ntiles = (int)((a.n + 255) / 256);
    const unsigned nwg = (unsigned)(((ntiles + 7) / 8) * 8);
    hipLaunchKernelGGL(k_backtra_step, dim3(nwg), dim3(256), 0, s, a, ntiles);
    return hipGetLastError();
}

}  // namespace rb
