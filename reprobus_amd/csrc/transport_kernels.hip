// Semi-Lagrangian transport interpolation on an nlon x nlat x niv grid (the model:
// 180 x 91 x 137): the cell indices and Lagrange weights of the departure points
// (posx/posy/posz, poidx/poidy/poidz, F:7371-7647, as backtra calls them at F:6444-6451)
// and the tricubic interpolation with its limiter (internew, F:8339-8618) of every tracer.
//
// One lane per grid point.  The weights kernel reproduces the reference's arithmetic
// operation for operation (FP64, built with -ffp-contract=off).  The interpolation kernel
// takes a tile of points per workgroup and, per tracer, copies the part of the field that
// the tile's 4 x 4 x 4 stencils cover into LDS, each value once; every lane reads its
// stencil from there and makes one coalesced store.  The reference's extended grid
// x(0:nlon+2, 0:nlat+1, 0:niv+1) (F:8374-8413) is not built: each stencil index is mapped
// to the field instead.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "chem_device.h"
#include "transport_device.h"

namespace rb {

// Cell of departure point p and its 12 weights (field-major: w[m * n + p], m = alfxm1
// alfxp0 alfxp1 alfxp2 alfym1 .. alfzp2; idx[m * n + p], m = ig0 kg0 lg0, 1-based).
__global__ void k_transport_weights(TransportGrid t, const double *__restrict__ aaa,
                                    const double *__restrict__ bbb, const double *__restrict__ x0,
                                    const double *__restrict__ y0, const double *__restrict__ z0,
                                    const double *__restrict__ p0, int64_t n,
                                    int32_t *__restrict__ idx, double *__restrict__ w,
                                    unsigned long long *__restrict__ bad)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int c[3];
    double wt[12];
    cell_weights(t, aaa, bbb, x0[p], y0[p], z0[p], p0[p], true, c, wt);
#pragma unroll
    for (int m = 0; m < 3; ++m) idx[m * n + p] = c[m];
#pragma unroll
    for (int m = 0; m < 12; ++m) w[m * n + p] = wt[m];
    if (!(c[0] >= 1 && c[0] <= t.nlon && c[1] >= 1 && c[1] <= t.nlat)) atomicAdd(bad, 1ull);
}

// Two doubles at an 8-byte aligned address: one global_load_dwordx4 (multi-dword global
// accesses need dword alignment only).
typedef double double2_a8 __attribute__((ext_vector_type(2), aligned(8)));

// The tricubic interpolation with the limiter of one point (cell i, k, l; weights wx, wy,
// wz) for tracers tr0 .. tr1 - 1, its stencil read from the fields: the extended grid of
// internew (F:8374-8413: periodic longitudes, the polar rows mirrored through the pole,
// x(i, 0) = x(ip(i), 2) and x(i, nlat+1) = x(ip(i), nlat-1) with ip(i) = mod(i + nlon/2 - 1,
// nlon) + 1 (F:6089-6091), levels clamped to 1..niv) as an index map, not a copy.  The
// level loop is not unrolled (126 VGPRs), and a stencil row of four longitudes that needs
// neither the periodic wrap nor the polar mirror is read as two 16-byte loads.
__device__ __forceinline__ void interp_point(const TransportGrid &t, int i, int k, int l,
                                             const double (&wx)[4], const double (&wy)[4],
                                             const double (&wz)[4], const double *__restrict__ q_in,
                                             int64_t sin_, double *__restrict__ q_out, int64_t sout,
                                             int tr0, int tr1, int64_t p)
{
    // kg0 = nlat: a point on the south-pole row.  Its stencil reaches one row past the
    // reference's extended grid (x(:, 0:nlat+1)), with weight 0 when the point is on the
    // row; that row is read as row nlat + 1 here.
    int32_t po[4][4];   // [lat offset][lon offset] -> offset in a level plane
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int kk = k + b - 1 > t.nlat + 1 ? t.nlat + 1 : k + b - 1;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            int ii = i + a - 1;
            if (ii == 0) ii = t.nlon;
            else if (ii > t.nlon) ii -= t.nlon;
            int is = ii, ks = kk;
            if (kk == 0 || kk == t.nlat + 1) {
                is = (ii + t.nlon / 2 - 1) % t.nlon + 1;
                ks = kk == 0 ? 2 : t.nlat - 1;
            }
            po[b][a] = (ks - 1) * t.nlon + (is - 1);
        }
    }
    // rows i-1 .. i+2 of lat rows k-1 .. k+2 all inside the field: po[b][a] = po[b][0] + a
    const bool rows = i >= 2 && i + 2 <= t.nlon && k >= 2 && k + 2 <= t.nlat;
    int32_t lo[4];
    const int32_t plane = t.nlon * t.nlat;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int lc = l + c - 1;
        lo[c] = ((lc < 1 ? 1 : (lc > t.niv ? t.niv : lc)) - 1) * plane;
    }
    for (int tr = tr0; tr < tr1; ++tr) {
        const double *q = q_in + (int64_t)tr * sin_;
        double z[4], in[2][2][2];   // in: [level lp0/lp1][lat kp0/kp1][lon ip0/ip1]
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            double xv[4][4];
            if (rows) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const double *r = q + lo[c] + po[b][0];
                    const double2_a8 u0 = *(const double2_a8 *)r;
                    const double2_a8 u1 = *(const double2_a8 *)(r + 2);
                    xv[b][0] = u0.x;
                    xv[b][1] = u0.y;
                    xv[b][2] = u1.x;
                    xv[b][3] = u1.y;
                }
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b)
#pragma unroll
                    for (int a = 0; a < 4; ++a) xv[b][a] = q[lo[c] + po[b][a]];
            }
            if (c == 1 || c == 2)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int a = 0; a < 2; ++a) in[c - 1][b][a] = xv[b + 1][a + 1];
            // F:8515-8555 along x, F:8589-8599 along y
            double y[4];
#pragma unroll
            for (int b = 0; b < 4; ++b)
                y[b] = wx[0] * xv[b][0] + wx[1] * xv[b][1] + wx[2] * xv[b][2] + wx[3] * xv[b][3];
            z[c] = wy[0] * y[0] + wy[1] * y[1] + wy[2] * y[2] + wy[3] * y[3];
        }
        // F:8603-8606 along z; the limiter (F:8560-8585, F:8611-8613): the range of the 8
        // nearest values
        const double v = wz[0] * z[0] + wz[1] * z[1] + wz[2] * z[2] + wz[3] * z[3];
        const double rmx = fmax(fmax(fmax(in[0][0][0], in[0][1][0]), fmax(in[0][0][1], in[0][1][1])),
                                fmax(fmax(in[1][0][0], in[1][1][0]), fmax(in[1][0][1], in[1][1][1])));
        const double rmn = fmin(fmin(fmin(in[0][0][0], in[0][1][0]), fmin(in[0][0][1], in[0][1][1])),
                                fmin(fmin(in[1][0][0], in[1][1][0]), fmin(in[1][0][1], in[1][1][1])));
        q_out[(int64_t)tr * sout + p] = fmax(fmin(v, rmx), rmn);
    }
}

// ntr fields (q_in + t * sin, t < ntr) interpolated at the departure points into
// q_out + t * sout; a point whose indices fall outside the grid is counted in *bad and not
// written.  A workgroup takes a tile of TX x TY points of one level, with the fields read
// once per tile into LDS.  Per tile, the window of the extended grid that the tile's
// stencils cover -- longitudes unwrapped around the tile (a departure cell across the
// periodic boundary counts from the tile's side), extended rows 0 .. nlat + 2 and levels
// 0 .. niv + 2 before the index map -- is found by a block reduction.  Each tracer's window
// is copied into LDS with coalesced loads (each value once, through interp_point's index
// map), and each lane reads its 64 stencil values from there.  The copy of tracer tr + 1 is
// in flight while tracer tr is interpolated (two buffers, one barrier per tracer).
//
// If the window of every stencil exceeds WMAX values, the window covers only the points
// whose departure cell lies within RX, RY, RZ cells of the arrival point (it then fits by
// construction), and the others -- next to the poles, where a cell spans few kilometres of
// longitude -- gather their stencils from the fields (interp_point).  The arithmetic is
// interp_point's in every case.
//
// Tiles go level-fastest within each XCD's contiguous range (workgroup j takes tile
// (j mod 8) * per + j / 8), so that the tiles above and below a tile, which copy most of
// the same rows, run on the same L2 close in time.
//
// The copies wait on memory latency, so occupancy decides the time: WMAX sets the LDS per
// workgroup (two buffers) and the register bound three workgroups per CU, twelve waves
// (16 x 16 tiles, WMAX 3200: 1.96 ms for 43 tracers on 180 x 91 x 137 points, against 2.30
// with two workgroups per CU and 3.04 with every stencil gathered from the fields).
template <int TX, int TY, int WMAX>
__global__ void __launch_bounds__(TX * TY, 3) k_transport_interp_tiled(
    TransportGrid t, const int32_t *__restrict__ idx, const double *__restrict__ w,
    const double *__restrict__ q_in, int64_t sin_, double *__restrict__ q_out, int64_t sout,
    int ntr, int64_t n, int group, int ntx, int ntiles, unsigned long long *__restrict__ bad)
{
    constexpr int NT = TX * TY;
    constexpr int EPT = (WMAX + NT - 1) / NT;   // window values per lane
    constexpr int RX = 2, RY = 2, RZ = 1;
    static_assert((TX + 2 * RX + 3) * (TY + 2 * RY + 3) * (2 * RZ + 4) <= WMAX,
                  "the window of the near points fits");
    __shared__ double s_win[2][WMAX];
    __shared__ int s_bnd[6];                    // x lo, x hi, y lo, y hi, z lo, z hi
    const int per = (ntiles + 7) / 8;
    const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (tile >= ntiles) return;                 // whole workgroup: no barrier is skipped
    const int l = tile % t.niv + 1;
    const int col = tile / t.niv;
    const int i = (col % ntx) * TX + (int)(threadIdx.x % TX) + 1;
    const int k = (col / ntx) * TY + (int)(threadIdx.x / TX) + 1;
    const int32_t plane = t.nlon * t.nlat;
    const bool valid = i <= t.nlon && k <= t.nlat;
    const int64_t p = (int64_t)(l - 1) * plane + (int64_t)(k - 1) * t.nlon + (i - 1);
    int ig = 0, kg = 0, lg = 0;
    bool good = false;
    if (valid) {
        ig = idx[p];
        kg = idx[n + p];
        lg = idx[2 * n + p];
        good = ig >= 1 && ig <= t.nlon && kg >= 1 && kg <= t.nlat && lg >= 1 && lg <= t.niv;
        if (!good && blockIdx.y == 0) atomicAdd(bad, 1ull);
    }
    // the departure longitude unwrapped to within half a circle of the arrival point
    int u = ig;
    if (2 * (ig - i) > t.nlon) u -= t.nlon;
    else if (2 * (ig - i) < -t.nlon) u += t.nlon;
    // the window of the stencils of the points with `in` set (uniform results)
    auto window = [&](bool in, int &x0, int &y0, int &z0, int &wxn, int &wyn, int &wzn) {
        __syncthreads();                        // s_bnd free
        if (threadIdx.x == 0) {
            s_bnd[0] = s_bnd[2] = s_bnd[4] = 0x7fffffff;
            s_bnd[1] = s_bnd[3] = s_bnd[5] = -0x7fffffff;
        }
        __syncthreads();
        if (in) {
            atomicMin(&s_bnd[0], u - 1);
            atomicMax(&s_bnd[1], u + 2);
            atomicMin(&s_bnd[2], kg - 1);
            atomicMax(&s_bnd[3], kg + 2);
            atomicMin(&s_bnd[4], lg - 1);
            atomicMax(&s_bnd[5], lg + 2);
        }
        __syncthreads();
        // no point contributed: the bounds are still their initial values (and their
        // difference would overflow)
        const bool any = s_bnd[1] >= s_bnd[0];
        x0 = s_bnd[0];
        y0 = s_bnd[2];
        z0 = s_bnd[4];
        wxn = any ? s_bnd[1] - x0 + 1 : 0;
        wyn = any ? s_bnd[3] - y0 + 1 : 0;
        wzn = any ? s_bnd[5] - z0 + 1 : 0;
    };
    int x0, y0, z0, wxn, wyn, wzn;
    window(good, x0, y0, z0, wxn, wyn, wzn);
    if (wxn <= 0) return;                       // no point of the tile inside the grid
    bool in = good;
    if ((int64_t)wxn * wyn * wzn > WMAX) {
        in = good && abs(u - i) <= RX && abs(kg - k) <= RY && abs(lg - l) <= RZ;
        window(in, x0, y0, z0, wxn, wyn, wzn);
    }
    const bool far = good && !in;
    const int tr0 = (int)blockIdx.y * group;
    const int tr1 = tr0 + group < ntr ? tr0 + group : ntr;
    double wx[4], wy[4], wz[4];
    if (good) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            wx[m] = w[m * n + p];
            wy[m] = w[(4 + m) * n + p];
            wz[m] = w[(8 + m) * n + p];
        }
    }
    if (wxn <= 0) {
        // every point of the tile is far: the whole workgroup gathers, no barrier follows
        if (far) interp_point(t, ig, kg, lg, wx, wy, wz, q_in, sin_, q_out, sout, tr0, tr1, p);
        return;
    }
    const int wxy = wxn * wyn, wn = wxy * wzn;
    // the field offset of each window value this lane copies (-1: none), through the index
    // map of interp_point
    int32_t off[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int e = (int)threadIdx.x + m * NT;
        off[m] = -1;
        if (e < wn) {
            const int ez = e / wxy, er = e - ez * wxy, ey = er / wxn, ex = er - ey * wxn;
            const int uu = x0 + ex, row = y0 + ey, lev = z0 + ez;
            const int ii = ((uu - 1) % t.nlon + t.nlon) % t.nlon + 1;
            const int kk = row > t.nlat + 1 ? t.nlat + 1 : row;
            int is = ii, ks = kk;
            if (kk == 0 || kk == t.nlat + 1) {
                is = (ii + t.nlon / 2 - 1) % t.nlon + 1;
                ks = kk == 0 ? 2 : t.nlat - 1;
            }
            const int lc = lev < 1 ? 1 : (lev > t.niv ? t.niv : lev);
            off[m] = (lc - 1) * plane + (ks - 1) * t.nlon + (is - 1);
        }
    }
    // the lane's stencil origin in the window
    const int sb = (lg - 1 - z0) * wxy + (kg - 1 - y0) * wxn + (u - 1 - x0);
#pragma unroll
    for (int m = 0; m < EPT; ++m)
        if (off[m] >= 0) s_win[0][threadIdx.x + m * NT] = q_in[(int64_t)tr0 * sin_ + off[m]];
    __syncthreads();
    for (int tr = tr0; tr < tr1; ++tr) {
        const int cur = (tr - tr0) & 1;
        const bool more = tr + 1 < tr1;
        double nx[EPT];
        if (more) {
            const double *q = q_in + (int64_t)(tr + 1) * sin_;
#pragma unroll
            for (int m = 0; m < EPT; ++m) nx[m] = off[m] >= 0 ? q[off[m]] : 0.;
        }
        if (in) {
            const double *win = s_win[cur] + sb;
            double z[4], inn[2][2][2];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double xv[4][4];
#pragma unroll
                for (int b = 0; b < 4; ++b)
#pragma unroll
                    for (int a = 0; a < 4; ++a) xv[b][a] = win[c * wxy + b * wxn + a];
                if (c == 1 || c == 2)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int a = 0; a < 2; ++a) inn[c - 1][b][a] = xv[b + 1][a + 1];
                double y[4];
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    y[b] = wx[0] * xv[b][0] + wx[1] * xv[b][1] + wx[2] * xv[b][2] + wx[3] * xv[b][3];
                z[c] = wy[0] * y[0] + wy[1] * y[1] + wy[2] * y[2] + wy[3] * y[3];
            }
            const double v = wz[0] * z[0] + wz[1] * z[1] + wz[2] * z[2] + wz[3] * z[3];
            const double rmx = fmax(fmax(fmax(inn[0][0][0], inn[0][1][0]), fmax(inn[0][0][1], inn[0][1][1])),
                                    fmax(fmax(inn[1][0][0], inn[1][1][0]), fmax(inn[1][0][1], inn[1][1][1])));
            const double rmn = fmin(fmin(fmin(inn[0][0][0], inn[0][1][0]), fmin(inn[0][0][1], inn[0][1][1])),
                                    fmin(fmin(inn[1][0][0], inn[1][1][0]), fmin(inn[1][0][1], inn[1][1][1])));
            q_out[(int64_t)tr * sout + p] = fmax(fmin(v, rmx), rmn);
        }
        if (more) {
#pragma unroll
            for (int m = 0; m < EPT; ++m)
                if (off[m] >= 0) s_win[cur ^ 1][threadIdx.x + m * NT] = nx[m];
        }
        __syncthreads();
    }
    // the far points, after the last barrier
    if (far) interp_point(t, ig, kg, lg, wx, wy, wz, q_in, sin_, q_out, sout, tr0, tr1, p);
}

TransportGrid transport_grid(int nlon, int nlat, int niv)
{
    // initransp F:6028-6034, 6041, 6051, 6075-6085 (host libm, as the model's)
    TransportGrid t{};
    const double xpi = 2. * asin(1.);
    const double xpi2 = 2. * xpi;
    t.xpih = xpi / 2.;
    t.ddlat = xpi / (double)(nlat - 1);
    t.ddlon = xpi2 / (double)nlon;
    const double coefx = 1. / (t.ddlon * t.ddlon * t.ddlon);
    t.xcoef[0] = -coefx / 6.;
    t.xcoef[1] = coefx / 2.;
    t.xcoef[2] = -coefx / 2.;
    t.xcoef[3] = coefx / 6.;
    const double coefy = 1. / (t.ddlat * t.ddlat * t.ddlat);
    t.ycoef[0] = -coefy / 6.;
    t.ycoef[1] = coefy / 2.;
    t.ycoef[2] = -coefy / 2.;
    t.ycoef[3] = coefy / 6.;
    t.nlon = nlon;
    t.nlat = nlat;
    t.niv = niv;
    return t;
}

hipError_t launch_transport_weights(const TransportGrid &t, const double *aaa, const double *bbb,
                                    const double *x0, const double *y0, const double *z0,
                                    const double *p0, int32_t *idx, double *w,
                                    unsigned long long *bad, hipStream_t s)
{
    const int64_t n = (int64_t)t.nlon * t.nlat * t.niv;
    hipLaunchKernelGGL(k_transport_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t,
                       aaa, bbb, x0, y0, z0, p0, n, idx, w, bad);
    return hipGetLastError();
}

hipError_t launch_transport_interp(const TransportGrid &t, const int32_t *idx, const double *w,
                                   const double *q_in, int64_t sin_, double *q_out, int64_t sout,
                                   int ntr, unsigned long long *bad, hipStream_t s, int group)
{
    // field offsets are 32-bit: grids up to 2^31 - 1 points (the entry points check it)
    const int64_t n = (int64_t)t.nlon * t.nlat * t.niv;
    if (n > INT32_MAX) return hipErrorInvalidValue;
    // 16 x 16 tiles of one level, every tracer in one workgroup (the tile's set-up --
    // window, index map, weights -- done once)
    constexpr int TX = 16, TY = 16;
    if (group < 1) group = ntr;
    const unsigned ng = (unsigned)((ntr + group - 1) / group);
    const int ntx = (t.nlon + TX - 1) / TX, nty = (t.nlat + TY - 1) / TY;
    const int64_t ntiles = (int64_t)ntx * nty * t.niv;
    const unsigned nwg = (unsigned)(((ntiles + 7) / 8) * 8);
    hipLaunchKernelGGL((k_transport_interp_tiled<TX, TY, TRANSPORT_WINDOW>), dim3(nwg, ng),
                       dim3(TX * TY), 0, s, t, idx, w, q_in, sin_, q_out, sout, ntr, n, group,
                       ntx, (int)ntiles, bad);
    return hipGetLastError();
}

}  // namespace rb
