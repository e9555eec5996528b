// REPROBUS chemistry step on gfx950: one box per lane, species-major SoA.
//
// Reference: src/reprobus_l137_20230723.f ("F") and src/jno.f90 ("J").
//   k_columns       : O3 column (total, F:5153-5186) and the O2 / NO column sums of
//                     jno (J:103-126) -- the only vertical coupling of the step; one
//                     lane per column walks its levels top-down.
//   k_setup         : the per-box set-up into a record: J-values (phot F:4996-5149, jno
//                     J:128-246), PSC/aerosol and heterogeneous rates (analytic
//                     F:2620-3361, hetero F:3365-4029, density F:4033-4073), thermal
//                     rates (F:920-1306), rain-out (F:1315-1338).
//   k_chem          : from the record, the fixed-sweep QSSA + linearised backward-Euler
//                     integration (F:1348-2378) and the write-back (F:2380-2410);
//                     optionally conserv + floor.
//   k_chem_conv     : the converged mode (sweeps to the fixed point, persistent, with a
//                     device work queue; optionally Anderson-accelerated).
//   k_conserv_floor : conserv (F:2526-2616) and the 1e-30 floor (F:431-440).
// Expressions keep the reference's evaluation order (the library is built with
// -ffp-contract=off) so the step reproduces the reference to a few ulps per box.
#include "chem_device.h"
#include "crmath.h"
#include "newton_gen.h"

#include <math.h>

namespace rb {

__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double cube(double x) { return x * x * x; }

// ------------------------------------------------------------------ rates
struct Rates {
    double a1, a1et, a2, a5, a6, a6b, a7, a17, a19, a26, a27, a36;
    double b3, b4, b6, b7, b9, b12, b22, b23, b24, b27, b28, b32, b38, b39;
    double c2, c4, c5, c7, c8, c9, c10, c14, c15, c17a, c17b, c18, c19;
    double d2, d3, d4, d5, d6, d7, d8, d9, d10, d11, d31, d32, d33, d34, d36, d37, d60, d61,
        d62, d94, d95, d97;
    double e2, e3, e4, e5a, e5b, e5c, e6, e7, e8, e10, e11, e13, e15, e16, e51, e54;
    double hk1, hk2, hk3, hk4, hk5;
};
// temperature-independent rates (F:964-972, 982-985, 1061-1068, 1082, 1163, 1196-1208,
// 1245, 1276-1281) and the chem3d scalars a3et, a23a/b/c
constexpr double A3ET = 1.2e-10, A23A = 7.2e-11, A23B = 6.9e-12, A23C = 1.6e-12;
constexpr double A30 = 1.8e-12, A31 = 1.8e-12, C1A = 1.31e-10, C1B = 0.35e-10, C1C = 0.09e-10,
                 C3 = 1.1e-10, C12 = 5.2e-12, D35 = 1.7e-13, D90 = 2.3e-10, D91 = 1.4e-10,
                 D92 = 2.0e-10, D93 = 3.3e-10, D96 = 1.0e-10, E9 = 1.5e-10, E50 = 1.8e-10,
                 E52 = 1.5e-10, E53 = 1.0e-10;

// JPL falloff, Fc = 0.6 (F:930-935 and its 8 siblings).
__device__ __forceinline__ double falloff(double a0, double a1, double n0, double n1,
                                          double lnt, double hnm, double xln6)
{
    const double k0 = a0 * exp(lnt * n0);
    const double ki = a1 * exp(lnt * n1);
    const double rate = (k0 * hnm) / (1. + k0 * hnm / ki);
    const double x = 1. / (1. + sq(log10((k0 * hnm) / ki)));
    return rate * exp(xln6 * x);
}

__device__ void thermal_rates(const Troe &K, double t, double hnm, double pm, Rates &r)
{
    const double xln6 = log(0.6);
    const double lnt = log(t);
    // one reciprocal for the Arrhenius arguments: c*rt is within an ulp of c/t and the
    // rates feed only well-conditioned parts of the step
    const double rt = 1. / t;
    r.a1 = falloff(K.a0hox, K.a1hox, -1.3, 0.2, lnt, hnm, xln6);
    r.a1et = 1.63e-10 * exp(60. * rt);
    r.a2 = 1.4e-10 * exp(-470. * rt);
    r.a5 = 1.8e-11 * exp(180. * rt);
    r.a6 = 1.7e-12 * exp(-940. * rt);
    r.a6b = 1.0e-14 * exp(-490. * rt);
    r.a19 = 2.8e-12 * exp(-1800. * rt);
    r.a27 = 1.5e-12 * exp(19. * rt) + 2.1e-33 * exp(920. * rt) * hnm;
    r.a36 = 1.5e-13 * (1. + .6 * pm / 1013.);

    r.b3 = 5.1e-12 * exp(210. * rt);
    r.b4 = 3.0e-12 * exp(-1500. * rt);
    r.b6 = 2.1e-11 * exp(100. * rt);
    r.b7 = 1.5e-11 * exp(-3600. * rt);
    r.b9 = 1.2e-13 * exp(-2450. * rt);
    r.b12 = falloff(K.a0nox1, K.a1nox1, -4.4, -0.7, lnt, hnm, xln6);
    r.b22 = falloff(K.a0nox2, K.a1nox2, -3.0, 0.0, lnt, hnm, xln6);
    r.b23 = falloff(K.a0nox3, K.a1nox3, -3.4, -1.1, lnt, hnm, xln6);
    r.b24 = r.b23 / (2.1e-27 * exp(10900. * rt) * hnm);          // via equilibrium constant
    {
        const double aux1 = 2.7e-17 * exp(2199. * rt);
        const double aux2 = 1. + aux1 / (6.5e-34 * exp(1335. * rt) * hnm);
        r.b27 = 2.4e-14 * exp(460. * rt) + aux1 / aux2;
    }
    r.b28 = 1.3e-12 * exp(380. * rt);
    r.b32 = r.b12 / (2.7e-27 * exp(11000. * rt) * hnm);
    r.b38 = 4.63e-11 * exp(20. * rt);
    r.b39 = 7.25e-11 * exp(20. * rt);

    r.c2 = 2.45e-12 * exp(-1775. * rt);
    r.c4 = falloff(K.a0c1, K.a1c1, -3.6, 1.1, lnt, hnm, xln6);
    r.c5 = 2.8e-12 * exp(300. * rt);
    r.c7 = 3.8e-13 * exp(800. * rt);
    r.c8 = 5.5e-12 * exp(125. * rt);
    r.c10 = 8.1e-11 * exp(-30. * rt);
    r.c14 = 9.5e-14 * exp(390. * rt);
    r.c15 = 3.9e-14 * exp(-900. * rt);
    {
        const double e200 = exp(200. * rt);
        r.c17a = 3.8e-12 * e200 * 0.7;
        r.c17b = 3.8e-12 * e200 * 0.3;
        r.e11 = 5.5e-12 * e200;
        r.a7 = 3.0e-11 * e200;
    }
    r.c18 = 3.3e-12 * exp(-115. * rt);
    r.c19 = 1.7e-11 * exp(-800. * rt);

    r.d2 = 2.3e-11 * exp(-200. * rt);
    r.d3 = 2.8e-11 * exp(85. * rt);
    {
        const double e290 = exp(290. * rt);
        r.d4 = 6.4e-12 * e290;
        r.d33 = 2.6e-12 * e290;
        r.e5c = 4.1e-13 * e290;
    }
    r.d5 = 7.3e-12 * exp(-1280. * rt);
    r.d6 = 3.05e-11 * exp(-2270. * rt);
    r.d7 = 1.4e-11 * exp(269. * rt);
    {
        const double e270 = exp(270. * rt);
        r.d8 = 7.4e-12 * e270;
        r.a26 = 3.3e-12 * e270;
    }
    {
        const double e230 = exp(230. * rt);
        r.d9 = 6.0e-13 * e230;
        r.e3 = 1.9e-11 * e230;
        r.e6 = 1.5e-12 * e230;
    }
    r.d10 = 3.6e-11 * exp(-375. * rt);
    r.d11 = 1.8e-12 * exp(-250. * rt);
    r.d31 = falloff(K.a0clx1, K.a1clx1, -3.4, -1.9, lnt, hnm, xln6);
    {
        const double em840 = exp(-840. * rt);
        r.d32 = 3.6e-12 * em840;
        r.e54 = 2.0e-12 * em840;
    }
    r.d34 = 3.0e-12 * exp(-500. * rt);
    r.d36 = falloff(K.a0clx2, K.a1clx2, -2., -1., lnt, hnm, xln6);
    r.d37 = 3.4e-12 * exp(-130. * rt);
    r.d60 = falloff(K.a0clx3, K.a1clx3, -4.5, -2.0, lnt, hnm, xln6);
    r.d61 = r.d60 / (1.72e-27 * exp(8649. * rt) * hnm);
    r.d62 = 1.0e-12 * exp(-1590. * rt) + 3.0e-11 * exp(-2450. * rt);
    r.d94 = 1.64e-12 * exp(-1520. * rt);
    r.d95 = 2.4e-12 * exp(-1250. * rt);
    {
        const double em1600 = exp(-1600. * rt);
        r.d97 = 1.05e-12 * em1600;
        r.c9 = 3.4e-11 * em1600;
    }

    r.e2 = 1.6e-11 * exp(-780. * rt);
    {
        const double e260 = exp(260. * rt);
        r.e4 = 8.8e-12 * e260;
        r.e5b = 2.3e-12 * e260;
    }
    r.e5a = 9.5e-13 * exp(550. * rt);
    r.e7 = 4.8e-12 * exp(-310. * rt);
    r.e8 = 1.9e-11 * exp(215. * rt);
    {
        const double e250 = exp(250. * rt);
        r.e10 = 1.7e-11 * e250;
        r.a17 = 4.8e-11 * e250;
    }
    r.e13 = falloff(K.a0brx, K.a1brx, -3.2, -2.9, lnt, hnm, xln6);
    {
        const double e460 = exp(460. * rt);
        r.e15 = 4.5e-12 * e460;
    }
    r.e16 = 1.2e-10 * exp(-430. * rt);
    r.e51 = 2.35e-12 * exp(-1300. * rt);

    r.hk1 = 4.23e-28 * hnm / (t * t);
    r.hk2 = 6.0e-34 * pow(300. / t, 2.4) * hnm;
    r.hk3 = 8.e-12 * exp(-2060. * rt);
    r.hk4 = 2.15e-11 * exp(110. * rt);
    r.hk5 = 3.3e-11 * exp(55. * rt);
}

// ------------------------------------------------------------------ photolysis
__constant__ double c_dsza[NSZA] = {0.,  5.,  10., 15., 20., 25., 30., 35., 40.,
                                     45., 50., 55., 60., 65., 70., 75., 80., 82.,
                                     84., 86., 88., 90., 91., 92., 93., 94., 95.};
__constant__ double c_xv3[NOZO] = {0.4, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6};

// Trilinear lookup in (altitude, SZA, O3-column ratio) (phot F:5043-5132); J = 0 above 95 deg.
__device__ void table_j(const double *__restrict__ ajl, const double *__restrict__ o3up,
                        double alt, double sza, double o3t, double *tj)
{
    if (!(sza <= 95.)) {
#pragma unroll
        for (int n = 0; n < NJTAB; ++n) tj[n] = 0.;
        return;
    }
    // SZA bracket: first node >= sza, minus one (F:5060-5068)
    int is = NSZA - 1;
#pragma unroll
    for (int i = NSZA; i >= 1; --i)
        if (c_dsza[i - 1] >= sza) is = i - 1;
    // alt < 0 or NaN would index o3up out of bounds (undefined in the reference too)
    const double zin = fmin(fmax(alt, 0.0), 78.9);
    const int iz = (int)zin + 1;
    const int i0 = (int)zin;
    const double v3std = o3up[i0] - (zin - (double)i0) * (o3up[i0] - o3up[i0 + 1]);
    const double stepo3 = (c_xv3[NOZO - 1] - c_xv3[0]) / (double)(NOZO - 1);
    double v3rat = o3t / v3std;
    v3rat = fmin(fmax(v3rat, c_xv3[0]), c_xv3[NOZO - 1]);
    int iv = (int)trunc((v3rat - c_xv3[0]) / stepo3) + 1;
    iv = iv < NOZO - 1 ? iv : NOZO - 1;
    const int j = iz > 1 ? iz : 1, k = is > 1 ? is : 1, l = iv > 1 ? iv : 1;
    const double cz = (zin - (double)(j - 1) * 1.) / 1.;
    const double cs = (sza - c_dsza[k - 1]) / (c_dsza[k] - c_dsza[k - 1]);
    const double co = (v3rat - c_xv3[l - 1]) / (c_xv3[l] - c_xv3[l - 1]);
    const double w0 = (1. - cz) * (1. - cs) * (1. - co), w1 = (1. - cz) * (1. - cs) * co;
    const double w2 = (1. - cz) * cs * (1. - co), w3 = (1. - cz) * cs * co;
    const double w4 = cz * (1. - cs) * (1. - co), w5 = cz * (1. - cs) * co;
    const double w6 = cz * cs * (1. - co), w7 = cz * cs * co;
    // 40 J-values of one node are contiguous (Fortran ajl(40,101,27,7))
    auto node = [&](int jj, int kk, int ll) {
        return ajl + NJTAB * ((jj - 1) + NZ * ((kk - 1) + NSZA * (ll - 1)));
    };
    const double *p0 = node(j, k, l), *p1 = node(j, k, l + 1), *p2 = node(j, k + 1, l),
                 *p3 = node(j, k + 1, l + 1), *p4 = node(j + 1, k, l),
                 *p5 = node(j + 1, k, l + 1), *p6 = node(j + 1, k + 1, l),
                 *p7 = node(j + 1, k + 1, l + 1);
#pragma unroll
    for (int n = 0; n < NJTAB; ++n)
        tj[n] = w0 * p0[n] + w1 * p1[n] + w2 * p2[n] + w3 * p3[n] + w4 * p4[n] + w5 * p5[n] +
                w6 * p6[n] + w7 * p7[n];
}

// NO photolysis, Minschwaner & Siskind (jno J:128-246), from the box's column sums.
__constant__ double c_sig_o2[3][6] = {
    {1.12e-23, 2.45e-23, 7.19e-23, 3.04e-22, 1.75e-21, 1.11e-20},
    {1.35e-22, 2.99e-22, 7.33e-22, 3.07e-21, 1.69e-20, 1.66e-19},
    {2.97e-22, 5.83e-22, 2.05e-21, 8.19e-21, 4.80e-20, 2.66e-19}};
__constant__ double c_sig_no1[3][6] = {
    {0.00e+00, 1.32e-18, 6.35e-19, 7.09e-19, 2.18e-19, 4.67e-19},
    {0.00e+00, 0.00e+00, 3.05e-21, 5.76e-19, 2.29e-18, 2.21e-18},
    {1.80e-18, 1.50e-18, 5.01e-19, 7.20e-20, 6.72e-20, 1.49e-21}};
__constant__ double c_sig_no2[3][6] = {
    {0.00e+00, 4.41e-17, 4.45e-17, 4.50e-17, 2.94e-17, 4.35e-17},
    {0.00e+00, 0.00e+00, 3.20e-21, 5.71e-17, 9.09e-17, 6.00e-17},
    {1.40e-16, 1.52e-16, 7.00e-17, 2.83e-17, 2.73e-17, 6.57e-18}};
__constant__ double c_w_no1[3][6] = {
    {0.00e+00, 5.12e-02, 1.36e-01, 1.65e-01, 1.41e-01, 4.50e-02},
    {0.00e+00, 0.00e+00, 1.93e-03, 9.73e-02, 9.75e-02, 3.48e-02},
    {4.50e-02, 1.80e-01, 2.25e-01, 2.25e-01, 1.80e-01, 4.50e-02}};
__constant__ double c_w_no2[3][6] = {
    {0.00e+00, 5.68e-03, 1.52e-02, 1.83e-02, 1.57e-02, 5.00e-03},
    {0.00e+00, 0.00e+00, 2.14e-04, 1.08e-02, 1.08e-02, 3.86e-03},
    {5.00e-03, 2.00e-02, 2.50e-02, 2.50e-02, 2.00e-02, 5.00e-03}};

__device__ double j_no(double colo2, double colno, double colo3, double hnm, double sza_top)
{
    if (!(sza_top <= 89.0)) return 0.;
    constexpr double dl[3] = {2.3, 1.5, 1.5}, fl[3] = {3.98e+11, 2.21e+11, 2.30e+11},
                     so3[3] = {4.6e-19, 6.7e-19, 7.1e-19};
    const double sec = 1. / cos(sza_top * (3.141592 / 180.));
    colo2 *= sec;
    colno *= sec;
    colo3 *= sec;
    const double prob = 1.65e+09 / (5.1e+07 + 1.65e+09 + 1.5e-09 * (hnm * 0.78));
    double band[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        double s = 0.;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double term =
                exp(-c_sig_o2[b][i] * colo2) *
                (c_w_no1[b][i] * c_sig_no1[b][i] * exp(-c_sig_no1[b][i] * colno) +
                 c_w_no2[b][i] * c_sig_no2[b][i] * exp(-c_sig_no2[b][i] * colno));
            s = i == 0 ? term : s + term;
        }
        band[b] = dl[b] * fl[b] * exp(-so3[b] * colo3) * prob * s;
    }
    return fmax(band[0] + band[1] + band[2], 1.e-30);
}

// ------------------------------------------------------------------ aerosol
// Correctly rounded values of the reference's constant libm calls (checked in
// tests/test_crmath.py): log(760), log(10) (F:2863), rnats**3, rnatl**3, rice**3.
constexpr double LOG760 = 6.633318433280377, LOG10 = 2.302585092994046;
constexpr double RNATS3 = 1.2500000000000002e-13, RNATL3 = 2.7462499999999997e-10, RICE3 = 1e-09;

// x**0.5: the correctly rounded power equals the correctly rounded square root.
__device__ __forceinline__ double sqrt_pow(double x) { return x == -INFINITY ? INFINITY : sqrt(x) + 0.; }

struct Het {
    double k[9];                    // het1..het9
    double pno3, pno3sl, ph2o, phcl; // gas fractions
    double surf, vsed;
};

__device__ __forceinline__ double solid_uptake(double g, double t, double mw, double r, double n,
                                               double ptot, double z)
{
    // R = g*pi*r^2*cbar*n/(1+3g r/(4L)) in the Turco form of F:3815-3880
    return g * 4.56E4 * sqrt(t / mw) * (r * r) * n / (1.0 + 3.3E4 * g * r * ptot / t) / z;
}

__device__ double ternary_density(double ws, double wn, double t)
{
    // density F:4059-4069
    const double w = ws + wn, wh = 1.0 - w, t2 = t * t, t3 = t * t * t;
    const double v1 = 2.393284E-02 + -4.359335E-05 * t + 7.961181E-08 * t2 + 0.0 * t3;
    const double vs = -0.198716351 + 1.39564574E-03 * t + -2.020633E-06 * t2 +
                      (0.51684706 + -3.0539E-03 * t + 4.505475E-06 * t2) * w +
                      (-0.30119511 + 1.840408E-03 * t + -2.7221253742E-06 * t2) * w * w;
    const double vn = -0.11331674116 + 8.47763E-04 * t + -1.22336185E-06 * t2 +
                      (0.3455282 + -2.2111E-03 * t + 3.503768245E-06 * t2) * w +
                      (-0.2315332 + 1.60074E-03 * t + -2.5827835E-06 * t2) * w * w;
    const double vm = wh / 18.0160 * v1 + vs * ws / 98.080 + vn * wn / 63.0160;
    return 1.0 / vm * 0.001;
}

// Elementary functions of the PSC block up to the liquid composition.  MathCR: the
// correctly rounded ones of crmath.h, for boxes that may take the ternary branch
// (T < 215 K, F:2994), whose solution is ill-conditioned near its binary limit and
// amplifies last-bit differences (see crmath.h).  MathLib: the device library, for the
// warmer boxes, where NAT, ice and the binary solution are well conditioned.
struct MathCR {
    static __device__ __forceinline__ double pow10(double x) { return crm::pow10_cr(x); }
    static __device__ __forceinline__ double log10(double x) { return crm::log10_fdlibm(x); }
    static __device__ __forceinline__ double log(double x) { return crm::log_cr(x); }
    static __device__ __forceinline__ double exp(double x) { return crm::exp_cr(x); }
    static __device__ __forceinline__ double pow(double x, double y) { return crm::pow_cr(x, y); }
    static __device__ __forceinline__ double atan(double x) { return crm::atan_cr(x); }
    static __device__ __forceinline__ double cos(double x) { return crm::cos_cr(x); }
};
struct MathLib {
    static __device__ __forceinline__ double pow10(double x) { return ::pow(10., x); }
    static __device__ __forceinline__ double log10(double x) { return ::log10(x); }
    static __device__ __forceinline__ double log(double x) { return ::log(x); }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
    static __device__ __forceinline__ double pow(double x, double y) { return ::pow(x, y); }
    static __device__ __forceinline__ double atan(double x) { return ::atan(x); }
    static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
};

// analytic (NAT F:2795-2857, ice F:2861-2913, liquid F:2921-3228, area F:3240, HNO3
// partition F:3252, settling F:3264-3348) then hetero (F:3436-4023) for one box, with
// the elementary functions of M up to the liquid composition (ws, wn, pn).  Area,
// solubilities, settling and the uptake rates downstream are well conditioned and use
// the library.
template <class M>
__device__ void psc_block(double temp, double ptot, double hnm, const double *ck, Het &o)
{
    constexpr double R = 8.205e-5, NDROP = 10., NNAT_SMALL = 1., RNATS = 0.5e-4,
                     RNATL = 6.5e-4, RICE = 10.e-4, AMU = 1.66e-24;
    const double pi = 3.141592653589793;   // 2*asin(1.0) (F:2769), exactly
    double nnats = 0., nnatl = 0., anat = 0., nice = 0., aice = 0., vnat = 0., vnatl = 0.;
    double pno3l = 1., pno3s = 1.;
    int lnat = 0, lice = 0, condliq;
    o.vsed = 0.;
    o.pno3 = 1.;
    o.pno3sl = 1.;
    o.ph2o = 1.;
    o.phcl = 1.;

    double pw = fmin(ck[H2O] / hnm * ptot / 1013.0, 2.e-3 / 1013.);
    double pn0 = ck[HNO3] / hnm * ptot / 1013.0;
    double pn;
    const double tm = fmin(temp, 273.);

    // NAT (Hanson & Mauersberger)
    {
        const double zmt = -2.7836 - 0.00088 * tm;
        const double zbt = 38.9855 - 11397.0 / tm + 0.009179 * tm;
        const double eq = M::pow10(zmt * M::log10(pw * 1013. * 0.75) + zbt);
        if (pn0 * 1013. * 0.75 / eq >= 1.) {
            lnat = 1;
            pn = eq / (0.75 * 1013.);
            const double hno3s = fmax((pn0 - pn) * 1013. * hnm / ptot, 0.);
            vnat = hno3s * AMU * 117. / 1.35;
            const double vsmall = NNAT_SMALL * (4. / 3. * pi * RNATS3);
            if (vnat > vsmall) {
                nnats = NNAT_SMALL;
                nnatl = (vnat - vsmall) / (4. / 3. * pi * RNATL3);
                vnatl = vnat - vsmall;
            } else {
                nnats = vnat / (4. / 3. * pi * RNATS3);
                nnatl = 0.;
                vnatl = 0.;
            }
            anat = nnats * 4 * pi * (RNATS * RNATS) + nnatl * 4 * pi * (RNATL * RNATL);
            pno3s = fmin(fmax(1.0 - (pn0 - pn) / pn0, 0.), 1.);
            o.pno3sl = fmin(fmax(1.0 - (pn0 - pn) * vnatl / vnat / pn0, 0.), 1.);
        }
    }
    // ice
    const double tice = 2668.70 / (10.4310 - (M::log(pw) + LOG760) / LOG10);
    if (tm <= tice) {
        lice = 1;
        lnat = 0;
        anat = 0.;
        const double ph2o = M::pow10(-2668.7 / tm + 10.4310) / 760.0;
        const double h2os = fmax(pw - ph2o, 0.) * 1013. * hnm / ptot;
        const double vice = h2os * AMU * 18. / 0.928;
        nice = vice / (4. / 3. * pi * RICE3);
        aice = nice * 4 * pi * (RICE * RICE);
        o.ph2o = fmin(fmax(1.0 - fmax(pw - ph2o, 0.) / pw, 0.), 1.);
        const double zmt = -2.7836 - 0.00088 * tm;
        const double zbt = 38.9855 - 11397.0 / tm + 0.009179 * tm;
        const double eq = M::pow10(zmt * M::log10(pw * o.ph2o * 1013. * 0.75) + zbt);
        pn = eq / (0.75 * 1013.);
        pno3s = fmin(fmax(1.0 - (pn0 - pn) / pn0, 0.), 1.);
        o.pno3sl = pno3s;
    }
    // liquid binary H2SO4/H2O then ternary (Carslaw et al. 1995)
    const double t = fmax(temp, fmax(tice - 3.0, 185.));
    pw = fmax(fmin(ck[H2O] * o.ph2o / hnm * ptot / 1013.0, 2.e-3 / 1013.), 2.e-5 / 1013.);
    const double lt = M::log(t), lpw = M::log(pw);   // shared by binary and ternary
    const double qh2so4 = fmax(ck[H2SO4] / hnm, 1.e-18);
    const double ns = ptot * 100.0 * qh2so4 / 8.314 / t;
    pn0 = ptot / 1013.0 * (ck[HNO3] * pno3s / hnm);
    double ms, ws, wn = 0.0, msb;
    condliq = 1;
    {
        const double a = 51.81 + -15732.0 / t;
        const double b = -21.661 + 2724.2 / t;
        const double c = 47.004 + -6969.0 / t + -4.6183 * lt - lpw;
        const double det = (b * b) - 4. * a * c;
        if (det > 0.) {
            const double xsb = (-b - sqrt(det)) / (2. * a);
            msb = 55.51 * xsb / (1.0 - xsb);
        } else {
            msb = 0.;
            condliq = 0;
        }
    }
    ms = msb;
    ws = msb * 0.098076 / (1.0 + msb * 0.098076);
    {
        // H2SO4 saturation pressure (Ayers et al. 1980, Giauque et al. 1960)
        const double wts = fmax(41., ws * 100.);
        const double mu = 4.184 * (1.514e4 - 286. * (wts - 40.) + 1.08 * sq(wts - 40.) -
                                   3941. / M::pow(wts - 40., 0.1));
        const double ph2so4 = M::exp(-10156. / t + 16.2590 - mu / (8.314 * t));
        if (qh2so4 * ptot / 1013. <= ph2so4) {
            ms = 0.;
            ws = 0.;
            condliq = 0;
        }
    }
    if (condliq == 1 && t < 215.) {
        constexpr double qn[10] = {14.5734, 0.0615994, -1.14895, 0.691693, -0.098863,
                                   0.0051579, 0.123472, -0.115574, 0.0110113, 0.0097914};
        constexpr double qs[10] = {14.4700, 0.0638795, -3.29597, 1.778224, -0.223244,
                                   0.0086486, 0.536695, -0.335164, 0.0265153, 0.0157550};
        constexpr double kn[7] = {-39.136, 6358.4, 83.29, -17650.0, 198.53, -11948, -28.469};
        const double tt = R * t * ns, tr = 1.0e4 / t - 43.4782608, pr = lpw + 18.4;
        const double tr2 = tr * tr, tr3 = tr * tr * tr, pr2 = pr * pr, pr3 = pr * pr * pr;
        const double hsb = M::exp(qs[0] + qs[1] * tr2 +
                               (qs[2] + qs[3] * tr + qs[4] * tr2 + qs[5] * tr3) * pr +
                               (qs[6] + qs[7] * tr + qs[8] * tr2) * pr2 + qs[9] * tr * pr3);
        const double kk = kn[0] + kn[1] / t;
        const double xnb = 1.0 / (2.0 * (kn[2] + kn[3] / t)) *
                           (-kn[0] - kn[1] / t -
                            sqrt_pow(kk * kk - 4.0 * (kn[2] + kn[3] / t) *
                                                   (kn[4] + kn[5] / t + kn[6] * lt - lpw)));
        const double mnb = 55.51 * xnb / (1.0 - xnb);
        const double hnb = M::exp(qn[0] + qn[1] * tr2 +
                               (qn[2] + qn[3] * tr + qn[4] * tr2 + qn[5] * tr3) * pr +
                               (qn[6] + qn[7] * tr + qn[8] * tr2) * pr2 + qn[9] * tr * pr3);
        const double A = (tt * hnb * (mnb * mnb) - tt * hsb * mnb * msb - 2.0 * (mnb * mnb) * msb +
                          mnb * (msb * msb) + hnb * mnb * msb * pn0 - hsb * (msb * msb) * pn0) /
                         ((mnb * mnb) - mnb * msb);
        const double B =
            msb * (-2.0 * tt * hnb * mnb + tt * hsb * msb + mnb * msb - hnb * msb * pn0) /
            (mnb - msb);
        const double C = (tt * hnb * mnb * (msb * msb)) / (mnb - msb);
        const double u = (A * A) - 3.0 * B;
        const double v = -2.0 * (A * A * A) + 9.0 * A * B - 27.0 * C;
        double phi = M::atan(sqrt(fmax(4.0 * (u * u * u) - (v * v), 0.)) / v);
        if (phi < 0.) phi = phi + pi;
        const double msf = fmax(-1.0 / 3.0 * (A + 2.0 * sqrt((A * A) - 3.0 * B) * M::cos((pi + phi) / 3.0)), 0.);
        const double mnf = fmax(mnb * (1.0 - msf / msb), 0.);
        pn = mnf / (hnb * mnf / (mnf + msf) + hsb * msf / (mnf + msf));
        ms = msf;
        ws = msf * 0.098076 / (1.0 + msf * 0.098076 + mnf * 0.063012);
        wn = mnf * 0.063012 / (1.0 + msf * 0.098076 + mnf * 0.063012);
        pno3l = fmin(1.0 - (pn0 - pn) / pn0, 1.);
    }
    ws = fmax(ws, 0.005);
    const double dens = ternary_density(ws, wn, temp);

    double aliq = 0., rmean = 0., aw = 0., wt = 0., hhcl = 0., hclono2 = 0., hhocl = 0.,
           hhobr = 0.;
    if (condliq == 1) {
        const double ls2 = sq(log(1.8));
        const double vliq = ns * 98.076e-6 / ws / dens;
        const double rmode = pow(3.0 * vliq / (4. * pi * NDROP) * exp(-9.0 / 2.0 * ls2), 1. / 3.);
        rmean = rmode * exp(0.5 * ls2);
        aliq = 4.0 * pi * (rmode * rmode) * NDROP * exp(2.0 * ls2);
        const double ph2o0 = exp(18.452406985 - 3505.1578807 / t - 330918.55082 / (t * t) +
                                 12725068.262 / (t * t * t));
        aw = fmin(pw * 1013. / ph2o0, 1.);
        double y1, y2;
        if (aw <= 0.05) {
            y1 = 12.37208932 * pow(aw, -0.16125516114) + -30.490657554 * aw + -2.1133114241;
            y2 = 13.455394705 * pow(aw, -0.1921312255) + -34.285174607 * aw + -1.7620073078;
        } else if (aw < 0.85) {
            y1 = 11.820654354 * pow(aw, -0.20786404244) + -4.807306373 * aw + -5.1727540348;
            y2 = 12.891938068 * pow(aw, -0.232338847708) + -6.4261237757 * aw + -4.9005471319;
        } else {
            y1 = -180.06541028 * pow(aw, -0.38601102592) + -93.317846778 * aw + 273.88132245;
            y2 = -176.95814097 * pow(aw, -0.36257048154) + -90.469744201 * aw + 267.45509988;
        }
        const double m = y1 + (t - 190.) * (y2 - y1) / 70.;
        wt = fmax(9800. * m / (98. * m + 1000.), 0.5);
        const double mh2so4 = dens * wt / 9.8;
        const double bigx = wt / (wt + (100. - wt) * 98. / 18.);
        hhcl = (0.094 - 0.61 * bigx + 1.2 * (bigx * bigx)) *
               exp(-8.68 + (8515. - 10718. * pow(bigx, 0.7)) / t);
        const double phcl0 = ck[HCL] / hnm * ptot / 1013.0;
        const double phcl1 = (phcl0 / (R * t)) / (1. / (R * t) + hhcl * ns / ms);
        o.phcl = 1.0 - (phcl0 - phcl1) / phcl0;
        hclono2 = 1.6e-6 * exp(4710. / t) * exp(-(0.306 + 24. / t) * mh2so4);
        hhocl = 1.91e-6 * exp(5862.4 / t) * exp(-(0.0776 + 59.18 / t) * mh2so4);
        hhobr = 4.6e-4 * exp(4.5e3 / t);
    }
    o.surf = aliq + anat * (double)lnat + aice * (double)lice;
    o.pno3 = fmin(pno3l * pno3s, 1.);

    // settling velocity of NAT / ice particles (Pruppacher & Klett; Larsen 1991)
    if (lice == 1 || lnat == 1) {
        const double rpart = fmin(((double)lice * RICE + (double)lnat * RNATL) * 1.e-2, 535.e-6);
        if (rpart >= 0.5e-6) {
            constexpr double bn[7] = {-3.18657, 0.992696, -0.153193e-2, -0.987059e-3,
                                      -0.578878e-3, 0.855176e-4, -0.327815e-5};
            const double rhopart = (double)lice * (0.928 * 1000.0) + (double)lnat * (1.350 * 1000.0);
            const double rhoair = (hnm * 28.97 / 6.022e23) * 1000.0;
            const double tcel = temp - 273.15;
            const double eta = (1.718 + 0.0049 * tcel - 1.2e-5 * tcel * tcel) * 1.e-5;
            const double cd = 32.0 * (rpart * rpart * rpart) * (rhopart - rhoair) * rhoair * 9.80665 /
                              (3.0 * (eta * eta));
            const double lx = log(cd), l2 = lx * lx, l4 = l2 * l2;
            const double nre = exp(bn[0] + bn[1] * lx + bn[2] * l2 + bn[3] * (lx * lx * lx) +
                                   bn[4] * l4 + bn[5] * (l4 * lx) + bn[6] * (l4 * l2));
            if (nre >= 1.e-2) {
                o.vsed = eta * nre / (2 * rhoair * rpart);
            } else {
                const double lam = 6.6e-8 * (1013.25 / ptot) * (temp / 293.15);
                const double us = 2.0 * (rpart * rpart) * 9.80665 * (rhopart - rhoair) / (9.0 * eta);
                o.vsed = us / 1.2 *
                         (1.0 + 1.246 * lam / rpart + 0.42 * lam / rpart * exp(-0.87 * rpart / lam));
            }
        }
    }

    // heterogeneous second-order rates: liquid (JPL 2000 / Shi et al.), NAT, ice
    double kl[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    double ks[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    if (condliq == 1) {
        const double T = temp, sT = sqrt(T);
        const double zh2o = fmax(ck[H2O], 1.E-08 * hnm);
        const double zhcl = fmax(ck[HCL], 1.E-13 * hnm);
        const double zclono2 = fmax(ck[CLONO2], 1.E-13 * hnm);
        const double zhobr = fmax(ck[HOBR], 1.E-15 * hnm);
        const double phcl = (zhcl / hnm) * ptot / 1013.;
        const double pclono2 = (zclono2 / hnm) * ptot / 1013.;
        const double phobr = (zhobr / hnm) * ptot / 1013.;
        const double wt2 = wt * wt, wt3 = wt * wt * wt;
        const double t0 = 144.11 + 0.166 * wt - 0.015 * wt2 + 2.18e-4 * wt3;
        const double biga = 169.5 + 5.18 * wt - 0.0825 * wt2 + 3.27e-3 * wt3;
        const double eta = biga * pow(T, -1.43) * exp(448. / (T - t0));
        const double ah = exp(60.51 - 0.095 * wt + 0.0077 * wt2 - 1.61e-5 * wt3 -
                              (1.76 + 2.52e-4 * wt2) * sT + (-805.89 + 253.05 * pow(wt, 0.076)) / sT);
        // ClONO2 + H2O / HCl
        const double dclono2 = 5.e-8 * T / eta;
        const double cclono2 = 1474. * sT;
        const double khydr = 1.95e10 * exp(-2800. / T) * aw + 1.22e12 * exp(-6200. / T) * ah * aw;
        const double gbh2o = 4 * hclono2 * 0.082 * T * sqrt(dclono2 * khydr) / cclono2;
        const double mhcl = hhcl * phcl;
        const double khcl = 7.9e11 * ah * dclono2 * mhcl;
        const double lcl = sqrt(dclono2 / (khydr + khcl));
        const double fcl = 1. / tanh(rmean / lcl) - lcl / rmean;
        double grxn = fcl * gbh2o * sqrt(1. + khcl / khydr);
        const double gbhcl = grxn * khcl / (khcl + khydr);
        const double gs = 66.12 * hclono2 * mhcl * exp(-1374. / T);
        const double fhcl = 1. / (1 + 0.612 * (gs + gbhcl) * pclono2 / phcl);
        const double gps = fhcl * gs, gbhclp = fhcl * gbhcl;
        const double gb = gbhclp + grxn * khydr / (khcl + khydr);
        const double g = 1. / (1. + 1. / (gps + gb));
        const double ghcl = g * (gps + gbhclp) / (gps + gb);
        kl[0] = 0.25 * (g - ghcl) * cclono2 * aliq / zh2o;
        kl[1] = 0.25 * ghcl * cclono2 * aliq / zhcl;
        // HOCl + HCl
        const double chocl = 2009. * sT;
        const double dhocl = 6.4e-8 * T / eta;
        const double khh = 1.25e9 * ah * dhocl * mhcl;
        const double lhocl = sqrt(dhocl / khh);
        const double fhocl = 1. / tanh(rmean / lhocl) - lhocl / rmean;
        grxn = 4. * hhocl * 0.082 * T * sqrt(dhocl * khh) / chocl;
        kl[4] = 0.25 * (1. / (1. + 1. / (fhocl * grxn * fhcl))) * chocl * aliq / zhcl;
        // HOBr + HCl (Waschewsky & Abbatt 1999), limiting reagent switch
        const double mhobr = hhobr * phobr;
        const double k2 = exp(0.542 * wt - 6.44e3 / T + 10.3);
        if (mhobr / mhcl < 1.) {
            const double k1 = k2 * mhcl, dd = 6.2e-8 * T / eta, ll = sqrt(dd / k1);
            const double ff = 1. / tanh(rmean / ll) - ll / rmean, cb = 1477. * sT;
            const double gg = 4. * hhobr * 0.082 * T * sqrt(dd * k1) * ff / cb;
            kl[6] = 0.25 * gg * cb * aliq / zhcl;
        } else {
            const double k1 = k2 * mhobr, dd = 7.8e-8 * T / eta, ll = sqrt(dd / k1);
            const double ff = 1. / tanh(rmean / ll) - ll / rmean, cb = 2408. * sT;
            const double gg = 4. * hhcl * 0.082 * T * sqrt(dd * k1) * ff / cb;
            kl[6] = 0.25 * gg * cb * aliq / zhobr;
        }
        // N2O5 + H2O (gamma 0.1), BrONO2 + H2O (JPL 2000)
        kl[2] = 0.25 * 0.1 * (1400.1 * sT) * aliq / zh2o;
        const double gxrn = exp(29.24 - 0.396 * 100. * ws) + 0.114;
        kl[5] = 0.25 * (1. / (1. / 0.805 + 1. / gxrn)) * (1221.4 * sT) * aliq / zh2o;
    }
    if (lnat == 1) {
        const double T = temp;
        const double zh2o = fmax(ck[H2O], 1.E-08 * hnm), zhcl = fmax(ck[HCL], 1.E-13 * hnm);
        auto nat = [&](double g, double mw, double z) {
            return solid_uptake(g, T, mw, RNATS, nnats, ptot, z) +
                   solid_uptake(g, T, mw, RNATL, nnatl, ptot, z);
        };
        ks[0] = nat(0.004, 97.45, zh2o);
        ks[1] = nat(0.2, 97.45, zhcl);
        ks[2] = nat(4.0E-4, 108.0, zh2o);
        ks[3] = nat(3.0E-3, 108.0, zhcl);
        ks[4] = nat(0.1, 52.5, zhcl);
        ks[5] = nat(0., 142.0, zh2o);
        ks[6] = nat(0., 96.9, zhcl);
    }
    if (lice == 1) {
        const double T = temp;
        const double zh2o = fmax(ck[H2O], 1.E-08 * hnm), zhcl = fmax(ck[HCL], 1.E-13 * hnm),
                     zhobr = fmax(ck[HOBR], 1.E-13 * hnm);
        ks[0] = solid_uptake(0.3, T, 97.45, RICE, nice, ptot, zh2o);
        ks[1] = solid_uptake(0.3, T, 97.45, RICE, nice, ptot, zhcl);
        ks[2] = solid_uptake(0.02, T, 108.0, RICE, nice, ptot, zh2o);
        ks[3] = solid_uptake(0.03, T, 108.0, RICE, nice, ptot, zhcl);
        ks[4] = solid_uptake(0.2, T, 52.5, RICE, nice, ptot, zhcl);
        ks[5] = solid_uptake(0.3, T, 142.0, RICE, nice, ptot, zh2o);
        ks[6] = solid_uptake(0.3, T, 96.9, RICE, nice, ptot, zhcl);
        ks[7] = solid_uptake(0.1, T, 96.9, RICE, nice, ptot, zhobr);
    }
    const double nsol = (double)(lnat + lice);
#pragma unroll
    for (int i = 0; i < 8; ++i) o.k[i] = kl[i] + nsol * ks[i];
    o.k[8] = 0.;
}

// The PSC block with correctly rounded math where the ternary branch can be taken.
__device__ __forceinline__ void psc_and_het(double temp, double ptot, double hnm, const double *ck,
                                            Het &o)
{
    if (temp < 215.)
        psc_block<MathCR>(temp, ptot, hnm, ck, o);
    else
        psc_block<MathLib>(temp, ptot, hnm, ck, o);
}

// ------------------------------------------------------------------ the sweeps
// Fixed-count family/QSSA iteration of chem3d (F:1348-2378). ck0 is the start-of-step
// state, ck/cc the iterates (number densities), updated in place.
struct BoxEnv {
    double hnm, n2, o2, co2, h2, dt;
    double het1, het2, het3, het4, het5, het6, het7, het8, het9;
    double pno3, ph2o, phcl, pbr;
    double rain_hno3, rain_hno4, rain_hcl, rain_hbr, rain_h2o2, rain_ch2o, rain_ch3o2h;
    bool day;
};

// The converged mode's fixed-point test, without divisions: a sweep has converged when
// every long- and short-lived species changed by less than tol relative to its new value
// (values under the floor fl compared at fl): |v - old| / max(|v|, fl) < tol; `fin` stays
// true while every change is finite.  A NaN fails both.  Each species is set once per
// sweep, so `old` is its value at the start of the sweep.
// dmax is the largest relative change of the sweep, also the input of the distance-to-
// fixed-point estimate; it only steers when to stop, so the hardware reciprocal (not the
// IEEE division) forms it.
struct Track {
    bool conv, fin;            // set at the sweep's end (track_end)
    double tol, fl;
    double dmax;               // the sweep's largest relative change (set by track_end)
    uint32_t hmax = 0;         // the largest high word of the relative changes
};

// One species' change r = |v - old| / max(|v|, fl).  For r >= 0 the high 32 bits of the
// double, read as an unsigned integer, order like r to 2^-20 relative, and every NaN or
// infinity reads above every finite r: one integer maximum per species keeps both the
// largest change and whether every change was finite.  A new value that is infinite or NaN
// makes r NaN (|v - old| / max(|v|, fl) is NaN or inf * 0), and so does a NaN start.
__device__ __forceinline__ void track(Track &t, double v, double old)
{
    const double dv = fabs(v - old), sc = fmax(fabs(v), t.fl);
    // a reciprocal, not a division: off the chain of maxima that bounds a lone lane's sweep
    // (the largest ratio by cross-multiplication, divided once per sweep, put two
    // multiplications per species on that chain: the plain c2 solve, bound by its slowest
    // lane, took 42-45 ms against 37)
    const uint32_t h = (uint32_t)__double2hiint(dv * __builtin_amdgcn_rcp(sc));
    t.hmax = t.hmax > h ? t.hmax : h;
}

// The sweep's largest change, its high word with the low word zero (below the true maximum
// by less than 2^-20 relative: the test against tol and the distance estimate do not depend
// on such a difference), and the finiteness of every change.
__device__ __forceinline__ void track_end(Track &t)
{
    t.fin = t.hmax < 0x7ff00000u;
    t.dmax = __hiloint2double((int)t.hmax, 0);
    t.conv = t.dmax < t.tol;
}

// Successive plain sweeps the converged mode needs before it may stop: its estimate of
// the distance to the fixed point, d rho / (1 - rho) after a sweep of largest relative
// change d, takes rho as the larger of the last two ratios of successive changes (the
// larger, so that an oscillating sequence -- the damped radicals -- cannot show a small
// rate at a low point), capped at CONV_RHOMAX: changes at rounding level have ratios near 1
// at random, and the cap accepts a change below ~1e-3 tol whatever its ratios.
// A change that stays below tol for CONV_STALL successive plain sweeps without decreasing
// monotonically (it rose at least once) is rounding noise -- near the terminator the OH
// quadratic's cancellation leaves ~1e-14 -- and the box has converged.
constexpr int CONV_NPLAIN = 3;
constexpr double CONV_RHOMAX = 0.999;
constexpr int CONV_STALL = 8;

__device__ __forceinline__ bool near_fixed_point(double d, double d1, double d2, double tol)
{
    const double rho = fmin(fmax(d / d1, d1 / d2), CONV_RHOMAX);
    return d == 0. || d * rho < tol * (1. - rho);
}

// J-values that the sweep kernels keep in LDS beside ck0 (lane-private columns, stride
// BLK): 42 + 38 doubles per lane fill the 160 KB of LDS at one wave per SIMD.  The three
// the sweep reads most often stay in registers with the box state and the 82 rates.  At
// block 64 this took the fixed-sweep kernel's spill from 106 to 46 registers and the plain
// converged kernel's from 163 to 95; an LDS read also replaces an AGPR read, which is a
// VALU instruction.
// JL_SET orders the J-values by how often the sweep reads them; the first NJL go to LDS.
constexpr int NJL = 38;
constexpr int JL_SET[NJ] = {
    JBRO, JCO2, JN2O, JNO2, JNO, JO2, JO3_O3P,                                 // 1 use
    JCH2BR2, JCH2O_CO, JCH3BR, JCL2, JH1301, JH2O2, JH2O, JHCL, JN2O5, JO3_O1D,
    JOCLO,                                                                     // 2 uses
    JCCL4, JCFC113, JCFC11, JCFC12, JCH2O_HCO, JCH3CCL3, JCH3CL, JCH3O2H, JCLNO2,
    JH1211, JHBR, JHCFC22, JHNO3, JHNO4, JNO3_NO2,                             // 3 uses
    JCLONO2_CLO, JBRONO2, JCL2O2, JCLONO2_CL, JHOBR, JHOCL, JNO3_NO, JBRCL};
__host__ __device__ constexpr int jl_slot(int j)
{
    for (int i = 0; i < NJL; ++i)
        if (JL_SET[i] == j) return i;
    return -1;
}
constexpr bool jl_set_is_permutation()
{
    for (int j = 0; j < NJ; ++j) {
        int n = 0;
        for (int i = 0; i < NJ; ++i) n += JL_SET[i] == j;
        if (n != 1) return false;
    }
    return true;
}
static_assert(jl_set_is_permutation(), "JL_SET lists every J-value once");
static_assert(NJL >= 0 && NJL <= NJ, "J-values in LDS");

// ck0 is the start-of-step state in LDS, one column per lane: ck0[s * CK0_STRIDE].
// With TRACK, tr accumulates the fixed-point test over this sweep.
template <int CK0_STRIDE, bool TRACK, bool JLDS = false>
__device__ void sweep(const Rates &R, const double *tj, const BoxEnv &e,
                                      const double *ck0, double *ck, double *cc, double &ch3o,
                                      double &hco, Track &tr, const double *tjl = nullptr)
{
// J-value X: from LDS (tjl, stride CK0_STRIDE) if JLDS and X is one of JL_SET, else tj[X]
#define TJ(X) ((JLDS && jl_slot(X) >= 0) ? tjl[jl_slot(X) * CK0_STRIDE] : tj[X])
#define CCSET(sl, expr)                                                              \
    do {                                                                             \
        const double v_ = (expr);                                                    \
        if (TRACK) track(tr, v_, cc[sl]);                                            \
        cc[sl] = v_;                                                                 \
    } while (0)
    const double hnm = e.hnm, n2 = e.n2, o2 = e.o2, co2 = e.co2, h2 = e.h2;
    double day;
    if (e.day) {
        day = 1.;
        CCSET(CH3, ((R.c2 * cc[OH] + C1A * cc[O1D] + R.d5 * cc[CL]) * ck[CH4]) /
                  (R.c4 * o2 + C3 * cc[O3P]));
        ch3o = (C1B * cc[O1D] * ck[CH4] +
                (R.c5 * cc[NO] + 2. * R.c14 * ck[CH3O2] + R.c18 * cc[CLO]) * ck[CH3O2] +
                TJ(JCH3O2H) * ck[CH3O2H]) /
               (R.c15 * o2);
        hco = (R.c8 * cc[OH] + R.c9 * cc[O3P] + R.c10 * cc[CL] + R.c19 * cc[BR] + TJ(JCH2O_HCO)) *
              ck[CH2O] / (C12 * o2);
        const double r1 = (TJ(JO3_O1D) + TJ(JO3_O3P) + R.a2 * cc[H] + R.a6 * cc[OH] +
                           R.a6b * cc[HO2] + R.b4 * cc[NO] + R.d2 * cc[CL]) /
                          (R.hk2 * o2);
        CCSET(O3, 1. / (1. + r1) * ck[OX]);
        CCSET(O3P, r1 * cc[O3]);
        CCSET(O1D, TJ(JO3_O1D) / (R.hk4 * n2 + R.hk5 * o2) * cc[O3]);
        const double r4 = (R.a5 * cc[O3P] + R.a36 * ck[CO] + R.a19 * h2) /
                          (R.a1 * o2 + R.a2 * cc[O3] + (A23A + A23B + A23C) * cc[HO2]);
        CCSET(H, r4 * cc[OH]);
        const double r7 = TJ(JNO) / (R.b7 * o2 + R.b6 * cc[NO]);
        const double r8 = (R.b9 * cc[O3]) / (TJ(JNO3_NO) + TJ(JNO3_NO2) + R.b12 * cc[NO2]);
        const double r5 = (TJ(JNO2) + R.b3 * cc[O3P] + TJ(JNO3_NO) * r8) /
                          (R.b4 * cc[O3] + R.a26 * cc[HO2] + R.d4 * cc[CLO] + R.e4 * cc[BRO] +
                           R.c5 * ck[CH3O2]);
        CCSET(NO2, ck[NOX] / (1. + r5 + r7 * r5 + r8));
        CCSET(NO, r5 * cc[NO2]);
        CCSET(N, r7 * cc[NO]);
        CCSET(NO3, r8 * cc[NO2]);
        const double r6 = (R.a6 * cc[O3] + A30 * ck[H2O2] + R.c8 * ck[CH2O] + R.d8 * cc[CLO] +
                           R.e10 * cc[BRO] + R.a1 * o2 * r4) /
                          ((A23A + A23B + A23C) * cc[H] + R.a6b * cc[O3] + R.a7 * cc[O3P] +
                           R.a26 * cc[NO] + 2. * R.a27 * cc[HO2] + R.a17 * cc[OH] +
                           R.d7 * cc[CL] + R.d10 * cc[CL] + R.d33 * cc[CLO] + R.e7 * cc[BR] +
                           R.e15 * cc[BRO]);
        CCSET(HO2, .5 * cc[HO2] + .5 * cc[OH] * r6);        // 0.5 under-relaxation
        const double r13 = R.d31 * cc[NO2] /
                           (TJ(JCLONO2_CL) + TJ(JCLONO2_CLO) + R.d32 * cc[O3P] +
                            e.het1 * ck[H2O] + e.het2 * ck[HCL]);
        const double r9 = (R.d33 * cc[HO2] + e.het1 * ck[H2O] * r13) /
                          (e.het5 * ck[HCL] + D35 * cc[O3P] + TJ(JHOCL));
        const double r10 = (R.d60 * cc[CLO]) / (R.d61 * hnm + TJ(JCL2O2));
        const double r12 = R.e5c * cc[BRO] / TJ(JBRCL);
        const double r11 = (R.d3 * cc[O3P] + R.d4 * cc[NO] + R.d8 * cc[OH] + R.c18 * ck[CH3O2] +
                            R.e5b * cc[BRO] + TJ(JCLONO2_CL) * r13 + TJ(JHOCL) * r9 +
                            2. * TJ(JCL2O2) * r10 + TJ(JBRCL) * r12) /
                           (R.d2 * cc[O3] + R.d5 * ck[CH4] + R.d7 * cc[HO2] + R.d10 * cc[HO2] +
                            R.d37 * ck[HOCL]);
        CCSET(CLO, .5 * cc[CLO] + .5 * ck[CLOX] / (1. + r11));
        CCSET(CL, r11 * cc[CLO]);
        const double r27 = R.e13 * cc[NO2] / (TJ(JBRONO2) + R.e8 * cc[O3P]);
        const double r28 = R.e15 * cc[HO2] / (TJ(JHOBR) + R.e16 * cc[O3P]);
        const double r29 = R.e5c * cc[CLO] / TJ(JBRCL);
        const double r30 = (R.e3 * cc[O3P] + R.e4 * cc[NO] + R.e5a * cc[CLO] + R.e5b * cc[CLO] +
                            2. * R.e6 * cc[BRO] + R.e10 * cc[OH] + TJ(JBRO) +
                            TJ(JBRONO2) * 0.85 * r27 + TJ(JHOBR) * r28 + TJ(JBRCL) * r29) /
                           (R.e2 * cc[O3] + R.e7 * cc[HO2] + R.c19 * ck[CH2O]);
        CCSET(BRO, .5 * cc[BRO] + .5 * ck[BROX] / (1. + r30));
        CCSET(BR, r30 * cc[BRO]);
        // OH from its quadratic budget (F:1600-1665)
        const double coea = 2. * A31 + 2. * (R.a17 + R.a27 * r6 + (A23B + A23C) * r4) * r6;
        const double coeb = R.b27 * ck[HNO3] * e.pno3 + R.b22 * cc[NO2] + R.b28 * ck[HO2NO2] +
                            R.b23 * cc[NO2] * r6 + R.c2 * ck[CH4] + R.c7 * ck[CH3O2] * r6 +
                            R.c8 * ck[CH2O] + R.c17a * ck[CH3O2H] + R.d7 * cc[CL] * r6 +
                            R.d9 * cc[CLO] + R.d11 * ck[HCL] * e.phcl + R.d33 * cc[CLO] * r6 +
                            R.d34 * ck[HOCL] + R.e11 * ck[HBR] * e.pbr + R.e7 * cc[BR] * r6 +
                            R.e15 * cc[BRO] * r6;
        const double coec =
            2. * (R.a1et * cc[O1D] + TJ(JH2O)) * ck[H2O] * e.ph2o +
            ((C1A + C1B) * ck[CH4] + 2. * A3ET * h2) * cc[O1D] + C3 * cc[CH3] * cc[O3P] +
            TJ(JCH3O2H) * ck[CH3O2H] + R.c15 * ch3o * o2 +
            (TJ(JCH2O_HCO) + R.c9 * cc[O3P]) * ck[CH2O] + C12 * hco * o2 +
            TJ(JHNO3) * ck[HNO3] * e.pno3 + (R.b24 * hnm + TJ(JHNO4)) * ck[HO2NO2] +
            (D35 * cc[O3P] + R.d37 * cc[CL] + TJ(JHOCL)) * ck[HOCL] + 2. * TJ(JH2O2) * ck[H2O2] +
            R.d6 * h2 * cc[CL] + TJ(JHOBR) * ck[HOBR] + R.e16 * cc[O3P] * ck[HOBR] +
            TJ(JHBR) * ck[HBR];
        const double argum = fmax(coeb * coeb + 4. * coea * coec, 0.);
        CCSET(OH, .5 * cc[OH] + .5 * (sqrt(argum) - coeb) / (2. * coea));
    } else {
        // night (F:1671-1746)
        day = 0.;
        CCSET(H, 0.);
        CCSET(CH3, 0.);
        ch3o = 0.;
        hco = 0.;
        CCSET(O1D, 0.);
        CCSET(O3P, 0.);
        CCSET(O3, ck[OX]);
        CCSET(OH, 0.);
        CCSET(CL, 0.);
        CCSET(HO2, 0.);
        CCSET(NO, 0.);
        const double r8 = R.b9 * cc[O3] / (R.b12 * fmax(cc[NO2], 1.e-30));
        CCSET(NO2, ck[NOX] / (1. + r8));
        CCSET(NO3, r8 * cc[NO2]);
        CCSET(N, 0.);
        CCSET(CLO, fmax(ck[CLOX], 0.));
        CCSET(BR, 0.);
        CCSET(BRO, fmax(ck[BROX], 0.));
    }

    // production P and loss L (F:1758-2358), then ck = (ck0 + dt P)/(1 + dt L)
    const double dt = e.dt;
    double nk[NLONG];
#define UPDATE(s, P, Lr) nk[s] = (ck0[(s) * CK0_STRIDE] + dt * (P)) / (1. + dt * (Lr))
    UPDATE(N2O, 0.e0, TJ(JN2O) + (R.b38 + R.b39) * cc[O1D]);
    UPDATE(CH4, 0.e0, R.c2 * cc[OH] + C1A * cc[O1D] + C1B * cc[O1D] + C1C * cc[O1D] + R.d5 * cc[CL]);
    UPDATE(H2O,
           R.c2 * cc[OH] * ck[CH4] + R.c8 * cc[OH] * ck[CH2O] + R.c17a * ck[CH3O2H] * cc[OH] +
               R.c17b * ck[CH3O2H] * cc[OH] + (R.a17 * cc[OH] + A23C * cc[H]) * cc[HO2] +
               R.a19 * cc[OH] * h2 + A30 * cc[OH] * ck[H2O2] + A31 * cc[OH] * cc[OH] +
               R.b27 * cc[OH] * ck[HNO3] * e.pno3 + R.d11 * cc[OH] * ck[HCL] * e.phcl,
           (TJ(JH2O) + R.a1et * cc[O1D]) * e.ph2o);
    UPDATE(NOY, 2. * R.b39 * cc[O1D] * ck[N2O],
           (2. * R.b6 * cc[N] * cc[NO] + e.rain_hno3 * ck[HNO3] + e.rain_hno4 * ck[HO2NO2]) / ck[NOY]);
    UPDATE(HNO3,
           R.b22 * cc[OH] * cc[NO2] + e.het1 * ck[CLONO2] * ck[H2O] + e.het2 * ck[CLONO2] * ck[HCL] +
               2. * e.het3 * ck[N2O5] * ck[H2O] + e.het4 * ck[N2O5] * ck[HCL] +
               e.het6 * ck[BRONO2] * ck[H2O],
           (TJ(JHNO3) + R.b27 * cc[OH]) * e.pno3 + e.rain_hno3);
    UPDATE(N2O5, R.b12 * cc[NO3] * cc[NO2],
           TJ(JN2O5) + R.b32 * hnm + e.het3 * ck[H2O] + e.het4 * ck[HCL]);
    UPDATE(CLY,
           3. * (TJ(JCFC11) + D90 * cc[O1D]) * ck[CFC11] + 2. * (TJ(JCFC12) + D91 * cc[O1D]) * ck[CFC12] +
               3. * (TJ(JCFC113) + D92 * cc[O1D]) * ck[CFC113] +
               4. * (TJ(JCCL4) + D93 * cc[O1D]) * ck[CCL4] +
               3. * (TJ(JCH3CCL3) + R.d94 * cc[OH]) * ck[CH3CCL3] +
               1. * (TJ(JCH3CL) + R.d95 * cc[OH]) * ck[CH3CL] +
               1. * (TJ(JHCFC22) + D96 * cc[O1D] + R.d97 * cc[OH]) * ck[HCFC22] +
               1. * (TJ(JH1211) + E52 * cc[O1D]) * ck[H1211],
           e.rain_hcl * ck[HCL] / ck[CLY]);
    {
        // odd oxygen (F:1864-1950)
        const double ratio = TJ(JCL2O2) / (TJ(JCL2O2) + R.d61 * hnm);
        const double fno3 = TJ(JNO3_NO) / fmax(TJ(JNO3_NO) + TJ(JNO3_NO2), 1.e-30);
        const double loss =
            2. * R.hk1 * cc[O3P] * cc[O3P] + 2. * R.hk3 * cc[O3] * cc[O3P] +
            2. * R.b3 * cc[NO2] * cc[O3P] + 2. * R.d3 * cc[CLO] * cc[O3P] +
            2. * R.d60 * ratio * cc[CLO] * cc[CLO] +
            fmin(TJ(JHOCL) * ck[HOCL], R.d33 * cc[HO2] * cc[CLO]) +
            2. * fmin(R.e5c * cc[BRO] * cc[CLO], TJ(JBRCL) * ck[BRCL]) +
            2. * R.e5b * cc[BRO] * cc[CLO] * day + fmin(R.e15 * cc[BRO] * cc[HO2], TJ(JHOBR) * ck[HOBR]) +
            2. * R.e6 * cc[BRO] * cc[BRO] * day + 2. * fno3 * TJ(JCLONO2_CL) * ck[CLONO2] +
            ((R.a2 * cc[H] + R.a6 * cc[OH] + R.a6b * cc[HO2]) * cc[O3] +
             (R.a5 * cc[OH] + R.a7 * cc[HO2]) * cc[O3P]) +
            2. * R.b9 * cc[NO2] * cc[O3] * TJ(JNO3_NO) / fmax(TJ(JNO3_NO) + TJ(JNO3_NO2), 1.e-30);
        UPDATE(OX, 2. * TJ(JO2) * o2 + R.a26 * cc[HO2] * cc[NO] + R.c5 * cc[NO] * ck[CH3O2],
               loss / ck[OX]);
    }
    UPDATE(CO, TJ(JCO2) * co2 + TJ(JCH2O_CO) * ck[CH2O] + C12 * hco * o2, R.a36 * cc[OH]);
    UPDATE(OCLO, R.e5a * cc[BRO] * cc[CLO], TJ(JOCLO));
    UPDATE(OXP, 0., 0.);
    UPDATE(H2SO4, 0., 0.);
    UPDATE(HCL,
           (R.c10 * ck[CH2O] + R.d5 * ck[CH4] + R.d7 * cc[HO2] + R.d6 * h2) * cc[CL] +
               R.d9 * cc[OH] * cc[CLO],
           (TJ(JHCL) + R.d11 * cc[OH]) * e.phcl + e.het2 * ck[CLONO2] + e.het4 * ck[N2O5] +
               e.het5 * ck[HOCL] + e.het7 * ck[HOBR] + e.rain_hcl);
    UPDATE(CLONO2, R.d31 * cc[NO2] * cc[CLO],
           TJ(JCLONO2_CL) + TJ(JCLONO2_CLO) + R.d32 * cc[O3P] + e.het1 * ck[H2O] + e.het2 * ck[HCL]);
    UPDATE(HOCL, R.d33 * cc[HO2] * cc[CLO] + e.het1 * ck[CLONO2] * ck[H2O],
           TJ(JHOCL) + R.d34 * cc[OH] + D35 * cc[O3P] + R.d37 * cc[CL] + e.het5 * ck[HCL] +
               e.het9 * ck[HBR]);
    UPDATE(CL2,
           R.d37 * ck[HOCL] * cc[CL] + R.d62 * cc[CLO] * cc[CLO] + e.het2 * ck[CLONO2] * ck[HCL] +
               e.het5 * ck[HOCL] * ck[HCL],
           TJ(JCL2));
    UPDATE(H2O2, R.a27 * cc[HO2] * cc[HO2], TJ(JH2O2) + A30 * cc[OH] + e.rain_h2o2);
    UPDATE(CLNO2, R.d36 * cc[CL] * cc[NO2] + e.het4 * ck[N2O5] * ck[HCL], TJ(JCLNO2));
    UPDATE(HBR, R.e7 * cc[HO2] * cc[BR] + R.c19 * ck[CH2O] * cc[BR],
           (R.e11 * cc[OH] + E9 * cc[O1D] + TJ(JHBR)) * e.pbr + e.het8 * ck[HOBR] +
               e.het9 * ck[HOCL] + e.rain_hbr);
    UPDATE(BRONO2, R.e13 * cc[BRO] * cc[NO2], TJ(JBRONO2) + R.e8 * cc[O3P] + e.het6 * ck[H2O]);
    UPDATE(NOX,
           (TJ(JHNO4) + R.b28 * cc[OH] + R.b24 * hnm) * ck[HO2NO2] + TJ(JHNO3) * ck[HNO3] * e.pno3 +
               2. * (TJ(JN2O5) + R.b32 * hnm) * ck[N2O5] + 2. * R.b39 * cc[O1D] * ck[N2O] +
               TJ(JCLNO2) * ck[CLNO2] + R.b27 * ck[HNO3] * e.pno3 * cc[OH] +
               R.d32 * ck[CLONO2] * cc[O3P] + (TJ(JCLONO2_CL) + TJ(JCLONO2_CLO)) * ck[CLONO2] +
               TJ(JBRONO2) * ck[BRONO2] + R.e8 * ck[BRONO2] * cc[O3P],
           ((R.d31 * cc[CLO] + R.b22 * cc[OH] + R.b23 * cc[HO2] + 2. * R.b12 * cc[NO3] +
             R.d36 * cc[CL] + R.e13 * cc[BRO]) * cc[NO2] +
            2. * R.b6 * cc[N] * cc[NO]) / ck[NOX]);
    UPDATE(HO2NO2, R.b23 * cc[HO2] * cc[NO2], TJ(JHNO4) + R.b28 * cc[OH] + R.b24 * hnm + e.rain_hno4);
    UPDATE(CLOX,
           2. * (R.d61 * hnm + TJ(JCL2O2)) * ck[CL2O2] +
               (TJ(JCLONO2_CL) + TJ(JCLONO2_CLO) + R.d32 * cc[O3P]) * ck[CLONO2] +
               (TJ(JHOCL) + R.d34 * cc[OH] + D35 * cc[O3P]) * ck[HOCL] +
               (TJ(JOCLO) * ck[OCLO] + (TJ(JHCL) + R.d11 * cc[OH]) * ck[HCL] * e.phcl) +
               2. * TJ(JCL2) * ck[CL2] + TJ(JCLNO2) * ck[CLNO2] + TJ(JBRCL) * ck[BRCL] +
               3. * (TJ(JCFC11) + D90 * cc[O1D]) * ck[CFC11] +
               2. * (TJ(JCFC12) + D91 * cc[O1D]) * ck[CFC12] +
               3. * (TJ(JCFC113) + D92 * cc[O1D]) * ck[CFC113] +
               4. * (TJ(JCCL4) + D93 * cc[O1D]) * ck[CCL4] +
               3. * (TJ(JCH3CCL3) + R.d94 * cc[OH]) * ck[CH3CCL3] +
               1. * (TJ(JCH3CL) + R.d95 * cc[OH]) * ck[CH3CL] +
               1. * (TJ(JHCFC22) + D96 * cc[O1D] + R.d97 * cc[OH]) * ck[HCFC22],
           ((R.d31 * cc[NO2] + 2. * R.d60 * cc[CLO] + 2. * R.d62 * cc[CLO] + R.e5a * cc[BRO] +
             R.e5c * cc[BRO] + R.d9 * cc[OH] + R.d33 * cc[HO2]) * cc[CLO] +
            (R.c10 * ck[CH2O] + R.d5 * ck[CH4] + R.d7 * cc[HO2] + R.d6 * h2 + R.d36 * cc[NO2] +
             R.d37 * ck[HOCL]) * cc[CL]) / ck[CLOX]);
    UPDATE(BROX,
           TJ(JHOBR) * ck[HOBR] + TJ(JBRONO2) * ck[BRONO2] + R.e8 * ck[BRONO2] * cc[O3P] +
               R.e16 * ck[HOBR] * cc[O3P] + TJ(JBRCL) * ck[BRCL] +
               (R.e11 * cc[OH] + E9 * cc[O1D] + TJ(JHBR)) * ck[HBR] * e.pbr +
               2. * e.het8 * ck[HBR] * ck[HOBR],
           ((R.e13 * cc[NO2] + R.e5c * cc[CLO] + R.e15 * cc[HO2]) * cc[BRO] +
            R.c19 * cc[BR] * ck[CH2O] + R.e7 * cc[BR] * cc[HO2]) / ck[BROX]);
    UPDATE(CL2O2, R.d60 * cc[CLO] * cc[CLO], TJ(JCL2O2) + R.d61 * hnm);
    UPDATE(HOBR, R.e15 * cc[BRO] * cc[HO2] + e.het6 * ck[BRONO2] * ck[H2O],
           TJ(JHOBR) + R.e16 * cc[O3P] + e.het7 * ck[HCL] + e.het8 * ck[HBR]);
    UPDATE(BRCL,
           R.e5c * cc[CLO] * cc[BRO] + e.het7 * ck[HOBR] * ck[HCL] + e.het9 * ck[HBR] * ck[HOCL],
           TJ(JBRCL));
    UPDATE(CH2O,
           (C1C * cc[O1D] * ck[CH4] + C3 * cc[O3P] * cc[CH3] + R.c15 * ch3o * o2 +
            R.c17b * cc[OH] * ck[CH3O2H]),
           R.c8 * cc[OH] + R.c9 * cc[O3P] + R.c10 * cc[CL] + R.c19 * cc[BR] + TJ(JCH2O_HCO) +
               TJ(JCH2O_CO) + e.rain_ch2o);
    UPDATE(CH3O2, R.c4 * cc[CH3] * o2 + R.c17a * cc[OH] * ck[CH3O2H],
           2. * R.c14 * ck[CH3O2] + R.c7 * cc[HO2] + R.c5 * cc[NO] + R.c18 * cc[CLO]);
    UPDATE(CH3O2H, R.c7 * ck[CH3O2] * cc[HO2],
           R.c17a * cc[OH] + R.c17b * cc[OH] + TJ(JCH3O2H) + e.rain_ch3o2h);
    UPDATE(CFC11, 0., TJ(JCFC11) + D90 * cc[O1D]);
    UPDATE(CFC12, 0., TJ(JCFC12) + D91 * cc[O1D]);
    UPDATE(CFC113, 0., TJ(JCFC113) + D92 * cc[O1D]);
    UPDATE(CCL4, 0., TJ(JCCL4) + D93 * cc[O1D]);
    UPDATE(CH3CCL3, 0., TJ(JCH3CCL3) + R.d94 * cc[OH]);
    UPDATE(CH3CL, 0., TJ(JCH3CL) + R.d95 * cc[OH]);
    UPDATE(HCFC22, 0., TJ(JHCFC22) + D96 * cc[O1D] + R.d97 * cc[OH]);
    UPDATE(CH3BR, 0., TJ(JCH3BR) + E50 * cc[O1D] + R.e51 * cc[OH]);
    UPDATE(H1211, 0., TJ(JH1211) + E52 * cc[O1D]);
    UPDATE(H1301, 0., TJ(JH1301) + E53 * cc[O1D]);
    UPDATE(BRY,
           (TJ(JCH3BR) + E50 * cc[O1D] + R.e51 * cc[OH]) * ck[CH3BR] +
               (TJ(JH1211) + E52 * cc[O1D]) * ck[H1211] + (TJ(JH1301) + E53 * cc[O1D]) * ck[H1301] +
               2. * (R.e54 * cc[OH] + TJ(JCH2BR2)) * ck[CH2BR2],
           e.rain_hbr * ck[HBR] / ck[BRY]);
    UPDATE(CH2BR2, 0., TJ(JCH2BR2) + R.e54 * cc[OH]);
#undef UPDATE
#pragma unroll
    for (int s = 0; s < NLONG - 1; ++s) {
        if (TRACK) track(tr, nk[s], ck[s]);
        ck[s] = nk[s];
    }
    if (TRACK) track_end(tr);
#undef CCSET
#undef TJ
}

// ------------------------------------------------------------------ kernels
__global__ void k_columns(const double *__restrict__ p, const double *__restrict__ T,
                          const double *__restrict__ hc, double *__restrict__ col, int64_t plane,
                          int64_t field, int32_t nivbas)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    const int64_t nb = plane * (int64_t)nivbas;
    const double zcmt_t = 2.120e20, zcmt_j = 287.0 / (1.38e-23 * 9.81), r_o2 = 0.2095;
    const double *o3 = hc + (int64_t)O3 * field, *no = hc + (int64_t)NO * field;
    double o3t = 0., co2 = 0., cno = 0., hprev = 0., nprev = 0., pprev = 0., o3prev = 0.;
    // the levels in groups of COL_GROUP: a group's loads are issued together, then the
    // scan runs through it in level order (one memory round trip per group, not per level)
    constexpr int COL_GROUP = 8;
    for (int k0 = 0; k0 < nivbas; k0 += COL_GROUP) {
        double pv[COL_GROUP], tv[COL_GROUP], nv[COL_GROUP], ov[COL_GROUP];
#pragma unroll
        for (int u = 0; u < COL_GROUP; ++u) {
            const int64_t b = (int64_t)(k0 + u < nivbas ? k0 + u : nivbas - 1) * plane + c;
            pv[u] = p[b];
            tv[u] = T[b];
            nv[u] = no[b];
            ov[u] = o3[b];
        }
#pragma unroll
        for (int u = 0; u < COL_GROUP; ++u) {
            const int k = k0 + u;
            if (k >= nivbas) break;
            const int64_t b = (int64_t)k * plane + c;
            const double pk = pv[u], hnm = pk / (tv[u] * 1.38e-19), ccno = nv[u] * hnm, o3k = ov[u];
            if (k == 0) {
                o3t = o3k * hnm * 6.8e5;                       // total, F:5170
                co2 = 6.8e+05 * r_o2 * hnm;                     // jno, J:109
                cno = 6.8e+05 * ccno;                           // jno, J:110
            } else {
                const double dp = fmax((pk - pprev) * 100., 0.);
                o3t = o3t + zcmt_t * (o3k + o3prev) * .5 * dp;  // F:5176-5181
                co2 = co2 + zcmt_j * (r_o2 + r_o2) * 0.5 * dp * 1.e-4;
                cno = cno + zcmt_j * (nprev / hprev + ccno / hnm) * 0.5 * dp * 1.e-4;
            }
            col[b] = o3t;
            col[nb + b] = co2;
            col[2 * nb + b] = cno;
            hprev = hnm;
            nprev = ccno;
            pprev = pk;
            o3prev = o3k;
        }
    }
}

// Per-box state between set-up, sweeps and write-back.
struct Box {
    BoxEnv e;
    Rates R;
    double tj[NJ];
    double ck[NLONG], cc[NSHORT];
    double ch3o, hco;
    Het h;
    bool het_level;
};

// Write-back (F:2380-2410), optionally conserv (F:2541-2598) and the floor (F:431-440).
__device__ __forceinline__ void box_writeback(const StepArgs &a, int64_t b, const Box &x)
{
    const double hnm = x.e.hnm;
    double q[NLONG], hs[NSHORT];
#pragma unroll
    for (int s = 0; s < NLONG - 1; ++s) q[s] = x.ck[s] / hnm;
    q[HNO3G] = x.ck[HNO3] * x.e.pno3 / hnm;
#pragma unroll
    for (int s = 0; s < NSHORT; ++s) hs[s] = x.cc[s] / hnm;
    if (a.fuse_conserv) {
        const double rcly = (q[CLY] - q[BRCL]) /
                            (q[OCLO] + q[HCL] + q[CLONO2] + q[HOCL] + 2. * q[CL2] + q[CLNO2] +
                             q[CLOX] + 2. * q[CL2O2]);
        hs[CL] *= rcly; hs[CLO] *= rcly;
        q[OCLO] *= rcly; q[HCL] *= rcly; q[CLONO2] *= rcly; q[HOCL] *= rcly; q[CL2] *= rcly;
        q[CLNO2] *= rcly; q[CLOX] *= rcly; q[CL2O2] *= rcly;
        const double rnoy = fmax((q[NOY] - q[CLONO2] - q[CLNO2] - q[BRONO2]), 0.) /
                            (q[HNO3] + 2. * q[N2O5] + q[NOX] + q[HO2NO2]);
        hs[NO2] *= rnoy; hs[NO] *= rnoy; hs[N] *= rnoy; hs[NO3] *= rnoy;
        q[HNO3] *= rnoy; q[N2O5] *= rnoy; q[NOX] *= rnoy; q[HO2NO2] *= rnoy;
        const double rbry = q[BRY] / (q[HBR] + q[BRONO2] + q[BROX] + q[HOBR] + q[BRCL]);
        hs[BR] *= rbry; hs[BRO] *= rbry;
        q[HBR] *= rbry; q[BRONO2] *= rbry; q[BROX] *= rbry; q[HOBR] *= rbry; q[BRCL] *= rbry;
#pragma unroll
        for (int s = 0; s < NLONG - 1; ++s) q[s] = fmax(q[s], 1.e-30);
    }
#pragma unroll
    for (int s = 0; s < NLONG; ++s) a.qj1[(int64_t)s * a.field + b] = q[s];
#pragma unroll
    for (int s = 0; s < NSHORT; ++s) a.hc[(int64_t)s * a.field + b] = hs[s];
    if (x.het_level) a.surfarea[b] = x.h.surf;
    a.vsed[b] = x.h.vsed;
    a.p2o[b] = x.h.ph2o;
    a.pno3[b] = x.h.pno3sl;
}


// Per-box set-up record: everything the set-up computes that the sweeps read, once each,
// stored field-major (rec[f * nbox + b]) so that a wave's loads and stores of one field
// coalesce -- the air density (the other densities are fixed fractions of it), the
// rain-out rate, the day and level flags, the 82 thermal rates, the 41 J-values and the 15
// aerosol terms: 141 doubles per box.  Every field is written and read unconditionally:
// with the J-values and aerosol terms stored only where they are not zero, the loads sat
// behind branches and the sweep kernel took 0.2 ms longer.
constexpr int REC_HNM = 0, REC_RAIN = 1, REC_FLAGS = 2, REC_R = 3, REC_J = REC_R + 82,
              REC_H = REC_J + NJ, NREC = REC_H + 15;
static_assert(NJ == 41, "record layout assumes 41 J-values");
static_assert(NREC == CONV_RECORD, "record size");
static_assert(sizeof(Rates) == 82 * sizeof(double), "record layout assumes 82 rates");

// Fields LO <= f < HI of the record (one stage of the set-up at a time).
template <int LO, int HI>
__device__ __forceinline__ void rec_store(const Box &x, double *__restrict__ rec, int64_t nbox,
                                          int64_t b)
{
#define REC_PUT(f, v) \
    if ((f) >= LO && (f) < HI) rec[(int64_t)(f) * nbox + b] = (v)
    REC_PUT(REC_HNM, x.e.hnm);
    REC_PUT(REC_RAIN, x.e.rain_hno3);
    REC_PUT(REC_FLAGS, (x.e.day ? 1. : 0.) + (x.het_level ? 2. : 0.));
    REC_PUT(REC_R + 0, x.R.a1);
    REC_PUT(REC_R + 1, x.R.a1et);
    REC_PUT(REC_R + 2, x.R.a2);
    REC_PUT(REC_R + 3, x.R.a5);
    REC_PUT(REC_R + 4, x.R.a6);
    REC_PUT(REC_R + 5, x.R.a6b);
    REC_PUT(REC_R + 6, x.R.a7);
    REC_PUT(REC_R + 7, x.R.a17);
    REC_PUT(REC_R + 8, x.R.a19);
    REC_PUT(REC_R + 9, x.R.a26);
    REC_PUT(REC_R + 10, x.R.a27);
    REC_PUT(REC_R + 11, x.R.a36);
    REC_PUT(REC_R + 12, x.R.b3);
    REC_PUT(REC_R + 13, x.R.b4);
    REC_PUT(REC_R + 14, x.R.b6);
    REC_PUT(REC_R + 15, x.R.b7);
    REC_PUT(REC_R + 16, x.R.b9);
    REC_PUT(REC_R + 17, x.R.b12);
    REC_PUT(REC_R + 18, x.R.b22);
    REC_PUT(REC_R + 19, x.R.b23);
    REC_PUT(REC_R + 20, x.R.b24);
    REC_PUT(REC_R + 21, x.R.b27);
    REC_PUT(REC_R + 22, x.R.b28);
    REC_PUT(REC_R + 23, x.R.b32);
    REC_PUT(REC_R + 24, x.R.b38);
    REC_PUT(REC_R + 25, x.R.b39);
    REC_PUT(REC_R + 26, x.R.c2);
    REC_PUT(REC_R + 27, x.R.c4);
    REC_PUT(REC_R + 28, x.R.c5);
    REC_PUT(REC_R + 29, x.R.c7);
    REC_PUT(REC_R + 30, x.R.c8);
    REC_PUT(REC_R + 31, x.R.c9);
    REC_PUT(REC_R + 32, x.R.c10);
    REC_PUT(REC_R + 33, x.R.c14);
    REC_PUT(REC_R + 34, x.R.c15);
    REC_PUT(REC_R + 35, x.R.c17a);
    REC_PUT(REC_R + 36, x.R.c17b);
    REC_PUT(REC_R + 37, x.R.c18);
    REC_PUT(REC_R + 38, x.R.c19);
    REC_PUT(REC_R + 39, x.R.d2);
    REC_PUT(REC_R + 40, x.R.d3);
    REC_PUT(REC_R + 41, x.R.d4);
    REC_PUT(REC_R + 42, x.R.d5);
    REC_PUT(REC_R + 43, x.R.d6);
    REC_PUT(REC_R + 44, x.R.d7);
    REC_PUT(REC_R + 45, x.R.d8);
    REC_PUT(REC_R + 46, x.R.d9);
    REC_PUT(REC_R + 47, x.R.d10);
    REC_PUT(REC_R + 48, x.R.d11);
    REC_PUT(REC_R + 49, x.R.d31);
    REC_PUT(REC_R + 50, x.R.d32);
    REC_PUT(REC_R + 51, x.R.d33);
    REC_PUT(REC_R + 52, x.R.d34);
    REC_PUT(REC_R + 53, x.R.d36);
    REC_PUT(REC_R + 54, x.R.d37);
    REC_PUT(REC_R + 55, x.R.d60);
    REC_PUT(REC_R + 56, x.R.d61);
    REC_PUT(REC_R + 57, x.R.d62);
    REC_PUT(REC_R + 58, x.R.d94);
    REC_PUT(REC_R + 59, x.R.d95);
    REC_PUT(REC_R + 60, x.R.d97);
    REC_PUT(REC_R + 61, x.R.e2);
    REC_PUT(REC_R + 62, x.R.e3);
    REC_PUT(REC_R + 63, x.R.e4);
    REC_PUT(REC_R + 64, x.R.e5a);
    REC_PUT(REC_R + 65, x.R.e5b);
    REC_PUT(REC_R + 66, x.R.e5c);
    REC_PUT(REC_R + 67, x.R.e6);
    REC_PUT(REC_R + 68, x.R.e7);
    REC_PUT(REC_R + 69, x.R.e8);
    REC_PUT(REC_R + 70, x.R.e10);
    REC_PUT(REC_R + 71, x.R.e11);
    REC_PUT(REC_R + 72, x.R.e13);
    REC_PUT(REC_R + 73, x.R.e15);
    REC_PUT(REC_R + 74, x.R.e16);
    REC_PUT(REC_R + 75, x.R.e51);
    REC_PUT(REC_R + 76, x.R.e54);
    REC_PUT(REC_R + 77, x.R.hk1);
    REC_PUT(REC_R + 78, x.R.hk2);
    REC_PUT(REC_R + 79, x.R.hk3);
    REC_PUT(REC_R + 80, x.R.hk4);
    REC_PUT(REC_R + 81, x.R.hk5);
#pragma unroll
    for (int n = 0; n < NJ; ++n) REC_PUT(REC_J + n, x.tj[n]);
    REC_PUT(REC_H + 0, x.h.k[0]);
    REC_PUT(REC_H + 1, x.h.k[1]);
    REC_PUT(REC_H + 2, x.h.k[2]);
    REC_PUT(REC_H + 3, x.h.k[3]);
    REC_PUT(REC_H + 4, x.h.k[4]);
    REC_PUT(REC_H + 5, x.h.k[5]);
    REC_PUT(REC_H + 6, x.h.k[6]);
    REC_PUT(REC_H + 7, x.h.k[7]);
    REC_PUT(REC_H + 8, x.h.k[8]);
    REC_PUT(REC_H + 9, x.h.pno3);
    REC_PUT(REC_H + 10, x.h.pno3sl);
    REC_PUT(REC_H + 11, x.h.ph2o);
    REC_PUT(REC_H + 12, x.h.phcl);
    REC_PUT(REC_H + 13, x.h.surf);
    REC_PUT(REC_H + 14, x.h.vsed);
#undef REC_PUT
}

// The set-up of box b from its record, as the set-up left it (rec_load is exact: the
// densities are the same products of the air density, the flags small integers).
template <bool LOAD_R = true, bool LOAD_ENV = true, bool LOAD_J = true>
__device__ __forceinline__ void rec_load(Box &x, const double *__restrict__ rec, int64_t nbox,
                                         int64_t b, double dt)
{
#define REC_GET(f) rec[(int64_t)(f) * nbox + b]
    BoxEnv &e = x.e;
    if (LOAD_ENV) {
    e.hnm = REC_GET(REC_HNM);
    e.n2 = 0.78 * e.hnm;
    e.o2 = 0.21 * e.hnm;
    e.co2 = 360.e-6 * e.hnm;
    e.h2 = 5.e-7 * e.hnm;
    e.dt = dt;
    const double rr = REC_GET(REC_RAIN);
    e.rain_hno3 = rr; e.rain_hcl = rr; e.rain_hbr = rr; e.rain_h2o2 = rr;
    e.rain_hno4 = 0.; e.rain_ch2o = 0.; e.rain_ch3o2h = 0.;
    const double fl = REC_GET(REC_FLAGS);
    e.day = fl == 1. || fl == 3.;
    x.het_level = fl >= 2.;
    }
    if (LOAD_R) {
    x.R.a1 = REC_GET(REC_R + 0);
    x.R.a1et = REC_GET(REC_R + 1);
    x.R.a2 = REC_GET(REC_R + 2);
    x.R.a5 = REC_GET(REC_R + 3);
    x.R.a6 = REC_GET(REC_R + 4);
    x.R.a6b = REC_GET(REC_R + 5);
    x.R.a7 = REC_GET(REC_R + 6);
    x.R.a17 = REC_GET(REC_R + 7);
    x.R.a19 = REC_GET(REC_R + 8);
    x.R.a26 = REC_GET(REC_R + 9);
    x.R.a27 = REC_GET(REC_R + 10);
    x.R.a36 = REC_GET(REC_R + 11);
    x.R.b3 = REC_GET(REC_R + 12);
    x.R.b4 = REC_GET(REC_R + 13);
    x.R.b6 = REC_GET(REC_R + 14);
    x.R.b7 = REC_GET(REC_R + 15);
    x.R.b9 = REC_GET(REC_R + 16);
    x.R.b12 = REC_GET(REC_R + 17);
    x.R.b22 = REC_GET(REC_R + 18);
    x.R.b23 = REC_GET(REC_R + 19);
    x.R.b24 = REC_GET(REC_R + 20);
    x.R.b27 = REC_GET(REC_R + 21);
    x.R.b28 = REC_GET(REC_R + 22);
    x.R.b32 = REC_GET(REC_R + 23);
    x.R.b38 = REC_GET(REC_R + 24);
    x.R.b39 = REC_GET(REC_R + 25);
    x.R.c2 = REC_GET(REC_R + 26);
    x.R.c4 = REC_GET(REC_R + 27);
    x.R.c5 = REC_GET(REC_R + 28);
    x.R.c7 = REC_GET(REC_R + 29);
    x.R.c8 = REC_GET(REC_R + 30);
    x.R.c9 = REC_GET(REC_R + 31);
    x.R.c10 = REC_GET(REC_R + 32);
    x.R.c14 = REC_GET(REC_R + 33);
    x.R.c15 = REC_GET(REC_R + 34);
    x.R.c17a = REC_GET(REC_R + 35);
    x.R.c17b = REC_GET(REC_R + 36);
    x.R.c18 = REC_GET(REC_R + 37);
    x.R.c19 = REC_GET(REC_R + 38);
    x.R.d2 = REC_GET(REC_R + 39);
    x.R.d3 = REC_GET(REC_R + 40);
    x.R.d4 = REC_GET(REC_R + 41);
    x.R.d5 = REC_GET(REC_R + 42);
    x.R.d6 = REC_GET(REC_R + 43);
    x.R.d7 = REC_GET(REC_R + 44);
    x.R.d8 = REC_GET(REC_R + 45);
    x.R.d9 = REC_GET(REC_R + 46);
    x.R.d10 = REC_GET(REC_R + 47);
    x.R.d11 = REC_GET(REC_R + 48);
    x.R.d31 = REC_GET(REC_R + 49);
    x.R.d32 = REC_GET(REC_R + 50);
    x.R.d33 = REC_GET(REC_R + 51);
    x.R.d34 = REC_GET(REC_R + 52);
    x.R.d36 = REC_GET(REC_R + 53);
    x.R.d37 = REC_GET(REC_R + 54);
    x.R.d60 = REC_GET(REC_R + 55);
    x.R.d61 = REC_GET(REC_R + 56);
    x.R.d62 = REC_GET(REC_R + 57);
    x.R.d94 = REC_GET(REC_R + 58);
    x.R.d95 = REC_GET(REC_R + 59);
    x.R.d97 = REC_GET(REC_R + 60);
    x.R.e2 = REC_GET(REC_R + 61);
    x.R.e3 = REC_GET(REC_R + 62);
    x.R.e4 = REC_GET(REC_R + 63);
    x.R.e5a = REC_GET(REC_R + 64);
    x.R.e5b = REC_GET(REC_R + 65);
    x.R.e5c = REC_GET(REC_R + 66);
    x.R.e6 = REC_GET(REC_R + 67);
    x.R.e7 = REC_GET(REC_R + 68);
    x.R.e8 = REC_GET(REC_R + 69);
    x.R.e10 = REC_GET(REC_R + 70);
    x.R.e11 = REC_GET(REC_R + 71);
    x.R.e13 = REC_GET(REC_R + 72);
    x.R.e15 = REC_GET(REC_R + 73);
    x.R.e16 = REC_GET(REC_R + 74);
    x.R.e51 = REC_GET(REC_R + 75);
    x.R.e54 = REC_GET(REC_R + 76);
    x.R.hk1 = REC_GET(REC_R + 77);
    x.R.hk2 = REC_GET(REC_R + 78);
    x.R.hk3 = REC_GET(REC_R + 79);
    x.R.hk4 = REC_GET(REC_R + 80);
    x.R.hk5 = REC_GET(REC_R + 81);
    }
    if (LOAD_J) {
#pragma unroll
    for (int n = 0; n < NJ; ++n) x.tj[n] = REC_GET(REC_J + n);
    }
    x.h.k[0] = REC_GET(REC_H + 0);
    x.h.k[1] = REC_GET(REC_H + 1);
    x.h.k[2] = REC_GET(REC_H + 2);
    x.h.k[3] = REC_GET(REC_H + 3);
    x.h.k[4] = REC_GET(REC_H + 4);
    x.h.k[5] = REC_GET(REC_H + 5);
    x.h.k[6] = REC_GET(REC_H + 6);
    x.h.k[7] = REC_GET(REC_H + 7);
    x.h.k[8] = REC_GET(REC_H + 8);
    x.h.pno3 = REC_GET(REC_H + 9);
    x.h.pno3sl = REC_GET(REC_H + 10);
    x.h.ph2o = REC_GET(REC_H + 11);
    x.h.phcl = REC_GET(REC_H + 12);
    x.h.surf = REC_GET(REC_H + 13);
    x.h.vsed = REC_GET(REC_H + 14);
    e.het1 = x.h.k[0]; e.het2 = x.h.k[1]; e.het3 = x.h.k[2]; e.het4 = x.h.k[3];
    e.het5 = x.h.k[4]; e.het6 = x.h.k[5]; e.het7 = x.h.k[6]; e.het8 = x.h.k[7];
    e.het9 = x.h.k[8];
    e.pno3 = x.h.pno3;
    e.ph2o = x.h.ph2o;
    e.phcl = x.h.phcl;
    e.pbr = 1.;
#undef REC_GET
}

// Set-up of every box (F:800-1338 without the sweeps): thermal rates, then J-values, then
// the PSC/aerosol block, densities and rain-out, each stage stored as soon as it is
// computed, so that no stage's registers outlive it.  At two waves per SIMD (~250 VGPRs)
// the transcendental chains and table gathers of one wave overlap another's.
// Boxes box0 .. box0 + nr - 1; the record of box box0 + r at index r (stride nr).
// The air density and the fixed fractions of it, the day flag and the level flag of box b
// (chem3d F:800-851: hnm = p / (1.38e-19 T), N2, O2, CO2, H2; day where SZA <= 95;
// heterogeneous chemistry on levels nivhau..nivbas).
__device__ __forceinline__ void box_env(const StepArgs &a, int64_t b, double t, double pm,
                                        Box &x)
{
    BoxEnv &e = x.e;
    e.hnm = pm / (t * 1.38e-19);
    e.n2 = 0.78 * e.hnm;
    e.o2 = 0.21 * e.hnm;
    e.co2 = 360.e-6 * e.hnm;
    e.h2 = 5.e-7 * e.hnm;
    e.dt = a.dt;
    e.day = a.sza[b] <= 95.;
    const int k = (int)(b / a.plane);
    x.het_level = (k + 1 >= a.nivhau) && (k + 1 <= a.nivbas);
}

// rain-out above 100 ppmv water (F:1315-1338), from the box's water density
__device__ __forceinline__ void box_rain(Box &x)
{
    BoxEnv &e = x.e;
    const bool wet = x.ck[H2O] / e.hnm >= 100.e-6;
    const double rr = wet ? 1.e-6 : 0.;
    e.rain_hno3 = rr; e.rain_hcl = rr; e.rain_hbr = rr; e.rain_h2o2 = rr;
    e.rain_hno4 = 0.; e.rain_ch2o = 0.; e.rain_ch3o2h = 0.;
}

// the 41 J-values of box b (phot F:4996-5149 on the column's O3, jno J:1-249 on the column
// sums and the column top's SZA)
__device__ __forceinline__ void box_j(const StepArgs &a, int64_t b, Box &x)
{
    const int64_t c = b % a.plane;
    table_j(a.ajl, a.o3up, a.alt[b], a.sza[b], a.col[b], x.tj);
    x.tj[JNO] = j_no(a.col[a.nbox + b], a.col[2 * a.nbox + b], a.col[b], x.e.hnm, a.sza[c]);
}

// What a set-up record holds: everything (the converged mode, whose persistent kernels set
// a box up at each refill), or what the fixed-sweep step's k_chem does not compute itself:
// the J-values and the aerosol terms.  (With the J-values computed in k_chem too, the c1
// pair took 0.910 / 0.948 ms at block 64 / 128 against 0.914 / 0.940: no gain for 320 table
// gathers more in the kernel that waits on its loads.)
enum RecMode : int { REC_FULL = 0, REC_STEP = 1 };

template <int BLK, int MODE>
__global__ void __launch_bounds__(BLK, 2) k_setup(StepArgs a, double *__restrict__ rec,
                                                  int64_t box0, int64_t nr)
{
    constexpr bool FULL = MODE == REC_FULL;
    const int64_t r = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (r >= nr) return;
    const int64_t b = box0 + r;
    const double t = a.T[b], pm = a.p[b];
    Box x;
    BoxEnv &e = x.e;
    box_env(a, b, t, pm, x);
    if (FULL) {
        thermal_rates(a.K, t, e.hnm, pm, x.R);
        rec_store<REC_R, REC_J>(x, rec, nr, r);
    }
    box_j(a, b, x);
    rec_store<REC_J, REC_H>(x, rec, nr, r);
#pragma unroll
    for (int s = 0; s < NLONG; ++s) x.ck[s] = a.qj1[(int64_t)s * a.field + b] * e.hnm;
    Het &h = x.h;
    if (x.het_level) {
        psc_and_het(t, pm, e.hnm, x.ck, h);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) h.k[i] = 0.;
        h.pno3 = h.pno3sl = h.ph2o = h.phcl = 1.;
        h.vsed = 0.;
        h.surf = 0.;
    }
    e.het1 = h.k[0]; e.het2 = h.k[1]; e.het3 = h.k[2]; e.het4 = h.k[3]; e.het5 = h.k[4];
    e.het6 = h.k[5]; e.het7 = h.k[6]; e.het8 = h.k[7]; e.het9 = h.k[8];
    e.pno3 = h.pno3;
    e.ph2o = h.ph2o;
    e.phcl = h.phcl;
    e.pbr = 1.;
    if (FULL) {
        box_rain(x);
        rec_store<REC_HNM, REC_R>(x, rec, nr, r);
    }
    rec_store<REC_H, NREC>(x, rec, nr, r);
}

// The state of box b from its set-up record (chem3d before its sweep loop, F:800-1338),
// record r of stride nr: the iterate from the incoming state (or, with use_guess and a
// guess, from the guess), the start-of-step state to the LDS column ck0 (stride BLK).
template <int BLK, int MODE = REC_FULL>
__device__ __forceinline__ void box_load(const StepArgs &a, const double *__restrict__ rec,
                                         int64_t nr, int64_t r, int64_t b, Box &x, double *ck0,
                                         bool use_guess)
{
    constexpr bool FULL = MODE == REC_FULL;
    if (!FULL) {
        // what the step's record leaves out: k_setup's own computations from the same
        // inputs, so the same bits
        const double t = a.T[b], pm = a.p[b];
        box_env(a, b, t, pm, x);
        thermal_rates(a.K, t, x.e.hnm, pm, x.R);
    }
    rec_load<FULL, FULL, true>(x, rec, nr, r, a.dt);
    const double hnm = x.e.hnm;
#pragma unroll
    for (int s = 0; s < NLONG; ++s) {
        x.ck[s] = a.qj1[(int64_t)s * a.field + b] * hnm;
        if (s < NLONG - 1) ck0[s * BLK] = x.ck[s];
    }
#pragma unroll
    for (int s = 0; s < NSHORT; ++s) x.cc[s] = a.hc[(int64_t)s * a.field + b] * hnm;
    x.ch3o = 0.;
    x.hco = 0.;
    if (!FULL) box_rain(x);
    if (use_guess && a.guess_q) {
#pragma unroll
        for (int s = 0; s < NLONG - 1; ++s) x.ck[s] = a.guess_q[(int64_t)s * a.field + b] * hnm;
#pragma unroll
        for (int s = 0; s < NSHORT; ++s) x.cc[s] = a.guess_h[(int64_t)s * a.field + b] * hnm;
    }
}

// Fixed-sweep step (the reference scheme, F:1348 niter = 5) from the set-up record: one
// lane per box; conserv and the floor fused into the write-back when asked.  The set-up
// (k_setup) runs as its own launch at two waves per SIMD: in one kernel with the sweeps it
// ran at the sweeps' one wave per SIMD (the box state needs ~2 KB of registers) and took
// three quarters of the step.  The 82 thermal rates, the air density and the flags are
// the exception: this kernel computes them from T, p and SZA (the step's record leaves
// them out) -- its VALU is idle three quarters of the time while its loads are in flight,
// and 85 record fields, written and read back, cost more (the c1 pair at block 64: 0.914 ms
// against 1.004 with the full record).
// Boxes box0 .. box0 + nr - 1 from records 0 .. nr - 1 (k_setup's range).
template <int BLK>
__global__ void __launch_bounds__(BLK) k_chem(StepArgs a, const double *__restrict__ rec,
                                              int64_t box0, int64_t nr)
{
    // start-of-step state of the block's boxes, species-major: lane-private columns,
    // conflict-free, and off the register file that the sweep needs for rates and J
    // (and NJL of its J-values, the same way)
    static_assert((NLONG - 1 + NJL) * BLK * sizeof(double) <= 160 * 1024, "LDS per workgroup");
    __shared__ double s_ck0[(NLONG - 1 + NJL) * BLK];
    double *ck0 = s_ck0 + threadIdx.x;
    double *tjl = s_ck0 + (NLONG - 1) * BLK + threadIdx.x;
    // last blocks first: k_setup wrote the record in box order, so its most recent part is
    // still in the Infinity Cache when this launch starts (1.4-2.8 % faster than in order)
    const int64_t r = (int64_t)(gridDim.x - 1 - blockIdx.x) * BLK + threadIdx.x;
    if (r >= nr) return;
    const int64_t b = box0 + r;
    Box x;
    box_load<BLK, REC_STEP>(a, rec, nr, r, b, x, ck0, false);
#pragma unroll
    for (int i = 0; i < NJL; ++i) tjl[i * BLK] = x.tj[JL_SET[i]];
    Track notrack{true, true, 0., 0., 0.};
    for (int it = 0; it < a.niter; ++it) {
        // keeps the LDS reads of ck0 and the J-values inside the loop (not hoisted into
        // registers)
        asm volatile("" ::: "memory");
        sweep<BLK, false, true>(x.R, x.tj, x.e, ck0, x.ck, x.cc, x.ch3o, x.hco, notrack, tjl);
    }
    box_writeback(a, b, x);
}

// Anderson acceleration (Walker & Ni 2011, type II) of depth M on x = (ck, cc), the same
// arithmetic in the same order as the restatement's aa_update.  Per lane, a ring of M + 1
// (f, g) pairs lives in device memory: F(slot, i) = ws[((2 slot) * NX + i) * stride],
// G(slot, i) = ws[((2 slot + 1) * NX + i) * stride], and one more row for the new iterate:
// 2 (M + 1) + 1 rows.  The kernel passes stride 1 and a contiguous block per lane: every
// address is then the lane's base plus a constant, an immediate offset of the load or
// store.  (Interleaving the lanes, stride = lanes, coalesces a wave's accesses but needs a
// 64-bit offset per species; the compiler hoisted the 57 offsets out of the sweep loop and
// spilled them, 1.2 KB of scratch per lane against 0.9 KB, and the Anderson phase's tail,
// where it matters, runs one or two lanes per wave.)  On entry x holds f = F(x_k) and G(head) the iterate x_k
// (aa_save before the sweep), which becomes the residual g_k = f - x_k; on exit x holds
// x_{k+1} = f - dF gamma, gamma
// minimising || w (g_k - dG gamma) ||, w_i = 1 / max(|f_i|, fl)^2; components that would
// turn non-positive keep f_i.
constexpr int NX = NLONG - 1 + NSHORT;

// x_k = (ck, cc) into G(head) before an accelerated sweep (one block of coalesced stores
// instead of a conditional store per species inside the sweep).
__device__ __forceinline__ void aa_save(const double *ck, const double *cc, double *ws,
                                        int64_t stride, int head)
{
    double *g = ws + (int64_t)(2 * head + 1) * NX * stride;
#pragma unroll
    for (int i = 0; i < NLONG - 1; ++i) g[i * stride] = ck[i];
#pragma unroll
    for (int i = 0; i < NSHORT; ++i) g[(NLONG - 1 + i) * stride] = cc[i];
}

template <int M>
__device__ void aa_update(double *ck, double *cc, double *ws, int64_t stride, int &head, int &n,
                          double fl)
{
    auto F = [&](int slot, int i) -> double & { return ws[((int64_t)(2 * slot) * NX + i) * stride]; };
    auto G = [&](int slot, int i) -> double & { return ws[((int64_t)(2 * slot + 1) * NX + i) * stride]; };
    auto XN = [&](int i) -> double & { return ws[((int64_t)(2 * (M + 1)) * NX + i) * stride]; };
    auto X = [&](int i) -> double & { return i < NLONG - 1 ? ck[i] : cc[i - (NLONG - 1)]; };
    const int slot = head;
    // f and g = f - x_k into the ring (x_k was saved into G(slot) before the sweep)
#pragma unroll
    for (int i = 0; i < NX; ++i) {
        F(slot, i) = X(i);
        G(slot, i) = X(i) - G(slot, i);
    }
    head = (head + 1) % (M + 1);
    if (n < M + 1) ++n;
    const int mk = n - 1;
    if (mk < 1) return;
    int ia[M], ib[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
        ia[j] = (slot - j + (M + 1)) % (M + 1);
        ib[j] = (slot - j - 1 + 2 * (M + 1)) % (M + 1);
    }
    // the normal equations, streaming over the species from the ring (a runtime loop, so
    // the box state stays in registers; unrolled by 19 = 57 / 3, so that the loads of 19
    // species are in flight together: fully unrolled, the loop spills the box state)
    double Mx[M][M + 1];
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int c = 0; c <= M; ++c) Mx[r][c] = 0.;
#pragma unroll 19
    for (int i = 0; i < NX; ++i) {
        const double sc = fmax(fabs(F(slot, i)), fl);
        const double w = 1. / (sc * sc);
        const double gk = G(slot, i);
        double dg[M];
#pragma unroll
        for (int j = 0; j < M; ++j) dg[j] = (j < mk) ? G(ia[j], i) - G(ib[j], i) : 0.;
#pragma unroll
        for (int r = 0; r < M; ++r) {
#pragma unroll
            for (int c = 0; c < M; ++c) Mx[r][c] += dg[r] * w * dg[c];
            Mx[r][M] += dg[r] * w * gk;
        }
    }
    // restrict to the mk x mk system; the right-hand side moves to column mk
    double A[M][M + 1];
#pragma unroll
    for (int r = 0; r < M; ++r) {
#pragma unroll
        for (int c = 0; c < M; ++c) A[r][c] = Mx[r][c];
        A[r][M] = Mx[r][M];
    }
#pragma unroll
    for (int r = 0; r < M; ++r)
        if (r < mk) A[r][r] *= (1. + 1.e-10);
    bool ok = true;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        if (c >= mk || !ok) continue;
        int piv = c;
#pragma unroll
        for (int r = c + 1; r < M; ++r)
            if (r < mk && fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        if (!(fabs(A[piv][c]) > 0.)) { ok = false; continue; }
        if (piv != c) {
#pragma unroll
            for (int r = c + 1; r < M; ++r)
                if (r == piv) {
#pragma unroll
                    for (int k = 0; k <= M; ++k) {
                        const double t_ = A[c][k]; A[c][k] = A[r][k]; A[r][k] = t_;
                    }
                }
        }
#pragma unroll
        for (int r = c + 1; r < M; ++r) {
            if (r >= mk) continue;
            const double f_ = A[r][c] / A[c][c];
#pragma unroll
            for (int k = 0; k <= M; ++k) {
                if (k < c) continue;
                const int kk = (k == M) ? M : k;
                if (k < mk || k == M) A[r][kk] -= f_ * A[c][kk];
            }
        }
    }
    if (!ok) return;
    double gam[M];
#pragma unroll
    for (int r = M - 1; r >= 0; --r) {
        gam[r] = 0.;
        if (r >= mk) continue;
        double acc = A[r][M];
#pragma unroll
        for (int k = r + 1; k < M; ++k)
            if (k < mk) acc -= A[r][k] * gam[k];
        gam[r] = acc / A[r][r];
        if (!(fabs(gam[r]) <= 1.e6)) ok = false;
    }
    if (!ok) return;
    // x_{k+1} = f - dF gamma into the XN row (components that would turn non-positive
    // keep f), then back into the registers
#pragma unroll 19
    for (int i = 0; i < NX; ++i) {
        const double fi = F(slot, i);
        double v = fi;
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < mk) v -= gam[j] * (F(ia[j], i) - F(ib[j], i));
        XN(i) = (v > 0. && v <= 1.7976931348623157e308) ? v : fi;
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) X(i) = XN(i);
}

// Converged mode: sweeps to the fixed point of the sweep map.  Persistent: each lane
// takes boxes from a device work queue (one atomic per wave for all lanes that need work)
// and sweeps its box until every species changed by less than tol relative; a converged
// lane holds its result until the wave is served again, so a wave never waits for its
// slowest box.  A box whose iteration goes non-finite or reaches maxiter sweeps without
// converging is set up again and given the fixed-sweep answer (a.niter sweeps from the
// incoming state), its count stored negated.  Every lane leaves once its queue is
// exhausted, so every wave reaches the exit.
//
// MODE CONV_PLAIN: plain sweeps.  With Anderson acceleration (from sweep cv.aa_start) the
// solve is two launches, so that the wide history code never runs masked in a wave whose
// other lanes do plain sweeps: MODE CONV_SWEEP sweeps like CONV_PLAIN but hands a box that
// reaches sweep aa_start unconverged, with its iterate and count, to the spill list (at the
// wave's next serve, so that the sweep loop holds no spill-list code);
// MODE CONV_ANDERSON takes its boxes from that list and continues them with accelerated
// sweeps.  The sequence of operations per box is the one-launch sequence.
//
// Two loops: the inner one holds only the sweep, the fixed-point test and the Anderson
// update, and runs until cv.refill lanes of the wave wait for work (or none is busy, or a
// lane needs its fallback set-up); the outer one serves the wave -- write-back of the
// finished boxes, one queue claim for every waiting lane, set-up from the records.  Keeping
// the set-up and write-back out of the sweep loop keeps the box state in registers there.
// MODE CONV_RESUME: plain sweeps that continue the boxes the Newton solve hands over on the
// spill list (boxes it did not bring to the fixed point), like the Anderson phase without
// the mixing.
constexpr int CONV_PLAIN = 0, CONV_SWEEP = 1, CONV_ANDERSON = 2, CONV_RESUME = 3;

template <int BLK, int MODE>
__global__ void __launch_bounds__(BLK) k_chem_conv(StepArgs a, ConvArgs cv)
{
    constexpr int AAM = MODE == CONV_ANDERSON ? AA_DEPTH : 0;
    // the second launch of a two-phase solve: its boxes come from the spill list
    constexpr bool FROM_SPILL = MODE == CONV_ANDERSON || MODE == CONV_RESUME;
    // ck0 and NJL J-values in LDS, as in k_chem
    static_assert((NLONG - 1 + NJL) * BLK * sizeof(double) <= 160 * 1024, "LDS per workgroup");
    __shared__ double s_ck0[(NLONG - 1 + NJL) * BLK];
    double *ck0 = s_ck0 + threadIdx.x;
    double *tjl = s_ck0 + (NLONG - 1) * BLK + threadIdx.x;
    const int lane = threadIdx.x & (warpSize - 1);
    // the lane's Anderson history: one contiguous block of 2 (AAM + 1) + 1 rows of NX
    double *ws = AAM > 0 ? cv.aa_ws + ((int64_t)blockIdx.x * BLK + threadIdx.x) *
                                          (2 * (AAM + 1) + 1) * NX
                         : nullptr;
    // the spill list: the iterate (NX fields), box index, sweep count and the last two
    // sweeps' largest relative changes per entry
    const int64_t cap = a.nbox;
    int64_t *sp_box = MODE != CONV_PLAIN ? (int64_t *)(cv.spill + NX * cap) : nullptr;
    int32_t *sp_it = MODE != CONV_PLAIN ? (int32_t *)(cv.spill + (NX + 1) * cap) : nullptr;
    double *sp_d = MODE != CONV_PLAIN ? cv.spill + (NX + 2) * cap : nullptr;
    unsigned long long *queue = cv.counter + (FROM_SPILL ? 2 : 0);
    // the Anderson phase starts after the sweep phase has ended: the list is complete
    const int64_t nwork = FROM_SPILL
                              ? (int64_t)__atomic_load_n(cv.counter + 1, __ATOMIC_RELAXED)
                              : a.nbox;
    // lane states: active (sweeping), done (holding a result not yet written), restart
    // (to be set up again for the fallback), finished (queue exhausted); none = waiting
    bool active = false, done = false, finished = false, fallback = false, restart = false;
    // CONV_SWEEP: the box reached sweep aa_start unconverged and goes to the spill list at
    // the next serve (its lane counts as waiting until then)
    bool handoff = false;
    int64_t b = 0;
    int it = 0, tried = 0, head = 0, nstored = 0, result = 0;
    // the last two sweeps' largest relative changes and the number of successive plain
    // (unmixed) sweeps: the stopping estimate's inputs
    double d1 = 0., d2 = 0.;
    int nplain = 0;
    int crun = 0;          // successive plain sweeps whose change was below tol ...
    bool rose = false;     // ... and whether the change rose during that run
    Box x;
    for (;;) {
        // serve the wave
        if (done) {
            if (result == 0) {
                // the surrogate's answer passed its residual test: it is the result
                const double hnm = x.e.hnm;
#pragma unroll
                for (int s = 0; s < NLONG - 1; ++s) x.ck[s] = a.guess_q[(int64_t)s * a.field + b] * hnm;
#pragma unroll
                for (int s = 0; s < NSHORT; ++s) x.cc[s] = a.guess_h[(int64_t)s * a.field + b] * hnm;
            }
            box_writeback(a, b, x);
            if (cv.iters) cv.iters[b] = result;
            done = false;
        }
        if (MODE == CONV_SWEEP) {
            // the boxes that reached sweep aa_start unconverged: to the spill list, with their
            // iterate, count and last two changes (one atomic per wave)
            const unsigned long long gv = __ballot(handoff);
            if (gv) {
                const int leader = __ffsll((long long)gv) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(cv.counter + 1, (unsigned long long)__popcll(gv));
                base = __shfl(base, leader);
                if (handoff) {
                    const int64_t s = (int64_t)(base + (unsigned long long)__popcll(gv & ((1ull << lane) - 1ull)));
#pragma unroll
                    for (int i = 0; i < NLONG - 1; ++i) cv.spill[i * cap + s] = x.ck[i];
#pragma unroll
                    for (int i = 0; i < NSHORT; ++i) cv.spill[(NLONG - 1 + i) * cap + s] = x.cc[i];
                    sp_box[s] = b;
                    sp_it[s] = it;
                    sp_d[s] = d1;
                    sp_d[cap + s] = d2;
                    handoff = false;
                }
            }
        }
        const unsigned long long waiting = __ballot(!active && !finished && !restart);
        bool setup = restart, resume = false;
        int64_t w = 0;
        if (waiting) {
            const int leader = __ffsll((long long)waiting) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(queue, (unsigned long long)__popcll(waiting));
            base = __shfl(base, leader);
            if ((waiting >> lane) & 1ull) {
                const unsigned long long below = waiting & ((1ull << lane) - 1ull);
                w = (int64_t)(base + (unsigned long long)__popcll(below));
                if (w < nwork) {
                    b = FROM_SPILL ? sp_box[w] : (cv.order ? (int64_t)cv.order[w] : w);
                    setup = true;
                    resume = FROM_SPILL;
                    fallback = false;
                } else {
                    finished = true;
                }
            }
        }
        if (setup) {
            box_load<BLK>(a, cv.rec, a.nbox, b, b, x, ck0, !fallback && !resume);
#pragma unroll
            for (int i = 0; i < NJL; ++i) tjl[i * BLK] = x.tj[JL_SET[i]];
            active = true;
            restart = false;
            it = 0;
            head = 0;
            nstored = 0;
            d1 = d2 = 0.;
            nplain = 0;
            crun = 0;
            rose = false;
            if (resume) {
                // the iterate, sweep count and last changes the sweep phase handed over
                // (every sweep there was plain)
#pragma unroll
                for (int i = 0; i < NLONG - 1; ++i) x.ck[i] = cv.spill[i * cap + w];
#pragma unroll
                for (int i = 0; i < NSHORT; ++i) x.cc[i] = cv.spill[(NLONG - 1 + i) * cap + w];
                it = sp_it[w];
                d1 = sp_d[w];
                d2 = sp_d[cap + w];
                // the sweep phase's boxes arrive after plain sweeps; a box from the Newton
                // solve (no changes recorded) starts its estimate afresh
                nplain = d2 > 0. ? it : 0;
            }
        }
        if (__ballot(active) == 0ull) break;
        // sweep until the wave needs serving
        for (;;) {
            if (MODE == CONV_SWEEP && active && !fallback && it + 1 >= cv.aa_start) {
                // a box about to take its first accelerated sweep leaves the sweeps; the
                // serve puts it on the spill list
                handoff = true;
                active = false;
            }
            if (active) {
                const bool acc = AAM > 0 && !fallback && it + 1 >= cv.aa_start;
                if (AAM > 0 && acc) aa_save(x.ck, x.cc, ws, 1, head);
                // the first sweep from a surrogate's answer is its acceptance test
                const bool accept = cv.accept_tol > 0. && it == 0 && !fallback;
                Track tr{true, true, accept ? cv.accept_tol : cv.tol, 1.e-30 * x.e.hnm, 0.};
                // keeps the LDS reads inside the loop (not hoisted into registers)
                asm volatile("" ::: "memory");
                sweep<BLK, true, true>(x.R, x.tj, x.e, ck0, x.ck, x.cc, x.ch3o, x.hco, tr, tjl);
                ++it;
                if (fallback) {
                    if (it >= a.niter) {
                        result = -tried;
                        done = true;
                        active = false;
                    }
                } else if (accept && tr.conv && tr.fin) {
                    result = 0;
                    done = true;
                    active = false;
                } else if (tr.conv && tr.fin &&
                           ((nplain >= CONV_NPLAIN && near_fixed_point(tr.dmax, d1, d2, cv.tol)) ||
                            (crun + 1 >= CONV_STALL && (rose || tr.dmax >= d1) && nplain >= 1))) {
                    result = it + cv.iters_offset;
                    done = true;
                    active = false;
                } else if (!tr.fin || it >= cv.maxiter) {
                    // diverged (a non-finite value) or out of sweeps
                    tried = it;
                    fallback = true;
                    restart = true;
                    active = false;
                } else {
                    // with Anderson mixing, a box whose change is already below tol takes
                    // plain steps until the estimate decides
                    const bool mix = AAM > 0 && acc && !tr.conv;
                    if (AAM > 0 && mix)
                        aa_update<(AAM > 0 ? AAM : 1)>(x.ck, x.cc, ws, 1, head, nstored,
                                                       1.e-30 * x.e.hnm);
                    if (tr.conv && nplain >= 1) {
                        rose = rose || (crun > 0 && tr.dmax >= d1);
                        ++crun;
                    } else {
                        crun = 0;
                        rose = false;
                    }
                    nplain = mix ? 0 : nplain + 1;
                    d2 = d1;
                    d1 = tr.dmax;
                }
            }
            const unsigned long long waiting = __ballot(!active && !finished && !restart);
            const unsigned long long busy = __ballot(active);
            if (__popcll(waiting) >= (unsigned)cv.refill || busy == 0ull || __ballot(restart))
                break;
        }
    }
}

// Newton solve of the converged mode's fixed point (the same answer as the sweeps to the
// fixed point, reached in a few iterations whatever the sweep map's contraction), per box
// and lane.  Its boxes come either from the whole grid (FROM_SPILL false: set-up from the
// record and one plain sweep from the initial iterate -- the guess or the incoming state --
// which makes the short-lived species consistent with the long-lived ones and sets the
// night zeros exactly) or from a spill list (FROM_SPILL true: the boxes a plain sweep phase
// left unconverged after its first sweeps, with their iterates).  Then
//   1. Newton iterations on the local form of the fixed-point conditions, with the
//      generated residual, Jacobian and sparse LU (newton_gen.h: 55 unknowns, the day or
//      the night system), each step projected onto non-negative species, until a step
//      moves every species by less than tol relative (or the steps stop shrinking below
//      1e-8: the conditioning's floor, where the check decides);
//   2. one plain sweep from the Newton point as its check: every species must change by
//      less than tol, and the swept state is the result, as for the sweep solver.
// A box that fails (no convergence in max_newton iterations, a non-finite value, or the
// check) goes to the output spill list with its iterate, for the plain-sweep resume phase.
//
// The factors (514 doubles by day) do not fit the register file, so they live in the
// lane's scratch; WAVES resident waves per SIMD (1, 2, 4: a launch shape the autotuner
// chooses) overlap one wave's scratch traffic with another's arithmetic.  The start-of-step
// state and the J-values stay with the lane too (no LDS, which would cap the waves).  The
// persistent waves take 64 boxes per queue claim; every lane leaves once the queue is
// exhausted.
constexpr int NEWTON_BLK = 64;

// One Newton system of box x at its current iterate: the generated residual and Jacobian
// into (A, B), the sparse LU of A in place and the solve, B = J^-1 F (day or night system).
// ck0 at stride CK0_STRIDE; with JLDS the J-values of JL_SET in tjl (stride CK0_STRIDE), as
// the sweep reads them
template <int CK0_STRIDE, bool JLDS, class AT, class BT>
__device__ __forceinline__ void newton_system(const Box &x, const double *ck0, const double *tjl,
                                              AT &A, BT &B)
{
    const Rates &R = x.R;
    const BoxEnv &e = x.e;
    const double *ck = x.ck, *cc = x.cc;
#define TJ(X) ((JLDS && jl_slot(X) >= 0) ? tjl[jl_slot(X) * CK0_STRIDE] : x.tj[X])
#define CK0(S) ck0[(S) * CK0_STRIDE]
    if (e.day) {
        RB_NEWTON_DAY_JAC(A, B);
        RB_NEWTON_DAY_LU(A, B);
        RB_NEWTON_DAY_SOLVE(A, B);
    } else {
        RB_NEWTON_NIGHT_JAC(A, B);
        RB_NEWTON_NIGHT_LU(A, B);
        RB_NEWTON_NIGHT_SOLVE(A, B);
    }
#undef TJ
#undef CK0
    (void)R; (void)e; (void)ck; (void)cc; (void)tjl;
}

// A box's factors live in the lane's scratch, WAVES (1, 2, 4) resident waves per SIMD.  (In
// LDS instead, 569 doubles per box, a CU holds 32 boxes -- half a wave with nothing to hide
// its latency: the Newton-only c2 solve took 74.5 ms against 23.5 ms.)
template <int WAVES, bool FROM_SPILL>
__global__ void __launch_bounds__(NEWTON_BLK, WAVES) k_chem_newton(StepArgs a, ConvArgs cv,
                                                                   NewtonArgs na)
{
    const int lane = threadIdx.x & (warpSize - 1);
    // at one wave per SIMD the start-of-step state and NJL J-values live in LDS, as in the
    // sweep kernels (80 doubles per lane, the CU's 160 KB at four waves): registers for the
    // Jacobian's temporaries.  At more waves per SIMD the LDS would cap the occupancy.
    constexpr bool SL = WAVES == 1;
    constexpr int ST = SL ? NEWTON_BLK : 1;
    __shared__ double s_st[SL ? (NLONG - 1 + NJL) * NEWTON_BLK : 1];
    const int64_t cap = a.nbox;
    const int64_t *in_box = FROM_SPILL ? (const int64_t *)(na.in + NX * cap) : nullptr;
    const int32_t *in_it = FROM_SPILL ? (const int32_t *)(na.in + (NX + 1) * cap) : nullptr;
    int64_t *out_box = (int64_t *)(na.out + NX * cap);
    int32_t *out_it = (int32_t *)(na.out + (NX + 1) * cap);
    double *out_d = na.out + (NX + 2) * cap;
    // the input list is complete when this launch starts (the sweep phase has ended)
    const int64_t nwork = FROM_SPILL ? (int64_t)__atomic_load_n(na.n_in, __ATOMIC_RELAXED) : a.nbox;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(na.queue, (unsigned long long)warpSize);
        base = __shfl(base, 0);
        if ((int64_t)base >= nwork) break;
        const int64_t w = (int64_t)base + lane;
        if (w < nwork) {
            const int64_t b = FROM_SPILL ? in_box[w] : (cv.order ? (int64_t)cv.order[w] : w);
            Box x;
            double ck0r[SL ? 1 : NLONG - 1];
            double *ck0 = SL ? s_st + threadIdx.x : ck0r;
            double *tjl = SL ? s_st + (NLONG - 1) * NEWTON_BLK + threadIdx.x : nullptr;
            box_load<ST>(a, cv.rec, a.nbox, b, b, x, ck0, !FROM_SPILL);
            if (SL) {
#pragma unroll
                for (int i = 0; i < NJL; ++i) tjl[i * NEWTON_BLK] = x.tj[JL_SET[i]];
            }
            const double fl = 1.e-30 * x.e.hnm;
            int sweeps = 0;
            bool fin = true;
            if (FROM_SPILL) {
#pragma unroll
                for (int i = 0; i < NLONG - 1; ++i) x.ck[i] = na.in[i * cap + w];
#pragma unroll
                for (int i = 0; i < NSHORT; ++i) x.cc[i] = na.in[(NLONG - 1 + i) * cap + w];
                sweeps = in_it[w];
            } else {
                Track tr{true, true, cv.tol, fl, 0.};
                sweep<ST, true, SL>(x.R, x.tj, x.e, ck0, x.ck, x.cc, x.ch3o, x.hco, tr, tjl);
                sweeps = 1;
                fin = tr.fin;
            }
            int nit = 0;
            bool ok = false;
            double sprev = 1.e300;
            {
                for (nit = 1; fin && nit <= cv.max_newton; ++nit) {
                    double A[NEWTON_NNZ], B[NEWTON_N];
                    newton_system<ST, SL>(x, ck0, tjl, A, B);
                    // the full step, each species projected onto x >= 0: a species the step
                    // would take below zero (one heading for a tiny root: night NO2 below its
                    // clamp, HO2 or CH3OOH in weak sunlight) goes to 0 and the next step
                    // lifts it to its root.  (Shortening the whole step instead let one such
                    // species freeze every other one for dozens of iterations.)  A
                    // non-finite step -- a zero pivot -- fails the box: fmax would turn its
                    // NaN into 0.
                    double smax = 0.;
#pragma unroll
                    for (int i = 0; i < NEWTON_N; ++i) {
                        double &v = i < NEWTON_NLONG ? x.ck[NEWTON_SLOT[i]] : x.cc[NEWTON_SLOT[i]];
                        const double bi = B[i];
                        fin &= fabs(bi) <= 1.7976931348623157e308;
                        const double nv = fmax(v - bi, 0.);
                        smax = fmax(smax, fabs(nv - v) * __builtin_amdgcn_rcp(fmax(nv, fl)));
                        v = nv;
                    }
                    const bool done = smax < cv.tol || (smax < 1.e-8 && smax >= 0.5 * sprev);
                    sprev = smax;
                    if (done && fin) {
                        ok = true;
                        break;
                    }
                }
            }
            if (ok) {
                // the check: one plain sweep from the Newton point
                Track tc{true, true, cv.tol, fl, 0.};
                sweep<ST, true, SL>(x.R, x.tj, x.e, ck0, x.ck, x.cc, x.ch3o, x.hco, tc, tjl);
                ++sweeps;
                ok = tc.conv && tc.fin;
                fin = tc.fin;
            }
            if (ok) {
                box_writeback(a, b, x);
                if (cv.iters) cv.iters[b] = NEWTON_SOLVED + nit;
            } else {
                // hand the box to the sweeps: its iterate if finite, else its initial one
                if (!fin) {
                    const double hnm = x.e.hnm;
                    const double *gq = a.guess_q ? a.guess_q : a.qj1;
                    const double *gh = a.guess_q ? a.guess_h : a.hc;
#pragma unroll
                    for (int s = 0; s < NLONG - 1; ++s) x.ck[s] = gq[(int64_t)s * a.field + b] * hnm;
#pragma unroll
                    for (int s = 0; s < NSHORT; ++s) x.cc[s] = gh[(int64_t)s * a.field + b] * hnm;
                }
                const int64_t s = (int64_t)atomicAdd(na.n_out, 1ull);
#pragma unroll
                for (int i = 0; i < NLONG - 1; ++i) na.out[i * cap + s] = x.ck[i];
#pragma unroll
                for (int i = 0; i < NSHORT; ++i) na.out[(NLONG - 1 + i) * cap + s] = x.cc[i];
                out_box[s] = b;
                out_it[s] = sweeps;
                out_d[s] = 0.;
                out_d[cap + s] = 0.;
            }
        }
    }
}

__global__ void k_conserv_floor(double *__restrict__ qj1, double *__restrict__ hc, int64_t plane,
                                int64_t field, int32_t nivbas, int64_t start)
{
    const int64_t b = start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= field) return;
    double *q = qj1 + b, *h = hc + b;
#define Q(s) q[(int64_t)(s) * field]
#define HS(s) h[(int64_t)(s) * field]
    if (b < plane * (int64_t)nivbas) {
        const double rcly = (Q(CLY) - Q(BRCL)) /
                            (Q(OCLO) + Q(HCL) + Q(CLONO2) + Q(HOCL) + 2. * Q(CL2) + Q(CLNO2) +
                             Q(CLOX) + 2. * Q(CL2O2));
        HS(CL) *= rcly; HS(CLO) *= rcly;
        Q(OCLO) *= rcly; Q(HCL) *= rcly; Q(CLONO2) *= rcly; Q(HOCL) *= rcly; Q(CL2) *= rcly;
        Q(CLNO2) *= rcly; Q(CLOX) *= rcly; Q(CL2O2) *= rcly;
        const double rnoy = fmax((Q(NOY) - Q(CLONO2) - Q(CLNO2) - Q(BRONO2)), 0.) /
                            (Q(HNO3) + 2. * Q(N2O5) + Q(NOX) + Q(HO2NO2));
        HS(NO2) *= rnoy; HS(NO) *= rnoy; HS(N) *= rnoy; HS(NO3) *= rnoy;
        Q(HNO3) *= rnoy; Q(N2O5) *= rnoy; Q(NOX) *= rnoy; Q(HO2NO2) *= rnoy;
        const double rbry = Q(BRY) / (Q(HBR) + Q(BRONO2) + Q(BROX) + Q(HOBR) + Q(BRCL));
        HS(BR) *= rbry; HS(BRO) *= rbry;
        Q(HBR) *= rbry; Q(BRONO2) *= rbry; Q(BROX) *= rbry; Q(HOBR) *= rbry; Q(BRCL) *= rbry;
    }
#pragma unroll
    for (int s = 0; s < NLONG - 1; ++s) Q(s) = fmax(Q(s), 1.e-30);
    // HNO3 gas = HNO3 below the chemistry levels (F:444-450)
    if (b >= plane * (int64_t)nivbas) Q(HNO3G) = Q(HNO3);
#undef Q
#undef HS
}

// Sedimentation (F:2416-2522): one thread per column walks its levels bottom-up; each
// level's update reads the not yet updated values of itself and the level above, which
// is the reference's loop order.  Levels are 1-based as in the reference.
__global__ void k_sediment(const double *__restrict__ p, const double *__restrict__ T,
                           double *__restrict__ h2o, double *__restrict__ noy,
                           double *__restrict__ hno3, const double *__restrict__ vsed,
                           const double *__restrict__ p2o, const double *__restrict__ pno3,
                           int64_t plane, int32_t nivbas, int32_t nivhau, int32_t nivh2o,
                           double dt)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= plane) return;
    double *q3 = h2o + c, *q4 = noy + c, *q5 = hno3 + c;
    auto at = [&](int k) { return (int64_t)(k - 1) * plane; };
    auto hnm = [&](int k) { return p[at(k) + c] / (T[at(k) + c] * 1.38e-19); };
    for (int k = nivh2o - 1; k >= nivhau; --k) {
        const int64_t o = at(k), u = at(k - 1);
        const double dz = 7000. * log(p[o + c] / p[u + c]);
        const double fracup = fmin(vsed[u + c] * dt / dz, 1.0);
        const double fracdown = fmin(vsed[o + c] * dt / dz, 1.0);
        q3[o] = q3[o] + q3[u] * (1. - p2o[u + c]) * hnm(k - 1) / hnm(k) * fracup -
                q3[o] * (1. - p2o[o + c]) * fracdown;
    }
    for (int k = nivbas; k >= nivhau; --k) {
        const int64_t o = at(k), u = at(k - 1);
        const double dz = 7000. * log(p[o + c] / p[u + c]);
        const double fracup = fmin(vsed[u + c] * dt / dz, 1.0);
        const double fracdown = fmin(vsed[o + c] * dt / dz, 1.0);
        const double h5u = q5[u], h5 = q5[o];
        q4[o] = q4[o] + h5u * (1. - pno3[u + c]) * hnm(k - 1) / hnm(k) * fracup -
                h5 * (1. - pno3[o + c]) * fracdown;
        q5[o] = h5 + h5u * (1. - pno3[u + c]) * hnm(k - 1) / hnm(k) * fracup -
                h5 * (1. - pno3[o + c]) * fracdown;
    }
}

hipError_t launch_sediment(const double *p, const double *T, double *h2o, double *noy,
                           double *hno3, const double *vsed, const double *p2o, const double *pno3,
                           int64_t plane, int nivbas, int nivhau, int nivh2o, double dt,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_sediment, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, p, T,
                       h2o, noy, hno3, vsed, p2o, pno3, plane, nivbas, nivhau, nivh2o, dt);
    return hipGetLastError();
}

__global__ void k_count_nonfinite(const double *__restrict__ x, int64_t n,
                                  unsigned long long *__restrict__ count)
{
    unsigned long long local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        local += !isfinite(x[i]);
    // wave-level reduction, then one atomic per wave
    for (int off = warpSize / 2; off > 0; off >>= 1) local += __shfl_down(local, off);
    if ((threadIdx.x & (warpSize - 1)) == 0 && local) atomicAdd(count, local);
}

// ------------------------------------------------------------------ host launchers
Troe troe_constants()
{
    Troe c;
    c.a0hox = 4.4e-32 * pow(300., 1.3);
    c.a1hox = 7.5e-11 * pow(300., -0.2);
    c.a0nox1 = 2.0e-30 * pow(300., 4.4);
    c.a1nox1 = 1.4e-12 * pow(300., 0.7);
    c.a0nox2 = 1.8e-30 * pow(300., 3.0);
    c.a1nox2 = 2.8e-11 * pow(300., 0.0);
    c.a0nox3 = 2.0e-31 * pow(300., 3.4);
    c.a1nox3 = 2.9e-12 * pow(300., 1.1);
    c.a0c1 = 4.0e-31 * pow(300., 3.6);
    c.a1c1 = 1.2e-12 * pow(300., -1.1);
    c.a0clx1 = 1.8e-31 * pow(300., 3.4);
    c.a1clx1 = 1.5e-11 * pow(300., 1.9);
    c.a0clx2 = 1.8e-31 * pow(300., 2.0);
    c.a1clx2 = 1.0e-10 * pow(300., 1.0);
    c.a0clx3 = 1.6e-32 * pow(300., 4.5);
    c.a1clx3 = 3.0e-12 * pow(300., 2.0);
    c.a0brx = 5.2e-31 * pow(300., 3.2);
    c.a1brx = 6.9e-12 * pow(300., 2.9);
    return c;
}

hipError_t launch_columns(const StepArgs &a, double *col, hipStream_t s)
{
    const int blk = 64;
    const int64_t grid = (a.plane + blk - 1) / blk;
    hipLaunchKernelGGL(k_columns, dim3((unsigned)grid), dim3(blk), 0, s, a.p, a.T, a.hc, col,
                       a.plane, a.field, a.nivbas);
    return hipGetLastError();
}

template <int R>
static hipError_t launch_setup_t(const StepArgs &a, double *rec, int blk, hipStream_t s,
                                 int64_t box0, int64_t nr)
{
    const unsigned grid = (unsigned)((nr + blk - 1) / blk);
    switch (blk) {
    case 64: hipLaunchKernelGGL((k_setup<64, R>), dim3(grid), dim3(64), 0, s, a, rec, box0, nr); break;
    case 128: hipLaunchKernelGGL((k_setup<128, R>), dim3(grid), dim3(128), 0, s, a, rec, box0, nr); break;
    case 256: hipLaunchKernelGGL((k_setup<256, R>), dim3(grid), dim3(256), 0, s, a, rec, box0, nr); break;
    default: return hipErrorInvalidConfiguration;
    }
    return hipGetLastError();
}

hipError_t launch_setup(const StepArgs &a, double *rec, int block, hipStream_t s, int64_t box0,
                        int64_t nr, bool full)
{
    const int blk = block > 0 ? block : 256;
    if (nr < 0) nr = a.nbox - box0;
    return full ? launch_setup_t<REC_FULL>(a, rec, blk, s, box0, nr)
                 : launch_setup_t<REC_STEP>(a, rec, blk, s, box0, nr);
}

hipError_t launch_chem(const StepArgs &a, const double *rec, int block, hipStream_t s,
                       int64_t box0, int64_t nr)
{
    const int blk = block > 0 ? block : 256;
    if (nr < 0) nr = a.nbox - box0;
    const unsigned grid = (unsigned)((nr + blk - 1) / blk);
    switch (blk) {
    case 64: hipLaunchKernelGGL(k_chem<64>, dim3(grid), dim3(64), 0, s, a, rec, box0, nr); break;
    case 128: hipLaunchKernelGGL(k_chem<128>, dim3(grid), dim3(128), 0, s, a, rec, box0, nr); break;
    case 256: hipLaunchKernelGGL(k_chem<256>, dim3(grid), dim3(256), 0, s, a, rec, box0, nr); break;
    default: return hipErrorInvalidConfiguration;
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_conv_mode(const StepArgs &a, const ConvArgs &cv, int blk, int nblocks,
                                   hipStream_t s)
{
    switch (blk) {
    case 64: hipLaunchKernelGGL((k_chem_conv<64, MODE>), dim3(nblocks), dim3(64), 0, s, a, cv); break;
    case 128: hipLaunchKernelGGL((k_chem_conv<128, MODE>), dim3(nblocks), dim3(128), 0, s, a, cv); break;
    case 256: hipLaunchKernelGGL((k_chem_conv<256, MODE>), dim3(nblocks), dim3(256), 0, s, a, cv); break;
    default: return hipErrorInvalidConfiguration;
    }
    return hipGetLastError();
}

hipError_t launch_chem_conv(const StepArgs &a, const ConvArgs &cv, int block, int nblocks_sweep,
                            int nblocks_aa, hipStream_t s)
{
    const int blk = block > 0 ? block : 256;
    if (cv.accel == 0) return launch_conv_mode<CONV_PLAIN>(a, cv, blk, nblocks_sweep, s);
    if (cv.accel != AA_DEPTH || !cv.spill || !cv.aa_ws) return hipErrorInvalidValue;
    const hipError_t e = launch_conv_mode<CONV_SWEEP>(a, cv, blk, nblocks_sweep, s);
    if (e != hipSuccess) return e;
    return launch_conv_mode<CONV_ANDERSON>(a, cv, blk, nblocks_aa, s);
}

template <bool FS>
static hipError_t launch_newton_kernel(const StepArgs &a, const ConvArgs &cv, const NewtonArgs &na,
                                       int waves, int nblocks, hipStream_t s)
{
    switch (waves) {
    case 1: hipLaunchKernelGGL((k_chem_newton<1, FS>), dim3(nblocks), dim3(NEWTON_BLK), 0, s, a, cv, na); break;
    case 2: hipLaunchKernelGGL((k_chem_newton<2, FS>), dim3(nblocks), dim3(NEWTON_BLK), 0, s, a, cv, na); break;
    case 4: hipLaunchKernelGGL((k_chem_newton<4, FS>), dim3(nblocks), dim3(NEWTON_BLK), 0, s, a, cv, na); break;
    default: return hipErrorInvalidConfiguration;
    }
    return hipGetLastError();
}

hipError_t launch_chem_newton(const StepArgs &a, const ConvArgs &cv, const NewtonLaunch &nl,
                              hipStream_t s)
{
    if (!cv.spill || !cv.rec || !nl.spill2) return hipErrorInvalidValue;
    const int blk = nl.block > 0 ? nl.block : 256;
    unsigned long long *q = cv.counter;   // 5 counters, zero at launch
    NewtonArgs na{};
    na.out = nl.spill2;
    na.n_out = q + 3;
    hipError_t e;
    if (nl.sweeps_first > 0) {
        // plain sweeps first; the boxes still unconverged at sweep sweeps_first go to the
        // spill list (counters 0, 1), Newton takes them from there (counter 2)
        ConvArgs sc = cv;
        sc.accel = 0;
        sc.aa_ws = nullptr;
        sc.aa_start = nl.sweeps_first;
        if ((e = launch_conv_mode<CONV_SWEEP>(a, sc, blk, nl.nblocks_sweep, s)) != hipSuccess)
            return e;
        na.in = cv.spill;
        na.n_in = q + 1;
        na.queue = q + 2;
        e = launch_newton_kernel<true>(a, cv, na, nl.waves, nl.nblocks_newton, s);
    } else {
        na.queue = q;
        e = launch_newton_kernel<false>(a, cv, na, nl.waves, nl.nblocks_newton, s);
    }
    if (e != hipSuccess) return e;
    // plain sweeps for the boxes Newton handed over (list counters 3, 4)
    ConvArgs rc = cv;
    rc.accel = 0;
    rc.aa_ws = nullptr;
    rc.counter = q + 2;
    rc.spill = nl.spill2;
    rc.maxiter = nl.resume_maxiter;
    rc.iters_offset = NEWTON_RESUMED;
    return launch_conv_mode<CONV_RESUME>(a, rc, blk, nl.nblocks_resume, s);
}

int newton_blocks_per_cu(int waves)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (waves == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_chem_newton<1, true>, NEWTON_BLK, 0);
    if (waves == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_chem_newton<2, true>, NEWTON_BLK, 0);
    if (waves == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_chem_newton<4, true>, NEWTON_BLK, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
}

template <int MODE>
static int conv_occupancy(int block)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (block == 64) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_chem_conv<64, MODE>, 64, 0);
    if (block == 128) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_chem_conv<128, MODE>, 128, 0);
    if (block == 256) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_chem_conv<256, MODE>, 256, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
}

int conv_blocks_per_cu(int block, int accel, bool anderson_phase)
{
    if (accel == 0) return conv_occupancy<CONV_PLAIN>(block);
    return anderson_phase ? conv_occupancy<CONV_ANDERSON>(block) : conv_occupancy<CONV_SWEEP>(block);
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

hipError_t launch_zero_counter(unsigned long long *p, hipStream_t s, int n)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

// The converged mode's work order.  The persistent sweep kernels refill a wave's idle lanes
// from a queue; in grid order the lanes soon hold boxes from all over the grid, day and
// night mixed, and such a wave issues both branches of the sweep (F:1353-1670 by day,
// F:1671-1746 by night) on every sweep.  Queueing the night boxes first and the day boxes
// after keeps the two apart except while the queue passes from one part to the other.
// Within each part a block's boxes stay together in grid order (the day part reversed), so
// a wave's loads of the fields still fall on few cache lines.  The day test is k_setup's.
// A block takes ORDER_CHUNKS chunks of ORDER_BLK boxes: two global atomics per 4,096 boxes
// (one block per 256 boxes, 3,072 blocks on the c1 grid, spent ~80 us contending for the
// two counters).
constexpr int ORDER_BLK = 256, ORDER_CHUNKS = 16;

__global__ void __launch_bounds__(ORDER_BLK) k_order_day(const double *__restrict__ sza,
                                                         int64_t nbox, int32_t *__restrict__ order,
                                                         unsigned long long *cnt)
{
    constexpr int NW = ORDER_BLK / 64;
    __shared__ unsigned s_cnt[2][NW];
    __shared__ unsigned long long s_base[2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t first = (int64_t)blockIdx.x * ORDER_BLK * ORDER_CHUNKS;
    // the block's night and day counts, then one claim of each part
    unsigned cn = 0, cd = 0;
    for (int i = 0; i < ORDER_CHUNKS; ++i) {
        const int64_t b = first + (int64_t)i * ORDER_BLK + threadIdx.x;
        if (b < nbox) {
            if (sza[b] <= 95.) ++cd;
            else ++cn;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cn += __shfl_down(cn, off);
        cd += __shfl_down(cd, off);
    }
    if (lane == 0) {
        s_cnt[0][wave] = cn;
        s_cnt[1][wave] = cd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tn = 0, td = 0;
        for (int w = 0; w < NW; ++w) {
            tn += s_cnt[0][w];
            td += s_cnt[1][w];
        }
        s_base[0] = atomicAdd(cnt, (unsigned long long)tn);
        s_base[1] = atomicAdd(cnt + 1, (unsigned long long)td);
    }
    __syncthreads();
    unsigned long long pos[2] = {s_base[0], s_base[1]};
    // the chunks in order: each box's rank among its part's boxes of the chunk
    for (int i = 0; i < ORDER_CHUNKS; ++i) {
        const int64_t b = first + (int64_t)i * ORDER_BLK + threadIdx.x;
        const bool in = b < nbox;
        const bool day = in && sza[b] <= 95.;
        const bool night = in && !day;
        const unsigned long long mn = __ballot(night), md = __ballot(day);
        __syncthreads();   // every wave has read the previous chunk's counts
        if (lane == 0) {
            s_cnt[0][wave] = (unsigned)__popcll(mn);
            s_cnt[1][wave] = (unsigned)__popcll(md);
        }
        __syncthreads();
        unsigned before[2] = {0u, 0u}, tot[2] = {0u, 0u};
        for (int w = 0; w < NW; ++w) {
            if (w < wave) {
                before[0] += s_cnt[0][w];
                before[1] += s_cnt[1][w];
            }
            tot[0] += s_cnt[0][w];
            tot[1] += s_cnt[1][w];
        }
        if (in) {
            const int part = day ? 1 : 0;
            const unsigned rank = before[part] +
                                  (unsigned)__popcll((day ? md : mn) & ((1ull << lane) - 1ull));
            const int64_t p = (int64_t)(pos[part] + rank);
            order[day ? nbox - 1 - p : p] = (int32_t)b;
        }
        pos[0] += tot[0];
        pos[1] += tot[1];
    }
}

hipError_t launch_order_day(const double *sza, int64_t nbox, int32_t *order,
                            unsigned long long *cnt, hipStream_t s)
{
    if (nbox < 1 || nbox > INT32_MAX) return hipErrorInvalidValue;
    const int64_t per = (int64_t)ORDER_BLK * ORDER_CHUNKS;
    const unsigned grid = (unsigned)((nbox + per - 1) / per);
    hipLaunchKernelGGL(k_order_day, dim3(grid), dim3(ORDER_BLK), 0, s, sza, nbox, order, cnt);
    return hipGetLastError();
}

hipError_t launch_conserv_floor(double *qj1, double *hc, int64_t plane, int64_t field,
                                int32_t nivbas, int64_t start, hipStream_t s)
{
    const int blk = 256;
    const int64_t n = field - start;
    if (n <= 0) return hipSuccess;
    const int64_t grid = (n + blk - 1) / blk;
    hipLaunchKernelGGL(k_conserv_floor, dim3((unsigned)grid), dim3(blk), 0, s, qj1, hc, plane,
                       field, nivbas, start);
    return hipGetLastError();
}

hipError_t launch_count_nonfinite(const double *x, int64_t n, unsigned long long *count,
                                  hipStream_t s)
{
    const int blk = 256;
    int64_t grid = (n + blk - 1) / blk;
    if (grid > 4096) grid = 4096;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)grid), dim3(blk), 0, s, x, n, count);
    return hipGetLastError();
}

}  // namespace rb
