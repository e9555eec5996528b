// Correctly rounded FP64 elementary functions from double-double evaluation.
//
// The reference's PSC/aerosol block (analytic, src/reprobus_l137_20230723.f:2620-3361)
// solves the ternary HNO3/H2SO4/H2O cubic close to its binary limit, where
// mnf = mnb*(1 - msf/msb) cancels by up to ~1e10: a one-ulp difference in any
// libm result upstream of it changes the HNO3 gas fraction by up to ~1e-6.  The
// reference links glibc, whose exp/log/pow/atan/cos/tanh are correctly rounded
// on essentially every argument, so the aerosol block evaluates these functions
// to ~100 bits here and rounds once: the step then reproduces the reference in
// those boxes instead of amplifying the vendor library's last-bit errors.  The
// back-trajectories (backtra, F:6156-6470) use sin, cos, asin, acos and exp the
// same way, so that the device's departure points are the reference's.
//
// Every routine uses only IEEE add/mul/div/sqrt and fma, so the host build
// (used by the CPU tests) and the gfx950 build return identical bits.
#pragma once
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define CR_HD __host__ __device__ __forceinline__
#else
#define CR_HD static inline
#endif

namespace crm {

struct dd {
    double hi, lo;
};

CR_HD dd two_sum(double a, double b)
{
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
CR_HD dd quick_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}
CR_HD dd two_prod(double a, double b)
{
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
CR_HD dd add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
CR_HD dd add_d(dd a, double b)
{
    dd s = two_sum(a.hi, b);
    s.lo += a.lo;
    return quick_two_sum(s.hi, s.lo);
}
CR_HD dd neg(dd a) { return {-a.hi, -a.lo}; }
CR_HD dd mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}
CR_HD dd mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return quick_two_sum(p.hi, p.lo);
}
CR_HD dd div_d(dd a, double b)
{
    const double q1 = a.hi / b;
    const dd p = two_prod(q1, b);
    const double r = ((a.hi - p.hi) - p.lo) + a.lo;
    return quick_two_sum(q1, r / b);
}
CR_HD dd div(dd a, dd b)
{
    const double q1 = a.hi / b.hi;
    dd r = add(a, neg(mul_d(b, q1)));
    const double q2 = r.hi / b.hi;
    r = add(r, neg(mul_d(b, q2)));
    const double q3 = r.hi / b.hi;
    return add_d(quick_two_sum(q1, q2), q3);
}
CR_HD dd scale(dd a, int e) { return {ldexp(a.hi, e), ldexp(a.lo, e)}; }

// ln 2, 1/ln 2, pi/2 and 1/ln 10 to ~107 bits (hi + lo).
constexpr double LN2_HI = 6.93147180559945286227e-01, LN2_LO = 2.31904681384629955842e-17;
constexpr double INV_LN2 = 1.44269504088896338700e+00;
constexpr double PIO2_HI = 1.57079632679489655800e+00, PIO2_LO = 6.12323399573676603587e-17;
constexpr double PIO2_LO2 = -1.49738490485916983e-33;
constexpr double INV_LN10_HI = 4.34294481903251816668e-01, INV_LN10_LO = 1.09831965021676510e-17;
constexpr double LN10_HI = 2.302585092994046, LN10_LO = -2.1707562233822494e-16;
// 1/j (j = 1..9) and 1/(j(j-1)) (j = 2..19) as double-double: the Taylor steps
// multiply instead of divide.
constexpr double INVJ_HI[10] = {0., 1.0, 0.5, 0.3333333333333333, 0.25, 0.2, 0.16666666666666666,
                                0.14285714285714285, 0.125, 0.1111111111111111};
constexpr double INVJ_LO[10] = {0., 0.0, 0.0, 1.850371707708594e-17, 0.0, -1.1102230246251566e-17,
                                9.25185853854297e-18, 7.93016446160826e-18, 0.0,
                                6.1679056923619804e-18};
constexpr double INVJJ_HI[20] = {
    0., 0., 0.5, 0.16666666666666666, 0.08333333333333333, 0.05, 0.03333333333333333,
    0.023809523809523808, 0.017857142857142856, 0.013888888888888888, 0.011111111111111112,
    0.00909090909090909, 0.007575757575757576, 0.00641025641025641, 0.005494505494505495,
    0.004761904761904762, 0.004166666666666667, 0.003676470588235294, 0.0032679738562091504,
    0.0029239766081871343};
constexpr double INVJJ_LO[20] = {
    0., 0., 0.0, 9.25185853854297e-18, 4.625929269271485e-18, -2.7755575615628915e-18,
    4.625929269271486e-19, 1.32169407693471e-18, 9.912705577010326e-19, 7.709882115452476e-19,
    -4.2404351634988616e-19, 4.415659757031872e-19, -2.1026951223961299e-19,
    2.2240044563805217e-19, -4.2891514515910067e-19, -4.295505750037808e-19,
    5.782411586589357e-20, 5.102127870520021e-20, -9.920804192677818e-20, 1.6231330769373633e-19};

// exp of a double-double, ~2^-100 relative.  Reduction x = n ln2 + r, then
// exp(r/256) by Taylor (9 terms, |r/256| < 1.4e-3) and 8 squarings.
CR_HD dd exp_dd(dd a)
{
    if (a.hi > 709.79) return {INFINITY, 0.};
    if (a.hi < -745.2) return {0., 0.};
    const double n = rint(a.hi * INV_LN2);
    const dd p = two_prod(n, LN2_HI);
    dd r = add(a, neg(p));
    r = add_d(r, -n * LN2_LO);
    r = scale(r, -8);
    dd s = {1., 0.};
#pragma unroll
    for (int j = 9; j >= 2; --j) s = add_d(mul(mul(s, r), dd{INVJ_HI[j], INVJ_LO[j]}), 1.);
    s = add_d(mul(s, r), 1.);
#pragma unroll
    for (int i = 0; i < 8; ++i) s = mul(s, s);
    // 2^n in two steps so that n near the exponent limits cannot overflow the factor
    const int ni = (int)n;
    return scale(scale(s, ni / 2), ni - ni / 2);
}

// log of a positive finite double: one Newton step in double-double from the
// library's value, y = y0 + (x e^-y0 - 1) - (x e^-y0 - 1)^2/2.
CR_HD dd log_dd(double x)
{
    const double y0 = log(x);
    const dd e = exp_dd({-y0, 0.});
    dd t = add_d(mul_d(e, x), -1.);
    t = add_d(t, -0.5 * t.hi * t.hi);
    return add_d(t, y0);
}

CR_HD double exp_cr(double x)
{
    if (x != x) return x;
    const dd r = exp_dd({x, 0.});
    return r.hi + r.lo;
}

CR_HD double log_cr(double x)
{
    if (!(x > 0.) || x == INFINITY) return log(x);
    if (x == 1.) return 0.;
    const dd r = log_dd(x);
    return r.hi + r.lo;
}

CR_HD double log10_cr(double x)
{
    if (!(x > 0.) || x == INFINITY) return log10(x);
    if (x == 1.) return 0.;
    const dd r = mul(log_dd(x), dd{INV_LN10_HI, INV_LN10_LO});
    return r.hi + r.lo;
}

// log10 as the reference's C library evaluates it: the classic fdlibm splitting
// (published with fdlibm 5.3, e_log10.c) x = 2^k x', log10 x = k log10(2) + log(x')/ln 10
// with log10(2) in two parts -- not correctly rounded (~3 % of arguments differ from the
// correctly rounded value by an ulp), so the reference's results follow this form and
// so do ours.  The natural log inside is the correctly rounded one.
CR_HD double log10_fdlibm(double x)
{
    const double ivln10 = 4.34294481903251816668e-01, log10_2hi = 3.01029995663611771306e-01,
                 log10_2lo = 3.69423907715893078616e-13, two54 = 1.80143985094819840000e+16;
    union { double d; unsigned long long u; } v;
    v.d = x;
    int hx = (int)(v.u >> 32);
    int k = 0;
    if (hx < 0x00100000) {
        if (((hx & 0x7fffffff) | (unsigned)(v.u & 0xffffffffu)) == 0) return -INFINITY;
        if (hx < 0) return NAN;
        k -= 54;
        v.d = x * two54;
        hx = (int)(v.u >> 32);
    }
    if (hx >= 0x7ff00000) return x + x;
    k += (hx >> 20) - 1023;
    const int i = (int)(((unsigned)k & 0x80000000u) >> 31);
    hx = (hx & 0x000fffff) | ((0x3ff - i) << 20);
    const double y = (double)(k + i);
    v.u = ((unsigned long long)(unsigned)hx << 32) | (v.u & 0xffffffffull);
    const double z = y * log10_2lo + ivln10 * log_cr(v.d);
    return z + y * log10_2hi;
}

// 10**y, correctly rounded: exp of y*ln10 with ln10 in double-double.
CR_HD double pow10_cr(double y)
{
    if (y != y) return y;
    const dd r = exp_dd(mul_d(dd{LN10_HI, LN10_LO}, y));
    return r.hi + r.lo;
}

CR_HD double pow_cr(double x, double y)
{
    if (y == 0. || x == 1.) return 1.;
    if (!(x > 0.) || x == INFINITY || y != y || y == INFINITY || y == -INFINITY) return pow(x, y);
    const dd r = exp_dd(mul_d(log_dd(x), y));
    return r.hi + r.lo;
}

// sin and cos of a double-double with |a| < ~100: quarter-period reduction,
// Taylor on r/8 (|r/8| < 0.1, terms to r^19), three double-angle steps.
CR_HD void sincos_dd(dd a, dd &s, dd &c)
{
    const double k = rint(a.hi * (1. / PIO2_HI));
    const dd p = two_prod(k, PIO2_HI);
    dd r = add(a, neg(p));
    r = add(r, neg(two_prod(k, PIO2_LO)));
    r = add_d(r, -k * PIO2_LO2);
    r = scale(r, -3);
    const dd r2 = mul(r, r);
    dd ts = {1., 0.}, tc = {1., 0.};
#pragma unroll
    for (int j = 19; j >= 3; j -= 2)
        ts = add_d(neg(mul(mul(ts, r2), dd{INVJJ_HI[j], INVJJ_LO[j]})), 1.);
    ts = mul(ts, r);
#pragma unroll
    for (int j = 18; j >= 2; j -= 2)
        tc = add_d(neg(mul(mul(tc, r2), dd{INVJJ_HI[j], INVJJ_LO[j]})), 1.);
    for (int i = 0; i < 3; ++i) {
        const dd s2 = mul_d(mul(ts, tc), 2.);
        const dd c2 = add_d(neg(mul_d(mul(ts, ts), 2.)), 1.);
        ts = s2;
        tc = c2;
    }
    const int q = ((int)k) & 3;
    if (q == 0) { s = ts; c = tc; }
    else if (q == 1) { s = tc; c = neg(ts); }
    else if (q == 2) { s = neg(ts); c = neg(tc); }
    else { s = neg(tc); c = ts; }
}

CR_HD double cos_cr(double x)
{
    if (x != x || x == INFINITY || x == -INFINITY || fabs(x) > 100.) return cos(x);
    dd s, c;
    sincos_dd({x, 0.}, s, c);
    return c.hi + c.lo;
}

CR_HD double sin_cr(double x)
{
    if (x != x || x == INFINITY || x == -INFINITY || fabs(x) > 100. || x == 0.) return sin(x);
    dd s, c;
    sincos_dd({x, 0.}, s, c);
    return s.hi + s.lo;
}

// sin and cos of one argument from one double-double evaluation (the same bits as
// sin_cr and cos_cr)
CR_HD void sincos_cr(double x, double &s, double &c)
{
    if (x != x || x == INFINITY || x == -INFINITY || fabs(x) > 100.) {
        s = sin(x);
        c = cos(x);
        return;
    }
    dd ss, cc;
    sincos_dd({x, 0.}, ss, cc);
    s = x == 0. ? x : ss.hi + ss.lo;
    c = cc.hi + cc.lo;
}

// asin: one Newton step on sin(y) = x from the library's value,
// y = y0 + (x - sin(y0)) / cos(y0); asin(+-1) = +-pi/2 rounded.
CR_HD double asin_cr(double x)
{
    if (x != x || x == 0.) return x;
    if (x == 1.) return PIO2_HI;
    if (x == -1.) return -PIO2_HI;
    if (fabs(x) > 1.) return asin(x);
    const double y0 = asin(x);
    dd s, c;
    sincos_dd({y0, 0.}, s, c);
    const dd y = add_d(div(add_d(neg(s), x), c), y0);
    return y.hi + y.lo;
}

// acos: one Newton step on cos(y) = x from the library's value,
// y = y0 + (cos(y0) - x) / sin(y0); acos(1) = 0, acos(-1) = pi rounded.
CR_HD double acos_cr(double x)
{
    if (x != x) return x;
    if (x == 1.) return 0.;
    if (x == -1.) return 2. * PIO2_HI;
    if (fabs(x) > 1.) return acos(x);
    const double y0 = acos(x);
    dd s, c;
    sincos_dd({y0, 0.}, s, c);
    const dd y = add_d(div(add_d(c, -x), s), y0);
    return y.hi + y.lo;
}

// atan: one Newton step on tan(y) = x from the library's value,
// y = y0 + cos(y0) (x cos(y0) - sin(y0)).
CR_HD double atan_cr(double x)
{
    if (x != x || x == 0.) return x;
    if (x == INFINITY) return PIO2_HI;
    if (x == -INFINITY) return -PIO2_HI;
    const double y0 = atan(x);
    dd s, c;
    sincos_dd({y0, 0.}, s, c);
    const dd corr = mul(c, add(mul_d(c, x), neg(s)));
    const dd y = add_d(corr, y0);
    return y.hi + y.lo;
}

CR_HD double tanh_cr(double x)
{
    if (x != x) return x;
    const double ax = fabs(x);
    double t;
    if (ax > 22.) {
        t = 1.;
    } else if (ax < 0.01) {
        // expm1(2ax) by Taylor (13 terms), tanh = em / (em + 2)
        const dd z = {2. * ax, 0.};
        dd s = {1., 0.};
#pragma unroll
        for (int j = 13; j >= 2; --j) s = add_d(div_d(mul(s, z), (double)j), 1.);
        const dd em = mul(s, z);
        const dd q = div(em, add_d(em, 2.));
        t = q.hi + q.lo;
    } else {
        const dd e = exp_dd({-2. * ax, 0.});
        const dd q = div(add_d(neg(e), 1.), add_d(e, 1.));
        t = q.hi + q.lo;
    }
    return x < 0. ? -t : t;
}

}  // namespace crm
