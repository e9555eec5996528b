/* Test infrastructure: CPU restatement of the REPROBUS chemistry step.
 *
 * Each function cites the reference lines it restates; "F" is
 * src/reprobus_l137_20230723.f and "J" is src/jno.f90.  Species slots keep the
 * reference's 1-based numbering (ck[1..43], cc[1..15], tj[1..41]) so every
 * expression can be read against the Fortran line by line.  Sums and products
 * are written in the reference's left-to-right order and the file is built
 * with -ffp-contract=off, so the only expected differences from the flang
 * build are the last-bit differences of libm calls.
 *
 * Parity: pinned against the flang-built reference (tests/golden fixtures
 * and test_oracle_vs_reference) and against closed forms in the unit tests.
 */
#define _GNU_SOURCE             /* sincos */
#include "rbchem_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NBCON 43
#define NCM 15
#define NPHOT 41

/* photolysis slots, F:754-794 */
enum {
    J_O2 = 1, J_NO2, J_O3_O1D, J_O3_O3P, J_NO3_NO, J_NO3_NO2, J_N2O5, J_HNO3, J_HNO4,
    J_H2O2, J_H2O, J_N2O, J_HOCL, J_CLONO2_CL, J_CLONO2_CLO, J_CL2O2, J_HCL, J_CL2,
    J_CLNO2, J_BRO, J_BRONO2, J_HOBR, J_OCLO, J_BRCL, J_CH3O2H, J_CH2O_HCO, J_CH2O_CO,
    J_CFC11, J_CFC12, J_CFC113, J_CCL4, J_CH3CCL3, J_CH3CL, J_HCFC22, J_CH3BR, J_H1211,
    J_H1301, J_CO2, J_CH2BR2, J_HBR, J_NO
};

/* ---------------------------------------------------------------- const */
/* COMMON /constant/ Troe prefactors, F:5927-5948 (k0*300^n, kinf*300^m). */
typedef struct {
    double a0hox, a1hox, a0nox1, a1nox1, a0nox2, a1nox2, a0nox3, a1nox3,
        a0clx1, a1clx1, a0clx2, a1clx2, a0clx3, a1clx3, a0brx, a1brx, a0c1, a1c1;
} troe_t;

static troe_t troe_constants(void)
{
    troe_t c;
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

void rbo_troe_constants(double out18[18])
{
    troe_t c = troe_constants();
    /* COMMON order, F:709-713 */
    double v[18] = {c.a0hox, c.a1hox, c.a0nox1, c.a1nox1, c.a0nox2, c.a1nox2, c.a0nox3,
                    c.a1nox3, c.a0clx1, c.a1clx1, c.a0clx2, c.a1clx2, c.a0clx3, c.a1clx3,
                    c.a0brx, c.a1brx, c.a0c1, c.a1c1};
    memcpy(out18, v, sizeof v);
}

/* JPL falloff with Fc = 0.6, as written at F:930-935 (and the 8 copies). */
static double troe(double a0, double a1, double n0, double n1, double lntemp, double hnm,
                   double xln6)
{
    double ak0 = a0 * exp(lntemp * n0);
    double ak1 = a1 * exp(lntemp * n1);
    double rate = (ak0 * hnm) / (1. + ak0 * hnm / ak1);
    double lg = log10((ak0 * hnm) / ak1);
    double xpo = 1. / (1. + lg * lg);
    return rate * exp(xln6 * xpo);
}

double rbo_troe(double a0, double a1, double n0, double n1, double t, double hnm)
{
    return troe(a0, a1, n0, n1, log(t), hnm, log(0.6));
}

/* ---------------------------------------------------------------- rates */
/* Thermal rate constants of one box, F:920-1306. */
typedef struct {
    double a1, a1et, a2, a5, a6, a6b, a7, a17, a19, a26, a27, a30, a31, a36;
    double b3, b4, b6, b7, b9, b12, b22, b23, b24, b27, b28, b32, b38, b39;
    double c1a, c1b, c1c, c2, c3, c4, c5, c7, c8, c9, c10, c12, c14, c15, c17a, c17b, c18, c19;
    double d2, d3, d4, d5, d6, d7, d8, d9, d10, d11, d31, d32, d33, d34, d35, d36, d37, d60,
        d61, d62, d90, d91, d92, d93, d94, d95, d96, d97;
    double e2, e3, e4, e5a, e5b, e5c, e6, e7, e8, e9, e10, e11, e13, e15, e16, e50, e51, e52,
        e53, e54;
    double hk1, hk2, hk3, hk4, hk5;
} rates_t;

static const char *RATE_NAMES[] = {
    "a1", "a1et", "a2", "a5", "a6", "a6b", "a7", "a17", "a19", "a26", "a27", "a30", "a31", "a36",
    "b3", "b4", "b6", "b7", "b9", "b12", "b22", "b23", "b24", "b27", "b28", "b32", "b38", "b39",
    "c1a", "c1b", "c1c", "c2", "c3", "c4", "c5", "c7", "c8", "c9", "c10", "c12", "c14", "c15",
    "c17a", "c17b", "c18", "c19",
    "d2", "d3", "d4", "d5", "d6", "d7", "d8", "d9", "d10", "d11", "d31", "d32", "d33", "d34",
    "d35", "d36", "d37", "d60", "d61", "d62", "d90", "d91", "d92", "d93", "d94", "d95", "d96",
    "d97",
    "e2", "e3", "e4", "e5a", "e5b", "e5c", "e6", "e7", "e8", "e9", "e10", "e11", "e13", "e15",
    "e16", "e50", "e51", "e52", "e53", "e54",
    "hk1", "hk2", "hk3", "hk4", "hk5"};
#define NRATES ((int)(sizeof RATE_NAMES / sizeof RATE_NAMES[0]))

/* constants folded in chem3d as scalars */
static const double A3ET = 1.2e-10, A23A = 7.2e-11, A23B = 6.9e-12, A23C = 1.6e-12;

static void rates_box(const troe_t *K, double t, double hnm, double pm, rates_t *r)
{
    const double xln6 = log(0.6);
    const double lntemp = log(t);
    double deq, aux1, aux2;
    /* hydrogen, F:924-989 */
    r->a1 = troe(K->a0hox, K->a1hox, -1.3, 0.2, lntemp, hnm, xln6);
    r->a1et = 1.63e-10 * exp(60. / t);
    r->a2 = 1.4e-10 * exp(-470. / t);
    r->a5 = 1.8e-11 * exp(180. / t);
    r->a6 = 1.7e-12 * exp(-940. / t);
    r->a6b = 1.0e-14 * exp(-490. / t);
    r->a7 = 3.0e-11 * exp(200. / t);
    r->a17 = 4.8e-11 * exp(250. / t);
    r->a19 = 2.8e-12 * exp(-1800. / t);
    r->a26 = 3.3e-12 * exp(270. / t);
    r->a27 = 1.5e-12 * exp(19. / t) + 2.1e-33 * exp(920. / t) * hnm;
    r->a30 = 1.8e-12;
    r->a31 = 1.8e-12;
    r->a36 = 1.5e-13 * (1. + .6 * pm / 1013.);
    /* nitrogen, F:991-1053 */
    r->b3 = 5.1e-12 * exp(210. / t);
    r->b4 = 3.0e-12 * exp((-1500. / t));
    r->b6 = 2.1e-11 * exp((100. / t));
    r->b7 = 1.5e-11 * exp((-3600. / t));
    r->b9 = 1.2e-13 * exp((-2450. / t));
    r->b12 = troe(K->a0nox1, K->a1nox1, -4.4, -0.7, lntemp, hnm, xln6);
    r->b22 = troe(K->a0nox2, K->a1nox2, -3.0, 0.0, lntemp, hnm, xln6);
    r->b23 = troe(K->a0nox3, K->a1nox3, -3.4, -1.1, lntemp, hnm, xln6);
    deq = 2.1e-27 * exp(10900. / t);
    r->b24 = r->b23 / (deq * hnm);
    aux1 = 2.7e-17 * exp(2199. / t);
    aux2 = 1. + aux1 / (6.5e-34 * exp(1335. / t) * hnm);
    r->b27 = 2.4e-14 * exp(460. / t) + aux1 / aux2;
    r->b28 = 1.3e-12 * exp(380. / t);
    deq = 2.7e-27 * exp(11000. / t);
    r->b32 = r->b12 / (deq * hnm);
    r->b38 = 4.63e-11 * exp(20. / t);
    r->b39 = 7.25e-11 * exp(20. / t);
    /* carbon, F:1055-1118 */
    r->c1a = 1.31e-10;
    r->c1b = 0.35e-10;
    r->c1c = 0.09e-10;
    r->c2 = 2.45e-12 * exp((-1775. / t));
    r->c3 = 1.1e-10;
    r->c4 = troe(K->a0c1, K->a1c1, -3.6, 1.1, lntemp, hnm, xln6);
    r->c5 = 2.8e-12 * exp((300. / t));
    r->c7 = 3.8e-13 * exp((800. / t));
    r->c8 = 5.5e-12 * exp(125. / t);
    r->c9 = 3.4e-11 * exp((-1600. / t));
    r->c10 = 8.1e-11 * exp((-30. / t));
    r->c12 = 5.2e-12;
    r->c14 = 9.5e-14 * exp((390. / t));
    r->c15 = 3.9e-14 * exp((-900. / t));
    r->c17a = 3.8e-12 * exp((200. / t)) * 0.7;
    r->c17b = 3.8e-12 * exp((200. / t)) * 0.3;
    r->c18 = 3.3e-12 * exp((-115. / t));
    r->c19 = 1.7e-11 * exp((-800. / t));
    /* chlorine, F:1120-1218 */
    r->d2 = 2.3e-11 * exp((-200. / t));
    r->d3 = 2.8e-11 * exp(85. / t);
    r->d4 = 6.4e-12 * exp(290. / t);
    r->d5 = 7.3e-12 * exp((-1280. / t));
    r->d6 = 3.05e-11 * exp((-2270. / t));
    r->d7 = 1.4e-11 * exp(269. / t);
    r->d8 = 7.4e-12 * exp(270. / t);
    r->d9 = 6.0e-13 * exp(230. / t);
    r->d10 = 3.6e-11 * exp(-375. / t);
    r->d11 = 1.8e-12 * exp((-250. / t));
    r->d31 = troe(K->a0clx1, K->a1clx1, -3.4, -1.9, lntemp, hnm, xln6);
    r->d32 = 3.6e-12 * exp((-840. / t));
    r->d33 = 2.6e-12 * exp(290. / t);
    r->d34 = 3.0e-12 * exp((-500. / t));
    r->d35 = 1.7e-13;
    r->d36 = troe(K->a0clx2, K->a1clx2, -2., -1., lntemp, hnm, xln6);
    r->d37 = 3.4e-12 * exp((-130. / t));
    r->d60 = troe(K->a0clx3, K->a1clx3, -4.5, -2.0, lntemp, hnm, xln6);
    deq = 1.72e-27 * exp(8649. / t);
    r->d61 = r->d60 / (deq * hnm);
    r->d62 = 1.0e-12 * exp(-1590. / t) + 3.0e-11 * exp(-2450. / t);
    r->d90 = 2.3e-10;
    r->d91 = 1.4e-10;
    r->d92 = 2.0e-10;
    r->d93 = 3.3e-10;
    r->d94 = 1.64e-12 * exp((-1520. / t));
    r->d95 = 2.4e-12 * exp((-1250. / t));
    r->d96 = 1.0e-10;
    r->d97 = 1.05e-12 * exp((-1600. / t));
    /* bromine, F:1220-1285 */
    r->e2 = 1.6e-11 * exp((-780. / t));
    r->e3 = 1.9e-11 * exp(230. / t);
    r->e4 = 8.8e-12 * exp(260. / t);
    r->e5a = 9.5e-13 * exp(550. / t);
    r->e5b = 2.3e-12 * exp(260. / t);
    r->e5c = 4.1e-13 * exp(290. / t);
    r->e6 = 1.5e-12 * exp(230. / t);
    r->e7 = 4.8e-12 * exp(-310. / t);
    r->e8 = 1.9e-11 * exp(215. / t);
    r->e9 = 1.5e-10;
    r->e10 = 1.7e-11 * exp(250. / t);
    r->e11 = 5.5e-12 * exp(200. / t);
    r->e13 = troe(K->a0brx, K->a1brx, -3.2, -2.9, lntemp, hnm, xln6);
    r->e15 = 4.5e-12 * exp(460. / t);
    r->e16 = 1.2e-10 * exp((-430. / t));
    r->e50 = 1.8e-10;
    r->e51 = 2.35e-12 * exp((-1300. / t));
    r->e52 = 1.5e-10;
    r->e53 = 1.0e-10;
    r->e54 = 2.0e-12 * exp(-840. / t);
    /* oxygen, F:1287-1303 */
    r->hk1 = 4.23e-28 * hnm / (t * t);
    r->hk2 = 6.0e-34 * pow(300. / t, 2.4) * hnm;
    r->hk3 = 8.e-12 * exp((-2060. / t));
    r->hk4 = 2.15e-11 * exp(110. / t);
    r->hk5 = 3.3e-11 * exp(55. / t);
}

int rbo_rates(double t, double hnm, double pm, double *out)
{
    troe_t K = troe_constants();
    rates_t r;
    rates_box(&K, t, hnm, pm, &r);
    memcpy(out, &r, sizeof r);
    return NRATES;
}

const char *rbo_rate_name(int i) { return (i >= 0 && i < NRATES) ? RATE_NAMES[i] : NULL; }

/* ---------------------------------------------------------------- phot */
static const double DSZA[27] = {0., 5., 10., 15., 20., 25., 30., 35., 40., 45., 50., 55., 60., 65.,
                                70., 75., 80., 82., 84., 86., 88., 90., 91., 92., 93., 94., 95.};
static const double XV3[7] = {0.4, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6};

#define AJL(nn, j, k, l) ajl[((nn) - 1) + 40 * (((j) - 1) + 101 * (((k) - 1) + 27 * ((l) - 1)))]

/* phot, F:5043-5132: zero J above 95 deg, trilinear (z, sza, o3 ratio) otherwise. */
void rbo_phot_box(const double *ajl, const double *o3up, double alt, double sza, double o3t,
                  double *tj40)
{
    const double dz = 1.;
    int is = 27 - 1, i, iz, iv, j, k, l, nn;
    double zin, v3std, v3rat, stepo3, cizin, cisza, cio3t, w[8];
    for (nn = 0; nn < 40; ++nn) tj40[nn] = 0.;
    for (i = 1; i <= 27; ++i)
        if (DSZA[i - 1] >= sza && i - 1 < is) is = i - 1;
    stepo3 = (XV3[6] - XV3[0]) / (double)(7 - 1);
    zin = alt;
    zin = fmin(zin, 78.9);
    iz = (int)(zin / dz) + 1;
    i = (int)zin;
    v3std = o3up[i] - (zin - (double)i) * (o3up[i] - o3up[i + 1]);
    v3rat = o3t / v3std;
    v3rat = fmax(v3rat, XV3[0]);
    v3rat = fmin(v3rat, XV3[6]);
    iv = (int)trunc((v3rat - XV3[0]) / stepo3) + 1;
    if (iv > 7 - 1) iv = 7 - 1;
    if (!(sza <= 95.)) return;
    j = iz > 1 ? iz : 1;
    k = is > 1 ? is : 1;
    l = iv > 1 ? iv : 1;
    cizin = (zin - (double)(j - 1) * dz) / dz;
    cisza = (sza - DSZA[k - 1]) / (DSZA[k] - DSZA[k - 1]);
    cio3t = (v3rat - XV3[l - 1]) / (XV3[l] - XV3[l - 1]);
    w[0] = (1. - cizin) * (1. - cisza) * (1. - cio3t);
    w[1] = (1. - cizin) * (1. - cisza) * cio3t;
    w[2] = (1. - cizin) * cisza * (1. - cio3t);
    w[3] = (1. - cizin) * cisza * cio3t;
    w[4] = cizin * (1. - cisza) * (1. - cio3t);
    w[5] = cizin * (1. - cisza) * cio3t;
    w[6] = cizin * cisza * (1. - cio3t);
    w[7] = cizin * cisza * cio3t;
    for (nn = 1; nn <= 40; ++nn)
        tj40[nn - 1] = w[0] * AJL(nn, j, k, l) + w[1] * AJL(nn, j, k, l + 1) +
                       w[2] * AJL(nn, j, k + 1, l) + w[3] * AJL(nn, j, k + 1, l + 1) +
                       w[4] * AJL(nn, j + 1, k, l) + w[5] * AJL(nn, j + 1, k, l + 1) +
                       w[6] * AJL(nn, j + 1, k + 1, l) + w[7] * AJL(nn, j + 1, k + 1, l + 1);
}

/* ---------------------------------------------------------------- column pass */
/* total (F:5167-5184) and the jno column sums (J:103-126) for one column. */
void rbo_column(int nlev, const double *p, const double *T, const double *hc_o3,
                const double *hc_no, int st, double *o3t, double *colo2, double *colno)
{
    const double zcmt_t = 2.120e20;
    const double zcmt_j = 287.0 / (1.38e-23 * 9.81);
    const double r_o2 = 0.2095;
    double hnm_prev = 0., ccno_prev = 0.;
    for (int k = 0; k < nlev; ++k) {
        double hnm = p[k * st] / (T[k * st] * 1.38e-19);
        double ccno = hc_no[k * st] * hnm;
        if (k == 0) {
            o3t[0] = hc_o3[0] * hnm * 6.8e5;
            colo2[0] = 6.8e+05 * r_o2 * hnm;
            colno[0] = 6.8e+05 * ccno;
        } else {
            double dp = (p[k * st] - p[(k - 1) * st]) * 100.;
            dp = fmax(dp, 0.);
            o3t[k] = o3t[k - 1] + zcmt_t * (hc_o3[k * st] + hc_o3[(k - 1) * st]) * .5 * dp;
            colo2[k] = colo2[k - 1] + zcmt_j * (r_o2 + r_o2) * 0.5 * dp * 1.e-4;
            colno[k] = colno[k - 1] +
                       zcmt_j * (ccno_prev / hnm_prev + ccno / hnm) * 0.5 * dp * 1.e-4;
        }
        hnm_prev = hnm;
        ccno_prev = ccno;
    }
}

/* jno per box, J:132-246. */
static const double SIG_O2[3][6] = {
    {1.12e-23, 2.45e-23, 7.19e-23, 3.04e-22, 1.75e-21, 1.11e-20},
    {1.35e-22, 2.99e-22, 7.33e-22, 3.07e-21, 1.69e-20, 1.66e-19},
    {2.97e-22, 5.83e-22, 2.05e-21, 8.19e-21, 4.80e-20, 2.66e-19}};
static const double SIG_NO1[3][6] = {
    {0.00e+00, 1.32e-18, 6.35e-19, 7.09e-19, 2.18e-19, 4.67e-19},
    {0.00e+00, 0.00e+00, 3.05e-21, 5.76e-19, 2.29e-18, 2.21e-18},
    {1.80e-18, 1.50e-18, 5.01e-19, 7.20e-20, 6.72e-20, 1.49e-21}};
static const double SIG_NO2[3][6] = {
    {0.00e+00, 4.41e-17, 4.45e-17, 4.50e-17, 2.94e-17, 4.35e-17},
    {0.00e+00, 0.00e+00, 3.20e-21, 5.71e-17, 9.09e-17, 6.00e-17},
    {1.40e-16, 1.52e-16, 7.00e-17, 2.83e-17, 2.73e-17, 6.57e-18}};
static const double W_NO1[3][6] = {
    {0.00e+00, 5.12e-02, 1.36e-01, 1.65e-01, 1.41e-01, 4.50e-02},
    {0.00e+00, 0.00e+00, 1.93e-03, 9.73e-02, 9.75e-02, 3.48e-02},
    {4.50e-02, 1.80e-01, 2.25e-01, 2.25e-01, 1.80e-01, 4.50e-02}};
static const double W_NO2[3][6] = {
    {0.00e+00, 5.68e-03, 1.52e-02, 1.83e-02, 1.57e-02, 5.00e-03},
    {0.00e+00, 0.00e+00, 2.14e-04, 1.08e-02, 1.08e-02, 3.86e-03},
    {5.00e-03, 2.00e-02, 2.50e-02, 2.50e-02, 2.00e-02, 5.00e-03}};
static const double DLAMBDA[3] = {2.3, 1.5, 1.5};
static const double I0[3] = {3.98e+11, 2.21e+11, 2.30e+11};
static const double SIG_O3[3] = {4.6e-19, 6.7e-19, 7.1e-19};

double rbo_jno_box(double colo2, double colno, double colo3, double hnm, double sza_top)
{
    const double factor = 3.141592 / 180.;
    const double d = 1.65e+09, a = 5.1e+07, kq = 1.5e-09;
    if (!(sza_top <= 89.0)) return 0.;
    double sec = 1. / cos(sza_top * factor);
    colo2 = colo2 * sec;
    colno = colno * sec;
    colo3 = colo3 * sec;
    double n2 = hnm * 0.78;
    double prob = d / (a + d + kq * n2);
    double bd[3];
    for (int b = 0; b < 3; ++b) {
        double t_o3 = exp(-SIG_O3[b] * colo3);
        double s = 0.;
        for (int i = 0; i < 6; ++i) {
            double term = exp(-SIG_O2[b][i] * colo2) *
                          (W_NO1[b][i] * SIG_NO1[b][i] * exp(-SIG_NO1[b][i] * colno) +
                           W_NO2[b][i] * SIG_NO2[b][i] * exp(-SIG_NO2[b][i] * colno));
            s = (i == 0) ? term : s + term;
        }
        bd[b] = DLAMBDA[b] * I0[b] * t_o3 * prob * s;
    }
    return fmax(bd[0] + bd[1] + bd[2], 1.e-30);
}

/* ---------------------------------------------------------------- density */
static const double DX[22] = {
    2.393284E-02, -4.359335E-05, 7.961181E-08, 0.0, -0.198716351, 1.39564574E-03,
    -2.020633E-06, 0.51684706, -3.0539E-03, 4.505475E-06, -0.30119511, 1.840408E-03,
    -2.7221253742E-06, -0.11331674116, 8.47763E-04, -1.22336185E-06, 0.3455282,
    -2.2111E-03, 3.503768245E-06, -0.2315332, 1.60074E-03, -2.5827835E-06};

/* F:4059-4069 */
double rbo_density(double ws, double wn, double t)
{
#define X(i) DX[(i) - 1]
    double w = ws + wn;
    double wh = 1.0 - w;
    double v1 = X(1) + X(2) * t + X(3) * (t * t) + X(4) * (t * t * t);
    double vs = X(5) + X(6) * t + X(7) * (t * t) + (X(8) + X(9) * t + X(10) * (t * t)) * w +
                (X(11) + X(12) * t + X(13) * (t * t)) * w * w;
    double vn = X(14) + X(15) * t + X(16) * (t * t) + (X(17) + X(18) * t + X(19) * (t * t)) * w +
                (X(20) + X(21) * t + X(22) * (t * t)) * w * w;
    double vmcal = wh / 18.0160 * v1 + vs * ws / 98.080 + vn * wn / 63.0160;
    return 1.0 / vmcal * 0.001;
#undef X
}

/* ---------------------------------------------------------------- analytic + hetero */
/* Elementary functions of the PSC block up to the liquid composition (NAT, ice,
 * binary and ternary liquid).  Default: the C library, as the reference.  Built
 * with -DRBO_CR_PSC: binary128 evaluation rounded once (correctly rounded), with
 * log10 in the fdlibm form the C library uses -- the variant that isolates the
 * reference's library rounding from the algorithm (the ternary solve amplifies
 * last-bit differences by up to 1e10). */
#ifdef RBO_CR_PSC
#include <quadmath.h>
static double psc_log(double x) { return (double)logq((__float128)x); }
static double psc_exp(double x) { return (double)expq((__float128)x); }
static double psc_pow(double x, double y) { return (double)powq((__float128)x, (__float128)y); }
static double psc_atan(double x) { return (double)atanq((__float128)x); }
static double psc_cos(double x) { return (double)cosq((__float128)x); }
static double psc_log10(double x)
{
    const double ivln10 = 4.34294481903251816668e-01, log10_2hi = 3.01029995663611771306e-01,
                 log10_2lo = 3.69423907715893078616e-13;
    union { double d; unsigned long long u; } v;
    int hx, k, i;
    double y, z;
    v.d = x;
    hx = (int)(v.u >> 32);
    if (!(x > 0.) || hx < 0x00100000 || hx >= 0x7ff00000) return log10(x);
    k = (hx >> 20) - 1023;
    i = (int)(((unsigned)k & 0x80000000u) >> 31);
    hx = (hx & 0x000fffff) | ((0x3ff - i) << 20);
    y = (double)(k + i);
    v.u = ((unsigned long long)(unsigned)hx << 32) | (v.u & 0xffffffffull);
    z = y * log10_2lo + ivln10 * psc_log(v.d);
    return z + y * log10_2hi;
}
#else
#define psc_log log
#define psc_exp exp
#define psc_pow pow
#define psc_atan atan
#define psc_cos cos
#define psc_log10 log10
#endif
typedef struct {
    double het[10];               /* het[1..9] */
    double parthno3, parthno3sl, parth2o, parthcl;
    double surf, vsed;
    int lnat, lice, condliq;
} aero_t;

/* Uptake on NAT/ice with the gas-diffusion limit (F:3815-3880, F:3930-3993). */
static double solid_rate(double g, double t, double mw, double r, double n, double ptot,
                         double z)
{
    return g * 4.56E4 * sqrt(t / mw) * (r * r) * n / (1.0 + 3.3E4 * g * r * ptot / t) / z;
}

/* analytic F:2620-3361 then hetero F:3365-4029, for one heterogeneous-level box. */
static void aerosol_box(double temp, double ptot, double hnm, const double *ck, aero_t *o)
{
    const double r = 8.205e-5, sigma = 1.8;
    static const double qn[10] = {14.5734, 0.0615994, -1.14895, 0.691693, -0.098863,
                                  0.0051579, 0.123472, -0.115574, 0.0110113, 0.0097914};
    static const double qs[10] = {14.4700, 0.0638795, -3.29597, 1.778224, -0.223244,
                                  0.0086486, 0.536695, -0.335164, 0.0265153, 0.0157550};
    static const double kn[7] = {-39.136, 6358.4, 83.29, -17650.0, 198.53, -11948, -28.469};
    static const double ks[7] = {-21.661, 2724.2, 51.81, -15732.0, 47.004, -6969.0, -4.6183};
    static const double a1c[3] = {12.37208932, 11.820654354, -180.06541028};
    static const double b1c[3] = {-0.16125516114, -0.20786404244, -0.38601102592};
    static const double c1c[3] = {-30.490657554, -4.807306373, -93.317846778};
    static const double d1c[3] = {-2.1133114241, -5.1727540348, 273.88132245};
    static const double a2c[3] = {13.455394705, 12.891938068, -176.95814097};
    static const double b2c[3] = {-0.1921312255, -0.232338847708, -0.36257048154};
    static const double c2c[3] = {-34.285174607, -6.4261237757, -90.469744201};
    static const double d2c[3] = {-1.7620073078, -4.9005471319, 267.45509988};
    static const double bnre[7] = {-3.18657, 0.992696, -0.153193e-2, -0.987059e-3,
                                   -0.578878e-3, 0.855176e-4, -0.327815e-5};
    const double ndrop = 10., nnat_small_max = 1., rnats = 0.5e-4, rnatl = 6.5e-4,
                 rice = 10.e-4;
    const double pi = 2. * asin(1.0), amu = 1.66e-24;

    double rmean = 0., aliq = 0., nnats = 0., nnatl = 0., anat = 0., nice = 0., aice = 0.;
    double vnat = 0., vnatl = 0., vice = 0.;
    double parthno3l = 1., parthno3s = 1.;
    int lnat = 0, lice = 0, condliq = 0;
    double t, qh2o, pw, qhno3, pn0, pn = 0., tice;
    double ms, ws, wn, mn, ns, msb = 0.;
    double aw = 0., wt = 0., mh2so4 = 0., hhcl = 0., hclono2 = 0., hhocl = 0., hhobr = 0.;
    double dens;

    o->vsed = 0.;
    o->parthno3 = 1.;
    o->parthno3sl = 1.;
    o->parth2o = 1.;
    o->parthcl = 1.;

    /* F:2781-2788 */
    qh2o = ck[3] / hnm;
    pw = qh2o * ptot / 1013.0;
    pw = fmin(pw, 2.e-3 / 1013.);
    qhno3 = ck[5] / hnm;
    pn0 = qhno3 * ptot / 1013.0;

    /* NAT, F:2795-2857 */
    {
        double tt = fmin(temp, 273.);
        double pn0t = pn0 * 1013. * 0.75;
        double zmt = -2.7836 - 0.00088 * tt;
        double zbt = 38.9855 - 11397.0 / tt + 0.009179 * tt;
        double zhno3eq = psc_pow(10.0, zmt * psc_log10(pw * 1013. * 0.75) + zbt);
        if (pn0t / zhno3eq >= 1.) {
            double hno3s, vnat_small_max;
            lnat = 1;
            pn = zhno3eq / (0.75 * 1013.);
            hno3s = (pn0 - pn) * 1013. * hnm / ptot;
            hno3s = fmax(hno3s, 0.);
            vnat = hno3s * amu * 117. / 1.35;
            vnat_small_max = nnat_small_max * (4. / 3. * pi * psc_pow(rnats, 3.));
            if (vnat > vnat_small_max) {
                nnats = nnat_small_max;
                nnatl = (vnat - vnat_small_max) / (4. / 3. * pi * psc_pow(rnatl, 3.));
                vnatl = vnat - vnat_small_max;
            } else {
                nnats = vnat / (4. / 3. * pi * psc_pow(rnats, 3.));
                nnatl = 0.;
                vnatl = 0.;
            }
            anat = nnats * 4 * pi * (rnats * rnats) + nnatl * 4 * pi * (rnatl * rnatl);
            parthno3s = 1.0 - (pn0 - pn) / pn0;
            parthno3s = fmax(parthno3s, 0.);
            parthno3s = fmin(parthno3s, 1.);
            o->parthno3sl = 1.0 - (pn0 - pn) * vnatl / vnat / pn0;
            o->parthno3sl = fmax(o->parthno3sl, 0.);
            o->parthno3sl = fmin(o->parthno3sl, 1.);
        }
    }

    /* ice, F:2861-2913 */
    {
        double tt = fmin(temp, 273.);
        tice = 2668.70 / (10.4310 - (psc_log(pw) + psc_log(760.0)) / psc_log(10.0));
        if (tt <= tice) {
            double zh2oeq, ph2o, h2os, zmt, zbt, zhno3eq;
            lice = 1;
            lnat = 0;
            anat = 0.;
            zh2oeq = psc_pow(10.0, -2668.7 / tt + 10.4310);
            ph2o = zh2oeq / 760.0;
            h2os = fmax(pw - ph2o, 0.) * 1013. * hnm / ptot;
            vice = h2os * amu * 18. / 0.928;
            nice = vice / (4. / 3. * pi * psc_pow(rice, 3.));
            aice = nice * 4 * pi * (rice * rice);
            o->parth2o = 1.0 - fmax(pw - ph2o, 0.) / pw;
            o->parth2o = fmax(o->parth2o, 0.);
            o->parth2o = fmin(o->parth2o, 1.);
            zmt = -2.7836 - 0.00088 * tt;
            zbt = 38.9855 - 11397.0 / tt + 0.009179 * tt;
            zhno3eq = psc_pow(10.0, zmt * psc_log10(pw * o->parth2o * 1013. * 0.75) + zbt);
            pn = zhno3eq / (0.75 * 1013.);
            parthno3s = 1.0 - (pn0 - pn) / pn0;
            parthno3s = fmax(parthno3s, 0.);
            parthno3s = fmin(parthno3s, 1.);
            o->parthno3sl = parthno3s;
        }
    }

    /* liquid aerosol, F:2921-3083 */
    {
        double a, b, c, det, qh2so4, wts, muh2so4, ph2so4;
        condliq = 1;
        t = fmax(temp, fmax(tice - 3.0, 185.));
        qh2o = ck[3] * o->parth2o / hnm;
        pw = qh2o * ptot / 1013.0;
        pw = fmin(pw, 2.e-3 / 1013.);
        pw = fmax(pw, 2.e-5 / 1013.);
        qh2so4 = fmax(ck[12] / hnm, 1.e-18);
        ns = ptot * 100.0 * qh2so4 / 8.314 / t;
        qhno3 = ck[5] * parthno3s / hnm;
        pn0 = ptot / 1013.0 * qhno3;

        a = ks[2] + ks[3] / t;
        b = ks[0] + ks[1] / t;
        c = ks[4] + ks[5] / t + ks[6] * psc_log(t) - psc_log(pw);
        det = (b * b) - 4. * a * c;
        if (det > 0.) {
            double xsb = (-b - sqrt(det)) / (2. * a);
            msb = 55.51 * xsb / (1.0 - xsb);
        } else {
            msb = 0.;
            condliq = 0;
        }
        ms = msb;
        ws = msb * 0.098076 / (1.0 + msb * 0.098076);
        mn = 0.0;
        wn = 0.0;

        wts = fmax(41., ws * 100.);
        muh2so4 = 4.184 * (1.514e4 - 286. * (wts - 40.) + 1.08 * ((wts - 40.) * (wts - 40.)) -
                           3941. / psc_pow(wts - 40., 0.1));
        ph2so4 = psc_exp(-10156. / t + 16.2590 - muh2so4 / (8.314 * t));
        if (qh2so4 * ptot / 1013. <= ph2so4) {
            ms = 0.;
            ws = 0.;
            condliq = 0;
        }

        if (condliq == 1 && t < 215.) {
            double tf = t, tt = r * tf * ns, tr = 1.0e4 / tf - 43.4782608, pr = psc_log(pw) + 18.4;
            double hsb, xnb, mnb, hnb, phi, msf, mnf, wsf, wnf, A, B, C, q;
            hsb = qs[0] + qs[1] * (tr * tr) +
                  (qs[2] + qs[3] * tr + qs[4] * (tr * tr) + qs[5] * (tr * tr * tr)) * pr +
                  (qs[6] + qs[7] * tr + qs[8] * (tr * tr)) * (pr * pr) +
                  qs[9] * tr * (pr * pr * pr);
            hsb = psc_exp(hsb);
            q = (kn[0] + kn[1] / tf);
            xnb = 1.0 / (2.0 * (kn[2] + kn[3] / tf)) *
                  (-kn[0] - kn[1] / tf -
                   psc_pow(q * q - 4.0 * (kn[2] + kn[3] / tf) *
                                   (kn[4] + kn[5] / tf + kn[6] * psc_log(tf) - psc_log(pw)),
                       0.5));
            mnb = 55.51 * xnb / (1.0 - xnb);
            hnb = qn[0] + qn[1] * (tr * tr) +
                  (qn[2] + qn[3] * tr + qn[4] * (tr * tr) + qn[5] * (tr * tr * tr)) * pr +
                  (qn[6] + qn[7] * tr + qn[8] * (tr * tr)) * (pr * pr) +
                  qn[9] * tr * (pr * pr * pr);
            hnb = psc_exp(hnb);

            A = (tt * hnb * (mnb * mnb) - tt * hsb * mnb * msb - 2.0 * (mnb * mnb) * msb +
                 mnb * (msb * msb) + hnb * mnb * msb * pn0 - hsb * (msb * msb) * pn0) /
                ((mnb * mnb) - mnb * msb);
            B = msb * (-2.0 * tt * hnb * mnb + tt * hsb * msb + mnb * msb - hnb * msb * pn0) /
                (mnb - msb);
            C = (tt * hnb * mnb * (msb * msb)) / (mnb - msb);

            {
                double u = (A * A) - 3.0 * B;
                double v = -2.0 * (A * A * A) + 9.0 * A * B - 27.0 * C;
                phi = psc_atan(sqrt(fmax(4.0 * (u * u * u) - (v * v), 0.)) / v);
            }
            if (phi < 0.) phi = phi + pi;
            msf = -1.0 / 3.0 * (A + 2.0 * sqrt((A * A) - 3.0 * B) * psc_cos((pi + phi) / 3.0));
            msf = fmax(msf, 0.);
            mnf = mnb * (1.0 - msf / msb);
            mnf = fmax(mnf, 0.);
            pn = mnf / (hnb * mnf / (mnf + msf) + hsb * msf / (mnf + msf));
            wsf = msf * 0.098076 / (1.0 + msf * 0.098076 + mnf * 0.063012);
            wnf = mnf * 0.063012 / (1.0 + msf * 0.098076 + mnf * 0.063012);
            ms = msf;
            mn = mnf;
            ws = wsf;
            wn = wnf;
            parthno3l = 1.0 - (pn0 - pn) / pn0;
            parthno3l = fmin(parthno3l, 1.);
        }
        ws = fmax(ws, 0.005);
        (void)mn;
    }

    /* density F:3089, volume/radius/area and solubilities F:3095-3228 */
    dens = rbo_density(ws, wn, temp);
    if (condliq == 1) {
        double vliq, rmode, ph2o0, y1, y2, m, bigx, qhcl, phcl0, phcl1, sclono2, shocl, mh2so4c,
            bh2so4;
        double ls = log(sigma);
        vliq = ns * 98.076e-6 / ws / dens;
        rmode = pow(3.0 * vliq / (4. * pi * ndrop) * exp(-9.0 / 2.0 * (ls * ls)), 1. / 3.);
        rmean = rmode * exp(0.5 * (ls * ls));
        aliq = 4.0 * pi * (rmode * rmode) * ndrop * exp(2.0 * (ls * ls));

        t = fmax(temp, fmax(tice - 3.0, 185.));
        ph2o0 = exp(18.452406985 - 3505.1578807 / t - 330918.55082 / (t * t) +
                    12725068.262 / (t * t * t));
        aw = pw * 1013. / ph2o0;
        aw = fmin(aw, 1.);
        if (aw <= 0.05) {
            y1 = a1c[0] * pow(aw, b1c[0]) + c1c[0] * aw + d1c[0];
            y2 = a2c[0] * pow(aw, b2c[0]) + c2c[0] * aw + d2c[0];
        } else if (aw < 0.85) {
            y1 = a1c[1] * pow(aw, b1c[1]) + c1c[1] * aw + d1c[1];
            y2 = a2c[1] * pow(aw, b2c[1]) + c2c[1] * aw + d2c[1];
        } else {
            y1 = a1c[2] * pow(aw, b1c[2]) + c1c[2] * aw + d1c[2];
            y2 = a2c[2] * pow(aw, b2c[2]) + c2c[2] * aw + d2c[2];
        }
        m = y1 + (t - 190.) * (y2 - y1) / 70.;
        wt = 9800. * m / (98. * m + 1000.);
        wt = fmax(wt, 0.5);
        mh2so4 = dens * wt / 9.8;
        bigx = wt / (wt + (100. - wt) * 98. / 18.);
        hhcl = (0.094 - 0.61 * bigx + 1.2 * (bigx * bigx)) *
               exp(-8.68 + (8515. - 10718. * pow(bigx, 0.7)) / t);
        qhcl = ck[13] / hnm;
        phcl0 = qhcl * ptot / 1013.0;
        phcl1 = (phcl0 / (r * t)) / (1. / (r * t) + hhcl * ns / ms);
        o->parthcl = 1.0 - (phcl0 - phcl1) / phcl0;
        sclono2 = 0.306 + 24. / t;
        hclono2 = 1.6e-6 * exp(4710. / t) * exp(-sclono2 * mh2so4);
        shocl = 0.0776 + 59.18 / t;
        hhocl = 1.91e-6 * exp(5862.4 / t) * exp(-shocl * mh2so4);
        hhobr = 4.6e-4 * exp(4.5e3 / t);
        mh2so4c = -1.977e-4 * pow(wt, 2.) - 2.096e-2 * wt + 4.445;
        bh2so4 = -8.979e-5 * pow(wt, 2.) + 2.141e-2 * wt - 6.067;
        (void)mh2so4c;
        (void)bh2so4;
    }

    /* surface, F:3240-3246; total HNO3 partition, F:3252-3258 */
    o->surf = aliq + anat * (double)lnat + aice * (double)lice;
    o->parthno3 = parthno3l * parthno3s;
    o->parthno3 = fmin(o->parthno3, 1.);

    /* settling velocity, F:3264-3348 */
    if (lice == 1 || lnat == 1) {
        const double rhoice = 0.928 * 1000.0, rhonat = 1.350 * 1000.0;
        double tt = temp;
        double rpart = ((double)lice * rice + (double)lnat * rnatl) * 1.e-2;
        rpart = fmin(rpart, 535.e-6);
        if (rpart >= 0.5e-6) {
            double rhopart = (double)lice * rhoice + (double)lnat * rhonat;
            double rhoair = (hnm * 28.97 / 6.022e23) * 1000.0;
            double tcel = tt - 273.15;
            double netha = (1.718 + 0.0049 * tcel - 1.2e-5 * tcel * tcel) * 1.e-5;
            double cdnre2 = 32.0 * (rpart * rpart * rpart) * (rhopart - rhoair) * rhoair *
                            9.80665 / (3.0 * (netha * netha));
            double lx = log(cdnre2);
            double ynre = bnre[0] + bnre[1] * lx + bnre[2] * (lx * lx) + bnre[3] * (lx * lx * lx) +
                          bnre[4] * ((lx * lx) * (lx * lx)) +
                          bnre[5] * ((lx * lx) * (lx * lx) * lx) +
                          bnre[6] * ((lx * lx) * (lx * lx) * (lx * lx));
            double nre = exp(ynre);
            if (nre >= 1.e-2) {
                o->vsed = netha * nre / (2 * rhoair * rpart);
            } else {
                double lambda0 = 6.6e-8;
                double lambda = lambda0 * (1013.25 / ptot) * (tt / 293.15);
                double us = 2.0 * (rpart * rpart) * 9.80665 * (rhopart - rhoair) / (9.0 * netha);
                double shapek = 1.2;
                o->vsed = us / shapek *
                          (1.0 + 1.246 * lambda / rpart +
                           0.42 * lambda / rpart * exp(-0.87 * rpart / lambda));
            }
        }
    }

    /* ---- hetero, F:3436-4027 ---- */
    {
        double kl[10] = {0}, ks_[10] = {0};
        if (condliq == 1) {
            double T = temp;
            double zh2o = fmax(ck[3], 1.E-08 * hnm);
            double zhcl = fmax(ck[13], 1.E-13 * hnm);
            double zclono2 = fmax(ck[14], 1.E-13 * hnm);
            double zhocl = fmax(ck[15], 1.E-13 * hnm);
            double zhobr = fmax(ck[26], 1.E-15 * hnm);
            double zhbr = fmax(ck[19], 1.E-15 * hnm);
            double phcl = (zhcl / hnm) * ptot / 1013.;
            double pclono2 = (zclono2 / hnm) * ptot / 1013.;
            double phocl = (zhocl / hnm) * ptot / 1013.;
            double phobr = (zhobr / hnm) * ptot / 1013.;
            double phbr = (zhbr / hnm) * ptot / 1013.;
            double t0, biga, eta, ah, dclono2, cclono2, kh2o, kh, khydr, gbh2o, mhcl, khcl,
                lclono2, fclono2, grxn, gbhcl, gs, fhcl, gps, gbhclp, gb, gclono2, gclono2hcl,
                gclono2h2o;
            double chocl, dhocl, khoclhcl, lhocl, fhocl, ghoclhcl, mhobr, k2hobrhcl, ratio,
                k1hobrhcl, rhobrhcl, cbar, gxrn, gbrno3h2o;
            (void)phocl;
            (void)phbr;
            t0 = 144.11 + 0.166 * wt - 0.015 * (wt * wt) + 2.18e-4 * (wt * wt * wt);
            biga = 169.5 + 5.18 * wt - 0.0825 * (wt * wt) + 3.27e-3 * (wt * wt * wt);
            eta = biga * pow(T, -1.43) * exp(448. / (T - t0));
            ah = exp(60.51 - 0.095 * wt + 0.0077 * (wt * wt) - 1.61e-5 * (wt * wt * wt) -
                     (1.76 + 2.52e-4 * (wt * wt)) * sqrt(T) +
                     (-805.89 + 253.05 * pow(wt, 0.076)) / sqrt(T));
            dclono2 = 5.e-8 * T / eta;
            cclono2 = 1474. * sqrt(T);
            kh2o = 1.95e10 * exp(-2800. / T);
            kh = 1.22e12 * exp(-6200. / T);
            khydr = kh2o * aw + kh * ah * aw;
            gbh2o = 4 * hclono2 * 0.082 * T * sqrt(dclono2 * khydr) / cclono2;
            mhcl = hhcl * phcl;
            khcl = 7.9e11 * ah * dclono2 * mhcl;
            lclono2 = sqrt(dclono2 / (khydr + khcl));
            fclono2 = 1. / tanh(rmean / lclono2) - lclono2 / rmean;
            grxn = fclono2 * gbh2o * sqrt(1. + khcl / khydr);
            gbhcl = grxn * khcl / (khcl + khydr);
            gs = 66.12 * hclono2 * mhcl * exp(-1374. / T);
            fhcl = 1. / (1 + 0.612 * (gs + gbhcl) * pclono2 / phcl);
            gps = fhcl * gs;
            gbhclp = fhcl * gbhcl;
            gb = gbhclp + grxn * khydr / (khcl + khydr);
            gclono2 = 1. / (1. + 1. / (gps + gb));
            gclono2hcl = gclono2 * (gps + gbhclp) / (gps + gb);
            gclono2h2o = gclono2 - gclono2hcl;
            kl[1] = 0.25 * gclono2h2o * cclono2 * aliq / zh2o;
            kl[2] = 0.25 * gclono2hcl * cclono2 * aliq / zhcl;

            chocl = 2009. * sqrt(T);
            dhocl = 6.4e-8 * T / eta;
            khoclhcl = 1.25e9 * ah * dhocl * mhcl;
            lhocl = sqrt(dhocl / khoclhcl);
            fhocl = 1. / tanh(rmean / lhocl) - lhocl / rmean;
            grxn = 4. * hhocl * 0.082 * T * sqrt(dhocl * khoclhcl) / chocl;
            ghoclhcl = 1. / (1. + 1. / (fhocl * grxn * fhcl));
            kl[5] = 0.25 * ghoclhcl * chocl * aliq / zhcl;

            mhobr = hhobr * phobr;
            k2hobrhcl = exp(0.542 * wt - 6.44e3 / T + 10.3);
            ratio = mhobr / mhcl;
            if (ratio < 1.) {
                double dhobr, lhobr, fhobr, chobr, ghobrhcl;
                k1hobrhcl = k2hobrhcl * mhcl;
                dhobr = 6.2e-8 * T / eta;
                lhobr = sqrt(dhobr / k1hobrhcl);
                fhobr = 1. / tanh(rmean / lhobr) - lhobr / rmean;
                chobr = 1477. * sqrt(T);
                ghobrhcl = 4. * hhobr * 0.082 * T * sqrt(dhobr * k1hobrhcl) * fhobr / chobr;
                rhobrhcl = 0.25 * ghobrhcl * chobr * aliq / zhcl;
            } else {
                double dhcl, lhcl, fhcl2, chcl, ghobrhcl;
                k1hobrhcl = k2hobrhcl * mhobr;
                dhcl = 7.8e-8 * T / eta;
                lhcl = sqrt(dhcl / k1hobrhcl);
                fhcl2 = 1. / tanh(rmean / lhcl) - lhcl / rmean;
                chcl = 2408. * sqrt(T);
                ghobrhcl = 4. * hhcl * 0.082 * T * sqrt(dhcl * k1hobrhcl) * fhcl2 / chcl;
                rhobrhcl = 0.25 * ghobrhcl * chcl * aliq / zhobr;
            }
            kl[7] = rhobrhcl;
            kl[8] = 0.;
            kl[9] = 0.;
            cbar = 1400.1 * sqrt(T);
            kl[3] = 0.25 * 0.1 * cbar * aliq / zh2o;
            kl[4] = 0.;
            gxrn = exp(29.24 - 0.396 * 100. * ws) + 0.114;
            gbrno3h2o = 1. / (1. / 0.805 + 1. / gxrn);
            cbar = 1221.4 * sqrt(T);
            kl[6] = 0.25 * gbrno3h2o * cbar * aliq / zh2o;
        }
        if (lnat == 1) {
            double T = temp;
            double zh2o = fmax(ck[3], 1.E-08 * hnm);
            double zhcl = fmax(ck[13], 1.E-13 * hnm);
            /* NAT gammas F:3791-3799; two size modes each */
#define NATR(g, mw, z) \
    (solid_rate(g, T, mw, rnats, nnats, ptot, z) + solid_rate(g, T, mw, rnatl, nnatl, ptot, z))
            ks_[1] = NATR(0.004, 97.45, zh2o);
            ks_[2] = NATR(0.2, 97.45, zhcl);
            ks_[3] = NATR(4.0E-4, 108.0, zh2o);
            ks_[6] = NATR(0., 142.0, zh2o);
            ks_[4] = NATR(3.0E-3, 108.0, zhcl);
            ks_[5] = NATR(0.1, 52.5, zhcl);
            ks_[7] = NATR(0., 96.9, zhcl);
            ks_[8] = 0.;
            ks_[9] = 0.;
#undef NATR
        }
        if (lice == 1) {
            double T = temp;
            double zh2o = fmax(ck[3], 1.E-08 * hnm);
            double zhcl = fmax(ck[13], 1.E-13 * hnm);
            double zhobr = fmax(ck[26], 1.E-13 * hnm);
            ks_[2] = solid_rate(0.3, T, 97.45, rice, nice, ptot, zhcl);
            ks_[1] = solid_rate(0.3, T, 97.45, rice, nice, ptot, zh2o);
            ks_[3] = solid_rate(0.02, T, 108.0, rice, nice, ptot, zh2o);
            ks_[4] = solid_rate(0.03, T, 108.0, rice, nice, ptot, zhcl);
            ks_[5] = solid_rate(0.2, T, 52.5, rice, nice, ptot, zhcl);
            ks_[6] = solid_rate(0.3, T, 142.0, rice, nice, ptot, zh2o);
            ks_[7] = solid_rate(0.3, T, 96.9, rice, nice, ptot, zhcl);
            ks_[8] = solid_rate(0.1, T, 96.9, rice, nice, ptot, zhobr);
        }
        /* F:3997-4023 */
        for (int i = 1; i <= 8; ++i) o->het[i] = kl[i] + (double)(lnat + lice) * ks_[i];
        o->het[9] = 0.;
    }
    o->lnat = lnat;
    o->lice = lice;
    o->condliq = condliq;
}

void rbo_aerosol_box(double t, double p, double hnm, const double *ck_1based, double het[9],
                     double part[5], double *surf, double *vsed, int flags[3])
{
    aero_t o;
    aerosol_box(t, p, hnm, ck_1based, &o);
    for (int i = 0; i < 9; ++i) het[i] = o.het[i + 1];
    part[0] = o.parthno3;
    part[1] = o.parthno3sl;
    part[2] = o.parth2o;
    part[3] = o.parthcl;
    part[4] = 1.;
    *surf = o.surf;
    *vsed = o.vsed;
    flags[0] = o.lnat;
    flags[1] = o.lice;
    flags[2] = o.condliq;
}

/* ---------------------------------------------------------------- one box */
typedef struct {
    double t, pm, sza, sza_top, alt, o3t, colo2, colno; /* inputs */
    int het_level;
} boxenv_t;

/* Anderson acceleration (Walker & Ni 2011, type II, depth m <= AA_MAX) of the sweep map F
 * on x = (ck[1..42], cc[1..15]): from x_k and f = F(x_k), the next iterate is
 * f - dF gamma, gamma minimising || w (g_k - dG gamma) || with g = F(x) - x, dF/dG the
 * differences of the last m pairs, weights w_i = 1/max(|f_i|, floor) (relative scale).
 * Components that would turn non-positive keep f_i. */
#define AA_MAX 4
#define NX (NBCON - 1 + NCM)
/* converged mode: successive plain sweeps before the distance-to-fixed-point estimate,
 * and the cap of its contraction rate (changes at rounding level have ratios near 1 at
 * random: the cap accepts a change below ~1e-3 tol whatever its ratios) */
#define RBO_CONV_NPLAIN 3
#define RBO_CONV_RHOMAX 0.999
/* ... and a change that stays below tol for this many successive plain sweeps without
 * decreasing monotonically (it rose at least once) is rounding noise: converged */
#define RBO_CONV_STALL 8
typedef struct {
    int m, n;                       /* depth; stored pairs (<= m + 1, ring) */
    int head;
    double f[AA_MAX + 1][NX], g[AA_MAX + 1][NX];
} aa_t;

static void aa_update(aa_t *A, double *x, const double *xprev, double fl)
{
    /* x holds f = F(x_k) on entry and x_{k+1} on exit */
    double w[NX], gk[NX];
    for (int i = 0; i < NX; ++i) {
        gk[i] = x[i] - xprev[i];
        const double s = fmax(fabs(x[i]), fl);
        w[i] = 1. / (s * s);
    }
    const int slot = A->head;
    memcpy(A->f[slot], x, sizeof(double) * NX);
    memcpy(A->g[slot], gk, sizeof(double) * NX);
    A->head = (A->head + 1) % (A->m + 1);
    if (A->n < A->m + 1) A->n++;
    const int mk = A->n - 1;
    if (mk < 1) return;
    /* columns j = 1..mk: newest minus the one before it */
    double dg[AA_MAX][NX], df[AA_MAX][NX];
    for (int j = 0; j < mk; ++j) {
        const int a = (slot - j + (A->m + 1)) % (A->m + 1);
        const int b = (slot - j - 1 + 2 * (A->m + 1)) % (A->m + 1);
        for (int i = 0; i < NX; ++i) {
            dg[j][i] = A->g[a][i] - A->g[b][i];
            df[j][i] = A->f[a][i] - A->f[b][i];
        }
    }
    /* normal equations (dG^T W dG) gamma = dG^T W g_k, Gaussian elimination with pivoting */
    double M[AA_MAX][AA_MAX + 1];
    for (int r = 0; r < mk; ++r) {
        for (int c = 0; c < mk; ++c) {
            double acc = 0.;
            for (int i = 0; i < NX; ++i) acc += dg[r][i] * w[i] * dg[c][i];
            M[r][c] = acc;
        }
        double acc = 0.;
        for (int i = 0; i < NX; ++i) acc += dg[r][i] * w[i] * gk[i];
        M[r][mk] = acc;
    }
    for (int r = 0; r < mk; ++r) M[r][r] *= (1. + 1.e-10);      /* Tikhonov nudge */
    for (int c = 0; c < mk; ++c) {
        int piv = c;
        for (int r = c + 1; r < mk; ++r)
            if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
        if (!(fabs(M[piv][c]) > 0.)) return;                    /* singular: plain step */
        if (piv != c)
            for (int k = 0; k <= mk; ++k) {
                const double t_ = M[c][k]; M[c][k] = M[piv][k]; M[piv][k] = t_;
            }
        for (int r = c + 1; r < mk; ++r) {
            const double f_ = M[r][c] / M[c][c];
            for (int k = c; k <= mk; ++k) M[r][k] -= f_ * M[c][k];
        }
    }
    double gam[AA_MAX];
    for (int r = mk - 1; r >= 0; --r) {
        double acc = M[r][mk];
        for (int k = r + 1; k < mk; ++k) acc -= M[r][k] * gam[k];
        gam[r] = acc / M[r][r];
        if (!(fabs(gam[r]) <= 1.e6)) return;                    /* wild: plain step */
    }
    for (int i = 0; i < NX; ++i) {
        double v = x[i];
        for (int j = 0; j < mk; ++j) v -= gam[j] * df[j][i];
        if (v > 0. && v <= 1.7976931348623157e308) x[i] = v;
    }
}

/* chem3d for one box: F:800-833 (densities), F:857-912 (J, aerosol), F:920-1306
 * (rates), F:1315-1338 (rain-out), F:1348-2378 (sweeps), F:2380-2410 (write-back).
 * q[1..43] and h[1..15] are mixing ratios, updated in place. */
static void box_step(const troe_t *K, const double *ajl, const double *o3up, const boxenv_t *e,
                     double dt, int niter, double *q, double *h, double *surf, double *vsed,
                     double *p2osed, double *pno3sed, int *surf_written, double tol, int *iters,
                     const double *gq, const double *gh, int accel, int aa_start,
                     double accept_tol)
{
    double ck0[NBCON + 1], ck[NBCON + 1], cc[NCM + 1], tj[NPHOT + 1], pr[NBCON + 1],
        pe[NBCON + 1];
    double het1, het2, het3, het4, het5, het6, het7, het8, het9;
    double parthno3 = 1., parthno3sl = 1., parth2o = 1., parthcl = 1., parthbr = 1.;
    double rain_hno3, rain_hno4, rain_hcl, rain_hbr, rain_h2o2, rain_ch2o, rain_ch3o2h;
    double day = 0., ch3o = 0., hco = 0.;
    rates_t R;
    const double t = e->t, pm = e->pm, sza = e->sza;
    const double hnm = pm / (t * 1.38e-19);
    const double n2 = 0.78 * hnm, o2 = 0.21 * hnm, co2 = 360.e-6 * hnm, h2 = 5.e-7 * hnm;
    int n;

    for (n = 1; n <= NBCON; ++n) {
        ck0[n] = q[n] * hnm;
        ck[n] = ck0[n];
    }
    for (n = 1; n <= NCM; ++n) cc[n] = h[n] * hnm;

    /* J-values: phot + jno */
    rbo_phot_box(ajl, o3up, e->alt, sza, e->o3t, &tj[1]);
    tj[J_NO] = rbo_jno_box(e->colo2, e->colno, e->o3t, hnm, e->sza_top);

    /* heterogeneous chemistry (chem3d init F:863-880, analytic call F:898) */
    het1 = het2 = het3 = het4 = het5 = het6 = het7 = het8 = het9 = 0.;
    *vsed = 0.;
    *surf_written = 0;
    if (e->het_level) {
        aero_t a;
        aerosol_box(t, pm, hnm, ck, &a);
        het1 = a.het[1]; het2 = a.het[2]; het3 = a.het[3]; het4 = a.het[4]; het5 = a.het[5];
        het6 = a.het[6]; het7 = a.het[7]; het8 = a.het[8]; het9 = a.het[9];
        parthno3 = a.parthno3;
        parthno3sl = a.parthno3sl;
        parth2o = a.parth2o;
        parthcl = a.parthcl;
        *surf = a.surf;
        *vsed = a.vsed;
        *surf_written = 1;
    }
    *p2osed = parth2o;
    *pno3sed = parthno3sl;

    rates_box(K, t, hnm, pm, &R);

    /* rain-out F:1315-1338 */
    {
        const double rain_h2o = 100.e-6, rain_rate = 1.e-6;
        if (ck[3] / hnm >= rain_h2o) {
            rain_hno3 = rain_rate; rain_hno4 = 0.; rain_hcl = rain_rate; rain_hbr = rain_rate;
            rain_h2o2 = rain_rate; rain_ch2o = 0.; rain_ch3o2h = 0.;
        } else {
            rain_hno3 = rain_hno4 = rain_hcl = rain_hbr = rain_h2o2 = rain_ch2o = rain_ch3o2h = 0.;
        }
    }

    /* converged mode with an initial guess: the iterate (not ck0) starts from it */
    if (gq) {
        for (n = 1; n <= NBCON - 1; ++n) ck[n] = gq[n] * hnm;
        for (n = 1; n <= NCM; ++n) cc[n] = gh[n] * hnm;
    }

    aa_t aa;
    aa.m = accel < AA_MAX ? accel : AA_MAX;
    aa.n = 0;
    aa.head = 0;
    *iters = niter;
    /* converged mode: largest relative change of the previous sweep, and whether that
     * sweep was a plain one (its iterate is the sweep map's image of the one before) */
    double dprev = 0., dprev2 = 0.;
    int nplain = 0, crun = 0, rose = 0;
    for (int iter = 1; iter <= niter; ++iter) {
        double prev[NBCON + NCM + 1];
        if (tol > 0.) {
            for (n = 1; n <= NBCON - 1; ++n) prev[n] = ck[n];
            for (n = 1; n <= NCM; ++n) prev[NBCON - 1 + n] = cc[n];
        }
        double r1, r2, r3, r4, r5, r6, r7, r8, r9, r10, r11, r12, r13, r27, r28, r29, r30;
        double a, b, coea, coeb, coec, argum;
        double t01, t02, t03, t04, t05, t06, t07, t08, t09, t10, t11, t12, t13, t14, t15;
        if (sza <= 95.) {
            /* day, F:1353-1669 */
            day = 1.;
            cc[15] = ((R.c2 * cc[2] + R.c1a * cc[1] + R.d5 * cc[3]) * ck[2]) /
                     (R.c4 * o2 + R.c3 * cc[4]);
            ch3o = (R.c1b * cc[1] * ck[2] + (R.c5 * cc[8] + 2. * R.c14 * ck[29] + R.c18 * cc[11]) * ck[29] +
                    tj[J_CH3O2H] * ck[30]) /
                   (R.c15 * o2);
            hco = (R.c8 * cc[2] + R.c9 * cc[4] + R.c10 * cc[3] + R.c19 * cc[9] + tj[J_CH2O_HCO]) * ck[28] /
                  (R.c12 * o2);
            r1 = (tj[J_O3_O1D] + tj[J_O3_O3P] + R.a2 * cc[14] + R.a6 * cc[2] + R.a6b * cc[6] +
                  R.b4 * cc[8] + R.d2 * cc[3]) /
                 (R.hk2 * o2);
            r2 = 1. / (1. + r1);
            cc[5] = r2 * ck[8];
            cc[4] = r1 * cc[5];
            r3 = tj[J_O3_O1D] / (R.hk4 * n2 + R.hk5 * o2);
            cc[1] = r3 * cc[5];
            r4 = (R.a5 * cc[4] + R.a36 * ck[9] + R.a19 * h2) /
                 (R.a1 * o2 + R.a2 * cc[5] + (A23A + A23B + A23C) * cc[6]);
            cc[14] = r4 * cc[2];
            r7 = tj[J_NO] / (R.b7 * o2 + R.b6 * cc[8]);
            r8 = (R.b9 * cc[5]) / (tj[J_NO3_NO] + tj[J_NO3_NO2] + R.b12 * cc[7]);
            r5 = (tj[J_NO2] + R.b3 * cc[4] + tj[J_NO3_NO] * r8) /
                 (R.b4 * cc[5] + R.a26 * cc[6] + R.d4 * cc[11] + R.e4 * cc[12] + R.c5 * ck[29]);
            cc[7] = ck[21] / (1. + r5 + r7 * r5 + r8);
            cc[8] = r5 * cc[7];
            cc[10] = r7 * cc[8];
            cc[13] = r8 * cc[7];
            r6 = (R.a6 * cc[5] + R.a30 * ck[17] + R.c8 * ck[28] + R.d8 * cc[11] + R.e10 * cc[12] +
                  R.a1 * o2 * r4) /
                 ((A23A + A23B + A23C) * cc[14] + R.a6b * cc[5] + R.a7 * cc[4] + R.a26 * cc[8] +
                  2. * R.a27 * cc[6] + R.a17 * cc[2] + R.d7 * cc[3] + R.d10 * cc[3] +
                  R.d33 * cc[11] + R.e7 * cc[9] + R.e15 * cc[12]);
            a = .5;
            b = 1. - a;
            cc[6] = a * cc[6] + b * cc[2] * r6;
            r13 = R.d31 * cc[7] /
                  (tj[J_CLONO2_CL] + tj[J_CLONO2_CLO] + R.d32 * cc[4] + het1 * ck[3] + het2 * ck[13]);
            r9 = (R.d33 * cc[6] + het1 * ck[3] * r13) / (het5 * ck[13] + R.d35 * cc[4] + tj[J_HOCL]);
            r10 = (R.d60 * cc[11]) / (R.d61 * hnm + tj[J_CL2O2]);
            r12 = R.e5c * cc[12] / tj[J_BRCL];
            r11 = (R.d3 * cc[4] + R.d4 * cc[8] + R.d8 * cc[2] + R.c18 * ck[29] + R.e5b * cc[12] +
                   tj[J_CLONO2_CL] * r13 + tj[J_HOCL] * r9 + 2. * tj[J_CL2O2] * r10 +
                   tj[J_BRCL] * r12) /
                  (R.d2 * cc[5] + R.d5 * ck[2] + R.d7 * cc[6] + R.d10 * cc[6] + R.d37 * ck[15]);
            a = .5;
            b = 1. - a;
            cc[11] = a * cc[11] + b * ck[23] / (1. + r11);
            cc[3] = r11 * cc[11];
            r27 = R.e13 * cc[7] / (tj[J_BRONO2] + R.e8 * cc[4]);
            r28 = R.e15 * cc[6] / (tj[J_HOBR] + R.e16 * cc[4]);
            r29 = R.e5c * cc[11] / tj[J_BRCL];
            r30 = (R.e3 * cc[4] + R.e4 * cc[8] + R.e5a * cc[11] + R.e5b * cc[11] + 2. * R.e6 * cc[12] +
                   R.e10 * cc[2] + tj[J_BRO] + tj[J_BRONO2] * 0.85 * r27 + tj[J_HOBR] * r28 +
                   tj[J_BRCL] * r29) /
                  (R.e2 * cc[5] + R.e7 * cc[6] + R.c19 * ck[28]);
            a = .5;
            b = 1. - a;
            cc[12] = a * cc[12] + b * ck[24] / (1. + r30);
            cc[9] = r30 * cc[12];
            coea = 2. * R.a31 + 2. * (R.a17 + R.a27 * r6 + (A23B + A23C) * r4) * r6;
            coeb = R.b27 * ck[5] * parthno3 + R.b22 * cc[7] + R.b28 * ck[22] + R.b23 * cc[7] * r6 +
                   R.c2 * ck[2] + R.c7 * ck[29] * r6 + R.c8 * ck[28] + R.c17a * ck[30] +
                   R.d7 * cc[3] * r6 + R.d9 * cc[11] + R.d11 * ck[13] * parthcl + R.d33 * cc[11] * r6 +
                   R.d34 * ck[15] + R.e11 * ck[19] * parthbr + R.e7 * cc[9] * r6 + R.e15 * cc[12] * r6;
            t01 = 2. * (R.a1et * cc[1] + tj[J_H2O]) * ck[3] * parth2o;
            t02 = ((R.c1a + R.c1b) * ck[2] + 2. * A3ET * h2) * cc[1];
            t03 = R.c3 * cc[15] * cc[4];
            t04 = tj[J_CH3O2H] * ck[30];
            t05 = R.c15 * ch3o * o2;
            t06 = (tj[J_CH2O_HCO] + R.c9 * cc[4]) * ck[28];
            t07 = R.c12 * hco * o2;
            t08 = tj[J_HNO3] * ck[5] * parthno3;
            t09 = (R.b24 * hnm + tj[J_HNO4]) * ck[22];
            t10 = (R.d35 * cc[4] + R.d37 * cc[3] + tj[J_HOCL]) * ck[15];
            t11 = 2. * tj[J_H2O2] * ck[17];
            t12 = R.d6 * h2 * cc[3];
            t13 = tj[J_HOBR] * ck[26];
            t14 = R.e16 * cc[4] * ck[26];
            t15 = tj[J_HBR] * ck[19];
            coec = t01 + t02 + t03 + t04 + t05 + t06 + t07 + t08 + t09 + t10 + t11 + t12 + t13 + t14 + t15;
            argum = coeb * coeb + 4. * coea * coec;
            argum = fmax(argum, 0.);
            a = .5;
            b = 1. - a;
            cc[2] = a * cc[2] + b * (sqrt(argum) - coeb) / (2. * coea);
        } else {
            /* night, F:1671-1746 */
            day = 0.;
            cc[14] = 0.;
            cc[15] = 0.;
            ch3o = 0.;
            hco = 0.;
            cc[1] = 0.;
            cc[4] = 0.;
            cc[5] = ck[8];
            cc[2] = 0.;
            cc[3] = 0.;
            cc[6] = 0.;
            cc[8] = 0.;
            r8 = R.b9 * cc[5] / (R.b12 * fmax(cc[7], 1.e-30));
            cc[7] = ck[21] / (1. + r8);
            cc[13] = r8 * cc[7];
            cc[10] = 0.;
            cc[11] = fmax(ck[23], 0.);
            cc[9] = 0.;
            cc[12] = fmax(ck[24], 0.);
        }

        /* production / loss, F:1758-2358 */
        pr[1] = 0.e0;
        pe[1] = tj[J_N2O] + (R.b38 + R.b39) * cc[1];
        pr[2] = 0.e0;
        pe[2] = R.c2 * cc[2] + R.c1a * cc[1] + R.c1b * cc[1] + R.c1c * cc[1] + R.d5 * cc[3];
        pr[3] = R.c2 * cc[2] * ck[2] + R.c8 * cc[2] * ck[28] + R.c17a * ck[30] * cc[2] +
                R.c17b * ck[30] * cc[2] + (R.a17 * cc[2] + A23C * cc[14]) * cc[6] +
                R.a19 * cc[2] * h2 + R.a30 * cc[2] * ck[17] + R.a31 * cc[2] * cc[2] +
                R.b27 * cc[2] * ck[5] * parthno3 + R.d11 * cc[2] * ck[13] * parthcl;
        pe[3] = (tj[J_H2O] + R.a1et * cc[1]) * parth2o;
        pr[4] = 2. * R.b39 * cc[1] * ck[1];
        pe[4] = 2. * R.b6 * cc[10] * cc[8] + rain_hno3 * ck[5] + rain_hno4 * ck[22];
        pe[4] = pe[4] / ck[4];
        pr[5] = R.b22 * cc[2] * cc[7] + het1 * ck[14] * ck[3] + het2 * ck[14] * ck[13] +
                2. * het3 * ck[6] * ck[3] + het4 * ck[6] * ck[13] + het6 * ck[20] * ck[3];
        pe[5] = (tj[J_HNO3] + R.b27 * cc[2]) * parthno3 + rain_hno3;
        pr[6] = R.b12 * cc[13] * cc[7];
        pe[6] = tj[J_N2O5] + R.b32 * hnm + het3 * ck[3] + het4 * ck[13];
        pr[7] = 3. * (tj[J_CFC11] + R.d90 * cc[1]) * ck[31] + 2. * (tj[J_CFC12] + R.d91 * cc[1]) * ck[32] +
                3. * (tj[J_CFC113] + R.d92 * cc[1]) * ck[33] + 4. * (tj[J_CCL4] + R.d93 * cc[1]) * ck[34] +
                3. * (tj[J_CH3CCL3] + R.d94 * cc[2]) * ck[35] + 1. * (tj[J_CH3CL] + R.d95 * cc[2]) * ck[36] +
                1. * (tj[J_HCFC22] + R.d96 * cc[1] + R.d97 * cc[2]) * ck[37] +
                1. * (tj[J_H1211] + R.e52 * cc[1]) * ck[39];
        pe[7] = rain_hcl * ck[13];
        pe[7] = pe[7] / ck[7];
        {
            /* Ox, F:1864-1950 */
            double ratio = tj[J_CL2O2] / (tj[J_CL2O2] + R.d61 * hnm);
            pr[8] = 2. * tj[J_O2] * o2 + R.a26 * cc[6] * cc[8] + R.c5 * cc[8] * ck[29];
            t01 = 2. * R.hk1 * cc[4] * cc[4];
            t02 = 2. * R.hk3 * cc[5] * cc[4];
            t03 = 2. * R.b3 * cc[7] * cc[4];
            t04 = 2. * R.d3 * cc[11] * cc[4];
            t05 = 2. * R.d60 * ratio * cc[11] * cc[11];
            t06 = fmin(tj[J_HOCL] * ck[15], R.d33 * cc[6] * cc[11]);
            t07 = 2. * fmin(R.e5c * cc[12] * cc[11], tj[J_BRCL] * ck[27]);
            t08 = 2. * R.e5b * cc[12] * cc[11] * day;
            t09 = fmin(R.e15 * cc[12] * cc[6], tj[J_HOBR] * ck[26]);
            t10 = 2. * R.e6 * cc[12] * cc[12] * day;
            t11 = 2. * (tj[J_NO3_NO] / fmax(tj[J_NO3_NO] + tj[J_NO3_NO2], 1.e-30)) * tj[J_CLONO2_CL] * ck[14];
            t12 = (R.a2 * cc[14] + R.a6 * cc[2] + R.a6b * cc[6]) * cc[5] + (R.a5 * cc[2] + R.a7 * cc[6]) * cc[4];
            t13 = 2. * R.b9 * cc[7] * cc[5] * tj[J_NO3_NO] / fmax(tj[J_NO3_NO] + tj[J_NO3_NO2], 1.e-30);
            pe[8] = t01 + t02 + t03 + t04 + t05 + t06 + t07 + t08 + t09 + t10 + t11 + t12 + t13;
            pe[8] = pe[8] / ck[8];
        }
        pr[9] = tj[J_CO2] * co2 + tj[J_CH2O_CO] * ck[28] + R.c12 * hco * o2;
        pe[9] = R.a36 * cc[2];
        pr[10] = R.e5a * cc[12] * cc[11];
        pe[10] = tj[J_OCLO];
        pr[11] = 0.;
        pe[11] = 0.;
        pr[12] = 0.;
        pe[12] = 0.;
        pr[13] = (R.c10 * ck[28] + R.d5 * ck[2] + R.d7 * cc[6] + R.d6 * h2) * cc[3] + R.d9 * cc[2] * cc[11];
        pe[13] = (tj[J_HCL] + R.d11 * cc[2]) * parthcl + het2 * ck[14] + het4 * ck[6] + het5 * ck[15] +
                 het7 * ck[26] + rain_hcl;
        pr[14] = R.d31 * cc[7] * cc[11];
        pe[14] = tj[J_CLONO2_CL] + tj[J_CLONO2_CLO] + R.d32 * cc[4] + het1 * ck[3] + het2 * ck[13];
        pr[15] = R.d33 * cc[6] * cc[11] + het1 * ck[14] * ck[3];
        pe[15] = tj[J_HOCL] + R.d34 * cc[2] + R.d35 * cc[4] + R.d37 * cc[3] + het5 * ck[13] + het9 * ck[19];
        pr[16] = R.d37 * ck[15] * cc[3] + R.d62 * cc[11] * cc[11] + het2 * ck[14] * ck[13] +
                 het5 * ck[15] * ck[13];
        pe[16] = tj[J_CL2];
        pr[17] = R.a27 * cc[6] * cc[6];
        pe[17] = tj[J_H2O2] + R.a30 * cc[2] + rain_h2o2;
        pr[18] = R.d36 * cc[3] * cc[7] + het4 * ck[6] * ck[13];
        pe[18] = tj[J_CLNO2];
        pr[19] = R.e7 * cc[6] * cc[9] + R.c19 * ck[28] * cc[9];
        pe[19] = (R.e11 * cc[2] + R.e9 * cc[1] + tj[J_HBR]) * parthbr + het8 * ck[26] + het9 * ck[15] +
                 rain_hbr;
        pr[20] = R.e13 * cc[12] * cc[7];
        pe[20] = tj[J_BRONO2] + R.e8 * cc[4] + het6 * ck[3];
        pr[21] = (tj[J_HNO4] + R.b28 * cc[2] + R.b24 * hnm) * ck[22] + tj[J_HNO3] * ck[5] * parthno3 +
                 2. * (tj[J_N2O5] + R.b32 * hnm) * ck[6] + 2. * R.b39 * cc[1] * ck[1] +
                 tj[J_CLNO2] * ck[18] + R.b27 * ck[5] * parthno3 * cc[2] + R.d32 * ck[14] * cc[4] +
                 (tj[J_CLONO2_CL] + tj[J_CLONO2_CLO]) * ck[14] + tj[J_BRONO2] * ck[20] +
                 R.e8 * ck[20] * cc[4];
        pe[21] = (R.d31 * cc[11] + R.b22 * cc[2] + R.b23 * cc[6] + 2. * R.b12 * cc[13] + R.d36 * cc[3] +
                  R.e13 * cc[12]) * cc[7] +
                 2. * R.b6 * cc[10] * cc[8];
        pe[21] = pe[21] / ck[21];
        pr[22] = R.b23 * cc[6] * cc[7];
        pe[22] = tj[J_HNO4] + R.b28 * cc[2] + R.b24 * hnm + rain_hno4;
        {
            /* ClOx, F:2115-2160 */
            double u01, u02, u03, u04, u05, u06, u07, u08, u09, u10, u11, u12, u13, u14;
            u01 = 2. * (R.d61 * hnm + tj[J_CL2O2]) * ck[25];
            u02 = (tj[J_CLONO2_CL] + tj[J_CLONO2_CLO] + R.d32 * cc[4]) * ck[14];
            u03 = (tj[J_HOCL] + R.d34 * cc[2] + R.d35 * cc[4]) * ck[15];
            u04 = tj[J_OCLO] * ck[10] + (tj[J_HCL] + R.d11 * cc[2]) * ck[13] * parthcl;
            u05 = 2. * tj[J_CL2] * ck[16];
            u06 = tj[J_CLNO2] * ck[18];
            u07 = tj[J_BRCL] * ck[27];
            u08 = 3. * (tj[J_CFC11] + R.d90 * cc[1]) * ck[31];
            u09 = 2. * (tj[J_CFC12] + R.d91 * cc[1]) * ck[32];
            u10 = 3. * (tj[J_CFC113] + R.d92 * cc[1]) * ck[33];
            u11 = 4. * (tj[J_CCL4] + R.d93 * cc[1]) * ck[34];
            u12 = 3. * (tj[J_CH3CCL3] + R.d94 * cc[2]) * ck[35];
            u13 = 1. * (tj[J_CH3CL] + R.d95 * cc[2]) * ck[36];
            u14 = 1. * (tj[J_HCFC22] + R.d96 * cc[1] + R.d97 * cc[2]) * ck[37];
            pr[23] = u01 + u02 + u03 + u04 + u05 + u06 + u07 + u08 + u09 + u10 + u11 + u12 + u13 + u14;
        }
        pe[23] = (R.d31 * cc[7] + 2. * R.d60 * cc[11] + 2. * R.d62 * cc[11] + R.e5a * cc[12] +
                  R.e5c * cc[12] + R.d9 * cc[2] + R.d33 * cc[6]) * cc[11] +
                 (R.c10 * ck[28] + R.d5 * ck[2] + R.d7 * cc[6] + R.d6 * h2 + R.d36 * cc[7] +
                  R.d37 * ck[15]) * cc[3];
        pe[23] = pe[23] / ck[23];
        pr[24] = tj[J_HOBR] * ck[26] + tj[J_BRONO2] * ck[20] + R.e8 * ck[20] * cc[4] +
                 R.e16 * ck[26] * cc[4] + tj[J_BRCL] * ck[27] +
                 (R.e11 * cc[2] + R.e9 * cc[1] + tj[J_HBR]) * ck[19] * parthbr +
                 2. * het8 * ck[19] * ck[26];
        pe[24] = (R.e13 * cc[7] + R.e5c * cc[11] + R.e15 * cc[6]) * cc[12] + R.c19 * cc[9] * ck[28] +
                 R.e7 * cc[9] * cc[6];
        pe[24] = pe[24] / ck[24];
        pr[25] = R.d60 * cc[11] * cc[11];
        pe[25] = tj[J_CL2O2] + R.d61 * hnm;
        pr[26] = R.e15 * cc[12] * cc[6] + het6 * ck[20] * ck[3];
        pe[26] = tj[J_HOBR] + R.e16 * cc[4] + het7 * ck[13] + het8 * ck[19];
        pr[27] = R.e5c * cc[11] * cc[12] + het7 * ck[26] * ck[13] + het9 * ck[19] * ck[15];
        pe[27] = tj[J_BRCL];
        pr[28] = (R.c1c * cc[1] * ck[2] + R.c3 * cc[4] * cc[15] + R.c15 * ch3o * o2 + R.c17b * cc[2] * ck[30]);
        pe[28] = R.c8 * cc[2] + R.c9 * cc[4] + R.c10 * cc[3] + R.c19 * cc[9] + tj[J_CH2O_HCO] +
                 tj[J_CH2O_CO] + rain_ch2o;
        pr[29] = R.c4 * cc[15] * o2 + R.c17a * cc[2] * ck[30];
        pe[29] = 2. * R.c14 * ck[29] + R.c7 * cc[6] + R.c5 * cc[8] + R.c18 * cc[11];
        pr[30] = R.c7 * ck[29] * cc[6];
        pe[30] = R.c17a * cc[2] + R.c17b * cc[2] + tj[J_CH3O2H] + rain_ch3o2h;
        pr[31] = 0.;
        pe[31] = tj[J_CFC11] + R.d90 * cc[1];
        pr[32] = 0.;
        pe[32] = tj[J_CFC12] + R.d91 * cc[1];
        pr[33] = 0.;
        pe[33] = tj[J_CFC113] + R.d92 * cc[1];
        pr[34] = 0.;
        pe[34] = tj[J_CCL4] + R.d93 * cc[1];
        pr[35] = 0.;
        pe[35] = tj[J_CH3CCL3] + R.d94 * cc[2];
        pr[36] = 0.;
        pe[36] = tj[J_CH3CL] + R.d95 * cc[2];
        pr[37] = 0.;
        pe[37] = tj[J_HCFC22] + R.d96 * cc[1] + R.d97 * cc[2];
        pr[38] = 0.;
        pe[38] = tj[J_CH3BR] + R.e50 * cc[1] + R.e51 * cc[2];
        pr[39] = 0.;
        pe[39] = tj[J_H1211] + R.e52 * cc[1];
        pr[40] = 0.;
        pe[40] = tj[J_H1301] + R.e53 * cc[1];
        pr[41] = (tj[J_CH3BR] + R.e50 * cc[1] + R.e51 * cc[2]) * ck[38] + (tj[J_H1211] + R.e52 * cc[1]) * ck[39] +
                 (tj[J_H1301] + R.e53 * cc[1]) * ck[40] + 2. * (R.e54 * cc[2] + tj[J_CH2BR2]) * ck[42];
        pe[41] = rain_hbr * ck[19];
        pe[41] = pe[41] / ck[41];
        pr[42] = 0.;
        pe[42] = tj[J_CH2BR2] + R.e54 * cc[2];
        pr[43] = 0.;
        pe[43] = 0.;

        /* F:2364-2372 */
        for (n = 1; n <= NBCON - 1; ++n) ck[n] = (ck0[n] + dt * pr[n]) / (1. + dt * pe[n]);

        /* converged mode: stop near the fixed point of the sweep map.  Every species
         * must have changed by less than tol relative to its new value (values under the
         * 1e-30 mole-fraction floor compared at the floor): |v - old| < tol max(|v|, fl);
         * and, since the sweeps converge linearly (after a sweep whose largest relative
         * change is d the distance to the fixed point is about d rho / (1 - rho), rho the
         * contraction per sweep), the estimate of that distance must be below tol too.
         * rho is estimated as the larger of the last two ratios of successive changes
         * (d_k / d_(k-1), d_(k-1) / d_(k-2)), so the estimate needs RBO_CONV_NPLAIN = 3
         * successive plain sweeps: the larger ratio keeps an oscillating sequence (the
         * damped radicals) from showing a small rate at a low point.  With Anderson mixing,
         * a box whose change is below tol takes unmixed steps until the estimate decides. */
        if (tol > 0.) {
            const double fl = 1.e-30 * hnm;
            /* surrogate step: the first sweep from the surrogate's answer measures its
             * residual as a fixed point of the sweep map; under accept_tol the answer
             * itself is kept (count 0), otherwise the sweeps go on at tol */
            const int accept = iter == 1 && accept_tol > 0. && gq;
            const double tt = accept ? accept_tol : tol;
            int conv = 1, fin = 1;
            double dk = 0.;
            for (n = 1; n <= NBCON - 1; ++n) {
                const double sc = fmax(fabs(ck[n]), fl), dv = fabs(ck[n] - prev[n]);
                conv = conv && (dv < tt * sc);
                fin = fin && (fabs(ck[n]) <= 1.7976931348623157e308);
                dk = fmax(dk, dv / sc);
            }
            for (n = 1; n <= NCM; ++n) {
                const double sc = fmax(fabs(cc[n]), fl), dv = fabs(cc[n] - prev[NBCON - 1 + n]);
                conv = conv && (dv < tt * sc);
                fin = fin && (fabs(cc[n]) <= 1.7976931348623157e308);
                dk = fmax(dk, dv / sc);
            }
            if (conv && fin && accept) {
                for (n = 1; n <= NBCON - 1; ++n) ck[n] = gq[n] * hnm;
                for (n = 1; n <= NCM; ++n) cc[n] = gh[n] * hnm;
                *iters = 0;
                break;
            }
            int near = 0;
            if (nplain >= RBO_CONV_NPLAIN) {
                const double rho = fmin(fmax(dk / dprev, dprev / dprev2), RBO_CONV_RHOMAX);
                near = dk == 0. || dk * rho < tol * (1. - rho);
            }
            /* the run of successive plain sweeps below tol, and whether the change rose in it */
            if (conv && nplain >= 1) {
                rose = rose || (crun > 0 && dk >= dprev);
                ++crun;
            } else {
                crun = 0;
                rose = 0;
            }
            near = near || (crun >= RBO_CONV_STALL && rose);
            if (conv && fin && near) {
                *iters = iter;
                break;
            }
            if (!fin || iter == niter) {
                *iters = -iter;       /* diverged or out of sweeps */
                break;
            }
            const int mix = accel > 0 && iter >= aa_start && !conv;
            if (mix) {
                double xv[NX];
                for (n = 1; n <= NBCON - 1; ++n) xv[n - 1] = ck[n];
                for (n = 1; n <= NCM; ++n) xv[NBCON - 2 + n] = cc[n];
                aa_update(&aa, xv, prev + 1, fl);
                for (n = 1; n <= NBCON - 1; ++n) ck[n] = xv[n - 1];
                for (n = 1; n <= NCM; ++n) cc[n] = xv[NBCON - 2 + n];
            }
            nplain = mix ? 0 : nplain + 1;
            dprev2 = dprev;
            dprev = dk;
        }
    }

    /* write-back F:2380-2410 */
    for (n = 1; n <= NBCON - 1; ++n) q[n] = ck[n] / hnm;
    q[NBCON] = ck[5] * parthno3 / hnm;
    for (n = 1; n <= NCM; ++n) h[n] = cc[n] / hnm;
}

/* ---------------------------------------------------------------- grid driver */
static int chem_grid(const rbo_grid *g, double dt, const double *T, const double *p,
                     const double *sza, const double *alt, const double *ajl, const double *o3up,
                     double *qj1, double *hc, double *surfarea, double *vsed3d,
                     double *parth2osed, double *parthno3sed, int nthreads, double tol,
                     int *iters, const double *guess_qj1, const double *guess_hc, int accel,
                     int aa_start, double accept_tol);

int rbo_chem_step(const rbo_grid *g, double dt, const double *T, const double *p,
                  const double *sza, const double *alt, const double *ajl, const double *o3up,
                  double *qj1, double *hc, double *surfarea, double *vsed3d,
                  double *parth2osed, double *parthno3sed, int nthreads)
{
    return chem_grid(g, dt, T, p, sza, alt, ajl, o3up, qj1, hc, surfarea, vsed3d, parth2osed,
                     parthno3sed, nthreads, 0., NULL, NULL, NULL, 0, 0, 0.);
}

int rbo_chem_converge(const rbo_grid *g, double dt, double tol, const double *T, const double *p,
                      const double *sza, const double *alt, const double *ajl, const double *o3up,
                      double *qj1, double *hc, double *surfarea, double *vsed3d,
                      double *parth2osed, double *parthno3sed, int *iters,
                      const double *guess_qj1, const double *guess_hc, int nthreads)
{
    return rbo_chem_converge_aa(g, dt, tol, 0, 0, 0, T, p, sza, alt, ajl, o3up, qj1, hc,
                                surfarea, vsed3d, parth2osed, parthno3sed, iters, guess_qj1,
                                guess_hc, nthreads);
}

/* accelerated solves: the sweep cap is min(maxiter, aa_start + aa_budget) */
static rbo_grid capped(const rbo_grid *g, int accel, int aa_start, int aa_budget)
{
    rbo_grid c = *g;
    if (accel > 0 && aa_budget > 0 && (long)aa_start + aa_budget < c.niter)
        c.niter = aa_start + aa_budget;
    return c;
}

int rbo_chem_converge_aa(const rbo_grid *g, double dt, double tol, int accel, int aa_start,
                         int aa_budget, const double *T, const double *p, const double *sza,
                         const double *alt, const double *ajl, const double *o3up, double *qj1,
                         double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                         double *parthno3sed, int *iters, const double *guess_qj1,
                         const double *guess_hc, int nthreads)
{
    if ((guess_qj1 == NULL) != (guess_hc == NULL) || accel < 0 || accel > AA_MAX) return -1;
    if (accel > 0 && guess_qj1) return -2;      /* refused, as the library refuses it */
    const rbo_grid c = capped(g, accel, aa_start, aa_budget);
    return chem_grid(&c, dt, T, p, sza, alt, ajl, o3up, qj1, hc, surfarea, vsed3d, parth2osed,
                     parthno3sed, nthreads, tol, iters, guess_qj1, guess_hc, accel, aa_start, 0.);
}

int rbo_chem_surrogate(const rbo_grid *g, double dt, double accept_tol, double tol, int accel,
                       int aa_start, int aa_budget, const double *T, const double *p,
                       const double *sza,
                       const double *alt, const double *ajl, const double *o3up, double *qj1,
                       double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                       double *parthno3sed, int *iters, const double *sur_qj1,
                       const double *sur_hc, int nthreads)
{
    if (!sur_qj1 || !sur_hc || !(accept_tol > 0.) || accel < 0 || accel > AA_MAX) return -1;
    if (accel > 0) return -2;                   /* refused, as the library refuses it */
    const rbo_grid c = capped(g, accel, aa_start, aa_budget);
    return chem_grid(&c, dt, T, p, sza, alt, ajl, o3up, qj1, hc, surfarea, vsed3d, parth2osed,
                     parthno3sed, nthreads, tol, iters, sur_qj1, sur_hc, accel, aa_start,
                     accept_tol);
}

static int chem_grid(const rbo_grid *g, double dt, const double *T, const double *p,
                     const double *sza, const double *alt, const double *ajl, const double *o3up,
                     double *qj1, double *hc, double *surfarea, double *vsed3d,
                     double *parth2osed, double *parthno3sed, int nthreads, double tol,
                     int *iters, const double *guess_qj1, const double *guess_hc, int accel,
                     int aa_start, double accept_tol)
{
    const size_t plane = (size_t)g->nlon * g->nlat;          /* one level of a field */
    const size_t field = plane * g->niv;                       /* one species field */
    const size_t fieldb = plane * g->nivbas;                   /* nivbas-level field */
    const int nlev = g->nivbas, ncol = g->nlon * g->nlat;
    const troe_t K = troe_constants();
    int bad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : bad)
#endif
    for (int c = 0; c < ncol; ++c) {
        double o3t[256], colo2[256], colno[256];
        double q[NBCON + 1], h[NCM + 1], gq[NBCON + 1], gh[NCM + 1];
        if (nlev > 256) { bad++; continue; }
        rbo_column(nlev, p + c, T + c, hc + 4 * field + c, hc + 7 * field + c, (int)plane, o3t,
                   colo2, colno);
        for (int k = 0; k < nlev; ++k) {
            size_t b = (size_t)k * plane + c;
            boxenv_t e;
            double surf = 0., vsed, p2o, pno3;
            int sw, it;
            e.t = T[b];
            e.pm = p[b];
            e.sza = sza[b];
            e.sza_top = sza[c];
            e.alt = alt[b];
            e.o3t = o3t[k];
            e.colo2 = colo2[k];
            e.colno = colno[k];
            e.het_level = (k + 1 >= g->nivhau && k + 1 <= g->nivbas);
            for (int n = 1; n <= NBCON; ++n) q[n] = qj1[(size_t)(n - 1) * field + b];
            for (int n = 1; n <= NCM; ++n) h[n] = hc[(size_t)(n - 1) * field + b];
            if (guess_qj1) {
                for (int n = 1; n <= NBCON; ++n) gq[n] = guess_qj1[(size_t)(n - 1) * field + b];
                for (int n = 1; n <= NCM; ++n) gh[n] = guess_hc[(size_t)(n - 1) * field + b];
            }
            box_step(&K, ajl, o3up, &e, dt, g->niter, q, h, &surf, &vsed, &p2o, &pno3, &sw, tol,
                     &it, guess_qj1 ? gq : NULL, guess_qj1 ? gh : NULL, accel, aa_start,
                     accept_tol);
            if (it < 0) {
                /* converged mode without convergence: the reference answer, RBO_NITER_REF
                 * sweeps from the incoming state */
                for (int n = 1; n <= NBCON; ++n) q[n] = qj1[(size_t)(n - 1) * field + b];
                for (int n = 1; n <= NCM; ++n) h[n] = hc[(size_t)(n - 1) * field + b];
                int it5;
                box_step(&K, ajl, o3up, &e, dt, RBO_NITER_REF, q, h, &surf, &vsed, &p2o, &pno3,
                         &sw, 0., &it5, NULL, NULL, 0, 0, 0.);
            }
            if (iters) iters[b] = it;
            for (int n = 1; n <= NBCON; ++n) qj1[(size_t)(n - 1) * field + b] = q[n];
            for (int n = 1; n <= NCM; ++n) hc[(size_t)(n - 1) * field + b] = h[n];
            if (sw) surfarea[b] = surf;
            {
                size_t bb = (size_t)k * plane + c; /* nivbas-level fields */
                vsed3d[bb] = vsed;
                parth2osed[bb] = p2o;
                parthno3sed[bb] = pno3;
            }
        }
    }
    (void)fieldb;
    return bad ? -1 : 0;
}

/* conserv F:2526-2616 then the floor F:431-440. */
void rbo_conserv_floor(const rbo_grid *g, double *qj1, double *hc, int nthreads)
{
    const size_t plane = (size_t)g->nlon * g->nlat, field = plane * g->niv;
    const size_t nb = plane * g->nivbas, na = field;
#define Q(n) qj1[(size_t)((n) - 1) * field + b]
#define H(n) hc[(size_t)((n) - 1) * field + b]
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (size_t b = 0; b < nb; ++b) {
        double totmas, rapcly, rapnoy, rapbry;
        totmas = Q(10) + Q(13) + Q(14) + Q(15) + 2. * Q(16) + Q(18) + Q(23) + 2. * Q(25);
        rapcly = (Q(7) - Q(27)) / totmas;
        H(3) = H(3) * rapcly;
        H(11) = H(11) * rapcly;
        Q(10) = Q(10) * rapcly;
        Q(13) = Q(13) * rapcly;
        Q(14) = Q(14) * rapcly;
        Q(15) = Q(15) * rapcly;
        Q(16) = Q(16) * rapcly;
        Q(18) = Q(18) * rapcly;
        Q(23) = Q(23) * rapcly;
        Q(25) = Q(25) * rapcly;
        totmas = Q(5) + 2. * Q(6) + Q(21) + Q(22);
        rapnoy = fmax((Q(4) - Q(14) - Q(18) - Q(20)), 0.) / totmas;
        H(7) = H(7) * rapnoy;
        H(8) = H(8) * rapnoy;
        H(10) = H(10) * rapnoy;
        H(13) = H(13) * rapnoy;
        Q(5) = Q(5) * rapnoy;
        Q(6) = Q(6) * rapnoy;
        Q(21) = Q(21) * rapnoy;
        Q(22) = Q(22) * rapnoy;
        totmas = Q(19) + Q(20) + Q(24) + Q(26) + Q(27);
        rapbry = Q(41) / totmas;
        H(9) = H(9) * rapbry;
        H(12) = H(12) * rapbry;
        Q(19) = Q(19) * rapbry;
        Q(20) = Q(20) * rapbry;
        Q(24) = Q(24) * rapbry;
        Q(26) = Q(26) * rapbry;
        Q(27) = Q(27) * rapbry;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (size_t b = 0; b < na; ++b)
        for (int n = 1; n <= NBCON - 1; ++n) Q(n) = fmax(Q(n), 1.e-30);
    /* HNO3 gas = HNO3 below the chemistry levels, F:444-450 */
    for (size_t b = nb; b < na; ++b) Q(NBCON) = Q(5);
#undef Q
#undef H
}

/* sedimentation F:2416-2522: H2O (slot 3) on levels nivh2o-1 .. nivhau and NOy, HNO3
 * (slots 4, 5) on nivbas .. nivhau, bottom-up, each level from the not yet updated
 * values of itself and the level above (so the update is the same in any order). */
void rbo_sediment(const rbo_grid *g, double dt, int nivh2o, const double *p, const double *T,
                  double *qj1, const double *vsed3d, const double *parth2osed,
                  const double *parthno3sed)
{
    const size_t plane = (size_t)g->nlon * g->nlat, field = plane * g->niv;
    double *q3 = qj1 + 2 * field, *q4 = qj1 + 3 * field, *q5 = qj1 + 4 * field;
    for (size_t c = 0; c < plane; ++c) {
#define HNM(k) (p[(size_t)((k) - 1) * plane + c] / (T[(size_t)((k) - 1) * plane + c] * 1.38e-19))
#define AT(a, k) a[(size_t)((k) - 1) * plane + c]
        for (int k = nivh2o - 1; k >= g->nivhau; --k) {
            double dz = 7000. * log(AT(p, k) / AT(p, k - 1));
            double fracup = AT(vsed3d, k - 1) * dt / dz;
            double fracdown = AT(vsed3d, k) * dt / dz;
            fracup = fmin(fracup, 1.0);
            fracdown = fmin(fracdown, 1.0);
            AT(q3, k) = AT(q3, k) + AT(q3, k - 1) * (1. - AT(parth2osed, k - 1)) * HNM(k - 1) / HNM(k) * fracup -
                        AT(q3, k) * (1. - AT(parth2osed, k)) * fracdown;
        }
        for (int k = g->nivbas; k >= g->nivhau; --k) {
            double dz = 7000. * log(AT(p, k) / AT(p, k - 1));
            double fracup = AT(vsed3d, k - 1) * dt / dz;
            double fracdown = AT(vsed3d, k) * dt / dz;
            fracup = fmin(fracup, 1.0);
            fracdown = fmin(fracdown, 1.0);
            AT(q4, k) = AT(q4, k) + AT(q5, k - 1) * (1. - AT(parthno3sed, k - 1)) * HNM(k - 1) / HNM(k) * fracup -
                        AT(q5, k) * (1. - AT(parthno3sed, k)) * fracdown;
            AT(q5, k) = AT(q5, k) + AT(q5, k - 1) * (1. - AT(parthno3sed, k - 1)) * HNM(k - 1) / HNM(k) * fracup -
                        AT(q5, k) * (1. - AT(parthno3sed, k)) * fracdown;
        }
#undef HNM
#undef AT
    }
}

/* ---------------------------------------------------------------- input producers */
#ifndef RBO_OPCOUNT
/* zenith F:5308-5406 for n trajectory points (degrees), day number `day` */
void rbo_zenith(int64_t n, double day, const double *lon, const double *lat, double *sza)
{
    const double pi = 2. * asin(1.0);
    const double ut = (day - trunc(day)) * 24.;
    const double tz = 2. * pi * day / 365.;
    const double sintz = sin(tz), costz = cos(tz);
    const double sin2tz = 2. * sintz * costz;
    const double cos2tz = costz * costz - sintz * sintz;
    const double sin3tz = sintz * cos2tz + costz * sin2tz;
    const double cos3tz = costz * cos2tz - sintz * sin2tz;
    const double rdecl = 0.006918 - 0.399912 * costz + 0.070257 * sintz - 0.006758 * cos2tz +
                         0.000907 * sin2tz - 0.002697 * cos3tz + 0.001480 * sin3tz;
    const double eqr = 0.000075 + 0.001868 * costz - 0.032077 * sintz - 0.014615 * cos2tz -
                       0.040849 * sin2tz;
    const double eqh = eqr * 24. / (2. * pi);
    for (int64_t i = 0; i < n; ++i) {
        const double rlt = lat[i] * pi / 180.;
        const double lbut = 12. - eqh - lon[i] * 24. / 360.;
        const double lzut = 15. * (ut - lbut);
        const double zpt = lzut * pi / 180.;
        const double csz = sin(rlt) * sin(rdecl) + cos(rlt) * cos(rdecl) * cos(zpt);
        const double zr = acos(csz);
        sza[i] = fmax(zr * 180. / pi, 0.);
    }
}

/* altitude A:1-148 on a (mlon, mlat, mniv) regular grid with spacing dlon x dlat degrees
 * (the model: 2 x 2); arrays in Fortran order, lon fastest.  Nearest-column indices are
 * clamped into the grid (the reference assumes they fall inside). */
void rbo_altitude(int mlon, int mlat, int mniv, double dlon, double dlat, const double *relief,
                  const double *psol, const double *aaa, const double *bbb, const double *t,
                  const double *h2o, const double *tlon, const double *tlat, const double *tpre,
                  double *alt)
{
    const size_t plane = (size_t)mlon * mlat, field = plane * mniv;
    const double ctegas = 287.04 / 9.80665;
    const double pi = 2. * asin(1.);
    double *p = (double *)malloc(field * sizeof(double));
    double *tv = (double *)malloc(field * sizeof(double));
    double *geo = (double *)malloc(field * sizeof(double));
#define IX(i, k) ((size_t)(k) * plane + (i))
    for (size_t c = 0; c < plane; ++c) {
        for (int k = 0; k < mniv; ++k) {
            p[IX(c, k)] = aaa[k] + bbb[k] * psol[c];
            tv[IX(c, k)] = t[IX(c, k)] * (1. + 0.608 * h2o[IX(c, k)] * 18. / 28.97);
        }
        geo[IX(c, mniv - 1)] = relief[c] + ctegas * tv[IX(c, mniv - 1)] * log(psol[c] / p[IX(c, mniv - 1)]) * 1.e-3;
        for (int k = mniv - 2; k >= 0; --k)
            geo[IX(c, k)] = geo[IX(c, k + 1)] + ctegas * ((tv[IX(c, k + 1)] + tv[IX(c, k)]) / 2.) *
                                                    log(p[IX(c, k + 1)] / p[IX(c, k)]) * 1.e-3;
    }
    for (size_t b = 0; b < field; ++b) {
        int il = (int)round(tlon[b] / dlon) + 1;
        if (il > mlon) il = 1;
        int jl = (int)round((90. - tlat[b]) / dlat) + 1;
        if (il < 1) il = 1;
        if (jl < 1) jl = 1;
        if (jl > mlat) jl = mlat;
        const size_t c = (size_t)(jl - 1) * mlon + (il - 1);
        const double pirr = fmin(tpre[b], p[IX(c, mniv - 1)]);
        int kr = mniv;   /* smallest 1-based level >= 2 with p >= pirr (the reference's scan) */
        for (int k = 2; k <= mniv; ++k)
            if (p[IX(c, k - 1)] >= pirr) { kr = k; break; }
        const double dz = ctegas * (tv[IX(c, kr - 1)] + tv[IX(c, kr - 2)]) / 2. *
                          log(p[IX(c, kr - 1)] / pirr) * 1.e-3;
        const double g = geo[IX(c, kr - 1)] + dz;
        const double xlat = tlat[b] * pi / 180.;
        alt[b] = (1. + 0.002644 * cos(2. * xlat)) * g + (1. + 0.0089 * cos(2. * xlat)) * (g * g) / 6245.;
    }
#undef IX
    free(p);
    free(tv);
    free(geo);
}

/* ---------------------------------------------------------------- transport interpolation */
/* The grid constants of initransp (F:6028-6034, 6041, 6051, 6075-6085). */
typedef struct {
    double ddlon, ddlat, xpih, xcoef[4], ycoef[4];
} tgrid_t;

static tgrid_t transport_grid(int nlon, int nlat)
{
    tgrid_t t;
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
    return t;
}

void rbo_transport_weights(int nlon, int nlat, int niv, const double *aaa, const double *bbb,
                           const double *x0, const double *y0, const double *z0,
                           const double *p0, int32_t *idx, double *w)
{
    const tgrid_t t = transport_grid(nlon, nlat);
    const int64_t n = (int64_t)nlon * nlat * niv;
#define A(l) aaa[(l) - 1]
#define B(l) bbb[(l) - 1]
    for (int64_t i = 0; i < n; ++i) {
        /* posx F:7402-7406 */
        const int ix = (int)(x0[i] / t.ddlon);
        const double dx = x0[i] - (double)ix * t.ddlon;
        /* posy F:7442-7446 */
        const int iy = (int)((t.xpih - y0[i]) / t.ddlat);
        const double dy = t.xpih - (double)iy * t.ddlat - y0[i];
        /* posz F:7485-7494: the last level l in 2..niv-1 whose pressure is <= z0 */
        const double zz = z0[i], pp = p0[i];
        int ll = 1;
        for (int l = 2; l <= niv - 1; ++l) {
            const double pres = (A(l) + B(l) * pp) * 100.;
            if (zz - pres >= 0.) ll = l;
        }
        idx[i] = ix + 1;
        idx[n + i] = iy + 1;
        idx[2 * n + i] = ll;
        /* poidx F:7529-7536 */
        {
            const double dxm1 = dx + t.ddlon, dxp0 = dx, dxp1 = dx - t.ddlon, dxp2 = dx - t.ddlon * 2.;
            w[i] = t.xcoef[0] * dxp0 * dxp1 * dxp2;
            w[n + i] = t.xcoef[1] * dxm1 * dxp1 * dxp2;
            w[2 * n + i] = t.xcoef[2] * dxm1 * dxp0 * dxp2;
            w[3 * n + i] = t.xcoef[3] * dxm1 * dxp0 * dxp1;
        }
        /* poidy F:7571-7578 */
        {
            const double dym1 = dy + t.ddlat, dyp0 = dy, dyp1 = dy - t.ddlat, dyp2 = dy - t.ddlat * 2.;
            w[4 * n + i] = t.ycoef[0] * dyp0 * dyp1 * dyp2;
            w[5 * n + i] = t.ycoef[1] * dym1 * dyp1 * dyp2;
            w[6 * n + i] = t.ycoef[2] * dym1 * dyp0 * dyp2;
            w[7 * n + i] = t.ycoef[3] * dym1 * dyp0 * dyp1;
        }
        /* poidz F:7614-7644 */
        {
            const double prem1 = ll > 1 ? (A(ll - 1) + B(ll - 1) * pp) * 100. : 5.e-3 * 100.;
            const double prep0 = (A(ll) + B(ll) * pp) * 100.;
            const double prep1 = (A(ll + 1) + B(ll + 1) * pp) * 100.;
            const double prep2 = ll < niv - 1 ? (A(ll + 2) + B(ll + 2) * pp) * 100. : 1200. * 100.;
            const double dzm1 = zz - prem1, dzp0 = zz - prep0, dzp1 = zz - prep1, dzp2 = zz - prep2;
            const double dzp2m1 = prep2 - prem1, dzp2p0 = prep2 - prep0, dzp2p1 = prep2 - prep1;
            const double dzp1m1 = prep1 - prem1, dzp1p0 = prep1 - prep0, dzp0m1 = prep0 - prem1;
            w[8 * n + i] = -(dzp0 * dzp1 * dzp2) / (dzp0m1 * dzp1m1 * dzp2m1);
            w[9 * n + i] = (dzm1 * dzp1 * dzp2) / (dzp0m1 * dzp1p0 * dzp2p0);
            w[10 * n + i] = -(dzm1 * dzp0 * dzp2) / (dzp1m1 * dzp1p0 * dzp2p1);
            w[11 * n + i] = (dzm1 * dzp0 * dzp1) / (dzp2m1 * dzp2p0 * dzp2p1);
        }
    }
#undef A
#undef B
}

int64_t rbo_internew(int nlon, int nlat, int niv, const int32_t *idx, const double *w,
                     const double *q, double *out)
{
    const int64_t n = (int64_t)nlon * nlat * niv, plane = (int64_t)nlon * nlat;
    /* the extended grid x(0:nlon+2, 0:nlat+1, 0:niv+1) of F:8374-8413, as a map to the
     * field: levels clamped (F:8408-8411), longitudes periodic (F:8399-8403), the polar
     * rows mirrored through the pole (F:8388-8393, x(i,0) = x(ip(i),2),
     * x(i,nlat+1) = x(ip(i),nlat-1), ip(i) = mod(i+nlon/2-1,nlon)+1, F:6089-6091) */
#define IP(i) (((i) + nlon / 2 - 1) % nlon + 1)
    int64_t bad = 0;
    for (int64_t p = 0; p < n; ++p) {
        const int i = idx[p], k = idx[n + p], l = idx[2 * n + p];
        /* kg0 = nlat (a point on the south-pole row): the stencil reaches one row past
         * the extended grid, with weight 0 when the point is on the row; that row is read
         * as row nlat + 1 */
        if (i < 1 || i > nlon || k < 1 || k > nlat || l < 1 || l > niv) {
            ++bad;
            continue;
        }
        double xv[4][4][4];   /* [level][lat][lon] offsets -1..2 */
        for (int c = 0; c < 4; ++c) {
            int lc = l + c - 1;
            if (c == 0 && lc < 0) lc = 0;               /* lm1 = max(l-1, 0) */
            if (c == 3 && lc > niv + 1) lc = niv + 1;   /* lp2 = min(l+2, niv+1) */
            const int lv = lc < 1 ? 1 : (lc > niv ? niv : lc);
            for (int b = 0; b < 4; ++b) {
                const int kk = k + b - 1 > nlat + 1 ? nlat + 1 : k + b - 1;
                for (int a = 0; a < 4; ++a) {
                    int ii = i + a - 1;
                    if (ii == 0) ii = nlon;
                    else if (ii > nlon) ii -= nlon;
                    int is = ii, ks = kk;
                    if (kk == 0) { is = IP(ii); ks = 2; }
                    else if (kk == nlat + 1) { is = IP(ii); ks = nlat - 1; }
                    xv[c][b][a] = q[(int64_t)(lv - 1) * plane + (int64_t)(ks - 1) * nlon + (is - 1)];
                }
            }
        }
        /* F:8515-8555: along x; F:8589-8599: along y; F:8603-8606: along z */
        double z[4];
        for (int c = 0; c < 4; ++c) {
            double y[4];
            for (int b = 0; b < 4; ++b)
                y[b] = w[p] * xv[c][b][0] + w[n + p] * xv[c][b][1] + w[2 * n + p] * xv[c][b][2] +
                       w[3 * n + p] * xv[c][b][3];
            z[c] = w[4 * n + p] * y[0] + w[5 * n + p] * y[1] + w[6 * n + p] * y[2] + w[7 * n + p] * y[3];
        }
        double v = w[8 * n + p] * z[0] + w[9 * n + p] * z[1] + w[10 * n + p] * z[2] + w[11 * n + p] * z[3];
        /* the limiter, F:8560-8585 and F:8611-8613: the range of the 8 nearest values */
        const double p00mx = fmax(xv[1][1][1], xv[1][2][1]), p00mn = fmin(xv[1][1][1], xv[1][2][1]);
        const double p10mx = fmax(xv[1][1][2], xv[1][2][2]), p10mn = fmin(xv[1][1][2], xv[1][2][2]);
        const double p01mx = fmax(xv[2][1][1], xv[2][2][1]), p01mn = fmin(xv[2][1][1], xv[2][2][1]);
        const double p11mx = fmax(xv[2][1][2], xv[2][2][2]), p11mn = fmin(xv[2][1][2], xv[2][2][2]);
        const double rmx = fmax(fmax(p00mx, p10mx), fmax(p01mx, p11mx));
        const double rmn = fmin(fmin(p00mn, p10mn), fmin(p01mn, p11mn));
        out[p] = fmax(fmin(v, rmx), rmn);
    }
#undef IP
    return bad;
}

/* ---------------------------------------------------------------- back-trajectories */
/* The extended grids of the trajectory code, x(0:nlon+2, 0:nlat+1, 0:niv+1) (inivit
 * F:6557-6722, intert F:7142-7177) and x(0:nlon+2, 0:nlat+1) (interp F:7029-7050). */
#define EX(i, k, l) ((((size_t)(l)) * (size_t)(nlat + 2) + (size_t)(k)) * (size_t)(nlon + 3) + (size_t)(i))
#define EX2(i, k) ((size_t)(k) * (size_t)(nlon + 3) + (size_t)(i))

typedef struct {
    double xpi, xpi2, xpih, conv, dlatlim, rter;
    double *dlon, *dlat;                       /* dlon(0:nlon+2), dlat(0:nlat+1) */
    double *slon, *clon, *slat;                /* [i-1], [k-1] */
    double *slonr, *clonr, *slatr, *clatr;     /* [(k-1) nlon + i-1]; zero where unset */
    int *ip;                                   /* [i-1] */
} bgrid_t;

/* initransp, F:6010-6097 (cstes F:6010-6025, grid F:6034-6073, ip F:6089-6091,
 * dlatlim F:6101); the rotated-frame tables are not set on the row where clat = 1 and
 * keep the zero the COMMON block starts with. */
static void bgrid_init(bgrid_t *b, int nlon, int nlat)
{
    b->rter = 6371229.;
    b->xpi = 2. * asin(1.);
    b->xpi2 = 2. * b->xpi;
    b->xpih = b->xpi / 2.;
    b->conv = 180. / b->xpi;
    b->dlatlim = 45. / b->conv;
    b->dlon = calloc((size_t)nlon + 3, sizeof(double));
    b->dlat = calloc((size_t)nlat + 2, sizeof(double));
    b->slon = calloc((size_t)nlon, sizeof(double));
    b->clon = calloc((size_t)nlon, sizeof(double));
    b->slat = calloc((size_t)nlat, sizeof(double));
    b->slonr = calloc((size_t)nlon * nlat, sizeof(double));
    b->clonr = calloc((size_t)nlon * nlat, sizeof(double));
    b->slatr = calloc((size_t)nlon * nlat, sizeof(double));
    b->clatr = calloc((size_t)nlon * nlat, sizeof(double));
    b->ip = calloc((size_t)nlon, sizeof(int));
    const double ddlat = b->xpi / (double)(nlat - 1), ddlon = b->xpi2 / (double)nlon;
    double *clat = calloc((size_t)nlat, sizeof(double));
    for (int k = 0; k <= nlat + 1; ++k) b->dlat[k] = b->xpih - ((double)k - 1.) * ddlat;
    /* the flang build evaluates each of these two loops' sin and cos of one argument with
     * one sincos call, whose last bit differs from sin's or cos's on ~0.07 % of arguments */
    for (int k = 1; k <= nlat; ++k) sincos(b->dlat[k], &b->slat[k - 1], &clat[k - 1]);
    for (int i = 0; i <= nlon + 2; ++i) b->dlon[i] = ddlon * ((double)i - 1.);
    for (int i = 1; i <= nlon; ++i) sincos(b->dlon[i], &b->slon[i - 1], &b->clon[i - 1]);
    for (int k = 1; k <= nlat; ++k) {
        if (clat[k - 1] == 1.) continue;
        for (int i = 1; i <= nlon; ++i) {
            const size_t c = (size_t)(k - 1) * nlon + (i - 1);
            b->slatr[c] = -b->clon[i - 1] * clat[k - 1];
            b->clatr[c] = sqrt(1. - b->slatr[c] * b->slatr[c]);
            b->clonr[c] = fmax(-1., fmin(b->slat[k - 1] / b->clatr[c], 1.));
            const double sl = sqrt(1. - b->clonr[c] * b->clonr[c]);
            b->slonr[c] = copysign(sl, b->slon[i - 1]);
        }
    }
    for (int i = 1; i <= nlon; ++i) b->ip[i - 1] = (i + nlon / 2 - 1) % nlon + 1;
    free(clat);
}

static void bgrid_free(bgrid_t *b)
{
    free(b->dlon); free(b->dlat); free(b->slon); free(b->clon); free(b->slat);
    free(b->slonr); free(b->clonr); free(b->slatr); free(b->clatr); free(b->ip);
}

/* sin and cos of the trajectory code as separate calls, as the flang build makes them
 * there (kept out of line so that the compiler does not merge a pair into sincos) */
__attribute__((noinline)) static double bt_sin(double x) { return sin(x); }
__attribute__((noinline)) static double bt_cos(double x) { return cos(x); }

/* instant1 / instant2 (F:6473-6554): the weight of the later record at day daynum */
static double instant_weight(double daynum, double dm1, int ntime, int nintecmwf)
{
    double day = daynum;
    if (day < 1.e-5) day = day + trunc(dm1) + 1.;
    double cp1 = (day - dm1) / ((double)nintecmwf / (double)(24 * ntime));
    cp1 = fmin(cp1, 1.);
    cp1 = fmax(cp1, 0.);
    return cp1;
}

/* posx/posy/posz + poidx/poidy/poidz (F:7371-7647) at one point: cell (1-based) and
 * weights w[12] */
static void cell_weights(const tgrid_t *t, int niv, const double *aaa, const double *bbb,
                         double xx, double yy, double zz, double pp, int want_z, int c[3],
                         double w[12])
{
    const int ix = (int)(xx / t->ddlon);
    const double dx = xx - (double)ix * t->ddlon;
    const int iy = (int)((t->xpih - yy) / t->ddlat);
    const double dy = t->xpih - (double)iy * t->ddlat - yy;
    c[0] = ix + 1;
    c[1] = iy + 1;
    {
        const double dxm1 = dx + t->ddlon, dxp0 = dx, dxp1 = dx - t->ddlon, dxp2 = dx - t->ddlon * 2.;
        w[0] = t->xcoef[0] * dxp0 * dxp1 * dxp2;
        w[1] = t->xcoef[1] * dxm1 * dxp1 * dxp2;
        w[2] = t->xcoef[2] * dxm1 * dxp0 * dxp2;
        w[3] = t->xcoef[3] * dxm1 * dxp0 * dxp1;
    }
    {
        const double dym1 = dy + t->ddlat, dyp0 = dy, dyp1 = dy - t->ddlat, dyp2 = dy - t->ddlat * 2.;
        w[4] = t->ycoef[0] * dyp0 * dyp1 * dyp2;
        w[5] = t->ycoef[1] * dym1 * dyp1 * dyp2;
        w[6] = t->ycoef[2] * dym1 * dyp0 * dyp2;
        w[7] = t->ycoef[3] * dym1 * dyp0 * dyp1;
    }
    if (!want_z) return;
    int ll = 1;
    for (int l = 2; l <= niv - 1; ++l) {
        const double pres = (aaa[l - 1] + bbb[l - 1] * pp) * 100.;
        if (zz - pres >= 0.) ll = l;
    }
    c[2] = ll;
    const double prem1 = ll > 1 ? (aaa[ll - 2] + bbb[ll - 2] * pp) * 100. : 5.e-3 * 100.;
    const double prep0 = (aaa[ll - 1] + bbb[ll - 1] * pp) * 100.;
    const double prep1 = (aaa[ll] + bbb[ll] * pp) * 100.;
    const double prep2 = ll < niv - 1 ? (aaa[ll + 1] + bbb[ll + 1] * pp) * 100. : 1200. * 100.;
    const double dzm1 = zz - prem1, dzp0 = zz - prep0, dzp1 = zz - prep1, dzp2 = zz - prep2;
    const double dzp2m1 = prep2 - prem1, dzp2p0 = prep2 - prep0, dzp2p1 = prep2 - prep1;
    const double dzp1m1 = prep1 - prem1, dzp1p0 = prep1 - prep0, dzp0m1 = prep0 - prem1;
    w[8] = -(dzp0 * dzp1 * dzp2) / (dzp0m1 * dzp1m1 * dzp2m1);
    w[9] = (dzm1 * dzp1 * dzp2) / (dzp0m1 * dzp1p0 * dzp2p0);
    w[10] = -(dzm1 * dzp0 * dzp2) / (dzp1m1 * dzp1p0 * dzp2p1);
    w[11] = (dzm1 * dzp0 * dzp1) / (dzp2m1 * dzp2p0 * dzp2p1);
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* A cell the extended grid holds: ig0 in 1..nlon, kg0 in 1..nlat.  A point on the
 * south-pole row (kg0 = nlat) reaches one row past the reference's extended grid, with
 * weight 0 when the point is on the row; that row is read as row nlat + 1.  Cells outside
 * are counted (and read clamped, so that every read stays inside the fields). */
static int cell_inside(const int c[3], int nlon, int nlat)
{
    return c[0] >= 1 && c[0] <= nlon && c[1] >= 1 && c[1] <= nlat;
}

/* the tricubic Lagrange sum of interv (F:6784-6842, no limiter) or intert
 * (F:7181-7250, limiter at every stage) on an extended field e */
static double tricubic(int nlon, int nlat, const double *e, const int c0[3], const double w[12],
                       int limit)
{
    const int c[3] = {clampi(c0[0], 1, nlon), c0[1], c0[2]};
    double z[4];
    for (int ls = -1; ls <= 2; ++ls) {
        double y[4];
        for (int ks = -1; ks <= 2; ++ks) {
            const int i = c[0], k = clampi(c[1] + ks, 0, nlat + 1), l = c[2] + ls;
            const double xm1 = e[EX(i - 1, k, l)], xp0 = e[EX(i, k, l)], xp1 = e[EX(i + 1, k, l)],
                         xp2 = e[EX(i + 2, k, l)];
            double v = w[0] * xm1 + w[1] * xp0 + w[2] * xp1 + w[3] * xp2;
            if (limit) v = fmax(fmin(v, fmax(xp0, xp1)), fmin(xp0, xp1));
            y[ks + 1] = v;
        }
        double v = w[4] * y[0] + w[5] * y[1] + w[6] * y[2] + w[7] * y[3];
        if (limit) v = fmax(fmin(v, fmax(y[1], y[2])), fmin(y[1], y[2]));
        z[ls + 1] = v;
    }
    double v = w[8] * z[0] + w[9] * z[1] + w[10] * z[2] + w[11] * z[3];
    if (limit) v = fmax(fmin(v, fmax(z[1], z[2])), fmin(z[1], z[2]));
    return v;
}

/* interp (F:6982-7108): the surface pressure ex2 (extended) at (xx, yy), bicubic with
 * the limiter */
static double interp_ps(const tgrid_t *t, int nlon, int nlat, const double *ex2, double xx,
                        double yy, int *outside)
{
    int c[3];
    double w[12];
    cell_weights(t, 0, NULL, NULL, xx, yy, 0., 0., 0, c, w);
    *outside |= !cell_inside(c, nlon, nlat);
    c[0] = clampi(c[0], 1, nlon);
    double y[4];
    for (int ks = -1; ks <= 2; ++ks) {
        const int i = c[0], k = clampi(c[1] + ks, 0, nlat + 1);
        const double xm1 = ex2[EX2(i - 1, k)], xp0 = ex2[EX2(i, k)], xp1 = ex2[EX2(i + 1, k)],
                     xp2 = ex2[EX2(i + 2, k)];
        const double v = w[0] * xm1 + w[1] * xp0 + w[2] * xp1 + w[3] * xp2;
        y[ks + 1] = fmax(fmin(v, fmax(xp0, xp1)), fmin(xp0, xp1));
    }
    const double v = w[4] * y[0] + w[5] * y[1] + w[6] * y[2] + w[7] * y[3];
    return fmax(fmin(v, fmax(y[1], y[2])), fmin(y[1], y[2]));
}

/* corlat (F:7283-7329) */
static void corlat1(const bgrid_t *b, double *x, double *y)
{
    *x = *x - b->xpi2 * (double)(int)(*x / b->xpi2);
    if (*x < 0.) *x = *x + b->xpi2;
    *y = *y - b->xpi2 * (double)(int)(*y / b->xpi2);
    const double temp0 = b->xpih - fabs(*y);
    const double temp1 = *x - b->xpi < 0. ? *x + b->xpi : *x - b->xpi;
    const double temp2 = copysign(b->xpi, *y) - *y;
    if (temp0 < 0.) {
        *x = temp1;
        *y = temp2;
    }
}

/* corniv (F:7333-7367) */
static double corniv1(int niv, const double *aaa, const double *bbb, double z, double p)
{
    const double pmin = (aaa[0] + bbb[0] * p) * 100.;
    const double pmax = (aaa[niv - 1] + bbb[niv - 1] * p) * 100.;
    return fmin(pmax, fmax(pmin, z));
}

/* extended copy of a base-grid field (lon fastest): polar rows mirrored through the pole
 * with sign sg (F:6626-6635), longitudes periodic (F:6648-6660), levels clamped
 * (F:6690-6700); on the rows of skip (the equator row of u1m, v1m: -1 for none) nothing is
 * written, as inivit leaves them */
static void extend3(int nlon, int nlat, int niv, const int *ip, double *e, double sg, int skip)
{
    for (int l = 1; l <= niv; ++l)
        for (int i = 1; i <= nlon; ++i) {
            e[EX(i, 0, l)] = sg * e[EX(ip[i - 1], 2, l)];
            e[EX(i, nlat + 1, l)] = sg * e[EX(ip[i - 1], nlat - 1, l)];
        }
    for (int l = 1; l <= niv; ++l)
        for (int k = 0; k <= nlat + 1; ++k) {
            if (k == skip) continue;
            e[EX(0, k, l)] = e[EX(nlon, k, l)];
            e[EX(nlon + 1, k, l)] = e[EX(1, k, l)];
            e[EX(nlon + 2, k, l)] = e[EX(2, k, l)];
        }
    for (int i = 0; i <= nlon + 2; ++i)
        for (int k = 0; k <= nlat + 1; ++k) {
            if (k == skip) continue;
            e[EX(i, k, 0)] = e[EX(i, k, 1)];
            e[EX(i, k, niv + 1)] = e[EX(i, k, niv)];
        }
}

int rbo_backtra(int nlon, int nlat, int niv, const double *aaa, const double *bbb,
                const rbo_forcing *m1, const rbo_forcing *p1, double daynum, double dtdyn,
                int nrapp, int ntime, int nintecmwf, double *x0, double *y0, double *z0,
                double *p0, int32_t *idx, double *w, double *trajlon, double *trajlat,
                double *trajpre, double *trajtem, int nthreads)
{
    if (nlon < 4 || nlat < 5 || niv < 3 || nrapp < 1 || ntime < 1 || nintecmwf < 1) return -1;
    const size_t n = (size_t)nlon * nlat * niv, plane = (size_t)nlon * nlat;
    const size_t ne = (size_t)(nlon + 3) * (nlat + 2) * (niv + 2);
    const tgrid_t t = transport_grid(nlon, nlat);
    bgrid_t b;
    bgrid_init(&b, nlon, nlat);
    double *u1s = calloc(ne, sizeof(double)), *v1s = calloc(ne, sizeof(double));
    double *w1s = calloc(ne, sizeof(double)), *u1m = calloc(ne, sizeof(double));
    double *v1m = calloc(ne, sizeof(double)), *te = calloc(ne, sizeof(double));
    double *pe = calloc((size_t)(nlon + 3) * (nlat + 2), sizeof(double));
    double *pj = malloc(plane * sizeof(double));
    const int eq = (nlat - 1) / 2 + 1;          /* the row inivit skips for u1m, v1m */
    int64_t bad = 0;                            /* points whose cell fell outside, per sub-step */
    /* dtn, the sub-step of the trajectories (F:6237) */
    const double dtn = dtdyn / (double)nrapp;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    for (int iter = 1; iter <= nrapp; ++iter) {
        /* F:6245: the time of this sub-step; instant1/instant2 at it (F:6249-6258) */
        const double daypre = daynum - dtn * (double)(iter - 1) / 86400.;
        const double cp1 = instant_weight(daypre, m1->day, ntime, nintecmwf), cm1 = 1. - cp1;
        for (size_t c = 0; c < plane; ++c) pj[c] = .01 * (cp1 * exp(p1->lnps[c]) + cm1 * exp(m1->lnps[c]));
        /* inivit (F:6557-6722): the winds of this time, scaled, on the extended grid, and
         * in the frame rotated about the y axis poleward of the equator row */
        const double coefh = dtn / b.rter, coefv = dtn;
        for (int l = 1; l <= niv; ++l)
            for (int k = 1; k <= nlat; ++k)
                for (int i = 1; i <= nlon; ++i) {
                    const size_t s = (size_t)(l - 1) * plane + (size_t)(k - 1) * nlon + (i - 1);
                    const size_t e = EX(i, k, l);
                    const double uj = cp1 * p1->u[s] + cm1 * m1->u[s];
                    const double vj = cp1 * p1->v[s] + cm1 * m1->v[s];
                    const double tj = cp1 * p1->t[s] + cm1 * m1->t[s];
                    const double wj = cp1 * p1->w[s] + cm1 * m1->w[s];
                    u1s[e] = uj * coefh;
                    v1s[e] = vj * coefh;
                    w1s[e] = wj * coefv;
                    te[e] = tj;
                    if (k != eq) {
                        const size_t r = (size_t)(k - 1) * nlon + (i - 1);
                        v1m[e] = (u1s[e] * b.slon[i - 1] + v1s[e] * b.clon[i - 1] * b.slat[k - 1]) / b.clatr[r];
                        u1m[e] = (u1s[e] * b.clon[i - 1] - v1s[e] * b.slon[i - 1] * b.slat[k - 1] +
                                  v1m[e] * b.slonr[r] * b.slatr[r]) / b.clonr[r];
                    }
                }
        extend3(nlon, nlat, niv, b.ip, u1s, -1., -1);
        extend3(nlon, nlat, niv, b.ip, v1s, -1., -1);
        extend3(nlon, nlat, niv, b.ip, w1s, 1., -1);
        extend3(nlon, nlat, niv, b.ip, u1m, 1., eq);
        extend3(nlon, nlat, niv, b.ip, v1m, 1., eq);
        extend3(nlon, nlat, niv, b.ip, te, 1., -1);         /* intert F:7142-7177 */
        for (int k = 1; k <= nlat; ++k)                      /* interp F:7029-7050 */
            for (int i = 1; i <= nlon; ++i) pe[EX2(i, k)] = pj[(size_t)(k - 1) * nlon + (i - 1)];
        for (int i = 1; i <= nlon; ++i) {
            pe[EX2(i, 0)] = pe[EX2(b.ip[i - 1], 2)];
            pe[EX2(i, nlat + 1)] = pe[EX2(b.ip[i - 1], nlat - 1)];
        }
        for (int k = 0; k <= nlat + 1; ++k) {
            pe[EX2(0, k)] = pe[EX2(nlon, k)];
            pe[EX2(nlon + 1, k)] = pe[EX2(1, k)];
            pe[EX2(nlon + 2, k)] = pe[EX2(2, k)];
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t q = 0; q < (int64_t)n; ++q) {
            const int i = (int)(q % nlon) + 1, k = (int)((q / nlon) % nlat) + 1, l = (int)(q / plane) + 1;
            int outside = 0;
            if (iter == 1) {                            /* F:6262-6276 */
                x0[q] = b.dlon[i];
                y0[q] = b.dlat[k];
                z0[q] = (aaa[l - 1] + bbb[l - 1] * pj[(size_t)(k - 1) * nlon + (i - 1)]) * 100.;
                p0[q] = pj[(size_t)(k - 1) * nlon + (i - 1)];
                if (trajlon) {
                    const size_t o = (size_t)(nrapp - 1) * n + (size_t)q;
                    trajlon[o] = x0[q] * 180. / b.xpi;
                    trajlat[o] = y0[q] * 180. / b.xpi;
                    trajpre[o] = z0[q] / 100.;
                    trajtem[o] = te[EX(i, k, l)];
                }
            }
            /* first guess of the midpoint, F:6285-6296; the rotated frame, F:6298-6315 */
            double xm = x0[q], ym = y0[q], zm = z0[q], pm = p0[q];
            const int ind = fabs(ym) >= b.dlatlim;
            double x0r = 0., y0r = 0.;
            if (ind) {
                const double sy0r = -bt_cos(x0[q]) * bt_cos(y0[q]);
                y0r = asin(sy0r);
                const double cy0r = sqrt(1. - sy0r * sy0r);
                const double cx0r = cy0r == 0. ? 0. : fmax(-1., fmin(1., bt_sin(y0[q]) / cy0r));
                x0r = copysign(acos(cx0r), bt_sin(x0[q]));
            }
            const double *ue = ind ? u1m : u1s, *ve = ind ? v1m : v1s;
            int c[3];
            double wt[12], u2, v2, w2;
            /* the wind at the first guess (F:6329-6352): the node's own at iter 1 */
            if (iter == 1) {
                u2 = ue[EX(i, k, l)];
                v2 = ve[EX(i, k, l)];
                w2 = w1s[EX(i, k, l)];
            } else {                                    /* interv F:6726-6978 */
                cell_weights(&t, niv, aaa, bbb, xm, ym, zm, pm, 1, c, wt);
                outside |= !cell_inside(c, nlon, nlat);
                u2 = tricubic(nlon, nlat, ue, c, wt, 0);
                v2 = tricubic(nlon, nlat, ve, c, wt, 0);
                w2 = tricubic(nlon, nlat, w1s, c, wt, 0);
            }
            /* the midpoint, F:6356-6384 */
            if (!ind) {
                xm = x0[q] - 0.5 * u2 / bt_cos(ym);
                ym = y0[q] - 0.5 * v2;
            } else {
                const double symr = -bt_cos(xm) * bt_cos(ym);
                const double cymr = sqrt(1. - symr * symr);
                const double xmr = x0r - 0.5 * u2 / cymr;
                const double ymr = y0r - 0.5 * v2;
                const double sym = bt_cos(xmr) * bt_cos(ymr);
                ym = asin(sym);
                const double cym = sqrt(1. - sym * sym);
                const double cxm = cym == 0. ? 0. : fmax(-1., fmin(1., -bt_sin(ymr) / cym));
                xm = copysign(acos(cxm), bt_sin(xmr));
            }
            zm = z0[q] - 0.5 * w2;
            corlat1(&b, &xm, &ym);                      /* F:6388-6394 */
            pm = interp_ps(&t, nlon, nlat, pe, xm, ym, &outside);
            zm = corniv1(niv, aaa, bbb, zm, pm);
            /* the wind at the midpoint (F:6402) and the departure point, F:6404-6430 */
            cell_weights(&t, niv, aaa, bbb, xm, ym, zm, pm, 1, c, wt);
            outside |= !cell_inside(c, nlon, nlat);
            u2 = tricubic(nlon, nlat, ue, c, wt, 0);
            v2 = tricubic(nlon, nlat, ve, c, wt, 0);
            w2 = tricubic(nlon, nlat, w1s, c, wt, 0);
            double xx = x0[q], yy = y0[q];
            if (!ind) {
                xx = xx - u2 / bt_cos(ym);
                yy = yy - v2;
            } else {
                const double symr = -bt_cos(xm) * bt_cos(ym);
                const double cymr = sqrt(1. - symr * symr);
                x0r = x0r - u2 / cymr;
                y0r = y0r - v2;
                const double sy0 = bt_cos(x0r) * bt_cos(y0r);
                yy = asin(sy0);
                const double cy0 = sqrt(1. - sy0 * sy0);
                const double cx0 = cy0 == 0. ? 0. : fmax(-1., fmin(1., -bt_sin(y0r) / cy0));
                xx = copysign(acos(cx0), bt_sin(x0r));
            }
            double zz = z0[q] - w2;
            corlat1(&b, &xx, &yy);                      /* F:6434-6440 */
            const double pp = interp_ps(&t, nlon, nlat, pe, xx, yy, &outside);
            zz = corniv1(niv, aaa, bbb, zz, pp);
            x0[q] = xx;
            y0[q] = yy;
            z0[q] = zz;
            p0[q] = pp;
            /* the departure cell and weights (F:6444-6451), the temperature there (F:6455) */
            cell_weights(&t, niv, aaa, bbb, xx, yy, zz, pp, 1, c, wt);
            outside |= !cell_inside(c, nlon, nlat);
            if (outside) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                ++bad;
            }
            for (int m = 0; m < 3; ++m) idx[(size_t)m * n + q] = c[m];
            for (int m = 0; m < 12; ++m) w[(size_t)m * n + q] = wt[m];
            const double tt = tricubic(nlon, nlat, te, c, wt, 1);
            if (trajlon && iter < nrapp) {              /* F:6459-6469 */
                const size_t o = (size_t)(nrapp - iter - 1) * n + (size_t)q;
                trajlon[o] = xx * 180. / b.xpi;
                trajlat[o] = yy * 180. / b.xpi;
                trajpre[o] = zz / 100.;
                trajtem[o] = tt;
            }
        }
    }
    free(u1s); free(v1s); free(w1s); free(u1m); free(v1m); free(te); free(pe); free(pj);
    bgrid_free(&b);
    return bad > 0x7fffffff ? 0x7fffffff : (int)bad;
}
#undef EX
#undef EX2
#endif /* RBO_OPCOUNT */
