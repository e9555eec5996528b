/* Test infrastructure: the restatement under AddressSanitizer and UndefinedBehaviorSanitizer
 * (built by `make -C oracle sanitize`, host code only).  Reads one input file, runs the
 * chemistry step, conserv + floor, the converged mode (plain and Anderson), the transport
 * weights, internew and the back-trajectories on it, and writes what they returned, so that
 * a test can compare the sanitized build's results with the ordinary build's.
 *
 * Input (native stream): int32 nlon, nlat, niv, nivbas, nivhau, niter; real64 dt;
 *   T, p, sza, alt (nivbas levels), qj1 (43 fields of niv levels), hc (15), ajl (Fortran
 *   order, 40 * 101 * 27 * 7), o3up (80); the transport part: int32 tlon, tlat, tniv, nrapp,
 *   ntime, nintecmwf; real64 daynum, dtdyn, aaa (tniv), bbb (tniv), then for two records
 *   u, v, w, t (tlon * tlat * tniv) and lnps (tlon * tlat) and the day; a tracer q (tlon *
 *   tlat * tniv).
 * Output (native stream): qj1 and hc after the step (before conserv), after conserv; the
 *   converged qj1 and hc and counts (plain, then Anderson); x0, y0, z0, p0, idx, w of the
 *   back-trajectories; q after internew. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rbchem_oracle.h"

static void *rd(FILE *f, size_t n, size_t sz)
{
    void *p = malloc(n * sz);
    if (!p || fread(p, sz, n, f) != n) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
    return p;
}

static double rd1(FILE *f)
{
    double v;
    if (fread(&v, 8, 1, f) != 1) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
    return v;
}

static void wr(FILE *f, const void *p, size_t n, size_t sz)
{
    if (fwrite(p, sz, n, f) != n) exit(3);
}

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t *hd = rd(f, 6, 4);
    rbo_grid g = {hd[0], hd[1], hd[2], hd[3], hd[4], hd[5]};
    const double dt = rd1(f);
    const size_t plane = (size_t)g.nlon * g.nlat, field = plane * g.niv, fb = plane * g.nivbas;
    double *T = rd(f, fb, 8), *p = rd(f, fb, 8), *sza = rd(f, fb, 8), *alt = rd(f, fb, 8);
    double *q0 = rd(f, 43 * field, 8), *h0 = rd(f, 15 * field, 8);
    double *ajl = rd(f, 40 * 101 * 27 * 7, 8), *o3up = rd(f, 80, 8);
    int32_t *th = rd(f, 6, 4);
    const int tlon = th[0], tlat = th[1], tniv = th[2], nrapp = th[3];
    const size_t tn = (size_t)tlon * tlat * tniv, tp = (size_t)tlon * tlat;
    double *tt = rd(f, 2, 8);
    double *aaa = rd(f, tniv, 8), *bbb = rd(f, tniv, 8);
    rbo_forcing rec[2];
    for (int r = 0; r < 2; ++r) {
        rec[r].u = rd(f, tn, 8);
        rec[r].v = rd(f, tn, 8);
        rec[r].w = rd(f, tn, 8);
        rec[r].t = rd(f, tn, 8);
        rec[r].lnps = rd(f, tp, 8);
        rec[r].day = rd1(f);
    }
    double *tq = rd(f, tn, 8);
    fclose(f);

    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    double *q = malloc(43 * field * 8), *h = malloc(15 * field * 8);
    double *surf = calloc(field, 8), *vs = calloc(fb, 8), *p2o = calloc(fb, 8), *pn = calloc(fb, 8);
    int *it = calloc(fb, sizeof(int));
    memcpy(q, q0, 43 * field * 8);
    memcpy(h, h0, 15 * field * 8);
    if (rbo_chem_step(&g, dt, T, p, sza, alt, ajl, o3up, q, h, surf, vs, p2o, pn, 1)) return 4;
    wr(o, q, 43 * field, 8);
    wr(o, h, 15 * field, 8);
    rbo_conserv_floor(&g, q, h, 1);
    wr(o, q, 43 * field, 8);
    wr(o, h, 15 * field, 8);
    for (int accel = 0; accel <= 2; accel += 2) {
        rbo_grid gc = g;
        gc.niter = 2000;
        memcpy(q, q0, 43 * field * 8);
        memcpy(h, h0, 15 * field * 8);
        if (rbo_chem_converge_aa(&gc, dt, 1e-11, accel, 10, 240, T, p, sza, alt, ajl, o3up, q, h,
                                 surf, vs, p2o, pn, it, NULL, NULL, 1))
            return 5;
        wr(o, q, 43 * field, 8);
        wr(o, h, 15 * field, 8);
        wr(o, it, fb, sizeof(int));
    }
    double *x0 = malloc(tn * 8), *y0 = malloc(tn * 8), *z0 = malloc(tn * 8), *p0 = malloc(tn * 8);
    int32_t *idx = malloc(3 * tn * 4);
    double *w = malloc(12 * tn * 8), *tr = malloc(4 * (size_t)nrapp * tn * 8);
    if (rbo_backtra(tlon, tlat, tniv, aaa, bbb, &rec[0], &rec[1], tt[0], tt[1], nrapp, th[4],
                    th[5], x0, y0, z0, p0, idx, w, tr, tr + nrapp * tn, tr + 2 * nrapp * tn,
                    tr + 3 * nrapp * tn, 1) < 0)
        return 6;
    wr(o, x0, tn, 8);
    wr(o, y0, tn, 8);
    wr(o, z0, tn, 8);
    wr(o, p0, tn, 8);
    wr(o, idx, 3 * tn, 4);
    wr(o, w, 12 * tn, 8);
    double *out = malloc(tn * 8);
    memcpy(out, tq, tn * 8);
    rbo_internew(tlon, tlat, tniv, idx, w, tq, out);
    wr(o, out, tn, 8);
    fclose(o);
    free(q); free(h); free(surf); free(vs); free(p2o); free(pn); free(it); free(x0); free(y0);
    free(z0); free(p0); free(idx); free(w); free(tr); free(out);
    free(T); free(p); free(sza); free(alt); free(q0); free(h0); free(ajl); free(o3up); free(hd);
    free(th); free(tt); free(aaa); free(bbb); free(tq);
    for (int r = 0; r < 2; ++r) {
        free((void *)rec[r].u); free((void *)rec[r].v); free((void *)rec[r].w);
        free((void *)rec[r].t); free((void *)rec[r].lnps);
    }
    return 0;
}
