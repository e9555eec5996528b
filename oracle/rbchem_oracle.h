/* Test infrastructure: CPU restatement of the REPROBUS per-box chemistry step.
 *
 * Follows src/reprobus_l137_20230723.f (chem3d F:601-2412, analytic F:2620-3361,
 * hetero F:3365-4029, density F:4033-4073, phot F:4996-5149, total F:5153-5186,
 * const Troe block F:5927-5948, conserv F:2526-2616, floor F:431-440) and
 * src/jno.f90 (J:1-249).  Used only by tests, smoke checks and the CPU baseline;
 * the product path never links it.
 *
 * Arrays use the model's Fortran layout: a field is (nlon, nlat, niv) with lon
 * fastest; species fields are stacked last.  Level 1 is the model top.
 */
#ifndef RBCHEM_ORACLE_H
#define RBCHEM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nlon, nlat, niv;   /* array extents */
    int nivbas;            /* chemistry acts on levels 1..nivbas */
    int nivhau;            /* heterogeneous chemistry on levels nivhau..nivbas */
    int niter;             /* sweeps per step (reference: 5) */
} rbo_grid;

/* One chemistry step over the whole grid (chem3d for every row). qj1 (43
 * fields) and hc (15 fields) are updated in place; surfarea (niv levels) is
 * written on levels nivhau..nivbas; vsed3d/parth2osed/parthno3sed (nivbas
 * levels) everywhere.  ajl is (40,101,27,7), o3up (80).  nthreads <= 0: all. */
int rbo_chem_step(const rbo_grid *g, double dt, const double *T, const double *p,
                  const double *sza, const double *alt, const double *ajl,
                  const double *o3up, double *qj1, double *hc, double *surfarea,
                  double *vsed3d, double *parth2osed, double *parthno3sed, int nthreads);

/* Converged mode: the same per-box algorithm, sweeping until the relative change of
 * every long- and short-lived species is below tol (values under the 1e-30 mole-fraction
 * floor compared at the floor); iters[box] receives the count.  A box whose change goes
 * non-finite, or that has not converged after g->niter sweeps, gets the reference answer
 * (RBO_NITER_REF sweeps from the incoming state) and iters[box] = -(sweeps tried).
 * guess_qj1/guess_hc (state layout, or both NULL) replace the initial iterate. */
#define RBO_NITER_REF 5
int rbo_chem_converge(const rbo_grid *g, double dt, double tol, const double *T, const double *p,
                      const double *sza, const double *alt, const double *ajl, const double *o3up,
                      double *qj1, double *hc, double *surfarea, double *vsed3d,
                      double *parth2osed, double *parthno3sed, int *iters,
                      const double *guess_qj1, const double *guess_hc, int nthreads);

/* Converged mode with Anderson acceleration of depth accel (0 = plain sweeps, <= 4) on the
 * sweep map from sweep aa_start on (plain sweeps before it): the same fixed point and
 * stopping test (the plain map's residual), a different sequence of iterates.  With
 * accel > 0 and aa_budget > 0 the sweep cap is min(g->niter, aa_start + aa_budget).
 * accel > 0 with a guess returns -2, as the library refuses it: from a guess the mixing
 * can carry a species whose fixed point is above the 1e-30 floor far under it, where the
 * stopping test no longer sees it move. */
int rbo_chem_converge_aa(const rbo_grid *g, double dt, double tol, int accel, int aa_start,
                         int aa_budget, const double *T, const double *p, const double *sza,
                         const double *alt, const double *ajl, const double *o3up, double *qj1,
                         double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                         double *parthno3sed, int *iters, const double *guess_qj1,
                         const double *guess_hc, int nthreads);

/* Surrogate step: sur_qj1/sur_hc (state layout) are a surrogate model's answer for the
 * step.  One sweep from it measures its residual as a fixed point of the sweep map (the
 * converged mode's test at accept_tol); a box under accept_tol keeps the surrogate's
 * answer and gets iters[box] = 0, every other box goes on sweeping to tol exactly as
 * rbo_chem_converge does from that guess (accel > 0 returns -2, as above). */
int rbo_chem_surrogate(const rbo_grid *g, double dt, double accept_tol, double tol, int accel,
                       int aa_start, int aa_budget, const double *T, const double *p,
                       const double *sza,
                       const double *alt, const double *ajl, const double *o3up, double *qj1,
                       double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                       double *parthno3sed, int *iters, const double *sur_qj1,
                       const double *sur_hc, int nthreads);

/* conserv (F:2526-2616) on levels 1..nivbas, then the 1e-30 floor of
 * species 1..42 on all niv levels (F:431-440), then HNO3 gas = HNO3 on levels
 * nivbas+1..niv (F:444-450). */
void rbo_conserv_floor(const rbo_grid *g, double *qj1, double *hc, int nthreads);

/* sedimentation (F:2416-2522) of H2O (levels nivh2o-1..nivhau) and NOy/HNO3 (levels
 * nivbas..nivhau) over a step dt; vsed3d, parth2osed, parthno3sed are the nivbas-level
 * outputs of the chemistry step; p, T as for rbo_chem_step.  Requires nivhau >= 2. */
void rbo_sediment(const rbo_grid *g, double dt, int nivh2o, const double *p, const double *T,
                  double *qj1, const double *vsed3d, const double *parth2osed,
                  const double *parthno3sed);

/* zenith (F:5308-5406): SZA (deg) of n trajectory points (lon, lat in deg) on day `day`. */
void rbo_zenith(int64_t n, double day, const double *lon, const double *lat, double *sza);
/* altitude (src/altitude.f90, A:1-148): geometric altitude (km) of trajectory points
 * (tlon, tlat, tpre on the same (mlon, mlat, mniv) shape) from the regular grid's
 * relief, surface pressure, hybrid coefficients p = aaa + bbb psol, T and H2O; the
 * regular grid has spacing dlon x dlat deg (the model: 2 x 2). */
void rbo_altitude(int mlon, int mlat, int mniv, double dlon, double dlat, const double *relief,
                  const double *psol, const double *aaa, const double *bbb, const double *t,
                  const double *h2o, const double *tlon, const double *tlat, const double *tpre,
                  double *alt);

/* Semi-Lagrangian interpolation of the transport on an nlon x nlat x niv grid (the model:
 * 180 x 91 x 137), arrays lon-fastest as the model's (nlon, nlat, niv).
 * rbo_transport_weights: departure points (x0 longitude and y0 latitude in radians, z0 and
 * p0 pressure and surface pressure of the departure column in Pa and hPa) to the cell
 * indices idx[3][n] (ig0, kg0, lg0, 1-based) and Lagrange weights w[12][n] (alfxm1 alfxp0
 * alfxp1 alfxp2 alfym1 .. alfzp2), as posx/posy/posz and poidx/poidy/poidz (F:7371-7647)
 * with the grid of initransp (F:6028-6090); aaa, bbb are the niv hybrid coefficients.
 * rbo_internew: one tracer field q interpolated at the departure points into out
 * (internew, F:8339-8618).  Returns the number of points whose indices fall outside the
 * grid (kg0 above nlat and the like; 0 for valid departure points; such points are left
 * untouched). */
void rbo_transport_weights(int nlon, int nlat, int niv, const double *aaa, const double *bbb,
                           const double *x0, const double *y0, const double *z0,
                           const double *p0, int32_t *idx, double *w);
int64_t rbo_internew(int nlon, int nlat, int niv, const int32_t *idx, const double *w,
                     const double *q, double *out);

/* One forcing record of the trajectory code (COMMON /forcm1/, /forcp1/, F:6197-6203):
 * winds u, v (m/s) and omega w (Pa/s), temperature t (K) on the (nlon, nlat, niv) grid,
 * log surface pressure lnps (ln Pa) on (nlon, nlat), and the record's day number. */
typedef struct {
    const double *u, *v, *w, *t, *lnps;
    double day;
} rbo_forcing;

/* backtra (F:6156-6470): the departure points of every node of the transport grid over
 * dtdyn seconds in nrapp sub-steps, the forcing interpolated in time between the records
 * m1 and p1 (instant1/instant2, ntime and nintecmwf as the model's), two-pass midpoint
 * trajectories in a frame rotated about the y axis poleward of 45 degrees.  Outputs the
 * departure points x0 y0 (rad) z0 (Pa) p0 (surface pressure there, hPa), their cells and
 * weights idx[3][n], w[12][n] as rbo_transport_weights gives them, and (trajlon non-NULL)
 * the positions and temperatures along the trajectory, traj*[s][n] for slot s + 1 =
 * 1..nrapp (slot nrapp: the arrival nodes).  Returns the number of point-sub-steps whose
 * cell fell outside the grid on the way (0 for a sound flow; reads stay clamped inside the
 * fields), or -1 for bad sizes. */
int rbo_backtra(int nlon, int nlat, int niv, const double *aaa, const double *bbb,
                const rbo_forcing *m1, const rbo_forcing *p1, double daynum, double dtdyn,
                int nrapp, int ntime, int nintecmwf, double *x0, double *y0, double *z0,
                double *p0, int32_t *idx, double *w, double *trajlon, double *trajlat,
                double *trajpre, double *trajtem, int nthreads);

/* ---- pieces exposed for unit tests ---- */
/* Troe prefactors of const (F:5927-5948), 18 values in COMMON order. */
void rbo_troe_constants(double out18[18]);
/* Falloff rate as chem3d evaluates it (e.g. F:930-935). */
double rbo_troe(double a0, double a1, double n0, double n1, double t, double hnm);
/* 103 thermal rates in the order of rbo_rate_names(); returns the count. */
int rbo_rates(double t, double hnm, double pm, double *out);
const char *rbo_rate_name(int i);
/* phot trilinear lookup for one box: tj[0..39]. */
void rbo_phot_box(const double *ajl, const double *o3up, double alt, double sza,
                  double o3t, double *tj40);
/* Column pre-pass for one column of nlev levels (stride between levels `st`):
 * o3t (total), colo2 and colno (jno prefix sums, before the sec(sza) factor). */
void rbo_column(int nlev, const double *p, const double *T, const double *hc_o3,
                const double *hc_no, int st, double *o3t, double *colo2, double *colno);
/* J(NO) for one box from its column sums (jno inner loop). */
double rbo_jno_box(double colo2, double colno, double colo3, double hnm, double sza_top);
/* density (F:4033-4073) */
double rbo_density(double ws, double wn, double t);
/* analytic + hetero for one box (het levels): outputs het[9], parthno3, parthno3sl,
 * parth2o, parthcl, surfarea, vsed, plus diagnostics flags[3] = lnat, lice, condliq. */
void rbo_aerosol_box(double t, double p, double hnm, const double *ck_1based,
                     double het[9], double part[5], double *surf, double *vsed, int flags[3]);

#ifdef __cplusplus
}
#endif
#endif
