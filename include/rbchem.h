/* rbchem -- MI355X (gfx950) chemistry step for REPROBUS, C ABI.
 *
 * Replaces the model's chemistry sub-step (src/reprobus_l137_20230723.f:398-440):
 * the row loop `do ilat; call chem3d(...)` (F:398-415, signature F:601-616),
 * `call conserv(qj1, hc)` (F:427, body F:2526-2616) and the 1e-30 floor
 * (F:431-440).  Plain pointers and sizes only; every entry returns 0 on success
 * or a negative RB_E* code, with a message from rb_last_error().  Nothing here
 * calls exit(); nothing retains caller pointers after return.
 *
 * Array layout is the model's Fortran layout: a field is (nlon, nlat, niv)
 * with lon fastest; species fields are stacked last (qj1: 43 fields, hc: 15).
 * T and p are the `trajtem` / `trajpre` slice of the current sub-step
 * (F:802-806); T, p, SZA and alt are read on levels 1..nivbas only, so any
 * (nlon, nlat, n >= nivbas) field works.  SZA is given per box in degrees (the
 * model computes it in `zenith`, F:845; pass `sza3d`).
 *
 * One context owns one device, one HIP stream and the device copies of the
 * photolysis table and work buffers.  Calls on one context are serialised on its
 * stream; a context is not re-entrant.
 */
#ifndef RBCHEM_H
#define RBCHEM_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define RB_OK 0
#define RB_EINVAL (-1)     /* bad argument (null pointer, bad grid) */
#define RB_EHIP (-2)       /* HIP runtime error */
#define RB_ENOTABLE (-3)   /* rb_load_jtable not called */
#define RB_ENONFINITE (-4) /* non-finite values in the output (count in rb_stats) */
#define RB_ENOMEM (-5)     /* device allocation failed */

typedef struct rb_ctx rb_ctx;

/* Grid extents (Fortran order) and the level ranges the reference hard-codes:
 * chemistry on levels 1..nivbas (F:17, nivbas = 94), heterogeneous chemistry on
 * levels nivhau..nivbas (F:2425, nivhau = 14).  Level 1 is the model top. */
typedef struct {
    int32_t nlon, nlat, niv;
    int32_t nivbas, nivhau;
} rb_grid;

typedef struct {
    int32_t niter;          /* fixed sweeps per step; the reference uses 5 (F:630) */
    int32_t check_finite;   /* 1: after rb_chem_step(_dev), rb_chem_step_noalloc_dev,
                               rb_chem_converge_dev, rb_chem_newton_dev and
                               rb_chem_surrogate_dev, count the non-finite values of qj1
                               and hc (rb_stats.nonfinite; synchronises the stream) and
                               return RB_ENONFINITE if there are any; skipped while a graph
                               is captured (rb_run_day_dev with graph = 1) */
    int32_t block;          /* threads per block of the per-box kernel: 64, 128, 256 (0: 256) */
    int32_t profile;        /* 1: time every step's per-box kernels (set-up and sweeps) with
                               HIP events */
    int32_t conv_refill;    /* converged mode: idle lanes of a wave that take new boxes
                               together (1..64; 0: 16) -- one set-up pass serves them all */
    int32_t conv_aa_start;  /* converged mode with accel: plain sweeps before this sweep
                               number, Anderson mixing from it on (0: from the first) */
    int32_t conv_aa_budget; /* converged mode with accel: accelerated sweeps a box gets
                               before it falls back, i.e. the sweep cap is
                               min(maxiter, conv_aa_start + conv_aa_budget) (0: 240) */
    int32_t newton_waves;   /* Newton solve: resident waves per SIMD of its kernel, 1, 2 or 4
                               (0: 2) -- more waves hide the latency of its factor storage,
                               fewer leave each more registers */
    int32_t conv_order;     /* converged mode (rb_chem_converge_dev, rb_chem_newton_dev,
                               rb_chem_surrogate_dev): the order its persistent kernels take
                               boxes in -- 0: night boxes (SZA > 95) first, then day boxes, so
                               that a wave's lanes mostly run one branch of the sweep; 1: grid
                               order.  Results are the same either way. */
} rb_config;

typedef struct {
    int64_t boxes;          /* boxes processed by the last step */
    int64_t nonfinite;      /* non-finite outputs of the last checked step */
    double last_kernel_ms;  /* device time of the last step's kernels */
    double h2d_ms, d2h_ms;  /* copy times of the last host-array call */
    double chem_ms_total;   /* profile mode: summed device time of the per-box kernels ... */
    int64_t chem_launches;  /* ... over this many launches since the previous rb_get_stats */
} rb_stats_t;

/* Create a context on HIP device `device`; cfg may be NULL (defaults).
 * Replaces the Troe set-up of `const` (F:5927-5948). */
int rb_init(rb_ctx **ctx, int device, const rb_config *cfg);

/* Load the photolysis table: ajl(40,101,27,7) and o3up(0:79), Fortran order.
 * Replaces the jstrato.txt read of `const` (F:5953-5973, COMMON /photodis/ F:714). */
int rb_load_jtable(rb_ctx *ctx, const double *ajl, const double *o3up);

/* Use `stream` (a hipStream_t) for subsequent calls; NULL restores the context's own. */
int rb_set_stream(rb_ctx *ctx, void *stream);

/* One chemistry step on host (caller-owned, Fortran-order) arrays: copies in,
 * runs, copies back.  qj1 and hc are updated in place; surfarea (niv levels)
 * is written on levels nivhau..nivbas only, like `analytic` (F:3240-3246);
 * vsed3d, parth2osed, parthno3sed are (nlon, nlat, nivbas) fields (F:906-912).
 * Replaces the chem3d row loop F:398-415. */
int rb_chem_step(rb_ctx *ctx, const rb_grid *g, double dt, const double *T, const double *p,
                 const double *sza, const double *alt, double *qj1, double *hc,
                 double *surfarea, double *vsed3d, double *parth2osed, double *parthno3sed);

/* conserv (F:2526-2616, levels 1..nivbas), the floor max(q, 1e-30) of species
 * 1..42 on all niv levels (F:431-440), and HNO3 gas (slot 43) = HNO3 (slot 5) on levels
 * nivbas+1..niv (F:444-450), on host arrays. */
int rb_conserv_floor(rb_ctx *ctx, const rb_grid *g, double *qj1, double *hc);

/* Device-resident variants: the same contracts on device pointers, enqueued on
 * the context's stream (asynchronous; use rb_sync).  `fuse_conserv` = 1 also
 * applies conserv + floor to the chemistry levels in the same pass (exact when
 * nothing -- e.g. sedimentation, F:419-423 -- runs in between). */
int rb_chem_step_dev(rb_ctx *ctx, const rb_grid *g, double dt, const double *T, const double *p,
                     const double *sza, const double *alt, double *qj1, double *hc,
                     double *surfarea, double *vsed3d, double *parth2osed, double *parthno3sed,
                     int fuse_conserv);
int rb_conserv_floor_dev(rb_ctx *ctx, const rb_grid *g, double *qj1, double *hc);

/* The device-resident step (rb_chem_step_dev's contract and results) without any device
 * allocation -- no counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Its work
 * buffers (the column pre-pass and the set-up record, 144 doubles per chemistry box) live in
 * a fixed device array that every GPU the library runs on holds until the process ends
 * (37.7 MB), so a grid may hold at most RB_NOALLOC_MAX_BOXES chemistry boxes
 * (nlon*nlat*nivbas; larger: RB_EINVAL, nothing launched).  Every call on one GPU shares
 * that array: calls that overlap on one GPU (from several contexts or streams) race.  The
 * per-launch timing of rb_config.profile is not taken (creating its events allocates). */
#define RB_NOALLOC_MAX_BOXES 32768
int rb_chem_step_noalloc_dev(rb_ctx *ctx, const rb_grid *g, double dt, const double *T,
                             const double *p, const double *sza, const double *alt, double *qj1,
                             double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                             double *parthno3sed, int fuse_conserv);

/* Input producers (replace zenith(trajlon, trajlat, irapp, ilat, daynum, sza), F:5308-5406,
 * called from chem3d at F:845, and altitude(nlon, nlat, niv, relief, psol, aaa, bbb, t,
 * h2o, trajlon, trajlat, trajpre, alt), src/altitude.f90:1-148, called at F:391), so the
 * driver hands over only the trajectory lon/lat/p, T and the day number.
 *
 * rb_zenith: SZA (deg) of levels 1..nivbas of the trajectory points (fields of the grid's
 * shape; sza receives nlon*nlat*nivbas values) on day number `day` (fractional part =
 * UT).  rb_altitude: geometric altitude (km) of every trajectory point (grid nlon x nlat x
 * niv, the same shape as the model's regular grid) from the regular grid's relief [km],
 * surface pressure psol [hPa], hybrid coefficients (p = aaa + bbb psol, niv each), T and
 * H2O mixing ratio; the regular grid has spacing dlon x dlat deg (the model: 2 x 2, with
 * nearest-column indices nint(lon/dlon)+1, nint((90-lat)/dlat)+1, clamped into the grid).
 * Host-array and device-array (_dev, stream-ordered) variants. */
int rb_zenith(rb_ctx *ctx, const rb_grid *g, double day, const double *trajlon,
              const double *trajlat, double *sza);
int rb_zenith_dev(rb_ctx *ctx, const rb_grid *g, double day, const double *trajlon,
                  const double *trajlat, double *sza);
int rb_altitude(rb_ctx *ctx, const rb_grid *g, double dlon, double dlat, const double *relief,
                const double *psol, const double *aaa, const double *bbb, const double *t,
                const double *h2o, const double *trajlon, const double *trajlat,
                const double *trajpre, double *alt);
int rb_altitude_dev(rb_ctx *ctx, const rb_grid *g, double dlon, double dlat,
                    const double *relief, const double *psol, const double *aaa,
                    const double *bbb, const double *t, const double *h2o,
                    const double *trajlon, const double *trajlat, const double *trajpre,
                    double *alt);

/* Semi-Lagrangian transport interpolation: what semilag does after the trajectories
 * (F:6121-6153), on the transport grid nlon x nlat x niv (g->nivbas, g->nivhau unused; the
 * model: 180 x 91 x 137, lon fastest as the model's (nlon, nlat, niv) arrays).  Departure
 * points: x0, y0 longitude and latitude (rad, x0 in [0, 2 pi), y0 inside the poles), z0
 * pressure (Pa) and p0 the surface pressure of the departure column (hPa), as backtra's
 * COMMON /depo/ (F:6215-6217); aaa, bbb the niv hybrid coefficients (COMMON /verniv/).
 * Each entry synchronises the stream (it reads back the count of departure points whose
 * cell falls outside the grid, RB_EINVAL when there is any). */
/* One forcing record of the trajectory code (COMMON /forcm1/ and /forcp1/, F:6197-6203),
 * device arrays: winds u, v (m/s) and omega w (Pa/s), temperature t (K), nlon*nlat*niv;
 * lnps the natural log of the surface pressure (Pa), nlon*nlat; day its day number. */
typedef struct {
    const double *u, *v, *w, *t, *lnps;
    double day;
} rb_forcing;
/* rb_transport_backtra_dev replaces `call backtra(daynum, dtdyn, nrapp, trajlon, trajlat,
 * trajpre, trajtem, ntime, nintecmwf)` (F:6142-6144, F:6156-6470): the back-trajectory of
 * every node over dtdyn seconds ending at day daynum, in nrapp sub-steps of two-pass
 * midpoint trajectories (in a frame rotated about the y axis poleward of 45 degrees), the
 * forcing interpolated in time between the records m1 and p1 (instant1/instant2; records
 * nintecmwf / ntime hours apart).  Outputs, device: the departure points x0, y0, z0, p0 (n
 * each, COMMON /depo/), their cells idx and weights w (as rb_transport_weights_dev gives
 * them, for rb_transport_interp_dev), and, with trajlon non-NULL, the positions (deg, deg,
 * hPa) and temperatures (K) along the trajectory: traj*[(s - 1) n + p] for sub-step slot
 * s = 1..nrapp, slot nrapp being the arrival nodes, as the model's trajlon(:,:,:,s).  The
 * four trajectory arrays are all set or all NULL.  Needs nlat >= 5; synchronises the
 * stream; RB_EINVAL if a point's cell falls outside the grid on the way. */
int rb_transport_backtra_dev(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                             const rb_forcing *m1, const rb_forcing *p1, double daynum,
                             double dtdyn, int32_t nrapp, int32_t ntime, int32_t nintecmwf,
                             double *x0, double *y0, double *z0, double *p0, int32_t *idx,
                             double *w, double *trajlon, double *trajlat, double *trajpre,
                             double *trajtem);
/* rb_transport_backtra: the same on host arrays (the forcing records' pointers are host
 * addresses; copies through the context's device buffers): the departure points x0 .. p0
 * and the trajectory slots trajlon .. trajtem (nrapp n each, all required), for the
 * model's backtra call with its COMMON arrays; rb_transport then interpolates. */
int rb_transport_backtra(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                         const rb_forcing *m1, const rb_forcing *p1, double daynum, double dtdyn,
                         int32_t nrapp, int32_t ntime, int32_t nintecmwf, double *x0, double *y0,
                         double *z0, double *p0, double *trajlon, double *trajlat,
                         double *trajpre, double *trajtem);
/* rb_transport_weights_dev replaces posx/posy/posz and poidx/poidy/poidz as backtra calls
 * them (F:6444-6451, F:7371-7647): idx (3 n int32: ig0, kg0, lg0, 1-based, COMMON
 * /indices/) and w (12 n doubles: alfxm1 .. alfzp2, COMMON /poids/ order), device. */
int rb_transport_weights_dev(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                             const double *x0, const double *y0, const double *z0,
                             const double *p0, int32_t *idx, double *w);
/* rb_transport_interp_dev replaces `do jc = 1, nbcon; call internew(qj1, jc)` (F:6148-6150,
 * F:8339-8618): ntr fields q_in (tracer stride n) interpolated at the departure points
 * described by idx and w into q_out (same layout, must not overlap q_in), device. */
int rb_transport_interp_dev(rb_ctx *ctx, const rb_grid *g, const int32_t *idx, const double *w,
                            int32_t ntr, const double *q_in, double *q_out);
/* rb_transport: both on host arrays, the ntr fields of qj1 updated in place (copies
 * through the context's device buffers). */
int rb_transport(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                 const double *x0, const double *y0, const double *z0, const double *p0,
                 int32_t ntr, double *qj1);

/* A day of sub-steps on device-resident state (the model's loop F:300-440 over one day of
 * chemistry, without transport): for k = 0..nsteps-1, rb_zenith_dev at day number
 * day0 + k dt / 86400 into sza (device, nlon*nlat*nivbas), then rb_chem_step_dev (or, with
 * converged = 1, rb_chem_converge_dev with tol/maxiter; iters may be NULL) with conserv and
 * the floor fused.  use_graph = 1 captures the whole day once into a HIP graph (kept in
 * the context and rebuilt when any argument or the configuration changes) and replays
 * it with one launch; results are identical to use_graph = 0.  No counterpart in the
 * reference. */
typedef struct {
    int32_t nsteps;
    int32_t use_graph;
    int32_t converged;
    int32_t maxiter;
    double day0, dt, tol;
    int32_t accel;          /* converged days: Anderson depth, 0 or 2 */
    int32_t pad_;
} rb_day_config;

int rb_run_day_dev(rb_ctx *ctx, const rb_grid *g, const rb_day_config *dc, const double *trajlon,
                   const double *trajlat, const double *T, const double *p, const double *alt,
                   double *sza, double *qj1, double *hc, double *surfarea, double *vsed3d,
                   double *parth2osed, double *parthno3sed, int32_t *iters);

/* Sedimentation (replaces sedimentation(trajpre, trajtem, irapp, qj1, vsed3d, nivh2oecmwf,
 * parth2osed, parthno3sed, dtsedi), F:2416-2522, called at F:419-423 between the
 * chemistry and rb_conserv_floor on the sub-steps where irapp == nrapp).  Moves H2O
 * (qj1 slot 3, levels nivh2o-1 .. nivhau) and NOy/HNO3 (slots 4, 5, levels nivbas ..
 * nivhau) down with the settling velocity vsed3d and the condensed fractions
 * (1 - parth2osed), (1 - parthno3sed) from rb_chem_step, over dt (the model's dtsedi).
 * T and p are the step's fields (levels 1..nivbas read).  Needs nivhau >= 2 and
 * nivh2o <= nivbas + 1 (the model: nivh2oecmwf = 59).  rb_sediment takes host arrays,
 * rb_sediment_dev device arrays (stream-ordered, no synchronisation). */
int rb_sediment(rb_ctx *ctx, const rb_grid *g, double dt, int32_t nivh2o, const double *T,
                const double *p, double *qj1, const double *vsed3d, const double *parth2osed,
                const double *parthno3sed);
int rb_sediment_dev(rb_ctx *ctx, const rb_grid *g, double dt, int32_t nivh2o, const double *T,
                    const double *p, double *qj1, const double *vsed3d,
                    const double *parth2osed, const double *parthno3sed);

/* Converged mode (no counterpart in the reference, which always stops after niter = 5
 * sweeps, F:630): the same per-box algorithm swept to its fixed point.  d_k is the largest
 * relative change of any long- or short-lived species over sweep k (values under the 1e-30
 * mole-fraction floor compared at the floor).  A box stops when every species changed by
 * less than tol and either
 *   - after at least 3 unmixed sweeps, the estimated distance to the fixed point is below
 *     tol: d_k rho / (1 - rho) < tol, rho = min(max(d_k / d_(k-1), d_(k-1) / d_(k-2)), 0.999);
 *   - or 8 sweeps in a row changed by less than tol and the change rose during them (the
 *     rounding-noise floor, or a 2-cycle across a min/max kink of the map);
 * or after maxiter sweeps.  A persistent kernel takes boxes from a device work queue, so
 * converged boxes free their lanes at once.
 *   - guess_qj1 / guess_hc (device, mixing ratios, state layout, or both NULL) give the
 *     initial iterate instead of the incoming state -- the learned initial guess; the
 *     start-of-step state ck0 is always qj1.  A guess that puts a species whose fixed point
 *     lies above the floor far under it can stop a box short of the fixed point, because
 *     changes under the floor are not seen; on the c2 grid the learned guess does not (its
 *     answer and the incoming state's differ by more than 1e-9 at 1 box of 786,240, by
 *     1.01e-9).
 *   - iters (device, nlon*nlat*nivbas int32, or NULL) receives the sweep count of each box.
 *   - accel = 2 accelerates the iteration (Anderson mixing of depth 2 over the last three
 *     sweeps, Walker & Ni 2011, from sweep rb_config.conv_aa_start on; a box whose change
 *     is below tol sweeps unmixed): the same fixed point and stopping test, fewer sweeps;
 *     0 sweeps plainly.  accel with a guess returns RB_EINVAL: from the learned guess the
 *     mixing carries night NO2 of ~150 c2 boxes far under the floor, as above.
 * At a few dozen boxes of 786,240 the fixed point is not defined to 1e-9: the map has
 * branches that switch within rounding of it, and its answer moves by up to 1.6e-8 when T
 * or p moves by one part in 1e15.  Different starts or mixing may end on either side there.
 * A box that goes non-finite or does not converge in maxiter sweeps (in min(maxiter,
 * conv_aa_start + conv_aa_budget) with accel) gets the fixed-sweep answer (niter sweeps
 * from the incoming state) and a negative count. */
int rb_chem_converge_dev(rb_ctx *ctx, const rb_grid *g, double dt, double tol, int32_t maxiter,
                         int32_t accel, const double *T, const double *p, const double *sza,
                         const double *alt, double *qj1, double *hc, double *surfarea,
                         double *vsed3d, double *parth2osed, double *parthno3sed,
                         const double *guess_qj1, const double *guess_hc, int32_t *iters,
                         int fuse_conserv);

/* Newton solve of the converged mode's fixed point (no counterpart in the reference): the
 * same answer as rb_chem_converge_dev -- the fixed point of the per-box sweep map -- from
 * Newton iterations on the fixed-point conditions in their local form (the QSSA ratios of
 * the short-lived species, the backward-Euler balance of the long-lived ones), whose
 * Jacobian is sparse (55 unknowns, at most 399 nonzeros) and is factorised by a generated
 * sparse LU with a fixed pivot order, in FP64.
 *   - With sweeps_first > 0, plain sweeps from the initial iterate (guess_qj1 / guess_hc, or
 *     the incoming state) come first, under rb_chem_converge_dev's stopping rule; the boxes
 *     still unconverged at sweep sweeps_first go on to Newton from their iterate.  With
 *     sweeps_first = 0 every box goes to Newton after one plain sweep.
 *   - Newton: at most max_newton iterations, each step projected onto non-negative species,
 *     until a step moves every species by less than tol relative; then one plain sweep as
 *     the check (every species must change by less than tol; the swept state is the result).
 *   - A box Newton does not solve gets plain sweeps again, at most 40 more (and maxiter in
 *     all), then the fixed-sweep answer, as in rb_chem_converge_dev.
 * iters (optional): k (1 <= k < 10000) for a box the first sweeps converged in k sweeps,
 * 10000 + n for a box Newton solved in n iterations, 20000 + k for a box the last sweeps
 * finished after k sweeps in all, negative for the fixed-sweep fallback.  The other
 * arguments are rb_chem_converge_dev's; 0 <= sweeps_first < maxiter. */
int rb_chem_newton_dev(rb_ctx *ctx, const rb_grid *g, double dt, double tol,
                       int32_t sweeps_first, int32_t max_newton, int32_t maxiter,
                       const double *T, const double *p, const double *sza, const double *alt,
                       double *qj1, double *hc, double *surfarea, double *vsed3d,
                       double *parth2osed, double *parthno3sed, const double *guess_qj1,
                       const double *guess_hc, int32_t *iters, int fuse_conserv);

/* Surrogate step (no counterpart in the reference): sur_qj1 / sur_hc (device, mixing
 * ratios, state layout) are a surrogate model's answer for the step -- the learned
 * network's, rb_mlp_guess_dev.  One FP64 sweep from it is its implicit-step residual
 * test: where every long- and short-lived species changes by less than accept_tol
 * relative (the converged mode's test), the surrogate's answer is kept and iters[box] = 0;
 * every other box goes on sweeping to tol from there, exactly as rb_chem_converge_dev with
 * that guess (the plain FP64 solver is the fallback and decides what is kept).
 * accept_tol must be positive; accel must be 0 (the surrogate's answer is a guess); the
 * other arguments are rb_chem_converge_dev's. */
int rb_chem_surrogate_dev(rb_ctx *ctx, const rb_grid *g, double dt, double accept_tol,
                          double tol, int32_t maxiter, int32_t accel, const double *T,
                          const double *p, const double *sza, const double *alt, double *qj1,
                          double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                          double *parthno3sed, const double *sur_qj1, const double *sur_hc,
                          int32_t *iters, int fuse_conserv);

int rb_sync(rb_ctx *ctx);
/* Waits for the timed launches, fills *out and resets the profile accumulators. */
int rb_get_stats(rb_ctx *ctx, rb_stats_t *out);
const char *rb_last_error(rb_ctx *ctx);
int rb_finalize(rb_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
