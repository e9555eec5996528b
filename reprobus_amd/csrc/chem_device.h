// Device-side declarations shared by the chemistry kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rb {

// Long-lived (transported) slots, src/reprobus_l137_20230723.f:720-730.
enum Long : int {
    N2O, CH4, H2O, NOY, HNO3, N2O5, CLY, OX, CO, OCLO, OXP, H2SO4, HCL, CLONO2, HOCL, CL2,
    H2O2, CLNO2, HBR, BRONO2, NOX, HO2NO2, CLOX, BROX, CL2O2, HOBR, BRCL, CH2O, CH3O2,
    CH3O2H, CFC11, CFC12, CFC113, CCL4, CH3CCL3, CH3CL, HCFC22, CH3BR, H1211, H1301, BRY,
    CH2BR2, HNO3G, NLONG
};
// Short-lived (photochemical-equilibrium) slots, F:743-746.
enum Short : int { O1D, OH, CL, O3P, O3, HO2, NO2, NO, BR, N, CLO, BRO, NO3, H, CH3, NSHORT };
// Photolysis slots, F:754-794 (0-based here).
enum Jslot : int {
    JO2, JNO2, JO3_O1D, JO3_O3P, JNO3_NO, JNO3_NO2, JN2O5, JHNO3, JHNO4, JH2O2, JH2O, JN2O,
    JHOCL, JCLONO2_CL, JCLONO2_CLO, JCL2O2, JHCL, JCL2, JCLNO2, JBRO, JBRONO2, JHOBR, JOCLO,
    JBRCL, JCH3O2H, JCH2O_HCO, JCH2O_CO, JCFC11, JCFC12, JCFC113, JCCL4, JCH3CCL3, JCH3CL,
    JHCFC22, JCH3BR, JH1211, JH1301, JCO2, JCH2BR2, JHBR, JNO, NJ
};

constexpr int NZ = 101, NSZA = 27, NOZO = 7, NJTAB = 40;

// Falloff prefactors k0*300^n and kinf*300^m (const, F:5927-5948).
struct Troe {
    double a0hox, a1hox, a0nox1, a1nox1, a0nox2, a1nox2, a0nox3, a1nox3, a0clx1, a1clx1,
        a0clx2, a1clx2, a0clx3, a1clx3, a0brx, a1brx, a0c1, a1c1;
};

struct StepArgs {
    const double *T, *p, *sza, *alt;   // (plane * niv) fields
    const double *col;                 // 3 fields of plane*nivbas: o3 column, O2 and NO sums
    double *qj1, *hc;                  // 43 / 15 fields of plane*niv
    double *surfarea;                  // plane*niv
    double *vsed, *p2o, *pno3;         // plane*nivbas
    const double *ajl, *o3up;
    int64_t plane, field, nbox;        // plane = nlon*nlat, field = plane*niv, nbox = plane*nivbas
    int32_t nivbas, nivhau, niter, fuse_conserv;
    double dt;
    Troe K;
    const double *guess_q, *guess_h;   // converged mode: optional initial iterate (mixing ratios)
};

struct ConvArgs {
    double tol;                        // fixed-point tolerance (relative change per sweep)
    int32_t maxiter;                   // sweep cap per box
    unsigned long long *counter;       // 3 counters, zero at launch: work-queue head of the
                                       // sweep phase, spill count, queue head of the
                                       // Anderson phase
    int32_t *iters;                    // per-box sweep count (nbox) or null; -n: the box
                                       // did not converge in n sweeps and holds the
                                       // fixed-sweep (niter) answer instead
    int32_t accel;                     // Anderson depth: 0 (plain sweeps) or 2
    double *aa_ws;                     // accel > 0: (2 (accel + 1) + 1) * 57 doubles per lane
    int32_t refill;                    // idle lanes of a wave before they take new boxes
    const double *rec;                 // per-box set-up records (launch_setup)
    int32_t aa_start;                  // accel > 0: plain sweeps before this sweep number
    double *spill;                     // accel > 0: boxes handed from the sweep phase to
                                       // the Anderson phase, CONV_SPILL doubles per box
                                       // (iterate, box index, sweep count), field-major
    double accept_tol;                 // > 0 (surrogate step, with a guess): the first
                                       // sweep from the guess tests it at accept_tol; a
                                       // box under it keeps the guess, count 0
    int32_t iters_offset;              // added to the sweep count of a converged box (the
                                       // Newton solve's resume phase: NEWTON_RESUMED)
    int32_t max_newton;                // Newton solve: iterations before a box is handed
                                       // to the sweeps
    const int32_t *order;              // or null: the box at queue position w is order[w]
                                       // (launch_order_day), else w
};
// doubles per box of the set-up record
constexpr int CONV_RECORD = 141;
// the set-up records of boxes box0 .. box0 + nr - 1 (nr < 0: to the last box), record r of
// box box0 + r at stride nr; full = false writes the fixed-sweep step's record, the J-values
// and the aerosol terms (launch_chem computes the rest itself)
hipError_t launch_setup(const StepArgs &a, double *rec, int block, hipStream_t s,
                        int64_t box0 = 0, int64_t nr = -1, bool full = true);
// Anderson depths the converged kernel is built for
constexpr int AA_DEPTH = 2;
constexpr int AA_NX = 57;
// doubles per box of the spill list between the two phases of the accelerated mode: the
// iterate, box index, sweep count and the last two sweeps' largest relative changes
constexpr int CONV_SPILL = AA_NX + 4;

// semi-Lagrangian transport interpolation (transport_kernels.hip): the grid constants of
// initransp and the launches of the departure-point weights and the tracer interpolation
struct TransportGrid {
    double ddlon, ddlat, xpih, xcoef[4], ycoef[4];
    int nlon, nlat, niv;
};
TransportGrid transport_grid(int nlon, int nlat, int niv);
hipError_t launch_transport_weights(const TransportGrid &t, const double *aaa, const double *bbb,
                                    const double *x0, const double *y0, const double *z0,
                                    const double *p0, int32_t *idx, double *w,
                                    unsigned long long *bad, hipStream_t s);
// values of a tile's window in each of the interpolation's two LDS buffers
constexpr int TRANSPORT_WINDOW = 3200;
hipError_t launch_transport_interp(const TransportGrid &t, const int32_t *idx, const double *w,
                                   const double *q_in, int64_t sin_, double *q_out, int64_t sout,
                                   int ntr, unsigned long long *bad, hipStream_t s, int group = 0);

// Back-trajectories (backtra): one forcing record on the device (winds and omega, T on
// the grid; ln surface pressure on the columns)
struct BtRecord {
    const double *u, *v, *w, *t, *lnps;
};
// initransp's tables on the device: dlon (nlon + 3), dlat (nlat + 2), slon, clon (nlon),
// slat (nlat), slonr, clonr, slatr, clatr (nlon * nlat; zero on the row where clat = 1)
struct BtTables {
    const double *dlon, *dlat, *slon, *clon, *slat, *slonr, *clonr, *slatr, *clatr;
};
struct BacktraArgs {
    TransportGrid t;
    const double *aaa, *bbb;
    BtRecord m1, p1;
    BtTables tb;
    double cp1, cm1, coefh, coefv;      // this sub-step's time weight, wind scalings
    double xpi, xpi2, dlatlim;
    double *f;                          // u1s v1s w1s u1m v1m T (n each), ps (nlon * nlat)
    double *x0, *y0, *z0, *p0;
    int32_t *idx;
    double *w;
    double *traj[4];                    // lon lat pre tem: slot s at s * n, or all null
    int iter, nrapp;
    int64_t n;
    unsigned long long *bad;            // points whose cell fell outside the grid
};
size_t backtra_table_size(int nlon, int nlat);
// initransp's tables into h (host libm, as the model's), in BtTables order
void backtra_tables(int nlon, int nlat, double *h);
BtTables backtra_table_ptrs(int nlon, int nlat, const double *d);
hipError_t launch_backtra_fields(const BacktraArgs &a, hipStream_t s);
hipError_t launch_backtra_step(const BacktraArgs &a, hipStream_t s);

Troe troe_constants();
hipError_t launch_columns(const StepArgs &a, double *col, hipStream_t s);
// the fixed-sweep step from records written by launch_setup with rates = false
hipError_t launch_chem(const StepArgs &a, const double *rec, int block, hipStream_t s,
                       int64_t box0 = 0, int64_t nr = -1);
// plain: one launch of nblocks_sweep blocks; accelerated: the sweep phase on nblocks_sweep
// blocks, then the Anderson phase on nblocks_aa (the Anderson workspace holds its lanes)
hipError_t launch_chem_conv(const StepArgs &a, const ConvArgs &cv, int block, int nblocks_sweep,
                            int nblocks_aa, hipStream_t s);
// Newton solve of the converged mode's fixed point (launch_chem_newton): optionally plain
// sweeps first (sweeps_first > 0: boxes unconverged at that sweep go on), Newton on those
// boxes (or on every box), plain sweeps (the resume phase, at most resume_maxiter) for the
// boxes Newton hands over.  cv.counter holds 5 counters (zero at launch), cv.spill and
// spill2 CONV_SPILL doubles per box each.  Per-box counts: a box converged by the first
// sweeps has its sweep count, a box Newton solved NEWTON_SOLVED + its Newton iterations,
// a box finished by the resume phase NEWTON_RESUMED + its sweeps; negative: fallback.
constexpr int32_t NEWTON_SOLVED = 10000, NEWTON_RESUMED = 20000;
struct NewtonArgs {
    const double *in;                  // input spill list (boxes from the sweep phase) or null
    double *out;                       // output spill list (boxes for the resume phase)
    unsigned long long *queue;         // work-queue head of the Newton launch
    const unsigned long long *n_in;    // entries of the input list
    unsigned long long *n_out;         // entries of the output list
};
struct NewtonLaunch {
    int32_t sweeps_first, waves, block, resume_maxiter;
    int32_t nblocks_sweep, nblocks_newton, nblocks_resume;
    double *spill2;
};
hipError_t launch_chem_newton(const StepArgs &a, const ConvArgs &cv, const NewtonLaunch &nl,
                              hipStream_t s);
int newton_blocks_per_cu(int waves);
// resident blocks per CU of the converged kernel (the Anderson phase's with accel and
// anderson_phase)
int conv_blocks_per_cu(int block, int accel, bool anderson_phase = false);
hipError_t launch_zero_counter(unsigned long long *p, hipStream_t s, int n = 1);
// the converged mode's work order: the night boxes (sza > 95, or not a number) of
// 0 .. nbox - 1 from order[0] on, the day boxes from order[nbox - 1] down; cnt holds 2
// counters, zero at launch
hipError_t launch_order_day(const double *sza, int64_t nbox, int32_t *order,
                            unsigned long long *cnt, hipStream_t s);
// conserv on boxes below plane*nivbas and the floor on every box, for boxes [start, field)
// day-dependent terms of zenith (F:5334-5380), computed once per call on the host
struct ZenithTerms {
    double pi, ut, eqh, sin_decl, cos_decl;
};
ZenithTerms zenith_terms(double day);
hipError_t launch_zenith(const double *lon, const double *lat, int64_t n, const ZenithTerms &z,
                         double *sza, hipStream_t s);
// work: 3 * mlon * mlat * niv doubles (level pressure, virtual temperature, geopotential)
hipError_t launch_altitude(const double *relief, const double *psol, const double *aaa,
                           const double *bbb, const double *t, const double *h2o,
                           const double *tlon, const double *tlat, const double *tpre, int mlon,
                           int mlat, int niv, double dlon, double dlat, double *work, double *alt,
                           hipStream_t s);

// h2o, noy, hno3: the qj1 fields of those slots
hipError_t launch_sediment(const double *p, const double *T, double *h2o, double *noy,
                           double *hno3, const double *vsed, const double *p2o, const double *pno3,
                           int64_t plane, int nivbas, int nivhau, int nivh2o, double dt,
                           hipStream_t s);
hipError_t launch_conserv_floor(double *qj1, double *hc, int64_t plane, int64_t field,
                                int32_t nivbas, int64_t start, hipStream_t s);
hipError_t launch_count_nonfinite(const double *x, int64_t n, unsigned long long *count,
                                  hipStream_t s);

}  // namespace rb
