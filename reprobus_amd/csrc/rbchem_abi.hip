// C ABI of the chemistry step (declared in include/rbchem.h).
#include "../../include/rbchem.h"

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "chem_device.h"

struct rb_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    rb_config cfg{5, 0, 256, 0, 16, 60, 240, 2, 0};
    rb::Troe K{};
    double *d_ajl = nullptr, *d_o3up = nullptr;
    bool table = false;
    double *d_col = nullptr;          // column pre-pass output: 3 * nbox
    size_t col_cap = 0;
    double *d_stage = nullptr;        // staging for the host-array entry points
    size_t stage_cap = 0;
    unsigned long long *d_count = nullptr;
    unsigned long long *d_queue = nullptr;   // converged mode: queue heads, spill counts (5),
                                             // work-order counts (2)
    double *d_work = nullptr;                // altitude: level pressure, Tv, geopotential
    size_t work_cap = 0;
    double *d_aa = nullptr;                  // converged mode: Anderson history per lane
    size_t aa_cap = 0;
    double *d_rec = nullptr;                 // converged mode: per-box set-up records
    size_t rec_cap = 0;
    double *d_spill = nullptr;               // accelerated mode: boxes between the phases
    size_t spill_cap = 0;
    int32_t *d_order = nullptr;              // converged mode: the boxes in queue order
    size_t order_cap = 0;
    double *d_tw = nullptr;                  // transport (host entry): indices, weights,
    size_t tw_cap = 0;                       // positions, levels
    double *d_tq = nullptr;                  // transport (host entry): fields in and out
    size_t tq_cap = 0;
    double *d_bt = nullptr;                  // back-trajectories: fields of a sub-step, tables
    size_t bt_cap = 0;
    double *d_bth = nullptr;                 // back-trajectories (host entry): inputs, outputs
    size_t bth_cap = 0;
    std::vector<double> bt_tables;           // host tables of the grid bt_grid
    int bt_grid[2] = {0, 0};
    int n_cu = 0;
    double *noalloc = nullptr;               // this device's copy of `scratch` (rb_init)
    bool capturing = false;                  // inside a graph capture: no event records
    hipGraph_t day_graph = nullptr;          // the captured day and what it was built for
    hipGraphExec_t day_exec = nullptr;
    std::vector<char> day_key;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_h2d = nullptr, ev_d2h = nullptr;
    bool timed = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof;   // profile-mode event pairs
    size_t prof_used = 0;
    rb_stats_t stats{};
    std::string err;
};

namespace {

// A named range around an entry point's launches, for tracers (rocprofv3 --marker-trace);
// a no-op without one.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePush(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

int fail(rb_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_fail(rb_ctx *c, hipError_t e, const char *what)
{
    (void)hipGetLastError();   // reported here: not again by the next call's launch check
    return fail(c, RB_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                          \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #call); \
    } while (0)

int check_grid(rb_ctx *ctx, const rb_grid *g)
{
    if (!g) return fail(ctx, RB_EINVAL, "grid is NULL");
    if (g->nlon < 1 || g->nlat < 1 || g->niv < 1)
        return fail(ctx, RB_EINVAL, "grid extents must be positive");
    if (g->nivbas < 1 || g->nivbas > g->niv)
        return fail(ctx, RB_EINVAL, "nivbas must be in 1..niv");
    if (g->nivhau < 1) return fail(ctx, RB_EINVAL, "nivhau must be >= 1");
    return RB_OK;
}

template <class T>
int ensure(rb_ctx *ctx, T **buf, size_t *cap, size_t n)
{
    if (*cap >= n) return RB_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc((void **)buf, n * sizeof(T));
    if (e != hipSuccess) {
        // the runtime keeps the failure as the thread's last error, which the next launch
        // check (hipGetLastError) would report: take it here, where it belongs
        (void)hipGetLastError();
        *buf = nullptr;
        return fail(ctx, RB_ENOMEM, std::string("device allocation of ") +
                                        std::to_string(n * sizeof(T)) + " bytes failed: " +
                                        hipGetErrorString(e));
    }
    *cap = n;
    return RB_OK;
}

// scratch of the converged mode for nbox boxes: set-up records and the work order; with
// Anderson acceleration also the per-lane history of the Anderson phase's grid and the
// spill list
int conv_buffers(rb_ctx *ctx, size_t nbox, int accel)
{
    int rc = ensure(ctx, &ctx->d_rec, &ctx->rec_cap, (size_t)rb::CONV_RECORD * nbox);
    if (!rc) rc = ensure(ctx, &ctx->d_order, &ctx->order_cap, nbox);
    if (rc || !accel) return rc;
    const int blk = ctx->cfg.block > 0 ? ctx->cfg.block : 256;
    const size_t lanes = (size_t)ctx->n_cu * rb::conv_blocks_per_cu(blk, accel, true) * blk;
    if ((rc = ensure(ctx, &ctx->d_aa, &ctx->aa_cap, (2 * (accel + 1) + 1) * rb::AA_NX * lanes)))
        return rc;
    return ensure(ctx, &ctx->d_spill, &ctx->spill_cap, (size_t)rb::CONV_SPILL * nbox);
}

bool any_null(std::initializer_list<const void *> ps)
{
    for (const void *p : ps)
        if (!p) return true;
    return false;
}

// Newton solve: the most plain sweeps a box that Newton hands back still gets (on top of
// the sweeps it had before Newton).  On the c2 grid (100 sweeps first) the handed-back boxes
// that converge at all do so within 22 such sweeps; the others (45 of 113) never do, and
// with a cap of 100 their sweeps set the phase's time (~1 ms).
#ifndef RB_NEWTON_RESUME_SWEEPS
#define RB_NEWTON_RESUME_SWEEPS 40
#endif
constexpr int32_t NEWTON_RESUME_SWEEPS = RB_NEWTON_RESUME_SWEEPS;

// rb_config.check_finite: count the non-finite values of the state after a step (the
// whole qj1 and hc fields), report the count in rb_stats and return RB_ENONFINITE when
// there is any.  It synchronises the stream, so it is skipped inside a graph capture.
int finite_check(rb_ctx *ctx, const double *qj1, const double *hc, int64_t field)
{
    if (!ctx->cfg.check_finite || ctx->capturing) return RB_OK;
    HIPCHK(rb::launch_zero_counter(ctx->d_count, ctx->stream));
    HIPCHK(rb::launch_count_nonfinite(qj1, 43 * field, ctx->d_count, ctx->stream));
    HIPCHK(rb::launch_count_nonfinite(hc, 15 * field, ctx->d_count, ctx->stream));
    unsigned long long cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, ctx->d_count, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->stats.nonfinite = (int64_t)cnt;
    if (cnt) return fail(ctx, RB_ENONFINITE, std::to_string(cnt) + " non-finite outputs");
    return RB_OK;
}

// The allocation-free step (rb_chem_step_noalloc_dev) keeps its work buffers -- the column
// pre-pass (3 doubles per box) and the set-up record (CONV_RECORD doubles per box) of at
// most RB_NOALLOC_MAX_BOXES boxes -- in this array, which every GPU the library runs on
// holds until the process ends.  Host code never passes `scratch` itself to a kernel (on the
// host the name is a stand-in, not a device address): rb_init looks up each device's copy
// once with hipGetSymbolAddress and the entry point passes that address.
constexpr size_t NOALLOC_DOUBLES = (size_t)(3 + rb::CONV_RECORD) * RB_NOALLOC_MAX_BOXES;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

std::mutex g_noalloc_mu;
std::vector<double *> g_noalloc;             // per device: its copy of scratch, or null

// the address of `device`'s copy of scratch, looked up once per device; the device must be
// current
hipError_t noalloc_scratch(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if ((size_t)device >= g_noalloc.size()) g_noalloc.resize((size_t)device + 1, nullptr);
    if (!g_noalloc[(size_t)device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return e != hipSuccess ? e : hipErrorInvalidSymbol;
        }
        g_noalloc[(size_t)device] = (double *)ptr;
    }
    *out = g_noalloc[(size_t)device];
    return hipSuccess;
}

// The launches of one chemistry step (column pre-pass, set-up, sweeps, conserv below the
// chemistry levels, the non-finite check) on caller-supplied work buffers: col (3 nbox
// doubles) and rec (CONV_RECORD nbox doubles).  profile = false skips the per-launch events
// (their creation allocates).  The boxes go in grid order: the converged mode's night-first
// order (launch_order_day) left the c1 step's two kernels at 1.017 against 1.021 ms
// (63 % of the grid-order waves hold day and night boxes) and its launch cost more.
int step_launch(rb_ctx *ctx, const rb_grid *g, double dt, const double *T, const double *p,
                const double *sza, const double *alt, double *qj1, double *hc, double *surfarea,
                double *vsed3d, double *parth2osed, double *parthno3sed, int fuse_conserv,
                double *col, double *rec, bool profile, int nranges = 1)
{
    const int64_t plane = (int64_t)g->nlon * g->nlat;
    const int64_t field = plane * g->niv;
    const int64_t nbox = plane * g->nivbas;
    rb::StepArgs a{};
    a.T = T; a.p = p; a.sza = sza; a.alt = alt; a.col = col;
    a.qj1 = qj1; a.hc = hc; a.surfarea = surfarea;
    a.vsed = vsed3d; a.p2o = parth2osed; a.pno3 = parthno3sed;
    a.ajl = ctx->d_ajl; a.o3up = ctx->d_o3up;
    a.plane = plane; a.field = field; a.nbox = nbox;
    a.nivbas = g->nivbas; a.nivhau = g->nivhau; a.niter = ctx->cfg.niter;
    a.fuse_conserv = fuse_conserv ? 1 : 0;
    a.dt = dt;
    a.K = ctx->K;
    // the boxes in nranges ranges of whole blocks, each set up then swept (a range's records
    // read back while still in the caches).  One range: on the c1 grid, 2, 3, 4, 6 and 8
    // ranges took 1.11, 1.21, 1.25, 1.21 and 1.48 ms against 1.03 ms, the fill and drain
    // of the extra launches costing more than the cache saves.
    const int blk = ctx->cfg.block;
    const int64_t per = ((nbox + nranges - 1) / nranges + blk - 1) / blk * blk;
    auto pair = [&]() -> hipError_t {
        for (int64_t b0 = 0; b0 < nbox; b0 += per) {
            const int64_t n = nbox - b0 < per ? nbox - b0 : per;
            // the step's record: k_chem computes the thermal rates, the air density and the
            // flags itself from T, p and SZA
            hipError_t e = rb::launch_setup(a, rec, blk, ctx->stream, b0, n, false);
            if (e == hipSuccess) e = rb::launch_chem(a, rec, blk, ctx->stream, b0, n);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };

    if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(rb::launch_columns(a, col, ctx->stream));
    if (profile && ctx->cfg.profile && !ctx->capturing) {
        if (ctx->prof_used == ctx->prof.size()) {
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            ctx->prof.emplace_back(e0, e1);
        }
        auto &pe = ctx->prof[ctx->prof_used++];
        HIPCHK(hipEventRecord(pe.first, ctx->stream));
        HIPCHK(pair());
        HIPCHK(hipEventRecord(pe.second, ctx->stream));
    } else {
        HIPCHK(pair());
    }
    if (fuse_conserv && field > nbox)   // floor of the levels below the chemistry
        HIPCHK(rb::launch_conserv_floor(qj1, hc, plane, field, g->nivbas, nbox, ctx->stream));
    if (!ctx->capturing) {
        HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        ctx->timed = true;
    }
    ctx->stats.boxes = nbox;
    return finite_check(ctx, qj1, hc, field);
}

int step_checks(rb_ctx *ctx, const rb_grid *g, double dt,
                std::initializer_list<const void *> arrays)
{
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!ctx->table) return fail(ctx, RB_ENOTABLE, "rb_load_jtable has not been called");
    if (any_null(arrays)) return fail(ctx, RB_EINVAL, "a device array is NULL");
    if (!(dt > 0.) || !std::isfinite(dt)) return fail(ctx, RB_EINVAL, "dt must be positive");
    return RB_OK;
}

}  // namespace

extern "C" {

int rb_init(rb_ctx **out, int device, const rb_config *cfg)
{
    if (!out) return RB_EINVAL;
    *out = nullptr;
    rb_ctx *ctx = new (std::nothrow) rb_ctx();
    if (!ctx) return RB_ENOMEM;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0 || device < 0 || device >= n) {
        delete ctx;
        return RB_EHIP;
    }
    ctx->device = device;
    if (cfg) ctx->cfg = *cfg;
    if (ctx->cfg.niter < 1) ctx->cfg.niter = 5;
    if (ctx->cfg.block <= 0) ctx->cfg.block = 256;
    if (ctx->cfg.conv_refill <= 0) ctx->cfg.conv_refill = 16;
    if (ctx->cfg.conv_refill > 64) ctx->cfg.conv_refill = 64;
    if (ctx->cfg.conv_aa_start < 0) ctx->cfg.conv_aa_start = 60;
    if (ctx->cfg.conv_aa_budget <= 0) ctx->cfg.conv_aa_budget = 240;
    if (ctx->cfg.newton_waves <= 0) ctx->cfg.newton_waves = 2;
    if (ctx->cfg.newton_waves != 1 && ctx->cfg.newton_waves != 2 && ctx->cfg.newton_waves != 4) {
        delete ctx;
        return RB_EINVAL;   // the Newton kernel is instantiated for 1, 2 and 4 waves per SIMD
    }
    if (ctx->cfg.block != 64 && ctx->cfg.block != 128 && ctx->cfg.block != 256) {
        delete ctx;
        return RB_EINVAL;   // the per-box kernel is instantiated for 64, 128 and 256 threads
    }
    ctx->K = rb::troe_constants();
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->own) != hipSuccess ||
        hipMalloc((void **)&ctx->d_count, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void **)&ctx->d_queue, 7 * sizeof(unsigned long long)) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreate(&ctx->ev_h2d) != hipSuccess || hipEventCreate(&ctx->ev_d2h) != hipSuccess) {
        rb_finalize(ctx);
        return RB_EHIP;
    }
    ctx->stream = ctx->own;
    // this device's copy of the allocation-free step's scratch (the device is current)
    if (noalloc_scratch(device, &ctx->noalloc) != hipSuccess) {
        rb_finalize(ctx);
        return RB_EHIP;
    }
    *out = ctx;
    return RB_OK;
}

int rb_load_jtable(rb_ctx *ctx, const double *ajl, const double *o3up)
{
    if (!ctx) return RB_EINVAL;
    if (any_null({ajl, o3up})) return fail(ctx, RB_EINVAL, "ajl/o3up is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)rb::NJTAB * rb::NZ * rb::NSZA * rb::NOZO;
    if (!ctx->d_ajl) HIPCHK(hipMalloc((void **)&ctx->d_ajl, n * sizeof(double)));
    if (!ctx->d_o3up) HIPCHK(hipMalloc((void **)&ctx->d_o3up, 80 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ctx->d_ajl, ajl, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_o3up, o3up, 80 * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->table = true;
    return RB_OK;
}

int rb_set_stream(rb_ctx *ctx, void *stream)
{
    if (!ctx) return RB_EINVAL;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own;
    return RB_OK;
}

int rb_chem_step_dev(rb_ctx *ctx, const rb_grid *g, double dt, const double *T, const double *p,
                     const double *sza, const double *alt, double *qj1, double *hc,
                     double *surfarea, double *vsed3d, double *parth2osed, double *parthno3sed,
                     int fuse_conserv)
{
    TraceRange trace("rb_chem_step_dev");
    if (!ctx) return RB_EINVAL;
    int rc = step_checks(ctx, g, dt, {T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                                      parthno3sed});
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t nbox = (int64_t)g->nlon * g->nlat * g->nivbas;
    if ((rc = ensure(ctx, &ctx->d_col, &ctx->col_cap, (size_t)(3 * nbox)))) return rc;
    if ((rc = conv_buffers(ctx, (size_t)nbox, 0))) return rc;        // the set-up record
    return step_launch(ctx, g, dt, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                       parthno3sed, fuse_conserv, ctx->d_col, ctx->d_rec, true);
}

int rb_chem_step_noalloc_dev(rb_ctx *ctx, const rb_grid *g, double dt, const double *T,
                             const double *p, const double *sza, const double *alt, double *qj1,
                             double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                             double *parthno3sed, int fuse_conserv)
{
    TraceRange trace("rb_chem_step_noalloc_dev");
    if (!ctx) return RB_EINVAL;
    int rc = step_checks(ctx, g, dt, {T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                                      parthno3sed});
    if (rc) return rc;
    const int64_t nbox = (int64_t)g->nlon * g->nlat * g->nivbas;
    if (nbox > RB_NOALLOC_MAX_BOXES)
        return fail(ctx, RB_EINVAL, "the grid has " + std::to_string(nbox) +
                                        " chemistry boxes; the allocation-free step takes at most " +
                                        std::to_string((long long)RB_NOALLOC_MAX_BOXES));
    if (!ctx->noalloc) return fail(ctx, RB_EHIP, "the allocation-free scratch is not available");
    HIPCHK(hipSetDevice(ctx->device));
    // both buffers are written in full before they are read: nothing to clear
    double *col = ctx->noalloc, *rec = ctx->noalloc + 3 * nbox;
    return step_launch(ctx, g, dt, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                       parthno3sed, fuse_conserv, col, rec, false);
}

// the converged mode; accept_tol > 0 makes it the surrogate step (guess = the surrogate)
static int converge_impl(rb_ctx *ctx, const rb_grid *g, double dt, double tol, int32_t maxiter,
                         int32_t accel, const double *T, const double *p, const double *sza,
                         const double *alt, double *qj1, double *hc, double *surfarea,
                         double *vsed3d, double *parth2osed, double *parthno3sed,
                         const double *guess_qj1, const double *guess_hc, int32_t *iters,
                         int fuse_conserv, double accept_tol)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!ctx->table) return fail(ctx, RB_ENOTABLE, "rb_load_jtable has not been called");
    if (any_null({T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed}))
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    if ((guess_qj1 == nullptr) != (guess_hc == nullptr))
        return fail(ctx, RB_EINVAL, "guess_qj1 and guess_hc must be both set or both NULL");
    if (!(dt > 0.) || !std::isfinite(dt)) return fail(ctx, RB_EINVAL, "dt must be positive");
    if (!(tol > 0.) || maxiter < 1) return fail(ctx, RB_EINVAL, "tol must be > 0, maxiter >= 1");
    if (accel != 0 && accel != rb::AA_DEPTH)
        return fail(ctx, RB_EINVAL, "accel must be 0 (plain sweeps) or 2 (Anderson depth 2)");
    // From a guess, the mixing extrapolates some boxes' night NO2 (F:1671-1746; fixed point
    // NOx - b9 O3 / b12, about 1e-11) down to ~1e-41, under the 1e-30 floor.  The stopping
    // test compares values there at the floor, so such a box stops while its NO2 still grows
    // by NOx b12 / (b9 O3) per sweep.  On the c2 grid the learned guess gives ~150 such boxes
    // of 786,240; from the incoming state, none.
    if (accel && guess_qj1)
        return fail(ctx, RB_EINVAL, "accel with a guess is refused: Anderson mixing from a "
                                    "guess can stop boxes under the 1e-30 floor short of the "
                                    "fixed point (use accel 0, or rb_chem_newton_dev)");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t plane = (int64_t)g->nlon * g->nlat;
    const int64_t field = plane * g->niv;
    const int64_t nbox = plane * g->nivbas;
    if ((rc = ensure(ctx, &ctx->d_col, &ctx->col_cap, (size_t)(3 * nbox)))) return rc;

    rb::StepArgs a{};
    a.T = T; a.p = p; a.sza = sza; a.alt = alt; a.col = ctx->d_col;
    a.qj1 = qj1; a.hc = hc; a.surfarea = surfarea;
    a.vsed = vsed3d; a.p2o = parth2osed; a.pno3 = parthno3sed;
    a.ajl = ctx->d_ajl; a.o3up = ctx->d_o3up;
    a.plane = plane; a.field = field; a.nbox = nbox;
    a.nivbas = g->nivbas; a.nivhau = g->nivhau; a.niter = ctx->cfg.niter;   // fallback
    a.fuse_conserv = fuse_conserv ? 1 : 0;
    a.dt = dt;
    a.K = ctx->K;
    a.guess_q = guess_qj1;
    a.guess_h = guess_hc;
    // one resident wave set per launch: blocks per CU from the occupancy calculator
    const int nblocks = ctx->n_cu * rb::conv_blocks_per_cu(ctx->cfg.block, accel);
    const int nblocks_aa = ctx->n_cu * rb::conv_blocks_per_cu(ctx->cfg.block, accel, true);
    if ((rc = conv_buffers(ctx, (size_t)nbox, accel))) return rc;
    // accelerated solves cap the sweeps at conv_aa_start + conv_aa_budget: a box that the
    // mixing has not brought to the fixed point by then only costs more accelerated sweeps
    const int32_t cap = accel ? (int32_t)std::min<int64_t>(maxiter, (int64_t)ctx->cfg.conv_aa_start +
                                                                          ctx->cfg.conv_aa_budget)
                              : maxiter;
    rb::ConvArgs cv{tol, cap, ctx->d_queue, iters, accel, accel ? ctx->d_aa : nullptr,
                    ctx->cfg.conv_refill, ctx->d_rec, ctx->cfg.conv_aa_start,
                    accel ? ctx->d_spill : nullptr, accept_tol};
    cv.order = ctx->cfg.conv_order == 0 ? ctx->d_order : nullptr;

    if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(rb::launch_columns(a, ctx->d_col, ctx->stream));
    HIPCHK(rb::launch_setup(a, ctx->d_rec, ctx->cfg.block, ctx->stream));
    HIPCHK(rb::launch_zero_counter(ctx->d_queue, ctx->stream, 7));
    if (cv.order) HIPCHK(rb::launch_order_day(sza, nbox, ctx->d_order, ctx->d_queue + 5, ctx->stream));
    if (ctx->cfg.profile && !ctx->capturing) {
        if (ctx->prof_used == ctx->prof.size()) {
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            ctx->prof.emplace_back(e0, e1);
        }
        auto &pe = ctx->prof[ctx->prof_used++];
        HIPCHK(hipEventRecord(pe.first, ctx->stream));
        HIPCHK(rb::launch_chem_conv(a, cv, ctx->cfg.block, nblocks, nblocks_aa, ctx->stream));
        HIPCHK(hipEventRecord(pe.second, ctx->stream));
    } else {
        HIPCHK(rb::launch_chem_conv(a, cv, ctx->cfg.block, nblocks, nblocks_aa, ctx->stream));
    }
    if (fuse_conserv && field > nbox)
        HIPCHK(rb::launch_conserv_floor(qj1, hc, plane, field, g->nivbas, nbox, ctx->stream));
    if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = !ctx->capturing;
    ctx->stats.boxes = nbox;
    return finite_check(ctx, qj1, hc, field);
}

int rb_chem_converge_dev(rb_ctx *ctx, const rb_grid *g, double dt, double tol, int32_t maxiter,
                         int32_t accel, const double *T, const double *p, const double *sza,
                         const double *alt, double *qj1, double *hc, double *surfarea,
                         double *vsed3d, double *parth2osed, double *parthno3sed,
                         const double *guess_qj1, const double *guess_hc, int32_t *iters,
                         int fuse_conserv)
{
    TraceRange trace("rb_chem_converge_dev");
    return converge_impl(ctx, g, dt, tol, maxiter, accel, T, p, sza, alt, qj1, hc, surfarea,
                         vsed3d, parth2osed, parthno3sed, guess_qj1, guess_hc, iters,
                         fuse_conserv, 0.);
}

int rb_chem_newton_dev(rb_ctx *ctx, const rb_grid *g, double dt, double tol, int32_t sweeps_first,
                       int32_t max_newton, int32_t maxiter, const double *T, const double *p,
                       const double *sza,
                       const double *alt, double *qj1, double *hc, double *surfarea,
                       double *vsed3d, double *parth2osed, double *parthno3sed,
                       const double *guess_qj1, const double *guess_hc, int32_t *iters,
                       int fuse_conserv)
{
    TraceRange trace("rb_chem_newton_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!ctx->table) return fail(ctx, RB_ENOTABLE, "rb_load_jtable has not been called");
    if (any_null({T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed}))
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    if ((guess_qj1 == nullptr) != (guess_hc == nullptr))
        return fail(ctx, RB_EINVAL, "guess_qj1 and guess_hc must be both set or both NULL");
    if (!(dt > 0.) || !std::isfinite(dt)) return fail(ctx, RB_EINVAL, "dt must be positive");
    if (!(tol > 0.) || maxiter < 1 || max_newton < 1 || sweeps_first < 0 ||
        sweeps_first >= maxiter)
        return fail(ctx, RB_EINVAL,
                    "tol must be > 0, maxiter >= 1, max_newton >= 1, 0 <= sweeps_first < maxiter");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t plane = (int64_t)g->nlon * g->nlat;
    const int64_t field = plane * g->niv;
    const int64_t nbox = plane * g->nivbas;
    if ((rc = ensure(ctx, &ctx->d_col, &ctx->col_cap, (size_t)(3 * nbox)))) return rc;
    if ((rc = conv_buffers(ctx, (size_t)nbox, 0))) return rc;
    // two spill lists: sweep phase -> Newton, Newton -> resume phase
    if ((rc = ensure(ctx, &ctx->d_spill, &ctx->spill_cap, 2 * (size_t)rb::CONV_SPILL * nbox)))
        return rc;

    rb::StepArgs a{};
    a.T = T; a.p = p; a.sza = sza; a.alt = alt; a.col = ctx->d_col;
    a.qj1 = qj1; a.hc = hc; a.surfarea = surfarea;
    a.vsed = vsed3d; a.p2o = parth2osed; a.pno3 = parthno3sed;
    a.ajl = ctx->d_ajl; a.o3up = ctx->d_o3up;
    a.plane = plane; a.field = field; a.nbox = nbox;
    a.nivbas = g->nivbas; a.nivhau = g->nivhau; a.niter = ctx->cfg.niter;   // fallback
    a.fuse_conserv = fuse_conserv ? 1 : 0;
    a.dt = dt;
    a.K = ctx->K;
    a.guess_q = guess_qj1;
    a.guess_h = guess_hc;
    const int blk = ctx->cfg.block;
    rb::ConvArgs cv{tol, maxiter, ctx->d_queue, iters, 0, nullptr, ctx->cfg.conv_refill,
                    ctx->d_rec, 0, ctx->d_spill, 0.};
    cv.max_newton = max_newton;
    cv.order = ctx->cfg.conv_order == 0 ? ctx->d_order : nullptr;
    rb::NewtonLaunch nl{};
    nl.sweeps_first = sweeps_first;
    nl.waves = ctx->cfg.newton_waves;
    nl.block = blk;
    // a box Newton hands over gets at most NEWTON_RESUME_SWEEPS more plain sweeps: such
    // boxes are the ones plain sweeps converge slowest or not at all, and one lane sweeping
    // one of them to maxiter would set the time of the whole solve
    nl.resume_maxiter = (int32_t)std::min<int64_t>(maxiter, (int64_t)sweeps_first + 2 +
                                                                NEWTON_RESUME_SWEEPS);
    nl.nblocks_sweep = ctx->n_cu * rb::conv_blocks_per_cu(blk, 2);
    nl.nblocks_newton = ctx->n_cu * rb::newton_blocks_per_cu(nl.waves);
    nl.nblocks_resume = ctx->n_cu * rb::conv_blocks_per_cu(blk, 0);
    nl.spill2 = ctx->d_spill + (size_t)rb::CONV_SPILL * nbox;

    if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(rb::launch_columns(a, ctx->d_col, ctx->stream));
    HIPCHK(rb::launch_setup(a, ctx->d_rec, blk, ctx->stream));
    HIPCHK(rb::launch_zero_counter(ctx->d_queue, ctx->stream, 7));
    if (cv.order) HIPCHK(rb::launch_order_day(sza, nbox, ctx->d_order, ctx->d_queue + 5, ctx->stream));
    if (ctx->cfg.profile && !ctx->capturing) {
        if (ctx->prof_used == ctx->prof.size()) {
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            ctx->prof.emplace_back(e0, e1);
        }
        auto &pe = ctx->prof[ctx->prof_used++];
        HIPCHK(hipEventRecord(pe.first, ctx->stream));
        HIPCHK(rb::launch_chem_newton(a, cv, nl, ctx->stream));
        HIPCHK(hipEventRecord(pe.second, ctx->stream));
    } else {
        HIPCHK(rb::launch_chem_newton(a, cv, nl, ctx->stream));
    }
    if (fuse_conserv && field > nbox)
        HIPCHK(rb::launch_conserv_floor(qj1, hc, plane, field, g->nivbas, nbox, ctx->stream));
    if (!ctx->capturing) HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = !ctx->capturing;
    ctx->stats.boxes = nbox;
    return finite_check(ctx, qj1, hc, field);
}

int rb_chem_surrogate_dev(rb_ctx *ctx, const rb_grid *g, double dt, double accept_tol,
                          double tol, int32_t maxiter, int32_t accel, const double *T,
                          const double *p, const double *sza, const double *alt, double *qj1,
                          double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                          double *parthno3sed, const double *sur_qj1, const double *sur_hc,
                          int32_t *iters, int fuse_conserv)
{
    TraceRange trace("rb_chem_surrogate_dev");
    if (!ctx) return RB_EINVAL;
    if (!sur_qj1 || !sur_hc)
        return fail(ctx, RB_EINVAL, "the surrogate's answer (sur_qj1, sur_hc) is NULL");
    if (!(accept_tol > 0.) || !std::isfinite(accept_tol))
        return fail(ctx, RB_EINVAL, "accept_tol must be positive and finite");
    return converge_impl(ctx, g, dt, tol, maxiter, accel, T, p, sza, alt, qj1, hc, surfarea,
                         vsed3d, parth2osed, parthno3sed, sur_qj1, sur_hc, iters, fuse_conserv,
                         accept_tol);
}

int rb_conserv_floor_dev(rb_ctx *ctx, const rb_grid *g, double *qj1, double *hc)
{
    TraceRange trace("rb_conserv_floor_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (any_null({qj1, hc})) return fail(ctx, RB_EINVAL, "a device array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t plane = (int64_t)g->nlon * g->nlat;
    HIPCHK(rb::launch_conserv_floor(qj1, hc, plane, plane * g->niv, g->nivbas, 0, ctx->stream));
    return RB_OK;
}

int rb_chem_step(rb_ctx *ctx, const rb_grid *g, double dt, const double *T, const double *p,
                 const double *sza, const double *alt, double *qj1, double *hc, double *surfarea,
                 double *vsed3d, double *parth2osed, double *parthno3sed)
{
    TraceRange trace("rb_chem_step");
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (any_null({T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed}))
        return fail(ctx, RB_EINVAL, "a host array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t field = (size_t)g->nlon * g->nlat * g->niv;
    const size_t nb = (size_t)g->nlon * g->nlat * g->nivbas;
    // T, p, SZA and alt are read on the chemistry levels only (nb boxes)
    const size_t total = 4 * nb + 43 * field + 15 * field + field + 3 * nb;
    if ((rc = ensure(ctx, &ctx->d_stage, &ctx->stage_cap, total))) return rc;
    double *d = ctx->d_stage;
    double *dT = d, *dp = dT + nb, *dsza = dp + nb, *dalt = dsza + nb;
    double *dq = dalt + nb, *dh = dq + 43 * field, *dsurf = dh + 15 * field;
    double *dv = dsurf + field, *dp2o = dv + nb, *dpno3 = dp2o + nb;
    const size_t B = sizeof(double);
    hipStream_t s = ctx->stream;
    HIPCHK(hipEventRecord(ctx->ev_h2d, s));
    HIPCHK(hipMemcpyAsync(dT, T, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dp, p, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dsza, sza, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dalt, alt, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dq, qj1, 43 * field * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dh, hc, 15 * field * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dsurf, surfarea, field * B, hipMemcpyHostToDevice, s));
    // a step with non-finite outputs still hands them back (then returns RB_ENONFINITE)
    const int step_rc = rb_chem_step_dev(ctx, g, dt, dT, dp, dsza, dalt, dq, dh, dsurf, dv, dp2o,
                                         dpno3, 0);
    if (step_rc && step_rc != RB_ENONFINITE) return step_rc;
    const std::string step_err = ctx->err;
    HIPCHK(hipMemcpyAsync(qj1, dq, 43 * field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hc, dh, 15 * field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(surfarea, dsurf, field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(vsed3d, dv, nb * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(parth2osed, dp2o, nb * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(parthno3sed, dpno3, nb * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ctx->ev_d2h, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev_h2d, ctx->ev0) == hipSuccess) ctx->stats.h2d_ms = ms;
    if (hipEventElapsedTime(&ms, ctx->ev1, ctx->ev_d2h) == hipSuccess) ctx->stats.d2h_ms = ms;
    if (step_rc) ctx->err = step_err;
    return step_rc;
}

int rb_conserv_floor(rb_ctx *ctx, const rb_grid *g, double *qj1, double *hc)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (any_null({qj1, hc})) return fail(ctx, RB_EINVAL, "a host array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t field = (size_t)g->nlon * g->nlat * g->niv;
    if ((rc = ensure(ctx, &ctx->d_stage, &ctx->stage_cap, 58 * field))) return rc;
    double *dq = ctx->d_stage, *dh = dq + 43 * field;
    const size_t B = sizeof(double);
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dq, qj1, 43 * field * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dh, hc, 15 * field * B, hipMemcpyHostToDevice, s));
    if ((rc = rb_conserv_floor_dev(ctx, g, dq, dh))) return rc;
    HIPCHK(hipMemcpyAsync(qj1, dq, 43 * field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hc, dh, 15 * field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RB_OK;
}

int rb_zenith_dev(rb_ctx *ctx, const rb_grid *g, double day, const double *trajlon,
                  const double *trajlat, double *sza)
{
    TraceRange trace("rb_zenith_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (any_null({trajlon, trajlat, sza})) return fail(ctx, RB_EINVAL, "a device array is NULL");
    if (!std::isfinite(day)) return fail(ctx, RB_EINVAL, "day must be finite");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t n = (int64_t)g->nlon * g->nlat * g->nivbas;
    HIPCHK(rb::launch_zenith(trajlon, trajlat, n, rb::zenith_terms(day), sza, ctx->stream));
    return RB_OK;
}

int rb_zenith(rb_ctx *ctx, const rb_grid *g, double day, const double *trajlon,
              const double *trajlat, double *sza)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (any_null({trajlon, trajlat, sza})) return fail(ctx, RB_EINVAL, "a host array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)g->nlon * g->nlat * g->nivbas;
    if ((rc = ensure(ctx, &ctx->d_stage, &ctx->stage_cap, 3 * n))) return rc;
    double *dl = ctx->d_stage, *dt = dl + n, *ds = dt + n;
    const size_t B = sizeof(double);
    HIPCHK(hipMemcpyAsync(dl, trajlon, n * B, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dt, trajlat, n * B, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = rb_zenith_dev(ctx, g, day, dl, dt, ds))) return rc;
    HIPCHK(hipMemcpyAsync(sza, ds, n * B, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return RB_OK;
}

static int check_altitude(rb_ctx *ctx, const rb_grid *g, double dlon, double dlat)
{
    if (!g || g->nlon < 1 || g->nlat < 1 || g->niv < 2) return fail(ctx, RB_EINVAL, "bad grid");
    if (!(dlon > 0.) || !(dlat > 0.)) return fail(ctx, RB_EINVAL, "dlon, dlat must be > 0");
    return RB_OK;
}

int rb_altitude_dev(rb_ctx *ctx, const rb_grid *g, double dlon, double dlat, const double *relief,
                    const double *psol, const double *aaa, const double *bbb, const double *t,
                    const double *h2o, const double *trajlon, const double *trajlat,
                    const double *trajpre, double *alt)
{
    TraceRange trace("rb_altitude_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_altitude(ctx, g, dlon, dlat);
    if (rc) return rc;
    if (any_null({relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre, alt}))
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t field = (size_t)g->nlon * g->nlat * g->niv;
    if ((rc = ensure(ctx, &ctx->d_work, &ctx->work_cap, 3 * field))) return rc;
    HIPCHK(rb::launch_altitude(relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre, g->nlon,
                               g->nlat, g->niv, dlon, dlat, ctx->d_work, alt, ctx->stream));
    return RB_OK;
}

int rb_altitude(rb_ctx *ctx, const rb_grid *g, double dlon, double dlat, const double *relief,
                const double *psol, const double *aaa, const double *bbb, const double *t,
                const double *h2o, const double *trajlon, const double *trajlat,
                const double *trajpre, double *alt)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_altitude(ctx, g, dlon, dlat);
    if (rc) return rc;
    if (any_null({relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre, alt}))
        return fail(ctx, RB_EINVAL, "a host array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t plane = (size_t)g->nlon * g->nlat, field = plane * g->niv;
    if ((rc = ensure(ctx, &ctx->d_stage, &ctx->stage_cap, 6 * field + 2 * plane + 2 * g->niv)))
        return rc;
    double *d = ctx->d_stage;
    double *dt_ = d, *dh = d + field, *dlo = d + 2 * field, *dla = d + 3 * field,
           *dpr = d + 4 * field, *dal = d + 5 * field, *dre = d + 6 * field, *dps = dre + plane,
           *da = dps + plane, *db = da + g->niv;
    const size_t B = sizeof(double);
    hipStream_t s = ctx->stream;
    const std::pair<double *, const double *> in[] = {
        {dt_, t}, {dh, h2o}, {dlo, trajlon}, {dla, trajlat}, {dpr, trajpre}};
    for (auto &io : in) HIPCHK(hipMemcpyAsync(io.first, io.second, field * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dre, relief, plane * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dps, psol, plane * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(da, aaa, g->niv * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(db, bbb, g->niv * B, hipMemcpyHostToDevice, s));
    if ((rc = rb_altitude_dev(ctx, g, dlon, dlat, dre, dps, da, db, dt_, dh, dlo, dla, dpr, dal)))
        return rc;
    HIPCHK(hipMemcpyAsync(alt, dal, field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RB_OK;
}

static int check_transport(rb_ctx *ctx, const rb_grid *g)
{
    if (!g) return fail(ctx, RB_EINVAL, "grid is NULL");
    if (g->nlon < 4 || g->nlat < 4 || g->niv < 3)
        return fail(ctx, RB_EINVAL, "transport needs nlon >= 4, nlat >= 4 and niv >= 3");
    if ((int64_t)g->nlon * g->nlat * g->niv >= (int64_t)1 << 31)
        return fail(ctx, RB_EINVAL, "transport grid above 2^31 points");
    return RB_OK;
}

// the transport kernels count departure points outside the grid; the count is read back
static int transport_outside(rb_ctx *ctx, const char *what)
{
    unsigned long long bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, ctx->d_count, sizeof bad, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (bad)
        return fail(ctx, RB_EINVAL, std::string(what) + ": " + std::to_string(bad) +
                                        " departure points outside the grid");
    return RB_OK;
}

int rb_transport_weights_dev(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                             const double *x0, const double *y0, const double *z0,
                             const double *p0, int32_t *idx, double *w)
{
    TraceRange trace("rb_transport_weights_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_transport(ctx, g);
    if (rc) return rc;
    if (any_null({aaa, bbb, x0, y0, z0, p0, w}) || !idx)
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const rb::TransportGrid t = rb::transport_grid(g->nlon, g->nlat, g->niv);
    HIPCHK(rb::launch_zero_counter(ctx->d_count, ctx->stream));
    HIPCHK(rb::launch_transport_weights(t, aaa, bbb, x0, y0, z0, p0, idx, w, ctx->d_count,
                                        ctx->stream));
    return transport_outside(ctx, "rb_transport_weights");
}

int rb_transport_interp_dev(rb_ctx *ctx, const rb_grid *g, const int32_t *idx, const double *w,
                            int32_t ntr, const double *q_in, double *q_out)
{
    TraceRange trace("rb_transport_interp_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_transport(ctx, g);
    if (rc) return rc;
    if (ntr < 1) return fail(ctx, RB_EINVAL, "ntr must be >= 1");
    if (any_null({w, q_in, q_out}) || !idx) return fail(ctx, RB_EINVAL, "a device array is NULL");
    const int64_t field = (int64_t)g->nlon * g->nlat * g->niv;
    if (q_out < q_in + ntr * field && q_in < q_out + ntr * field)
        return fail(ctx, RB_EINVAL, "q_out must not overlap q_in");
    HIPCHK(hipSetDevice(ctx->device));
    const rb::TransportGrid t = rb::transport_grid(g->nlon, g->nlat, g->niv);
    HIPCHK(rb::launch_zero_counter(ctx->d_count, ctx->stream));
    HIPCHK(rb::launch_transport_interp(t, idx, w, q_in, field, q_out, field, ntr, ctx->d_count,
                                       ctx->stream));
    return transport_outside(ctx, "rb_transport_interp");
}

int rb_transport_backtra_dev(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                             const rb_forcing *m1, const rb_forcing *p1, double daynum,
                             double dtdyn, int32_t nrapp, int32_t ntime, int32_t nintecmwf,
                             double *x0, double *y0, double *z0, double *p0, int32_t *idx,
                             double *w, double *trajlon, double *trajlat, double *trajpre,
                             double *trajtem)
{
    TraceRange trace("rb_transport_backtra_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_transport(ctx, g);
    if (rc) return rc;
    if (g->nlat < 5) return fail(ctx, RB_EINVAL, "backtra needs nlat >= 5");
    if (!m1 || !p1) return fail(ctx, RB_EINVAL, "a forcing record is NULL");
    if (any_null({aaa, bbb, x0, y0, z0, p0, w, m1->u, m1->v, m1->w, m1->t, m1->lnps, p1->u, p1->v,
                  p1->w, p1->t, p1->lnps}) || !idx)
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    const bool traj = trajlon || trajlat || trajpre || trajtem;
    if (traj && any_null({trajlon, trajlat, trajpre, trajtem}))
        return fail(ctx, RB_EINVAL, "the trajectory arrays must be all set or all NULL");
    if (nrapp < 1 || ntime < 1 || nintecmwf < 1)
        return fail(ctx, RB_EINVAL, "nrapp, ntime and nintecmwf must be >= 1");
    if (!(dtdyn > 0.) || !std::isfinite(dtdyn) || !std::isfinite(daynum) ||
        !std::isfinite(m1->day) || !std::isfinite(p1->day))
        return fail(ctx, RB_EINVAL, "dtdyn must be positive; days finite");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t n = (int64_t)g->nlon * g->nlat * g->niv, plane = (int64_t)g->nlon * g->nlat;
    const size_t ntab = rb::backtra_table_size(g->nlon, g->nlat);
    // the fields of a sub-step (6 n + plane) and the grid's tables
    if ((rc = ensure(ctx, &ctx->d_bt, &ctx->bt_cap, 6 * (size_t)n + (size_t)plane + ntab)))
        return rc;
    double *dtab = ctx->d_bt + 6 * (size_t)n + (size_t)plane;
    if (ctx->bt_grid[0] != g->nlon || ctx->bt_grid[1] != g->nlat) {
        ctx->bt_tables.assign(ntab, 0.);
        rb::backtra_tables(g->nlon, g->nlat, ctx->bt_tables.data());
        ctx->bt_grid[0] = g->nlon;
        ctx->bt_grid[1] = g->nlat;
    }
    // the tables are rewritten on every call: the buffer may have grown since
    HIPCHK(hipMemcpyAsync(dtab, ctx->bt_tables.data(), ntab * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
    rb::BacktraArgs a{};
    a.t = rb::transport_grid(g->nlon, g->nlat, g->niv);
    a.aaa = aaa;
    a.bbb = bbb;
    a.m1 = {m1->u, m1->v, m1->w, m1->t, m1->lnps};
    a.p1 = {p1->u, p1->v, p1->w, p1->t, p1->lnps};
    a.tb = rb::backtra_table_ptrs(g->nlon, g->nlat, dtab);
    // cstes (F:6010-6025) and dlatlim (F:6101)
    a.xpi = 2. * asin(1.);
    a.xpi2 = 2. * a.xpi;
    a.dlatlim = 45. / (180. / a.xpi);
    a.f = ctx->d_bt;
    a.x0 = x0;
    a.y0 = y0;
    a.z0 = z0;
    a.p0 = p0;
    a.idx = idx;
    a.w = w;
    if (traj) {
        a.traj[0] = trajlon;
        a.traj[1] = trajlat;
        a.traj[2] = trajpre;
        a.traj[3] = trajtem;
    }
    a.nrapp = nrapp;
    a.n = n;
    a.bad = ctx->d_count;
    // dtn, the sub-step (F:6237); inivit's scalings (F:6594-6595)
    const double dtn = dtdyn / (double)nrapp;
    a.coefh = dtn / 6371229.;
    a.coefv = dtn;
    HIPCHK(rb::launch_zero_counter(ctx->d_count, ctx->stream));
    for (int iter = 1; iter <= nrapp; ++iter) {
        // F:6245 and instant1/instant2's time weight (F:6495-6504)
        const double daypre = daynum - dtn * (double)(iter - 1) / 86400.;
        double day = daypre;
        if (day < 1.e-5) day = day + std::trunc(m1->day) + 1.;
        double cp1 = (day - m1->day) / ((double)nintecmwf / (double)(24 * ntime));
        cp1 = std::fmin(cp1, 1.);
        cp1 = std::fmax(cp1, 0.);
        a.cp1 = cp1;
        a.cm1 = 1. - cp1;
        a.iter = iter;
        HIPCHK(rb::launch_backtra_fields(a, ctx->stream));
        HIPCHK(rb::launch_backtra_step(a, ctx->stream));
    }
    return transport_outside(ctx, "rb_transport_backtra");
}

int rb_transport_backtra(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                         const rb_forcing *m1, const rb_forcing *p1, double daynum, double dtdyn,
                         int32_t nrapp, int32_t ntime, int32_t nintecmwf, double *x0, double *y0,
                         double *z0, double *p0, double *trajlon, double *trajlat,
                         double *trajpre, double *trajtem)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_transport(ctx, g);
    if (rc) return rc;
    if (!m1 || !p1) return fail(ctx, RB_EINVAL, "a forcing record is NULL");
    if (any_null({aaa, bbb, x0, y0, z0, p0, trajlon, trajlat, trajpre, trajtem, m1->u, m1->v,
                  m1->w, m1->t, m1->lnps, p1->u, p1->v, p1->w, p1->t, p1->lnps}))
        return fail(ctx, RB_EINVAL, "a host array is NULL");
    if (nrapp < 1) return fail(ctx, RB_EINVAL, "nrapp must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)g->nlon * g->nlat * g->niv, plane = (size_t)g->nlon * g->nlat;
    const size_t B = sizeof(double);
    // forcing 2 (4 n + plane), levels 2 niv, departure points 4 n, indices 3 n int32,
    // weights 12 n, trajectory 4 nrapp n
    const size_t need = 2 * (4 * n + plane) + 2 * (size_t)g->niv + 16 * n + (3 * n + 1) / 2 +
                        4 * (size_t)nrapp * n;
    if ((rc = ensure(ctx, &ctx->d_bth, &ctx->bth_cap, need))) return rc;
    double *q = ctx->d_bth;
    hipStream_t s = ctx->stream;
    rb_forcing dm[2];
    const rb_forcing *hm[2] = {m1, p1};
    for (int r = 0; r < 2; ++r) {
        double *f[5];
        const double *h[5] = {hm[r]->u, hm[r]->v, hm[r]->w, hm[r]->t, hm[r]->lnps};
        for (int k = 0; k < 5; ++k) {
            const size_t len = k < 4 ? n : plane;
            f[k] = q;
            HIPCHK(hipMemcpyAsync(q, h[k], len * B, hipMemcpyHostToDevice, s));
            q += len;
        }
        dm[r] = {f[0], f[1], f[2], f[3], f[4], hm[r]->day};
    }
    double *da = q, *db = q + g->niv;
    q += 2 * g->niv;
    HIPCHK(hipMemcpyAsync(da, aaa, g->niv * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(db, bbb, g->niv * B, hipMemcpyHostToDevice, s));
    double *dx = q, *dy = dx + n, *dz = dy + n, *dp = dz + n, *dw = dp + n;
    int32_t *didx = (int32_t *)(dw + 12 * n);
    double *dt = dw + 12 * n + (3 * n + 1) / 2;
    double *dtr[4] = {dt, dt + nrapp * n, dt + 2 * nrapp * n, dt + 3 * nrapp * n};
    if ((rc = rb_transport_backtra_dev(ctx, g, da, db, &dm[0], &dm[1], daynum, dtdyn, nrapp,
                                       ntime, nintecmwf, dx, dy, dz, dp, didx, dw, dtr[0], dtr[1],
                                       dtr[2], dtr[3])))
        return rc;
    const std::pair<double *, const double *> out[] = {{x0, dx}, {y0, dy}, {z0, dz}, {p0, dp}};
    for (auto &o : out) HIPCHK(hipMemcpyAsync(o.first, o.second, n * B, hipMemcpyDeviceToHost, s));
    double *htr[4] = {trajlon, trajlat, trajpre, trajtem};
    for (int k = 0; k < 4; ++k)
        HIPCHK(hipMemcpyAsync(htr[k], dtr[k], nrapp * n * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RB_OK;
}

int rb_transport(rb_ctx *ctx, const rb_grid *g, const double *aaa, const double *bbb,
                 const double *x0, const double *y0, const double *z0, const double *p0,
                 int32_t ntr, double *qj1)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_transport(ctx, g);
    if (rc) return rc;
    if (ntr < 1) return fail(ctx, RB_EINVAL, "ntr must be >= 1");
    if (any_null({aaa, bbb, x0, y0, z0, p0, qj1})) return fail(ctx, RB_EINVAL, "a host array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)g->nlon * g->nlat * g->niv, B = sizeof(double);
    // indices (3 n int32), weights 12 n, positions 4 n, levels 2 niv
    if ((rc = ensure(ctx, &ctx->d_tw, &ctx->tw_cap, (3 * n + 1) / 2 + 16 * n + 2 * g->niv))) return rc;
    if ((rc = ensure(ctx, &ctx->d_tq, &ctx->tq_cap, 2 * (size_t)ntr * n))) return rc;
    int32_t *didx = (int32_t *)ctx->d_tw;
    double *dw = ctx->d_tw + (3 * n + 1) / 2, *dx = dw + 12 * n, *dy = dx + n, *dz = dy + n,
           *dp = dz + n, *da = dp + n, *db = da + g->niv;
    double *qin = ctx->d_tq, *qout = ctx->d_tq + (size_t)ntr * n;
    hipStream_t s = ctx->stream;
    const std::pair<double *, const double *> in[] = {{dx, x0}, {dy, y0}, {dz, z0}, {dp, p0}};
    for (auto &io : in) HIPCHK(hipMemcpyAsync(io.first, io.second, n * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(da, aaa, g->niv * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(db, bbb, g->niv * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(qin, qj1, (size_t)ntr * n * B, hipMemcpyHostToDevice, s));
    if ((rc = rb_transport_weights_dev(ctx, g, da, db, dx, dy, dz, dp, didx, dw))) return rc;
    if ((rc = rb_transport_interp_dev(ctx, g, didx, dw, ntr, qin, qout))) return rc;
    HIPCHK(hipMemcpyAsync(qj1, qout, (size_t)ntr * n * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RB_OK;
}

static int check_sediment(rb_ctx *ctx, const rb_grid *g, double dt, int32_t nivh2o)
{
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (g->nivhau < 2) return fail(ctx, RB_EINVAL, "sedimentation needs nivhau >= 2");
    if (nivh2o > g->nivbas + 1) return fail(ctx, RB_EINVAL, "nivh2o must be <= nivbas + 1");
    if (!(dt >= 0.) || !std::isfinite(dt)) return fail(ctx, RB_EINVAL, "dt must be >= 0");
    return RB_OK;
}

int rb_sediment_dev(rb_ctx *ctx, const rb_grid *g, double dt, int32_t nivh2o, const double *T,
                    const double *p, double *qj1, const double *vsed3d, const double *parth2osed,
                    const double *parthno3sed)
{
    TraceRange trace("rb_sediment_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_sediment(ctx, g, dt, nivh2o);
    if (rc) return rc;
    if (any_null({T, p, qj1, vsed3d, parth2osed, parthno3sed}))
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t plane = (int64_t)g->nlon * g->nlat;
    const int64_t field = plane * g->niv;
    HIPCHK(rb::launch_sediment(p, T, qj1 + rb::H2O * field, qj1 + rb::NOY * field,
                               qj1 + rb::HNO3 * field, vsed3d, parth2osed, parthno3sed, plane,
                               g->nivbas, g->nivhau, nivh2o, dt, ctx->stream));
    return RB_OK;
}

int rb_sediment(rb_ctx *ctx, const rb_grid *g, double dt, int32_t nivh2o, const double *T,
                const double *p, double *qj1, const double *vsed3d, const double *parth2osed,
                const double *parthno3sed)
{
    if (!ctx) return RB_EINVAL;
    int rc = check_sediment(ctx, g, dt, nivh2o);
    if (rc) return rc;
    if (any_null({T, p, qj1, vsed3d, parth2osed, parthno3sed}))
        return fail(ctx, RB_EINVAL, "a host array is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t plane = (size_t)g->nlon * g->nlat, field = plane * g->niv;
    const size_t nb = plane * g->nivbas;
    // staging: H2O, NOy, HNO3 fields (slots 3-5, contiguous), T, p, vsed, the partitions
    if ((rc = ensure(ctx, &ctx->d_stage, &ctx->stage_cap, 3 * field + 5 * nb))) return rc;
    double *dq = ctx->d_stage, *dT = dq + 3 * field, *dp = dT + nb, *dv = dp + nb,
           *d2 = dv + nb, *d3 = d2 + nb;
    const size_t B = sizeof(double);
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dq, qj1 + 2 * field, 3 * field * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dT, T, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dp, p, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dv, vsed3d, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d2, parth2osed, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d3, parthno3sed, nb * B, hipMemcpyHostToDevice, s));
    HIPCHK(rb::launch_sediment(dp, dT, dq, dq + field, dq + 2 * field, dv, d2, d3, (int64_t)plane,
                               g->nivbas, g->nivhau, nivh2o, dt, s));
    HIPCHK(hipMemcpyAsync(qj1 + 2 * field, dq, 3 * field * B, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return RB_OK;
}

// one sub-step of a day: SZA of the step's time, then the chemistry step
static int day_step(rb_ctx *ctx, const rb_grid *g, const rb_day_config *dc, int k,
                    const double *trajlon, const double *trajlat, const double *T,
                    const double *p, const double *alt, double *sza, double *qj1, double *hc,
                    double *surfarea, double *vsed3d, double *p2o, double *pno3, int32_t *iters)
{
    int rc = rb_zenith_dev(ctx, g, dc->day0 + k * dc->dt / 86400., trajlon, trajlat, sza);
    if (rc) return rc;
    if (dc->converged)
        return rb_chem_converge_dev(ctx, g, dc->dt, dc->tol, dc->maxiter, dc->accel, T, p, sza, alt,
                                    qj1, hc, surfarea, vsed3d, p2o, pno3, nullptr, nullptr, iters, 1);
    return rb_chem_step_dev(ctx, g, dc->dt, T, p, sza, alt, qj1, hc, surfarea, vsed3d, p2o, pno3, 1);
}

int rb_run_day_dev(rb_ctx *ctx, const rb_grid *g, const rb_day_config *dc, const double *trajlon,
                   const double *trajlat, const double *T, const double *p, const double *alt,
                   double *sza, double *qj1, double *hc, double *surfarea, double *vsed3d,
                   double *parth2osed, double *parthno3sed, int32_t *iters)
{
    TraceRange trace("rb_run_day_dev");
    if (!ctx) return RB_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!dc || dc->nsteps < 1 || !(dc->dt > 0.) || !std::isfinite(dc->day0))
        return fail(ctx, RB_EINVAL, "bad day configuration");
    if (dc->converged && (!(dc->tol > 0.) || dc->maxiter < 1))
        return fail(ctx, RB_EINVAL, "converged day needs tol > 0 and maxiter >= 1");
    if (any_null({trajlon, trajlat, T, p, alt, sza, qj1, hc, surfarea, vsed3d, parth2osed,
                  parthno3sed}))
        return fail(ctx, RB_EINVAL, "a device array is NULL");
    if (!ctx->table) return fail(ctx, RB_ENOTABLE, "rb_load_jtable has not been called");
    HIPCHK(hipSetDevice(ctx->device));
    auto steps = [&]() -> int {
        for (int k = 0; k < dc->nsteps; ++k) {
            int r = day_step(ctx, g, dc, k, trajlon, trajlat, T, p, alt, sza, qj1, hc, surfarea,
                             vsed3d, parth2osed, parthno3sed, iters);
            if (r) return r;
        }
        return RB_OK;
    };
    if (!dc->use_graph) return steps();

    // scratch the steps use must exist before capture (no allocation inside it)
    const size_t nbox = (size_t)g->nlon * g->nlat * g->nivbas;
    if ((rc = ensure(ctx, &ctx->d_col, &ctx->col_cap, 3 * nbox))) return rc;
    if ((rc = conv_buffers(ctx, nbox, dc->converged ? dc->accel : 0))) return rc;
    // the graph is rebuilt when anything it was captured with changes
    std::vector<char> key;
    auto add = [&](const void *v, size_t n) {
        const char *c = (const char *)v;
        key.insert(key.end(), c, c + n);
    };
    const void *ptrs[] = {trajlon, trajlat, T, p, alt, sza, qj1, hc, surfarea, vsed3d,
                          parth2osed, parthno3sed, iters, ctx->stream, ctx->d_col, ctx->d_aa,
                          ctx->d_rec, ctx->d_spill, ctx->d_order};
    add(g, sizeof *g);
    add(dc, sizeof *dc);
    add(ptrs, sizeof ptrs);
    add(&ctx->cfg, sizeof ctx->cfg);
    if (!ctx->day_exec || key != ctx->day_key) {
        if (ctx->day_exec) (void)hipGraphExecDestroy(ctx->day_exec);
        if (ctx->day_graph) (void)hipGraphDestroy(ctx->day_graph);
        ctx->day_exec = nullptr;
        ctx->day_graph = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        ctx->capturing = true;
        rc = steps();
        ctx->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
        ctx->day_graph = graph;
        HIPCHK(hipGraphInstantiate(&ctx->day_exec, graph, nullptr, nullptr, 0));
        ctx->day_key = key;
    }
    HIPCHK(hipGraphLaunch(ctx->day_exec, ctx->stream));
    return RB_OK;
}

int rb_sync(rb_ctx *ctx)
{
    if (!ctx) return RB_EINVAL;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return RB_OK;
}

int rb_get_stats(rb_ctx *ctx, rb_stats_t *out)
{
    if (!ctx || !out) return RB_EINVAL;
    if (ctx->timed) {
        HIPCHK(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->stats.last_kernel_ms = ms;
    }
    double total = 0.;
    for (size_t i = 0; i < ctx->prof_used; ++i) {
        float ms = 0.f;
        HIPCHK(hipEventSynchronize(ctx->prof[i].second));
        HIPCHK(hipEventElapsedTime(&ms, ctx->prof[i].first, ctx->prof[i].second));
        total += ms;
    }
    ctx->stats.chem_ms_total = total;
    ctx->stats.chem_launches = (int64_t)ctx->prof_used;
    ctx->prof_used = 0;
    *out = ctx->stats;
    return RB_OK;
}

const char *rb_last_error(rb_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rb_finalize(rb_ctx *ctx)
{
    if (!ctx) return RB_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->own) (void)hipStreamSynchronize(ctx->own);
    for (double *b : {ctx->d_ajl, ctx->d_o3up, ctx->d_col, ctx->d_stage})
        if (b) (void)hipFree(b);
    if (ctx->d_count) (void)hipFree(ctx->d_count);
    if (ctx->d_queue) (void)hipFree(ctx->d_queue);
    if (ctx->d_work) (void)hipFree(ctx->d_work);
    if (ctx->d_aa) (void)hipFree(ctx->d_aa);
    if (ctx->d_rec) (void)hipFree(ctx->d_rec);
    if (ctx->d_spill) (void)hipFree(ctx->d_spill);
    if (ctx->d_order) (void)hipFree(ctx->d_order);
    if (ctx->d_tw) (void)hipFree(ctx->d_tw);
    if (ctx->d_tq) (void)hipFree(ctx->d_tq);
    if (ctx->d_bt) (void)hipFree(ctx->d_bt);
    if (ctx->d_bth) (void)hipFree(ctx->d_bth);
    if (ctx->day_exec) (void)hipGraphExecDestroy(ctx->day_exec);
    if (ctx->day_graph) (void)hipGraphDestroy(ctx->day_graph);
    for (hipEvent_t e : {ctx->ev0, ctx->ev1, ctx->ev_h2d, ctx->ev_d2h})
        if (e) (void)hipEventDestroy(e);
    for (auto &pe : ctx->prof) {
        (void)hipEventDestroy(pe.first);
        (void)hipEventDestroy(pe.second);
    }
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return RB_OK;
}

}  // extern "C"
