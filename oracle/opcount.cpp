// Test infrastructure: exact dynamic FP64 operation count of the chemistry step.
//
// Compiles the C restatement (rbchem_oracle.c) with `double` replaced by a type
// that counts every arithmetic operation and libm call, so the per-box work of
// the algorithm (not of any particular instruction selection) is known exactly.
// Used by the benchmark's roofline figure.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace opc {
struct Counts {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0, exp = 0, log = 0, pow = 0,
             trig = 0, minmax = 0;
};
static thread_local Counts C;

struct cd {
    double v;
    cd() : v(0.) {}
    cd(double x) : v(x) {}
    cd(int x) : v((double)x) {}
    explicit operator int() const { return (int)v; }
    explicit operator double() const { return v; }
    cd &operator+=(cd o) { ++C.add; v += o.v; return *this; }
    cd &operator-=(cd o) { ++C.add; v -= o.v; return *this; }
    cd &operator*=(cd o) { ++C.mul; v *= o.v; return *this; }
    cd &operator/=(cd o) { ++C.div; v /= o.v; return *this; }
};
inline cd operator+(cd a, cd b) { ++C.add; return cd(a.v + b.v); }
inline cd operator-(cd a, cd b) { ++C.add; return cd(a.v - b.v); }
inline cd operator*(cd a, cd b) { ++C.mul; return cd(a.v * b.v); }
inline cd operator/(cd a, cd b) { ++C.div; return cd(a.v / b.v); }
inline cd operator-(cd a) { return cd(-a.v); }
#define CMP(op) inline bool operator op(cd a, cd b) { ++C.cmp; return a.v op b.v; }
CMP(<) CMP(<=) CMP(>) CMP(>=) CMP(==) CMP(!=)
#undef CMP
inline cd exp(cd a) { ++C.exp; return cd(std::exp(a.v)); }
inline cd log(cd a) { ++C.log; return cd(std::log(a.v)); }
inline cd log10(cd a) { ++C.log; return cd(std::log10(a.v)); }
inline cd pow(cd a, cd b) { ++C.pow; return cd(std::pow(a.v, b.v)); }
inline cd sqrt(cd a) { ++C.sqrt; return cd(std::sqrt(a.v)); }
inline cd tanh(cd a) { ++C.trig; return cd(std::tanh(a.v)); }
inline cd atan(cd a) { ++C.trig; return cd(std::atan(a.v)); }
inline cd cos(cd a) { ++C.trig; return cd(std::cos(a.v)); }
inline cd asin(cd a) { ++C.trig; return cd(std::asin(a.v)); }
inline cd trunc(cd a) { return cd(std::trunc(a.v)); }
inline cd fabs(cd a) { return cd(std::fabs(a.v)); }
inline cd fmin(cd a, cd b) { ++C.minmax; return cd(std::fmin(a.v, b.v)); }
inline cd fmax(cd a, cd b) { ++C.minmax; return cd(std::fmax(a.v, b.v)); }
}  // namespace opc

using opc::cd;
using opc::exp; using opc::log; using opc::log10; using opc::pow; using opc::sqrt;
using opc::tanh; using opc::atan; using opc::cos; using opc::asin; using opc::trunc;
using opc::fmin; using opc::fmax; using opc::fabs;

// system headers first, so the macro below never reaches them
#include <math.h>
#include <stddef.h>
#include <string.h>

// The restatement, with every double counted.  Its public entry points take
// double* arrays; they are renamed so the counting build can wrap them.
#define double cd
#define RBO_OPCOUNT 1   /* the chemistry step only */
#define rbo_chem_step rbo_chem_step_counted_impl
#define rbo_conserv_floor rbo_conserv_floor_counted_impl
#define rbo_troe_constants opc_troe_constants
#define rbo_troe opc_troe
#define rbo_rates opc_rates
#define rbo_rate_name opc_rate_name
#define rbo_phot_box opc_phot_box
#define rbo_column opc_column
#define rbo_jno_box opc_jno_box
#define rbo_density opc_density
#define rbo_aerosol_box opc_aerosol_box
#undef _OPENMP
#include "rbchem_oracle.c"
#undef double

extern "C" {
// Counts for one chemistry step over the grid (single thread).  out[10] =
// add/sub, mul, div, sqrt, compare, exp, log/log10, pow, trig/tanh, min/max.
int rbo_count_ops(const rbo_grid *g, double dt, const double *T, const double *p,
                  const double *sza, const double *alt, const double *ajl, const double *o3up,
                  double *qj1, double *hc, double *surfarea, double *vsed3d, double *parth2osed,
                  double *parthno3sed, uint64_t out[10])
{
    static_assert(sizeof(cd) == sizeof(double), "layout");
    opc::C = opc::Counts{};
    int rc = rbo_chem_step_counted_impl(
        g, cd(dt), (const cd *)T, (const cd *)p, (const cd *)sza, (const cd *)alt,
        (const cd *)ajl, (const cd *)o3up, (cd *)qj1, (cd *)hc, (cd *)surfarea, (cd *)vsed3d,
        (cd *)parth2osed, (cd *)parthno3sed, 1);
    const opc::Counts &c = opc::C;
    uint64_t v[10] = {c.add, c.mul, c.div, c.sqrt, c.cmp, c.exp, c.log, c.pow, c.trig, c.minmax};
    std::memcpy(out, v, sizeof v);
    return rc;
}
}
