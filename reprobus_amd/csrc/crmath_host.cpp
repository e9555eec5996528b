// Host build of crmath.h for the CPU tests: the same source and IEEE operations
// as the device build, exported with C linkage.
#include "crmath.h"

extern "C" {
double rb_cr_exp(double x) { return crm::exp_cr(x); }
double rb_cr_log(double x) { return crm::log_cr(x); }
double rb_cr_log10(double x) { return crm::log10_cr(x); }
double rb_fdlibm_log10(double x) { return crm::log10_fdlibm(x); }
double rb_cr_pow(double x, double y) { return crm::pow_cr(x, y); }
double rb_cr_pow10(double y) { return crm::pow10_cr(y); }
double rb_cr_atan(double x) { return crm::atan_cr(x); }
double rb_cr_cos(double x) { return crm::cos_cr(x); }
double rb_cr_sin(double x) { return crm::sin_cr(x); }
double rb_cr_asin(double x) { return crm::asin_cr(x); }
double rb_cr_acos(double x) { return crm::acos_cr(x); }
void rb_cr_sincos(double x, double *s, double *c) { crm::sincos_cr(x, *s, *c); }
double rb_cr_tanh(double x) { return crm::tanh_cr(x); }
}
