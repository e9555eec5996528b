/* Test infrastructure: elementary functions evaluated in binary128 (libquadmath)
 * and rounded to double -- the correctly rounded value except for double-rounding
 * hard cases.  Checker for the product's double-double math. */
#include <quadmath.h>

double qr_exp(double x) { return (double)expq((__float128)x); }
double qr_log(double x) { return (double)logq((__float128)x); }
double qr_log10(double x) { return (double)log10q((__float128)x); }
double qr_pow(double x, double y) { return (double)powq((__float128)x, (__float128)y); }
double qr_atan(double x) { return (double)atanq((__float128)x); }
double qr_cos(double x) { return (double)cosq((__float128)x); }
double qr_sin(double x) { return (double)sinq((__float128)x); }
double qr_asin(double x) { return (double)asinq((__float128)x); }
double qr_acos(double x) { return (double)acosq((__float128)x); }
double qr_tanh(double x) { return (double)tanhq((__float128)x); }
