/* C ABI of the learned initial guess for the converged chemistry step.
 *
 * No counterpart in the reference (whose step always starts its sweeps from the incoming
 * state, F:820-833).  A 60 -> 256 -> 256 -> 256 -> 55 MLP (GELU) maps a box's state to the
 * log increments of its 55 species over one step; its prediction is handed to
 * rb_chem_converge_dev as guess_qj1/guess_hc, and the FP64 solver still decides the answer.
 * Its GEMMs run in BF16 on the matrix cores with FP32 accumulation; parameters, Adam
 * moments and gradients are FP32.
 *
 * Features (60, FP32, raw): log(max(x, 1e-30)) of the 40 integrated long-lived species
 * (qj1 slots 1..10 and 13..42, 1-based) and of the 15 short-lived species (hc), then T [K],
 * log p [hPa], cos SZA, log H2SO4 mixing ratio, log dt [s].  Outputs (55, FP32, raw):
 * log(max(x_out, 1e-30) / max(x_in, 1e-30)) for the same 55 species.  Normalisation
 * (set by rb_mlp_set_norm) is applied inside.
 *
 * Parameters (rb_mlp_param_count() FP32 values, this order): W1 [64][256], b1 [256],
 * W2 [256][256], b2 [256], W3 [256][256], b3 [256], W4 [256][64], b4 [64]; row-major,
 * W_l[in][out].  Rows 60..63 of W1 and columns 55..63 of W4 (padding) must be zero.
 *
 * All arrays marked "device" are device pointers; calls are ordered on the stream set by
 * rb_mlp_set_stream (default: the null stream).  Return codes are those of rbchem.h.
 */
#ifndef RBMLP_H
#define RBMLP_H
#include <stddef.h>
#include <stdint.h>

#include "rbchem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rb_mlp rb_mlp;

int rb_mlp_create(rb_mlp **out, int device);
void rb_mlp_destroy(rb_mlp *m);
int rb_mlp_set_stream(rb_mlp *m, void *hip_stream);
const char *rb_mlp_last_error(const rb_mlp *m);
int64_t rb_mlp_param_count(void);

/* host arrays; set_params also clears the Adam moments and step count */
int rb_mlp_set_params(rb_mlp *m, const float *params);
int rb_mlp_get_params(rb_mlp *m, float *params);
int rb_mlp_set_norm(rb_mlp *m, const float *in_mu60, const float *in_sd60,
                    const float *out_mu55, const float *out_sd55);

/* X: device n x 60 raw features; Y: device n x 55 raw outputs */
int rb_mlp_forward_dev(rb_mlp *m, const float *X, int64_t n, float *Y);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on the mean-squared error of the normalised
 * outputs against T (device n x 55 raw targets); *loss (host, optional) receives the
 * loss before the step (this synchronises the stream) */
int rb_mlp_train_step_dev(rb_mlp *m, const float *X, const float *T, int64_t n, float lr,
                          float *loss);
/* gradient of that loss with respect to all parameters (host, param_count values),
 * without updating; *loss (host, optional) */
int rb_mlp_gradients_dev(rb_mlp *m, const float *X, const float *T, int64_t n, float *grads,
                         float *loss);

/* features of every box on levels 1..nivbas of a state (device arrays as rb_chem_step_dev);
 * X: device nbox x 60 */
int rb_mlp_features_dev(rb_mlp *m, const rb_grid *g, double dt, const double *T, const double *p,
                        const double *sza, const double *qj1, const double *hc, float *X);
/* the initial iterate for rb_chem_converge_dev: features, forward, and
 * guess = max(x, 1e-30) * exp(y) for the 55 species (other qj1 slots copied); guess_qj1
 * and guess_hc are device arrays shaped like qj1 and hc (levels 1..nivbas written) */
int rb_mlp_guess_dev(rb_mlp *m, const rb_grid *g, double dt, const double *T, const double *p,
                     const double *sza, const double *qj1, const double *hc, double *guess_qj1,
                     double *guess_hc);

#ifdef __cplusplus
}
#endif
#endif
