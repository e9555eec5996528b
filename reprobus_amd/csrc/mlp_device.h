// Device-side declarations of the learned-guess MLP (mlp_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rbm {

constexpr int MLP_NIN = 60;      // features
constexpr int MLP_NOUT = 55;     // log increments: 40 integrated long-lived + 15 short-lived
constexpr int MLP_PAD = 64;      // input and output padded to one 64-column tile
constexpr int MLP_W = 256;       // hidden width
constexpr int MLP_NLONG = 40;
constexpr int MLP_NSHORT = 15;
// state slots (0-based) that the network neither reads as species nor predicts
constexpr int SLOT_OX_PASSIVE = 10, SLOT_H2SO4 = 11, SLOT_HNO3G = 42;

// the s-th integrated long-lived species: slots 0..9, then 12..41 (passive Ox and H2SO4
// have no production or loss)
__host__ __device__ constexpr int long_slot(int s) { return s < 10 ? s : s + 2; }

// parameter block (FP32 master copy, this order): W1 [64][256], b1 [256], W2 [256][256],
// b2 [256], W3 [256][256], b3 [256], W4 [256][64], b4 [64]
constexpr int64_t OFF_W1 = 0;
constexpr int64_t OFF_B1 = OFF_W1 + (int64_t)MLP_PAD * MLP_W;
constexpr int64_t OFF_W2 = OFF_B1 + MLP_W;
constexpr int64_t OFF_B2 = OFF_W2 + (int64_t)MLP_W * MLP_W;
constexpr int64_t OFF_W3 = OFF_B2 + MLP_W;
constexpr int64_t OFF_B3 = OFF_W3 + (int64_t)MLP_W * MLP_W;
constexpr int64_t OFF_W4 = OFF_B3 + MLP_W;
constexpr int64_t OFF_B4 = OFF_W4 + (int64_t)MLP_W * MLP_PAD;
constexpr int64_t NPARAM = OFF_B4 + MLP_PAD;
// BF16 weight copies, concatenated in the same order without the biases
constexpr int64_t BOFF_W1 = 0;
constexpr int64_t BOFF_W2 = BOFF_W1 + (int64_t)MLP_PAD * MLP_W;
constexpr int64_t BOFF_W3 = BOFF_W2 + (int64_t)MLP_W * MLP_W;
constexpr int64_t BOFF_W4 = BOFF_W3 + (int64_t)MLP_W * MLP_W;
constexpr int64_t NWEIGHT = BOFF_W4 + (int64_t)MLP_W * MLP_PAD;

// per-lane weight fragments of the fused guess kernel (k_pack_frags), BF16 count
constexpr int64_t NFRAG = NWEIGHT;

struct GuessArgs {
    const double *T, *p, *sza, *qj1, *hc;   // state and inputs (model layout)
    int64_t field, nbox;
    float log_dt;                            // log of the step (a feature)
    const __bf16 *frag;                      // packed weights (launch_pack_frags)
    const float *bias;                       // the FP32 parameter block (biases at OFF_B*)
    const float *norm;                       // in_mu, in_rsd, out_mu, out_sd (64 each)
    double *gq, *gh;                         // the guess (model layout)
};

enum Epilogue { EPI_F32 = 0, EPI_HIDDEN_FWD = 1, EPI_OUT_FWD = 2, EPI_HIDDEN_BWD = 3 };

struct GemmArgs {
    int64_t M;
    int N, K;
    const __bf16 *A;
    int64_t sam, sak;
    const __bf16 *B;
    int64_t sbk, sbn;
    int mode;
    const float *bias;   // EPI_HIDDEN_FWD, EPI_OUT_FWD
    float *Cf;           // FP32 output (pre-activation, gradient), may be null for EPI_HIDDEN_FWD
    __bf16 *Cb;          // BF16 output (activation or gradient)
    const float *Z;      // EPI_HIDDEN_BWD: pre-activation of the layer
    int64_t ldc;
    int n_valid;         // EPI_OUT_FWD: columns >= n_valid are written as 0
    int ksplit;          // EPI_F32 only: K split over this many workgroups (>= 1) ...
    float *part;         // ... each writing an M x N partial here, summed into Cf
};

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);
hipError_t launch_pack_inputs(const float *X, int64_t n, const float *mu, const float *rsd,
                              __bf16 *A0, hipStream_t s);
hipError_t launch_unpack_outputs(const float *Yn, int64_t n, const float *mu, const float *sd,
                                 float *Y, hipStream_t s);
hipError_t launch_loss_grad(const float *Yn, const float *T, int64_t n, const float *mu,
                            const float *rsd, __bf16 *dY, float *dYf, float *partial, float *loss,
                            hipStream_t s);
hipError_t launch_colsum(const float *D, int64_t n, int ncol, float *part, float *out, hipStream_t s);
constexpr int COLSUM_CHUNKS = 64;   // row chunks of launch_colsum (part: COLSUM_CHUNKS * ncol)
constexpr int MAX_KSPLIT = 32;
hipError_t launch_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, const int32_t *wbf, __bf16 *pb,
                       hipStream_t s);
hipError_t launch_to_bf16(const float *p, const int32_t *wbf, int64_t n, __bf16 *pb, hipStream_t s);
hipError_t launch_features(const double *T, const double *p, const double *sza, const double *qj1,
                           const double *hc, int64_t field, int64_t nbox, double dt, float *X,
                           hipStream_t s);
hipError_t launch_apply_guess(const double *qj1, const double *hc, const float *Y, int64_t field,
                              int64_t nbox, double *gq, double *gh, hipStream_t s);

hipError_t launch_pack_frags(const __bf16 *pb, __bf16 *frag, hipStream_t s);
// features, the network and the guess formula in one launch (as launch_features,
// launch_pack_inputs, the four GEMMs, launch_unpack_outputs and launch_apply_guess)
hipError_t launch_guess_fused(const GuessArgs &a, hipStream_t s);

}  // namespace rbm
