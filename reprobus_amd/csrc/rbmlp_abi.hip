// C ABI of the learned initial guess (include/rbmlp.h) over the kernels of mlp_kernels.hip.
#include "../../include/rbmlp.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mlp_device.h"

using namespace rbm;

struct rb_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // parameters, gradients, Adam moments (FP32) and BF16 weight copies
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    __bf16 *pb = nullptr;
    int32_t *wbf = nullptr;      // parameter index -> BF16 slot, -1 for biases
    float *norm = nullptr;       // in_mu[64] in_rsd[64] out_mu[64] out_sd[64] out_rsd[64]
    int64_t adam_t = 0;
    // work buffers, grown on demand (rows = batch or boxes)
    int64_t rows = 0;
    __bf16 *a0 = nullptr, *act[3] = {nullptr, nullptr, nullptr}, *db[2] = {nullptr, nullptr};
    float *z[3] = {nullptr, nullptr, nullptr}, *yn = nullptr, *df[2] = {nullptr, nullptr};
    __bf16 *dyb = nullptr;
    float *dyf = nullptr, *partial = nullptr, *loss = nullptr, *yraw = nullptr, *xf = nullptr;
    float *ws = nullptr;         // split-K and column-sum partials
    __bf16 *frag = nullptr;      // the BF16 weights as per-lane fragments of the fused guess
    bool train_bufs = false;
};

namespace {

int fail(rb_mlp *m, int code, const std::string &msg)
{
    if (m) m->err = msg;
    return code;
}

int hip_fail(rb_mlp *m, hipError_t e, const char *what)
{
    (void)hipGetLastError();   // reported here: not again by the next call's launch check
    return fail(m, e == hipErrorOutOfMemory ? RB_ENOMEM : RB_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return hip_fail(m, _e, #call); \
    } while (0)

template <class T>
hipError_t realloc_dev(T **ptr, size_t n)
{
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
    const hipError_t e = hipMalloc((void **)ptr, n * sizeof(T));
    if (e != hipSuccess) (void)hipGetLastError();   // not reported again by the next launch
    return e;
}

// grows the activation buffers to `rows` rows (inference needs no pre-activations or
// gradients; training needs all)
int ensure_rows(rb_mlp *m, int64_t rows, bool train)
{
    if (rows <= m->rows && (!train || m->train_bufs)) return RB_OK;
    const int64_t r = std::max(rows, m->rows);
    const size_t w = (size_t)r * MLP_W, pd = (size_t)r * MLP_PAD;
    HIPCHK(realloc_dev(&m->a0, pd));
    for (int i = 0; i < 3; ++i) HIPCHK(realloc_dev(&m->act[i], w));
    HIPCHK(realloc_dev(&m->yn, pd));
    HIPCHK(realloc_dev(&m->yraw, (size_t)r * MLP_NOUT));
    HIPCHK(realloc_dev(&m->xf, (size_t)r * MLP_NIN));
    if (train || m->train_bufs) {
        for (int i = 0; i < 3; ++i) HIPCHK(realloc_dev(&m->z[i], w));
        for (int i = 0; i < 2; ++i) {
            HIPCHK(realloc_dev(&m->db[i], w));
            HIPCHK(realloc_dev(&m->df[i], w));
        }
        HIPCHK(realloc_dev(&m->dyb, pd));
        HIPCHK(realloc_dev(&m->dyf, pd));
        HIPCHK(realloc_dev(&m->partial, pd / 256 + 1));
        m->train_bufs = true;
    }
    m->rows = r;
    return RB_OK;
}

GemmArgs gemm(int64_t M, int N, int K, const __bf16 *A, int64_t sam, int64_t sak, const __bf16 *B,
              int64_t sbk, int64_t sbn, int mode, int64_t ldc)
{
    GemmArgs g{};
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbn = sbn;
    g.mode = mode; g.ldc = ldc;
    g.n_valid = N;
    g.ksplit = 1;
    return g;
}

// normalised inputs in a0 -> normalised outputs in yn; keeps pre-activations if z[0] set
int forward(rb_mlp *m, int64_t n, bool keep)
{
    const __bf16 *W[4] = {m->pb + BOFF_W1, m->pb + BOFF_W2, m->pb + BOFF_W3, m->pb + BOFF_W4};
    const float *b[4] = {m->p + OFF_B1, m->p + OFF_B2, m->p + OFF_B3, m->p + OFF_B4};
    const __bf16 *in = m->a0;
    int kin = MLP_PAD;
    for (int l = 0; l < 3; ++l) {
        GemmArgs g = gemm(n, MLP_W, kin, in, kin, 1, W[l], MLP_W, 1, EPI_HIDDEN_FWD, MLP_W);
        g.bias = b[l];
        g.Cf = keep ? m->z[l] : nullptr;
        g.Cb = m->act[l];
        HIPCHK(launch_gemm(g, m->stream));
        in = m->act[l];
        kin = MLP_W;
    }
    GemmArgs g = gemm(n, MLP_PAD, MLP_W, in, MLP_W, 1, W[3], MLP_PAD, 1, EPI_OUT_FWD, MLP_PAD);
    g.bias = b[3];
    g.Cf = m->yn;
    g.n_valid = MLP_NOUT;
    HIPCHK(launch_gemm(g, m->stream));
    return RB_OK;
}

// forward + loss + backward into m->g
int backward(rb_mlp *m, const float *X, const float *T, int64_t n)
{
    int rc = ensure_rows(m, n, true);
    if (rc) return rc;
    const float *nm = m->norm;
    HIPCHK(launch_pack_inputs(X, n, nm, nm + 64, m->a0, m->stream));
    if ((rc = forward(m, n, true))) return rc;
    HIPCHK(launch_loss_grad(m->yn, T, n, nm + 128, nm + 256, m->dyb, m->dyf, m->partial, m->loss,
                            m->stream));
    const __bf16 *W[4] = {m->pb + BOFF_W1, m->pb + BOFF_W2, m->pb + BOFF_W3, m->pb + BOFF_W4};
    float *gW[4] = {m->g + OFF_W1, m->g + OFF_W2, m->g + OFF_W3, m->g + OFF_W4};
    float *gb[4] = {m->g + OFF_B1, m->g + OFF_B2, m->g + OFF_B3, m->g + OFF_B4};
    const __bf16 *acts[4] = {m->a0, m->act[0], m->act[1], m->act[2]};
    const int width[5] = {MLP_PAD, MLP_W, MLP_W, MLP_W, MLP_PAD};
    // layer 4: dW4 = a3^T dY, db4 = sum dY
    const __bf16 *dzb = m->dyb;
    const float *dzf = m->dyf;
    for (int l = 3; l >= 0; --l) {
        const int kin = width[l], nout = width[l + 1];
        // dW_l [kin x nout] = act_l^T [kin x n] * dz [n x nout]
        // (a kin x nout output is at most 16 tiles: split the batch over workgroups)
        GemmArgs gw = gemm(kin, nout, (int)n, acts[l], 1, kin, dzb, nout, 1, EPI_F32, nout);
        gw.Cf = gW[l];
        gw.ksplit = (int)std::min<int64_t>(MAX_KSPLIT, std::max<int64_t>(1, n / 256));
        gw.part = m->ws;
        HIPCHK(launch_gemm(gw, m->stream));
        HIPCHK(launch_colsum(dzf, n, nout, m->ws + (size_t)MAX_KSPLIT * MLP_W * MLP_W / 2, gb[l],
                             m->stream));
        if (l == 0) break;
        // d act_{l} [n x kin] = dz [n x nout] * W_l^T, times GELU'(z_{l-1})
        const int slot = l & 1;
        GemmArgs gd = gemm(n, kin, nout, dzb, nout, 1, W[l], 1, nout, EPI_HIDDEN_BWD, kin);
        gd.Z = m->z[l - 1];
        gd.Cf = m->df[slot];
        gd.Cb = m->db[slot];
        HIPCHK(launch_gemm(gd, m->stream));
        dzb = m->db[slot];
        dzf = m->df[slot];
    }
    return RB_OK;
}

int check_n(rb_mlp *m, int64_t n)
{
    if (n <= 0 || n > (int64_t)1 << 31) return fail(m, RB_EINVAL, "n must be in 1..2^31");
    return RB_OK;
}

}  // namespace

extern "C" {

int64_t rb_mlp_param_count(void) { return NPARAM; }

const char *rb_mlp_last_error(const rb_mlp *m) { return m ? m->err.c_str() : "null rb_mlp"; }

int rb_mlp_create(rb_mlp **out, int device)
{
    if (!out) return RB_EINVAL;
    *out = nullptr;
    rb_mlp *m = new rb_mlp();
    m->device = device;
    auto bad = [&](hipError_t e) {
        (void)hipGetLastError();
        rb_mlp_destroy(m);
        return e == hipErrorOutOfMemory ? RB_ENOMEM : RB_EHIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->p, NPARAM * sizeof(float))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->g, NPARAM * sizeof(float))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->m1, NPARAM * sizeof(float))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->m2, NPARAM * sizeof(float))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->pb, NWEIGHT * sizeof(__bf16))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->frag, NFRAG * sizeof(__bf16))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->wbf, NPARAM * sizeof(int32_t))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->norm, 5 * 64 * sizeof(float))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->loss, sizeof(float))) != hipSuccess) return bad(e);
    if ((e = hipMalloc((void **)&m->ws, (size_t)MAX_KSPLIT * MLP_W * MLP_W * sizeof(float))) != hipSuccess)
        return bad(e);
    std::vector<int32_t> map(NPARAM, -1);
    const int64_t woff[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4};
    const int64_t wlen[4] = {(int64_t)MLP_PAD * MLP_W, (int64_t)MLP_W * MLP_W,
                             (int64_t)MLP_W * MLP_W, (int64_t)MLP_W * MLP_PAD};
    const int64_t boff[4] = {BOFF_W1, BOFF_W2, BOFF_W3, BOFF_W4};
    for (int l = 0; l < 4; ++l)
        for (int64_t i = 0; i < wlen[l]; ++i) map[woff[l] + i] = (int32_t)(boff[l] + i);
    if ((e = hipMemcpy(m->wbf, map.data(), NPARAM * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
        return bad(e);
    std::vector<float> unit(5 * 64, 0.f);
    for (int i = 0; i < 64; ++i) unit[64 + i] = unit[192 + i] = unit[256 + i] = 1.f;
    if ((e = hipMemcpy(m->norm, unit.data(), unit.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return bad(e);
    std::vector<float> zero(NPARAM, 0.f);
    if ((e = hipMemcpy(m->p, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        return bad(e);
    *out = m;
    return rb_mlp_set_params(m, zero.data());
}

void rb_mlp_destroy(rb_mlp *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    void *bufs[] = {m->p, m->g, m->m1, m->m2, m->pb, m->wbf, m->norm, m->a0, m->act[0],
                    m->act[1], m->act[2], m->db[0], m->db[1], m->z[0], m->z[1], m->z[2],
                    m->yn, m->df[0], m->df[1], m->dyb, m->dyf, m->partial, m->loss, m->yraw,
                    m->xf, m->ws, m->frag};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete m;
}

int rb_mlp_set_stream(rb_mlp *m, void *stream)
{
    if (!m) return RB_EINVAL;
    m->stream = (hipStream_t)stream;
    return RB_OK;
}

int rb_mlp_set_params(rb_mlp *m, const float *params)
{
    if (!m || !params) return fail(m, RB_EINVAL, "null argument");
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipMemcpyAsync(m->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    std::vector<float> zero(NPARAM, 0.f);
    HIPCHK(hipMemcpyAsync(m->m1, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->m2, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(launch_to_bf16(m->p, m->wbf, NPARAM, m->pb, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    return RB_OK;
}

int rb_mlp_get_params(rb_mlp *m, float *params)
{
    if (!m || !params) return fail(m, RB_EINVAL, "null argument");
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipMemcpyAsync(params, m->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return RB_OK;
}

int rb_mlp_set_norm(rb_mlp *m, const float *in_mu, const float *in_sd, const float *out_mu,
                    const float *out_sd)
{
    if (!m || !in_mu || !in_sd || !out_mu || !out_sd) return fail(m, RB_EINVAL, "null argument");
    std::vector<float> h(5 * 64, 0.f);
    for (int i = 0; i < 64; ++i) h[64 + i] = h[192 + i] = h[256 + i] = 1.f;
    for (int i = 0; i < MLP_NIN; ++i) {
        if (!(in_sd[i] > 0.f) || !std::isfinite(in_mu[i])) return fail(m, RB_EINVAL, "bad input scale");
        h[i] = in_mu[i];
        h[64 + i] = 1.f / in_sd[i];
    }
    for (int i = 0; i < MLP_NOUT; ++i) {
        if (!(out_sd[i] > 0.f) || !std::isfinite(out_mu[i])) return fail(m, RB_EINVAL, "bad output scale");
        h[128 + i] = out_mu[i];
        h[192 + i] = out_sd[i];
        h[256 + i] = 1.f / out_sd[i];
    }
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(hipMemcpyAsync(m->norm, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return RB_OK;
}

int rb_mlp_forward_dev(rb_mlp *m, const float *X, int64_t n, float *Y)
{
    if (!m || !X || !Y) return fail(m, RB_EINVAL, "null argument");
    int rc = check_n(m, n);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->device));
    if ((rc = ensure_rows(m, n, false))) return rc;
    HIPCHK(launch_pack_inputs(X, n, m->norm, m->norm + 64, m->a0, m->stream));
    if ((rc = forward(m, n, false))) return rc;
    HIPCHK(launch_unpack_outputs(m->yn, n, m->norm + 128, m->norm + 192, Y, m->stream));
    return RB_OK;
}

int rb_mlp_gradients_dev(rb_mlp *m, const float *X, const float *T, int64_t n, float *grads,
                         float *loss)
{
    if (!m || !X || !T || !grads) return fail(m, RB_EINVAL, "null argument");
    int rc = check_n(m, n);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->device));
    if ((rc = backward(m, X, T, n))) return rc;
    HIPCHK(hipMemcpyAsync(grads, m->g, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (loss) *loss = l / (float)(n * MLP_NOUT);
    return RB_OK;
}

int rb_mlp_train_step_dev(rb_mlp *m, const float *X, const float *T, int64_t n, float lr, float *loss)
{
    if (!m || !X || !T) return fail(m, RB_EINVAL, "null argument");
    if (!(lr > 0.f)) return fail(m, RB_EINVAL, "lr must be > 0");
    int rc = check_n(m, n);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->device));
    if ((rc = backward(m, X, T, n))) return rc;
    const float b1 = 0.9f, b2 = 0.999f;
    m->adam_t += 1;
    const float c1 = (float)(1. / (1. - std::pow((double)b1, (double)m->adam_t)));
    const float c2 = (float)(1. / (1. - std::pow((double)b2, (double)m->adam_t)));
    HIPCHK(launch_adam(m->p, m->g, m->m1, m->m2, NPARAM, lr, b1, b2, 1e-8f, c1, c2, m->wbf, m->pb,
                       m->stream));
    if (loss) {
        float l = 0.f;
        HIPCHK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        *loss = l / (float)(n * MLP_NOUT);
    }
    return RB_OK;
}

static int grid_boxes(rb_mlp *m, const rb_grid *g, int64_t *field, int64_t *nbox)
{
    if (!g || g->nlon <= 0 || g->nlat <= 0 || g->niv <= 0 || g->nivbas <= 0 || g->nivbas > g->niv)
        return fail(m, RB_EINVAL, "bad grid");
    const int64_t plane = (int64_t)g->nlon * g->nlat;
    *field = plane * g->niv;
    *nbox = plane * g->nivbas;
    return RB_OK;
}

int rb_mlp_features_dev(rb_mlp *m, const rb_grid *g, double dt, const double *T, const double *p,
                        const double *sza, const double *qj1, const double *hc, float *X)
{
    if (!m || !T || !p || !sza || !qj1 || !hc || !X) return fail(m, RB_EINVAL, "null argument");
    if (!(dt > 0.)) return fail(m, RB_EINVAL, "dt must be > 0");
    int64_t field, nbox;
    int rc = grid_boxes(m, g, &field, &nbox);
    if (rc) return rc;
    HIPCHK(hipSetDevice(m->device));
    HIPCHK(launch_features(T, p, sza, qj1, hc, field, nbox, dt, X, m->stream));
    return RB_OK;
}

int rb_mlp_guess_dev(rb_mlp *m, const rb_grid *g, double dt, const double *T, const double *p,
                     const double *sza, const double *qj1, const double *hc, double *guess_qj1,
                     double *guess_hc)
{
    if (!m || !T || !p || !sza || !qj1 || !hc || !guess_qj1 || !guess_hc)
        return fail(m, RB_EINVAL, "null argument");
    if (!(dt > 0.)) return fail(m, RB_EINVAL, "dt must be > 0");
    int64_t field, nbox;
    int rc = grid_boxes(m, g, &field, &nbox);
    if (rc) return rc;
    if ((rc = check_n(m, nbox))) return rc;
    HIPCHK(hipSetDevice(m->device));
    // one launch for features, network and guess (the weights repacked first: they may
    // have changed since the last call); no per-box work buffers
    HIPCHK(launch_pack_frags(m->pb, m->frag, m->stream));
    GuessArgs a{T, p, sza, qj1, hc, field, nbox, (float)std::log(dt), m->frag, m->p, m->norm,
                guess_qj1, guess_hc};
    HIPCHK(launch_guess_fused(a, m->stream));
    return RB_OK;
}

}  // extern "C"
