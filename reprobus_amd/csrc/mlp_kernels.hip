// Learned initial guess for the converged chemistry step: a 60 -> 256 -> 256 -> 256 -> 55
// MLP with GELU activations.  Every GEMM runs on the gfx950 matrix cores in BF16 with FP32
// accumulation (v_mfma_f32_16x16x32_bf16); master weights, biases, Adam moments, losses and
// gradients are FP32.
//
// Layouts (row-major, batch rows): activations [n][64 or 256]; weights W_l [in][out] so a
// forward layer is act * W; the input 60 and the output 55 are zero-padded to 64 columns.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_device.h"

namespace rbm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// -------------------------------------------------------------------- GELU (tanh form)
// gelu(z) = 0.5 z (1 + tanh u), u = sqrt(2/pi) (z + 0.044715 z^3), evaluated as
// z sigma(2u) = z / (1 + exp(-2u)) (the same function: 1 + tanh u = 2 sigma(2u)), which
// has no cancellation and no branch (the library tanh branches on |u|, which splits an
// unrolled epilogue into one divergent region per value).  exp(-2u) = 2^(-2u log2 e) by
// the hardware exp2 and the reciprocal by the hardware rcp (1 ulp each, far below the
// BF16 rounding of the activations); exp2 overflows to inf for very negative z, where
// rcp gives 0 and the product -0, and underflows to 0 for large z, where it gives z.
__device__ __forceinline__ float gelu_sigma(float z)
{
    // -2 u log2(e) = z (c1 + c2 z^2), c1 = -2 sqrt(2/pi) log2(e), c2 = 0.044715 c1
    const float c1 = -2.302208198144325f, c2 = -0.10294324f;
    const float t = z * fmaf(c2, z * z, c1);
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(t));
}

__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }

// d/dz [z sigma(2u)] = sigma + 2 z sigma (1 - sigma) u'
__device__ __forceinline__ float gelu_grad(float z)
{
    const float sg = gelu_sigma(z);
    const float du = 0.7978845608028654f * (1.f + 3.f * 0.044715f * z * z);
    return sg + 2.f * z * sg * (1.f - sg) * du;
}

// -------------------------------------------------------------------- tiled GEMM
// C[M x N] = A[M x K] * B[K x N]; A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], so
// transposed operands are strides, not copies.  A 256-thread workgroup owns a 64 x 64
// tile of C; each of its four waves owns a 32 x 32 quarter = 2 x 2 MFMA 16x16 tiles.  The
// K loop stages 64 x 32 slabs of A and B through LDS, k-contiguous per row, so that each
// lane fetches its eight operand values with one 16-byte LDS read.
constexpr int TM = 64, TN = 64, TK = 32, LDS_K = TK + 8;   // +8: spreads rows over banks

// C tile epilogue of the GEMM kernels: lane l holds C rows 4 * (l / 16) + e, e = 0..3,
// column l % 16 of each of its wave's 2 x 2 tiles
__device__ __forceinline__ void gemm_epilogue(const GemmArgs &g, const f32x4 (&acc)[2][2],
                                              int64_t m0, int n0, int wm, int wn, int lane)
{
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t m = m0 + wm + i * 16 + (lane >> 4) * 4 + e;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][e];
                switch (g.mode) {
                case EPI_F32:
                    if (g.ksplit > 1)
                        g.part[((int64_t)blockIdx.z * g.M + m) * g.N + n] = v;
                    else
                        g.Cf[m * g.ldc + n] = v;
                    break;
                case EPI_HIDDEN_FWD: {
                    v += g.bias[n];
                    if (g.Cf) g.Cf[m * g.ldc + n] = v;
                    g.Cb[m * g.ldc + n] = (__bf16)gelu(v);
                    break;
                }
                case EPI_OUT_FWD:
                    g.Cf[m * g.ldc + n] = (n < g.n_valid) ? v + g.bias[n] : 0.f;
                    break;
                case EPI_HIDDEN_BWD: {
                    v *= gelu_grad(g.Z[m * g.ldc + n]);
                    g.Cf[m * g.ldc + n] = v;
                    g.Cb[m * g.ldc + n] = (__bf16)v;
                    break;
                }
                }
            }
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 sA[TM * LDS_K];
    __shared__ __attribute__((aligned(16))) __bf16 sB[TN * LDS_K];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t m0 = (int64_t)blockIdx.y * TM;
    const int n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const __bf16 zero = (__bf16)0.f;
    // split-K: workgroup z sums k in [kbeg, kend) (multiples of TK)
    const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + TK - 1) / TK * TK;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(g.K, kbeg + kchunk);
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        // stage: 64 x 32 = 2048 values of each operand, 8 per thread
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int idx = tid + r * 256;
            // A: consecutive threads walk k when A is k-contiguous, else m
            int mm, kk;
            if (g.sak == 1) { mm = idx >> 5; kk = idx & 31; } else { kk = idx >> 6; mm = idx & 63; }
            const int64_t gm = m0 + mm;
            const int gk = k0 + kk;
            sA[mm * LDS_K + kk] = (gm < g.M && gk < kend) ? g.A[gm * g.sam + (int64_t)gk * g.sak] : zero;
            int nn, kb;
            if (g.sbn == 1) { kb = idx >> 6; nn = idx & 63; } else { nn = idx >> 5; kb = idx & 31; }
            const int gn = n0 + nn;
            const int gkb = k0 + kb;
            sB[nn * LDS_K + kb] = (gn < g.N && gkb < kend) ? g.B[(int64_t)gkb * g.sbk + (int64_t)gn * g.sbn] : zero;
        }
        __syncthreads();
        // lane l supplies row (l % 16) of its A tile and column (l % 16) of its B tile,
        // k = 8 * (l / 16) .. + 7 of this 32-wide slab
        const int r16 = lane & 15, kq = (lane >> 4) * 8;
        bf16x8 va[2], vb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            va[i] = *reinterpret_cast<const bf16x8 *>(&sA[(wm + i * 16 + r16) * LDS_K + kq]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            vb[j] = *reinterpret_cast<const bf16x8 *>(&sB[(wn + j * 16 + r16) * LDS_K + kq]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va[i], vb[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    gemm_epilogue(g, acc, m0, n0, wm, wn, lane);
}

// The same product with staged operands: 16-byte global loads, 64-deep k slabs double
// buffered in LDS (one barrier per slab), for operands whose contiguous dimension is k
// (AK / BK) or m / n (the transposed operands of the gradients: each 16-byte load then
// carries 8 rows, written to the k-contiguous LDS rows one value at a time).  The tile, the
// wave layout, the MFMA sequence (k ascending in steps of 32) and the epilogue are k_gemm's,
// so the sums accumulate in the same order.  Needs 16-byte-aligned chunks: K a multiple of
// 8, and the strides and the m / n extent of a transposed operand multiples of 8
// (gemm_staged_ok).
constexpr int SK = 64, LDS_SK = SK + 8;

template <bool AK, bool BK>
__global__ void __launch_bounds__(256) k_gemm_staged(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 sA[2][TM * LDS_SK];
    __shared__ __attribute__((aligned(16))) __bf16 sB[2][TN * LDS_SK];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t m0 = (int64_t)blockIdx.y * TM;
    const int n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // k_gemm's split of K (multiples of TK): a last slab of a split may be half empty (zero)
    const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + TK - 1) / TK * TK;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(g.K, kbeg + kchunk);

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // a slab is 64 x 64 values of each operand = 512 chunks of 8: two per thread
    bf16x8 ra[2], rb[2];
    const bf16x8 zero8 = bf16x8{};
    auto load = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = tid + h * 256;
            const int r = c >> 3, q = (c & 7) * 8;   // r: row of the chunk grid, q: first of 8
            if (AK) {   // chunk: row m0 + r, k = k0 + q .. + 7
                const int64_t gm = m0 + r;
                const int gk = k0 + q;
                ra[h] = (gm < g.M && gk < kend)
                            ? *reinterpret_cast<const bf16x8 *>(g.A + gm * g.sam + gk)
                            : zero8;
            } else {    // chunk: k = k0 + r, rows m0 + q .. + 7
                const int64_t gm = m0 + q;
                const int gk = k0 + r;
                ra[h] = (gm < g.M && gk < kend)
                            ? *reinterpret_cast<const bf16x8 *>(g.A + (int64_t)gk * g.sak + gm)
                            : zero8;
            }
            if (BK) {   // chunk: column n0 + r, k = k0 + q .. + 7
                const int gn = n0 + r;
                const int gk = k0 + q;
                rb[h] = (gn < g.N && gk < kend)
                            ? *reinterpret_cast<const bf16x8 *>(g.B + (int64_t)gn * g.sbn + gk)
                            : zero8;
            } else {    // chunk: k = k0 + r, columns n0 + q .. + 7
                const int gn = n0 + q;
                const int gk = k0 + r;
                rb[h] = (gn < g.N && gk < kend)
                            ? *reinterpret_cast<const bf16x8 *>(g.B + (int64_t)gk * g.sbk + gn)
                            : zero8;
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = tid + h * 256;
            const int r = c >> 3, q = (c & 7) * 8;
            if (AK) {
                *reinterpret_cast<bf16x8 *>(&sA[buf][r * LDS_SK + q]) = ra[h];
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) sA[buf][(q + e) * LDS_SK + r] = ra[h][e];
            }
            if (BK) {
                *reinterpret_cast<bf16x8 *>(&sB[buf][r * LDS_SK + q]) = rb[h];
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) sB[buf][(q + e) * LDS_SK + r] = rb[h][e];
            }
        }
    };

    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    int buf = 0;
    if (kbeg < kend) {
        load(kbeg);
        store(0);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += SK) {
        const bool more = k0 + SK < kend;
        if (more) load(k0 + SK);   // in flight while this slab is multiplied
#pragma unroll
        for (int kk = 0; kk < SK; kk += 32) {
            bf16x8 va[2], vb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                va[i] = *reinterpret_cast<const bf16x8 *>(&sA[buf][(wm + i * 16 + r16) * LDS_SK + kk + kq]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                vb[j] = *reinterpret_cast<const bf16x8 *>(&sB[buf][(wn + j * 16 + r16) * LDS_SK + kk + kq]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va[i], vb[j], acc[i][j], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    gemm_epilogue(g, acc, m0, n0, wm, wn, lane);
}

// sums the ksplit partials (M x N each) into C (leading dimension ldc)
__global__ void k_splitk_reduce(const float *__restrict__ part, int ksplit, int64_t M, int N,
                                float *__restrict__ C, int64_t ldc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * N) return;
    float acc = 0.f;
    for (int z = 0; z < ksplit; ++z) acc += part[z * M * N + i];
    C[(i / N) * ldc + i % N] = acc;
}

// the operand layouts k_gemm_staged takes: 16-byte chunks along the contiguous dimension,
// 16-byte aligned (K, strides and transposed extents multiples of 8, aligned bases)
static bool gemm_staged_ok(const GemmArgs &g, bool &ak, bool &bk)
{
    const auto al = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    if (g.K % 8 != 0 || !al(g.A) || !al(g.B)) return false;
    if (g.sak == 1 && g.sam % 8 == 0) ak = true;
    else if (g.sam == 1 && g.sak % 8 == 0 && g.M % 8 == 0) ak = false;
    else return false;
    if (g.sbk == 1 && g.sbn % 8 == 0) bk = true;
    else if (g.sbn == 1 && g.sbk % 8 == 0 && g.N % 8 == 0) bk = false;
    else return false;
    return true;
}

hipError_t launch_gemm(const GemmArgs &gin, hipStream_t s)
{
    GemmArgs g = gin;
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    if (g.ksplit < 1 || g.mode != EPI_F32 || !g.part) g.ksplit = 1;
    if (g.ksplit > MAX_KSPLIT) return hipErrorInvalidValue;
    dim3 grid((unsigned)((g.N + TN - 1) / TN), (unsigned)((g.M + TM - 1) / TM), (unsigned)g.ksplit);
    bool ak = false, bk = false;
    if (gemm_staged_ok(g, ak, bk)) {
        if (ak && bk) hipLaunchKernelGGL((k_gemm_staged<true, true>), grid, dim3(256), 0, s, g);
        else if (ak) hipLaunchKernelGGL((k_gemm_staged<true, false>), grid, dim3(256), 0, s, g);
        else if (bk) hipLaunchKernelGGL((k_gemm_staged<false, true>), grid, dim3(256), 0, s, g);
        else hipLaunchKernelGGL((k_gemm_staged<false, false>), grid, dim3(256), 0, s, g);
    } else {
        hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, g);
    }
    if (g.ksplit > 1)
        hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)((g.M * g.N + 255) / 256)), dim3(256), 0, s,
                           g.part, g.ksplit, g.M, g.N, g.Cf, g.ldc);
    return hipGetLastError();
}

// -------------------------------------------------------------------- pack / loss / Adam
// Raw features (n x 60, FP32) -> normalised BF16 rows of 64 (zero pad).
__global__ void k_pack_inputs(const float *__restrict__ X, int64_t n, const float *__restrict__ mu,
                              const float *__restrict__ rsd, __bf16 *__restrict__ A0)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * MLP_PAD) return;
    const int64_t r = i / MLP_PAD;
    const int c = (int)(i % MLP_PAD);
    const float v = (c < MLP_NIN) ? (X[r * MLP_NIN + c] - mu[c]) * rsd[c] : 0.f;
    A0[i] = (__bf16)v;
}

// Normalised prediction (n x 64) -> raw outputs (n x 55).
__global__ void k_unpack_outputs(const float *__restrict__ Yn, int64_t n, const float *__restrict__ mu,
                                 const float *__restrict__ sd, float *__restrict__ Y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * MLP_NOUT) return;
    const int64_t r = i / MLP_NOUT;
    const int c = (int)(i % MLP_NOUT);
    Y[i] = Yn[r * MLP_PAD + c] * sd[c] + mu[c];
}

// Mean-squared error over the 55 normalised outputs: dY = 2 (y - t) / (n * 55) as BF16,
// and the per-block partial sums of the squared error.
__global__ void k_loss_grad(const float *__restrict__ Yn, const float *__restrict__ T, int64_t n,
                            const float *__restrict__ mu, const float *__restrict__ rsd,
                            __bf16 *__restrict__ dY, float *__restrict__ dYf,
                            float *__restrict__ partial)
{
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float se = 0.f;
    if (i < n * MLP_PAD) {
        const int64_t r = i / MLP_PAD;
        const int c = (int)(i % MLP_PAD);
        float d = 0.f;
        if (c < MLP_NOUT) d = Yn[i] - (T[r * MLP_NOUT + c] - mu[c]) * rsd[c];
        se = d * d;
        const float gd = 2.f * d / (float)(n * MLP_NOUT);
        dY[i] = (__bf16)gd;
        dYf[i] = gd;
    }
    red[threadIdx.x] = se;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Column sums of an FP32 [n x ncol] matrix (bias gradients): block (x, y) sums 64
// columns over row chunk y into part[y][ncol]; k_colsum_reduce adds the chunks in order.
__global__ void k_colsum(const float *__restrict__ D, int64_t n, int ncol, float *__restrict__ part)
{
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int q = threadIdx.x >> 6;
    const int64_t rows = (n + gridDim.y - 1) / gridDim.y;
    const int64_t r0 = blockIdx.y * rows, r1 = min(n, r0 + rows);
    float s = 0.f;
    if (c < ncol)
        for (int64_t r = r0 + q; r < r1; r += 4) s += D[r * ncol + c];
    red[q][threadIdx.x & 63] = s;
    __syncthreads();
    if (q == 0 && c < ncol)
        part[(int64_t)blockIdx.y * ncol + c] =
            red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

__global__ void k_colsum_reduce(const float *__restrict__ part, int chunks, int ncol,
                                float *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    float s = 0.f;
    for (int y = 0; y < chunks; ++y) s += part[(int64_t)y * ncol + c];
    out[c] = s;
}

__global__ void k_sum(const float *__restrict__ x, int n, float *__restrict__ out)
{
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; refreshes the BF16 copies of the
// weight matrices (wbf maps a parameter index to its BF16 slot, -1 for biases).
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, const int32_t *__restrict__ wbf, __bf16 *__restrict__ pb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    const int32_t j = wbf[i];
    if (j >= 0) pb[j] = (__bf16)pi;
}

__global__ void k_to_bf16(const float *__restrict__ p, const int32_t *__restrict__ wbf, int64_t n,
                          __bf16 *__restrict__ pb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t j = wbf[i];
    if (j >= 0) pb[j] = (__bf16)p[i];
}

// -------------------------------------------------------------------- chemistry features
// Feature vector of box b (FP32, 60): log mixing ratios (floored at 1e-30) of the 40
// integrated long-lived and 15 short-lived species, T, log p, cos SZA, log H2SO4 (the
// aerosol loading from which the liquid surface area follows), log dt.
__global__ void k_features(const double *__restrict__ T, const double *__restrict__ p,
                           const double *__restrict__ sza, const double *__restrict__ qj1,
                           const double *__restrict__ hc, int64_t field, int64_t nbox, double dt,
                           float *__restrict__ X)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    float *x = X + b * MLP_NIN;
    for (int s = 0; s < MLP_NLONG; ++s)
        x[s] = (float)log(fmax(qj1[(int64_t)long_slot(s) * field + b], 1.e-30));
    for (int s = 0; s < MLP_NSHORT; ++s)
        x[MLP_NLONG + s] = (float)log(fmax(hc[(int64_t)s * field + b], 1.e-30));
    x[55] = (float)T[b];
    x[56] = (float)log(p[b]);
    x[57] = (float)cos(sza[b] * 0.017453292519943295);
    x[58] = (float)log(fmax(qj1[(int64_t)SLOT_H2SO4 * field + b], 1.e-30));
    x[59] = (float)log(dt);
}

// Initial iterate from the predicted log increments y (raw, n x 55):
// guess = max(x, 1e-30) * exp(y) for the 55 species; the two inert slots and the
// diagnostic HNO3-gas slot are copied.
__global__ void k_apply_guess(const double *__restrict__ qj1, const double *__restrict__ hc,
                              const float *__restrict__ Y, int64_t field, int64_t nbox,
                              double *__restrict__ gq, double *__restrict__ gh)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const float *y = Y + b * MLP_NOUT;
    gq[(int64_t)SLOT_OX_PASSIVE * field + b] = qj1[(int64_t)SLOT_OX_PASSIVE * field + b];
    gq[(int64_t)SLOT_H2SO4 * field + b] = qj1[(int64_t)SLOT_H2SO4 * field + b];
    gq[(int64_t)SLOT_HNO3G * field + b] = qj1[(int64_t)SLOT_HNO3G * field + b];
    for (int s = 0; s < MLP_NLONG; ++s) {
        const int64_t o = (int64_t)long_slot(s) * field + b;
        gq[o] = fmax(qj1[o], 1.e-30) * exp((double)y[s]);
    }
    for (int s = 0; s < MLP_NSHORT; ++s) {
        const int64_t o = (int64_t)s * field + b;
        gh[o] = fmax(hc[o], 1.e-30) * exp((double)y[MLP_NLONG + s]);
    }
}

// -------------------------------------------------------------------- fused guess
// The guess of rb_mlp_guess_dev in one launch: features, the four layers and the guess
// formula for 32 boxes per wave (4 waves per workgroup), activations in registers from
// the features to the output.
//
// Each layer is computed transposed, C^T = W^T act^T, so that boxes are the MFMA columns:
// lane l then holds outputs 4 (l / 16) + e (e = 0..3) of box l % 16 in each 16 x 16 tile.
// Up to an order of k inside each 32-wide slab these are the values the next layer's B
// operand needs from lane l (8 values of box l % 16 at k = 8 (l / 16) .. + 7).  The order
// is absorbed into the weights: slot j of lane group g of slab s stands for input feature
// 32 s + kperm(g, j) on both operands, so no activation passes through LDS.  The weights
// are in LDS, prepacked as per-lane 16-byte fragments (k_pack_frags), 64 KB at a time (a
// small layer, or half a 256 x 256 one): the copy in and every read are straight vector
// accesses, conflict-free.
__host__ __device__ constexpr int kperm(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

// fragment block of layer l (KS slabs of 32 inputs, FT tiles of 16 outputs), BF16 offsets
constexpr int64_t FOFF_L1 = 0;
constexpr int64_t FOFF_L2 = FOFF_L1 + (int64_t)MLP_PAD * MLP_W;
constexpr int64_t FOFF_L3 = FOFF_L2 + (int64_t)MLP_W * MLP_W;
constexpr int64_t FOFF_L4 = FOFF_L3 + (int64_t)MLP_W * MLP_W;
static_assert(FOFF_L4 + (int64_t)MLP_W * MLP_PAD == NFRAG, "fragment block size");

// frag[((ft * KS + ks) * 64 + lane) * 8 + j] = W[32 ks + kperm(lane / 16, j)][16 ft + lane % 16]
// for the four weight matrices (W row-major [in][out], the BF16 copies).
__global__ void k_pack_frags(const __bf16 *__restrict__ pb, __bf16 *__restrict__ frag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NFRAG) return;
    int64_t base, wbase;
    int nin, nout;
    if (i < FOFF_L2) { base = FOFF_L1; wbase = BOFF_W1; nin = MLP_PAD; nout = MLP_W; }
    else if (i < FOFF_L3) { base = FOFF_L2; wbase = BOFF_W2; nin = MLP_W; nout = MLP_W; }
    else if (i < FOFF_L4) { base = FOFF_L3; wbase = BOFF_W3; nin = MLP_W; nout = MLP_W; }
    else { base = FOFF_L4; wbase = BOFF_W4; nin = MLP_W; nout = MLP_PAD; }
    const int64_t li = i - base;
    const int j = (int)(li & 7), lane = (int)((li >> 3) & 63);
    const int tile = (int)(li >> 9), ks_n = nin / 32;
    const int ks = tile % ks_n, ft = tile / ks_n;
    const int k = 32 * ks + kperm(lane >> 4, j), n = 16 * ft + (lane & 15);
    frag[i] = pb[wbase + (int64_t)k * nout + n];
}

constexpr int GW = 4;    // waves per workgroup (two workgroups per CU: one stages or waits
                         // at a barrier while the other computes)
constexpr int GB = 32;   // boxes per wave
constexpr int GTILES = 64;   // LDS holds 64 fragment tiles (64 KB): half of a 256 x 256 layer

// raw feature f of box b as k_features computes it, except that logarithms and the cosine
// are taken in FP32 (hardware log2 times ln 2 and the hardware cosine, a few ulp of FP32;
// the feature is rounded to BF16 right after normalisation); species are floored at 1e-30,
// a normal FP32 number
__device__ __forceinline__ float log_floor(double x)
{
    return __builtin_amdgcn_logf((float)fmax(x, 1.e-30)) * 0.6931471805599453f;
}

__device__ __forceinline__ float guess_feature(const GuessArgs &a, int f, int64_t b)
{
    if (f < MLP_NLONG) return log_floor(a.qj1[(int64_t)long_slot(f) * a.field + b]);
    if (f < MLP_NOUT) return log_floor(a.hc[(int64_t)(f - MLP_NLONG) * a.field + b]);
    if (f == 55) return (float)a.T[b];
    if (f == 56) return __builtin_amdgcn_logf((float)a.p[b]) * 0.6931471805599453f;
    if (f == 57) return __cosf((float)a.sza[b] * 0.017453292519943295f);
    if (f == 58) return log_floor(a.qj1[(int64_t)SLOT_H2SO4 * a.field + b]);
    return a.log_dt;
}

// copies a layer's fragments to LDS with direct global-to-LDS loads (16 bytes per lane,
// landing at the wave-uniform LDS base + 16 * lane, which is fragment i = base + lane):
// no registers held, every copy in flight at once; waits for them before returning
__device__ __forceinline__ void stage_frags(bf16x8 *sW, const bf16x8 *src, int nfrag)
{
    const int lane = threadIdx.x & 63;
#pragma unroll 16
    for (int i = threadIdx.x; i < nfrag; i += GW * 64)
        __builtin_amdgcn_global_load_lds(src + i, (__attribute__((address_space(3))) void *)(sW + (i - lane)),
                                         16, 0, 0);
    __builtin_amdgcn_s_waitcnt(0);
}

// bias + GELU of the FT = 16 output tiles into the next layer's operand (8 slabs)
__device__ __forceinline__ void hidden_epilogue(const f32x4 (&acc)[16][2], const float *bias,
                                                int g, bf16x8 (&act)[8][2])
{
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        // one slab at a time: its accumulator tiles die as its operand fills
        const float4 b0 = *reinterpret_cast<const float4 *>(bias + 32 * ks + 4 * g);
        const float4 b1 = *reinterpret_cast<const float4 *>(bias + 32 * ks + 16 + 4 * g);
        const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int bt = 0; bt < 2; ++bt)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int ft = 2 * ks + (j >> 2), e = j & 3;
                act[ks][bt][j] = (__bf16)gelu(acc[ft][bt][e] + bb[j]);
            }
    }
}

// MFMAs of output tiles [ft0, ft0 + NFT) over KS slabs, the tiles' fragments in LDS
// (tile (ft - ft0) * KS + ks)
template <int KS, int NFT>
__device__ __forceinline__ void layer_tiles(const bf16x8 *sW, const bf16x8 (&act)[8][2], int lane,
                                            int ft0, f32x4 (&acc)[16][2])
{
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int f = 0; f < NFT; ++f) {
            const bf16x8 w = sW[(f * KS + ks) * 64 + lane];
#pragma unroll
            for (int bt = 0; bt < 2; ++bt)
                acc[ft0 + f][bt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, act[ks][bt], acc[ft0 + f][bt], 0, 0, 0);
        }
}

__device__ __forceinline__ void zero_acc(f32x4 (&acc)[16][2])
{
#pragma unroll
    for (int ft = 0; ft < 16; ++ft)
#pragma unroll
        for (int bt = 0; bt < 2; ++bt) acc[ft][bt] = f32x4{0.f, 0.f, 0.f, 0.f};
}

// a 256 x 256 layer in two halves of 8 output tiles (64 KB of fragments each)
__device__ __forceinline__ void wide_layer(bf16x8 *sW, const bf16x8 *frag, const bf16x8 (&act)[8][2],
                                           int lane, f32x4 (&acc)[16][2])
{
    zero_acc(acc);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        stage_frags(sW, frag + h * GTILES * 64, GTILES * 64);
        __syncthreads();
        if (h == 0) layer_tiles<8, 8>(sW, act, lane, 0, acc);
        else layer_tiles<8, 8>(sW, act, lane, 8, acc);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(GW * 64, 2) k_guess_fused(GuessArgs a)
{
    __shared__ bf16x8 sW[GTILES * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t box0 = ((int64_t)blockIdx.x * GW + wave) * GB;
    const bf16x8 *frag = reinterpret_cast<const bf16x8 *>(a.frag);
    const float *in_mu = a.norm, *in_rsd = a.norm + 64, *out_mu = a.norm + 128, *out_sd = a.norm + 192;

    // features, normalised, as the layer-1 operand (slabs 0 and 1; inputs >= 60 are 0)
    bf16x8 act[8][2];
#pragma unroll
    for (int bt = 0; bt < 2; ++bt) {
        const int64_t b = box0 + 16 * bt + r;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int f = 32 * ks + kperm(g, j);
                float v = 0.f;
                if (b < a.nbox && f < MLP_NIN) v = (guess_feature(a, f, b) - in_mu[f]) * in_rsd[f];
                act[ks][bt][j] = (__bf16)v;
            }
    }
    f32x4 acc[16][2];
    // every wave takes part in the staging and barriers, boxes or not
    // layer 1: 16 output tiles x 2 slabs (32 KB)
    zero_acc(acc);
    stage_frags(sW, frag + FOFF_L1 / 8, 32 * 64);
    __syncthreads();
    layer_tiles<2, 16>(sW, act, lane, 0, acc);
    __syncthreads();
    hidden_epilogue(acc, a.bias + OFF_B1, g, act);
    wide_layer(sW, frag + FOFF_L2 / 8, act, lane, acc);
    hidden_epilogue(acc, a.bias + OFF_B2, g, act);
    wide_layer(sW, frag + FOFF_L3 / 8, act, lane, acc);
    hidden_epilogue(acc, a.bias + OFF_B3, g, act);
    // layer 4: 4 output tiles x 8 slabs (32 KB)
    zero_acc(acc);
    stage_frags(sW, frag + FOFF_L4 / 8, 32 * 64);
    __syncthreads();
    layer_tiles<8, 4>(sW, act, lane, 0, acc);
    // the guess: max(x, 1e-30) exp(y) for output y = f, as k_unpack_outputs and k_apply_guess
    const float *b4 = a.bias + OFF_B4;
#pragma unroll
    for (int bt = 0; bt < 2; ++bt) {
        const int64_t b = box0 + 16 * bt + r;
        if (b >= a.nbox) continue;
#pragma unroll
        for (int ft = 0; ft < 4; ++ft)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int f = 16 * ft + 4 * g + e;
                if (f >= MLP_NOUT) continue;
                const float yn = acc[ft][bt][e] + b4[f];
                const float y = yn * out_sd[f] + out_mu[f];
                if (f < MLP_NLONG) {
                    const int64_t o = (int64_t)long_slot(f) * a.field + b;
                    a.gq[o] = fmax(a.qj1[o], 1.e-30) * exp((double)y);
                } else {
                    const int64_t o = (int64_t)(f - MLP_NLONG) * a.field + b;
                    a.gh[o] = fmax(a.hc[o], 1.e-30) * exp((double)y);
                }
            }
        if (g == 0) {
            a.gq[(int64_t)SLOT_OX_PASSIVE * a.field + b] = a.qj1[(int64_t)SLOT_OX_PASSIVE * a.field + b];
            a.gq[(int64_t)SLOT_H2SO4 * a.field + b] = a.qj1[(int64_t)SLOT_H2SO4 * a.field + b];
            a.gq[(int64_t)SLOT_HNO3G * a.field + b] = a.qj1[(int64_t)SLOT_HNO3G * a.field + b];
        }
    }
}

// -------------------------------------------------------------------- launchers
static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_pack_inputs(const float *X, int64_t n, const float *mu, const float *rsd,
                              __bf16 *A0, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_inputs, dim3(blocks(n * MLP_PAD, 256)), dim3(256), 0, s, X, n, mu, rsd, A0);
    return hipGetLastError();
}

hipError_t launch_unpack_outputs(const float *Yn, int64_t n, const float *mu, const float *sd,
                                 float *Y, hipStream_t s)
{
    hipLaunchKernelGGL(k_unpack_outputs, dim3(blocks(n * MLP_NOUT, 256)), dim3(256), 0, s, Yn, n, mu, sd, Y);
    return hipGetLastError();
}

hipError_t launch_loss_grad(const float *Yn, const float *T, int64_t n, const float *mu,
                            const float *rsd, __bf16 *dY, float *dYf, float *partial, float *loss,
                            hipStream_t s)
{
    const unsigned nb = blocks(n * MLP_PAD, 256);
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, s, Yn, T, n, mu, rsd, dY, dYf, partial);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, partial, (int)nb, loss);
    return hipGetLastError();
}

hipError_t launch_colsum(const float *D, int64_t n, int ncol, float *part, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum, dim3(blocks(ncol, 64), COLSUM_CHUNKS), dim3(256), 0, s, D, n, ncol, part);
    hipLaunchKernelGGL(k_colsum_reduce, dim3(blocks(ncol, 256)), dim3(256), 0, s, part, COLSUM_CHUNKS,
                       ncol, out);
    return hipGetLastError();
}

hipError_t launch_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, const int32_t *wbf, __bf16 *pb,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps,
                       c1, c2, wbf, pb);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float *p, const int32_t *wbf, int64_t n, __bf16 *pb, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, s, p, wbf, n, pb);
    return hipGetLastError();
}

hipError_t launch_features(const double *T, const double *p, const double *sza, const double *qj1,
                           const double *hc, int64_t field, int64_t nbox, double dt, float *X,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_features, dim3(blocks(nbox, 256)), dim3(256), 0, s, T, p, sza, qj1, hc,
                       field, nbox, dt, X);
    return hipGetLastError();
}

hipError_t launch_apply_guess(const double *qj1, const double *hc, const float *Y, int64_t field,
                              int64_t nbox, double *gq, double *gh, hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_guess, dim3(blocks(nbox, 256)), dim3(256), 0, s, qj1, hc, Y, field,
                       nbox, gq, gh);
    return hipGetLastError();
}

hipError_t launch_pack_frags(const __bf16 *pb, __bf16 *frag, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_frags, dim3(blocks(NFRAG, 256)), dim3(256), 0, s, pb, frag);
    return hipGetLastError();
}

hipError_t launch_guess_fused(const GuessArgs &a, hipStream_t s)
{
    if (a.nbox <= 0) return hipSuccess;
    const int64_t per = (int64_t)GW * GB;
    hipLaunchKernelGGL(k_guess_fused, dim3((unsigned)((a.nbox + per - 1) / per)), dim3(GW * 64), 0, s, a);
    return hipGetLastError();
}

}  // namespace rbm
