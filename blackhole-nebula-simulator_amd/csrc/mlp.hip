// The learned starting point of the h-rho Newton solve: a small MLP (8 -> 64 -> 64 -> 1, tanh)
// that predicts the change of log h over a step from the previous step's state.
//
// Inference runs on the matrix cores: BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation. The layers are evaluated transposed (hidden x particles), so that the FP32
// result tile of one layer, converted to BF16, is already the B operand of the next layer in
// the MFMA register layout: no shuffles or LDS between layers. Training runs in FP32 on the
// vector units: a per-sample forward/backward kernel, a per-chunk weight-gradient kernel with
// LDS staging, and a deterministic reduction + Adam update (fixed summation order).
//
// Parameter vector (float32, kParams values): W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3.
// Normalisation vector (float32, 18 values): feature means [8], feature 1/std [8], target mean,
// target scale. The network maps normalised features to the normalised target.
#include <string>

#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

namespace nb {

static constexpr int kIn = 8;
static constexpr int kHid = 64;
static constexpr int kOffW1 = 0;
static constexpr int kOffB1 = kOffW1 + kIn * kHid;
static constexpr int kOffW2 = kOffB1 + kHid;
static constexpr int kOffB2 = kOffW2 + kHid * kHid;
static constexpr int kOffW3 = kOffB2 + kHid;
static constexpr int kOffB3 = kOffW3 + kHid;
static constexpr int kParams = kOffB3 + 1;  // 4801
static constexpr int kNorm = 2 * kIn + 2;
static constexpr int kTiles = 4 + 16;  // MFMA A tiles: W1^T (4 x 1), W2^T (4 x 4)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ short to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// ---------------------------------------------------------------- features and targets
// Raw features of particle i: log h, log rho, div v dt, |curl v| dt, log(neighbours), log r,
// log u, dt / dt_prev (row q of the output is particle q * stride).
__global__ void k_mlp_features(const double* __restrict__ h, const double* __restrict__ rho,
                               const double* __restrict__ div_v, const double* __restrict__ curl_v,
                               const double* __restrict__ nneigh, const double* __restrict__ x,
                               const double* __restrict__ u, int64_t rows, int64_t stride, double dt_host,
                               double dt_prev_host, const double* __restrict__ dts_dev, float* __restrict__ out)
{
    // dts_dev: (dt, dt_prev) on the device (a recorded step), else the host values
    const double dt = dts_dev ? dts_dev[0] : dt_host, dt_prev = dts_dev ? dts_dev[1] : dt_prev_host;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= rows) return;
    const int64_t i = q * stride;
    const double r = sqrt(x[3 * i] * x[3 * i] + x[3 * i + 1] * x[3 * i + 1] + x[3 * i + 2] * x[3 * i + 2]);
    float* f = out + q * kIn;
    f[0] = (float)log(h[i]);
    f[1] = (float)log(rho[i]);
    f[2] = (float)(div_v ? div_v[i] * dt : 0.0);
    f[3] = (float)(curl_v ? curl_v[i] * dt : 0.0);
    f[4] = (float)log(nneigh && nneigh[i] > 1.0 ? nneigh[i] : 1.0);
    f[5] = (float)log(r);
    f[6] = (float)log(u[i]);
    f[7] = (float)(dt_prev > 0.0 ? dt / dt_prev : 1.0);
}

// target of row q: log h_final - log h_prev of particle q * stride
__global__ void k_mlp_target(const double* __restrict__ h_prev, const double* __restrict__ h_final, int64_t rows,
                             int64_t stride, float* __restrict__ out)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= rows) return;
    const int64_t i = q * stride;
    out[q] = (float)(log(h_final[i]) - log(h_prev[i]));
}

// Newton starting h: h_prev exp(delta); the previous h when the prediction is not finite or
// changes h by more than a factor 2 (fallback, counted), and also when |delta| < gate: a change
// below the Newton tolerance converges in one step from h_prev anyway.
__global__ void k_mlp_h_start(const double* __restrict__ h_prev, const float* __restrict__ delta, int64_t n,
                              float gate, double* __restrict__ h_start, long long* __restrict__ part)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    long long fb = 0;
    if (i < n) {
        const float d = delta[i];
        if (isfinite(d) && fabsf(d) < gate) {
            h_start[i] = h_prev[i];
        } else if (isfinite(d) && fabsf(d) < 0.69314718f) {
            h_start[i] = h_prev[i] * exp((double)d);
        } else {
            h_start[i] = h_prev[i];
            fb = 1;
        }
    }
    block_partial<false>(fb, part);
}

// *out = sum of the per-workgroup counts (one workgroup, fixed order)
__global__ void __launch_bounds__(256) k_mlp_count(const long long* __restrict__ part, int64_t nb, long long* out)
{
    __shared__ long long sh[256];
    long long v = 0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) v += part[b];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ---------------------------------------------------------------- inference (BF16 MFMA)
// A operands of both layers in the MFMA lane layout: tile t, lane l, element j at
// ((t * 64) + l) * 4 + j. Tile mt (< 4) of W1^T holds W1[4 (l / 16) + j][16 mt + l % 16] (inputs
// past 8 are zero padding); tile 4 + 4 mt + kt of W2^T holds W2[16 kt + 4 (l / 16) + j][16 mt + l % 16].
__global__ void k_mlp_pack(const float* __restrict__ P, short* __restrict__ wpack)
{
    for (int e = threadIdx.x; e < kTiles * 64 * 4; e += blockDim.x) {
        const int t = e / 256, l = (e / 4) % 64, j = e % 4;
        float v;
        if (t < 4) {
            const int i = 4 * (l / 16) + j, o = 16 * t + l % 16;
            v = i < kIn ? P[kOffW1 + i * kHid + o] : 0.f;
        } else {
            const int mt = (t - 4) / 4, kt = (t - 4) % 4;
            const int i = 16 * kt + 4 * (l / 16) + j, o = 16 * mt + l % 16;
            v = P[kOffW2 + i * kHid + o];
        }
        wpack[e] = to_bf16(v);
    }
}

static constexpr int kInferTiles = 4;  // 16-particle tiles per wavefront

// tanh from the hardware exponential and reciprocal, 1 - 2 / (e^{2x} + 1) (absolute error about
// 1e-7, exact limits +-1): the inference rounds its activations to BF16 (8 significant bits), and
// the libm form costs several times the instructions of the two MFMA layers it feeds.
__device__ __forceinline__ float tanh_hw(float x)
{
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // 2^(2 x / ln 2) = e^(2x)
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

__global__ void __launch_bounds__(256)
k_mlp_infer(const short* __restrict__ wpack, const float* __restrict__ P, const float* __restrict__ norm,
            const float* __restrict__ feat, int64_t n, float* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const bf16x4* wp = (const bf16x4*)wpack;
    bf16x4 a1[4], a2[16];
#pragma unroll
    for (int t = 0; t < 4; t++) a1[t] = wp[t * 64 + lane];
#pragma unroll
    for (int t = 0; t < 16; t++) a2[t] = wp[(4 + t) * 64 + lane];
    // this lane's hidden units: 16 mt + 4 (lane / 16) + j
    float b1[16], b2[16], w3[16];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = 16 * mt + 4 * (lane / 16) + j;
            b1[4 * mt + j] = P[kOffB1 + o];
            b2[4 * mt + j] = P[kOffB2 + o];
            w3[4 * mt + j] = P[kOffW3 + o];
        }
    const float b3 = P[kOffB3];
    const float tmean = norm[2 * kIn], tscale = norm[2 * kIn + 1];
    for (int rt = 0; rt < kInferTiles; rt++) {
        const int64_t p = (wave * kInferTiles + rt) * 16 + (lane & 15);
        // B operand of layer 1: features 4 (lane / 16) + j of particle p (lanes >= 32: padding)
        bf16x4 bx;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * (lane / 16) + j;
            float v = 0.f;
            if (lane < 32 && p < n) v = (feat[p * kIn + k] - norm[k]) * norm[kIn + k];
            bx[j] = to_bf16(v);
        }
        bf16x4 h1[4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[mt], bx, acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++) h1[mt][j] = to_bf16(tanh_hw(acc[j] + b1[4 * mt + j]));
        }
        float part = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[4 * mt + kt], h1[kt], acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; j++) part += tanh_hw(acc[j] + b2[4 * mt + j]) * w3[4 * mt + j];
        }
        // lanes l, l + 16, l + 32, l + 48 hold the four quarters of particle l % 16's hidden layer
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (lane < 16 && p < n) out[p] = tmean + tscale * (part + b3);
    }
}

// ---------------------------------------------------------------- training (FP32)
static constexpr int kTrainBlock = 128;
static constexpr int kChunk = 64;  // samples per weight-gradient workgroup

// Per sample: forward pass, squared error, and the backward pass down to the pre-activations.
// Stored feature-major ([unit][batch]) for the weight-gradient kernel: h1, h2, dz1, dz2, and dy.
__global__ void __launch_bounds__(kTrainBlock)
k_mlp_fwd_bwd(const float* __restrict__ P, const float* __restrict__ norm, const float* __restrict__ feat,
              const float* __restrict__ target, int64_t B, float* __restrict__ h1s, float* __restrict__ h2s,
              float* __restrict__ dz1s, float* __restrict__ dz2s, float* __restrict__ dys, float* __restrict__ loss_part)
{
    __shared__ float w[kParams];
    for (int e = threadIdx.x; e < kParams; e += blockDim.x) w[e] = P[e];
    __syncthreads();
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    float se = 0.f;
    if (b < B) {
        float x[kIn];
#pragma unroll
        for (int k = 0; k < kIn; k++) x[k] = (feat[b * kIn + k] - norm[k]) * norm[kIn + k];
        const float t = (target[b] - norm[2 * kIn]) / norm[2 * kIn + 1];
        float h1[kHid];
#pragma unroll
        for (int o = 0; o < kHid; o++) {
            float z = w[kOffB1 + o];
#pragma unroll
            for (int k = 0; k < kIn; k++) z += x[k] * w[kOffW1 + k * kHid + o];
            h1[o] = tanhf(z);
            h1s[o * B + b] = h1[o];
        }
        float y = w[kOffB3];
        for (int o = 0; o < kHid; o++) {
            float z = w[kOffB2 + o];
#pragma unroll
            for (int k = 0; k < kHid; k++) z += h1[k] * w[kOffW2 + k * kHid + o];
            const float h2 = tanhf(z);
            h2s[o * B + b] = h2;
            y += h2 * w[kOffW3 + o];
        }
        const float e = y - t;
        se = e * e;
        const float dy = 2.f * e / (float)B;  // d(mean squared error)/dy
        dys[b] = dy;
        float dz2[kHid];
#pragma unroll
        for (int o = 0; o < kHid; o++) {
            const float h2 = h2s[o * B + b];
            dz2[o] = dy * w[kOffW3 + o] * (1.f - h2 * h2);
            dz2s[o * B + b] = dz2[o];
        }
        for (int k = 0; k < kHid; k++) {
            float dh = 0.f;
#pragma unroll
            for (int o = 0; o < kHid; o++) dh += w[kOffW2 + k * kHid + o] * dz2[o];
            const float hk = h1s[k * B + b];  // reloaded: a dynamic index would put h1 in scratch
            dz1s[k * B + b] = dh * (1.f - hk * hk);
        }
    }
    // block sum of squared errors (fixed tree order)
    __shared__ float sh[kTrainBlock];
    sh[threadIdx.x] = se;
    __syncthreads();
    for (int s = kTrainBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0];
}

// Per chunk of kChunk samples: the chunk's share of every parameter gradient (partials per chunk).
__global__ void __launch_bounds__(256)
k_mlp_wgrad(const float* __restrict__ norm, const float* __restrict__ feat, int64_t B, const float* __restrict__ h1s,
            const float* __restrict__ h2s, const float* __restrict__ dz1s, const float* __restrict__ dz2s,
            const float* __restrict__ dys, float* __restrict__ part)
{
    __shared__ float sx[kChunk][kIn], sh1[kChunk][kHid], sh2[kChunk][kHid], sd1[kChunk][kHid], sd2[kChunk][kHid];
    __shared__ float sdy[kChunk];
    const int64_t b0 = blockIdx.x * (int64_t)kChunk;
    const int cnt = (int)(B - b0 < kChunk ? B - b0 : kChunk);
    for (int e = threadIdx.x; e < kChunk * kHid; e += blockDim.x) {
        const int s = e / kHid, o = e % kHid;
        const bool in = s < cnt;
        sh1[s][o] = in ? h1s[o * B + b0 + s] : 0.f;
        sh2[s][o] = in ? h2s[o * B + b0 + s] : 0.f;
        sd1[s][o] = in ? dz1s[o * B + b0 + s] : 0.f;
        sd2[s][o] = in ? dz2s[o * B + b0 + s] : 0.f;
    }
    for (int e = threadIdx.x; e < kChunk * kIn; e += blockDim.x) {
        const int s = e / kIn, k = e % kIn;
        sx[s][k] = s < cnt ? (feat[(b0 + s) * kIn + k] - norm[k]) * norm[kIn + k] : 0.f;
    }
    for (int s = threadIdx.x; s < kChunk; s += blockDim.x) sdy[s] = s < cnt ? dys[b0 + s] : 0.f;
    __syncthreads();
    float* out = part + blockIdx.x * (int64_t)kParams;
    for (int p = threadIdx.x; p < kParams; p += blockDim.x) {
        float g = 0.f;
        if (p < kOffB1) {
            const int k = p / kHid, o = p % kHid;
            for (int s = 0; s < kChunk; s++) g += sx[s][k] * sd1[s][o];
        } else if (p < kOffW2) {
            const int o = p - kOffB1;
            for (int s = 0; s < kChunk; s++) g += sd1[s][o];
        } else if (p < kOffB2) {
            const int k = (p - kOffW2) / kHid, o = (p - kOffW2) % kHid;
            for (int s = 0; s < kChunk; s++) g += sh1[s][k] * sd2[s][o];
        } else if (p < kOffW3) {
            const int o = p - kOffB2;
            for (int s = 0; s < kChunk; s++) g += sd2[s][o];
        } else if (p < kOffB3) {
            const int o = p - kOffW3;
            for (int s = 0; s < kChunk; s++) g += sh2[s][o] * sdy[s];
        } else {
            for (int s = 0; s < kChunk; s++) g += sdy[s];
        }
        out[p] = g;
    }
}

// gradient = sum of the chunk partials (fixed order); optional Adam step; loss = mean squared
// error of the normalised target (from the forward blocks' partials)
__global__ void k_mlp_reduce_adam(const float* __restrict__ part, int64_t nchunks, const float* __restrict__ loss_part,
                                  int64_t nloss, int64_t B, float* __restrict__ P, float* __restrict__ m,
                                  float* __restrict__ v, float lr, float beta1, float beta2, float eps, float bc1,
                                  float bc2, float* __restrict__ grad_out, float* __restrict__ loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_out) {
        float s = 0.f;
        for (int64_t k = 0; k < nloss; k++) s += loss_part[k];
        loss_out[0] = s / (float)B;
    }
    if (p >= kParams) return;
    float g = 0.f;
    for (int64_t c = 0; c < nchunks; c++) g += part[c * kParams + p];
    if (grad_out) grad_out[p] = g;
    if (lr > 0.f) {
        const float mm = beta1 * m[p] + (1.f - beta1) * g;
        const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
}

static inline unsigned blocks_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

static int32_t mlp_launched(nb_ctx* ctx, const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return NB_OK;
    const std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    return ctx_fail(ctx, msg.c_str());
}

// forward/backward + gradient reduction (+ Adam when lr > 0)
static int32_t mlp_step(nb_ctx* ctx, hipStream_t s, float* params, float* m, float* v, int64_t step,
                        const float* norm, const float* feat, const float* target, int64_t B, float lr,
                        float* grad_out, float* loss_out, const char* what)
{
    if (!ctx || !params || !norm || !feat || !target || B <= 0 || (lr > 0.f && (!m || !v || step < 1)))
        return ctx_fail(ctx, what);
    const int64_t nfwd = blocks_of(B, kTrainBlock), nch = blocks_of(B, kChunk);
    const size_t need = sizeof(float) * ((size_t)4 * kHid * B + B + nfwd + (size_t)nch * kParams);
    if (ctx_reserve(ctx, (void**)&ctx->mlp_scratch, &ctx->mlp_scratch_bytes, need)) return NB_ERR_HIP;
    float* h1 = (float*)ctx->mlp_scratch;
    float* h2 = h1 + (size_t)kHid * B;
    float* d1 = h2 + (size_t)kHid * B;
    float* d2 = d1 + (size_t)kHid * B;
    float* dy = d2 + (size_t)kHid * B;
    float* lp = dy + B;
    float* part = lp + nfwd;
    NB_TIMED(ctx, s, "k_mlp_fwd_bwd", k_mlp_fwd_bwd, dim3((unsigned)nfwd), dim3(kTrainBlock), 0, s, params, norm, feat,
             target, B, h1, h2, d1, d2, dy, lp);
    NB_TIMED(ctx, s, "k_mlp_wgrad", k_mlp_wgrad, dim3((unsigned)nch), dim3(256), 0, s, norm, feat, B, h1, h2, d1, d2,
             dy, part);
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = lr > 0.f ? 1.f - powf(b1, (float)step) : 1.f;
    const float bc2 = lr > 0.f ? 1.f - powf(b2, (float)step) : 1.f;
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(blocks_of(kParams, 256)), dim3(256), 0, s, part, nch, lp, nfwd, B,
                       params, m, v, lr, b1, b2, 1e-8f, bc1, bc2, grad_out, loss_out);
    return mlp_launched(ctx, what);
}

}  // namespace nb

using namespace nb;

extern "C" int32_t nb_mlp_param_count(void) { return kParams; }

extern "C" int32_t nb_mlp_features(nb_ctx* ctx, void* stream, const double* h_prev, const double* rho_prev,
                                   const double* div_v, const double* curl_v, const double* nneigh, const double* x,
                                   const double* u, int64_t n, int64_t stride, double dt, double dt_prev, float* out)
{
    if (!ctx || n < 0 || stride < 1 || !h_prev || !rho_prev || !x || !u || !out)
        return ctx_fail(ctx, "nb_mlp_features: bad arguments");
    const int64_t rows = (n + stride - 1) / stride;
    if (rows == 0) return NB_OK;
    hipLaunchKernelGGL(k_mlp_features, dim3(blocks_of(rows, 256)), dim3(256), 0, (hipStream_t)stream, h_prev, rho_prev,
                       div_v, curl_v, nneigh, x, u, rows, stride, dt, dt_prev, (const double*)nullptr, out);
    return mlp_launched(ctx, "nb_mlp_features");
}

extern "C" int32_t nb_mlp_features_dev(nb_ctx* ctx, void* stream, const double* h_prev, const double* rho_prev,
                                       const double* div_v, const double* curl_v, const double* nneigh,
                                       const double* x, const double* u, int64_t n, int64_t stride,
                                       const double* dts_dev, float* out)
{
    if (!ctx || n < 0 || stride < 1 || !h_prev || !rho_prev || !x || !u || !out || !dts_dev)
        return ctx_fail(ctx, "nb_mlp_features_dev: bad arguments");
    const int64_t rows = (n + stride - 1) / stride;
    if (rows == 0) return NB_OK;
    hipLaunchKernelGGL(k_mlp_features, dim3(blocks_of(rows, 256)), dim3(256), 0, (hipStream_t)stream, h_prev, rho_prev,
                       div_v, curl_v, nneigh, x, u, rows, stride, 0.0, 0.0, dts_dev, out);
    return mlp_launched(ctx, "nb_mlp_features_dev");
}

extern "C" int32_t nb_mlp_target(nb_ctx* ctx, void* stream, const double* h_prev, const double* h_final, int64_t n,
                                 int64_t stride, float* out)
{
    if (!ctx || n < 0 || stride < 1 || !h_prev || !h_final || !out) return ctx_fail(ctx, "nb_mlp_target: bad arguments");
    const int64_t rows = (n + stride - 1) / stride;
    if (rows == 0) return NB_OK;
    hipLaunchKernelGGL(k_mlp_target, dim3(blocks_of(rows, 256)), dim3(256), 0, (hipStream_t)stream, h_prev, h_final,
                       rows, stride, out);
    return mlp_launched(ctx, "nb_mlp_target");
}

extern "C" int32_t nb_mlp_infer(nb_ctx* ctx, void* stream, const float* params, const float* norm, const float* feat,
                                int64_t n, float* out)
{
    if (!ctx || n < 0 || !params || !norm || !feat || !out) return ctx_fail(ctx, "nb_mlp_infer: bad arguments");
    if (n == 0) return NB_OK;
    hipStream_t s = (hipStream_t)stream;
    if (ctx_reserve(ctx, (void**)&ctx->mlp_pack, &ctx->mlp_pack_bytes, sizeof(short) * kTiles * 64 * 4))
        return NB_ERR_HIP;
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(256), 0, s, params, (short*)ctx->mlp_pack);
    const int64_t waves = (n + 16 * kInferTiles - 1) / (16 * kInferTiles);
    NB_TIMED(ctx, s, "k_mlp_infer", k_mlp_infer, dim3(blocks_of(waves, 4)), dim3(256), 0, s,
             (const short*)ctx->mlp_pack, params, norm, feat, n, out);
    return mlp_launched(ctx, "nb_mlp_infer");
}

extern "C" int32_t nb_mlp_h_start(nb_ctx* ctx, void* stream, const double* h_prev, const float* delta, int64_t n,
                                  float gate, double* h_start, long long* fallbacks_dev)
{
    if (!ctx || n < 0 || !h_prev || !delta || !h_start || !fallbacks_dev)
        return ctx_fail(ctx, "nb_mlp_h_start: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = blocks_of(n, 256);
    if (ctx_reserve(ctx, (void**)&ctx->mlp_part, &ctx->mlp_part_bytes, sizeof(long long) * (size_t)(nb + 1)))
        return NB_ERR_HIP;
    if (n > 0)
        hipLaunchKernelGGL(k_mlp_h_start, dim3((unsigned)nb), dim3(256), 0, s, h_prev, delta, n, gate, h_start,
                           ctx->mlp_part);
    hipLaunchKernelGGL(k_mlp_count, dim3(1), dim3(256), 0, s, ctx->mlp_part, nb, fallbacks_dev);
    return mlp_launched(ctx, "nb_mlp_h_start");
}

extern "C" int32_t nb_mlp_grad(nb_ctx* ctx, void* stream, const float* params, const float* norm, const float* feat,
                               const float* target, int64_t batch, float* grad_out, float* loss_out)
{
    if (!grad_out || !loss_out) return ctx_fail(ctx, "nb_mlp_grad: bad arguments");
    return mlp_step(ctx, (hipStream_t)stream, (float*)params, nullptr, nullptr, 0, norm, feat, target, batch, 0.f,
                    grad_out, loss_out, "nb_mlp_grad: bad arguments");
}

extern "C" int32_t nb_mlp_train_step(nb_ctx* ctx, void* stream, float* params, float* adam_m, float* adam_v,
                                     int64_t step, const float* norm, const float* feat, const float* target,
                                     int64_t batch, float lr, float* loss_out)
{
    if (!(lr > 0.f)) return ctx_fail(ctx, "nb_mlp_train_step: learning rate must be positive");
    return mlp_step(ctx, (hipStream_t)stream, params, adam_m, adam_v, step, norm, feat, target, batch, lr, nullptr,
                    loss_out, "nb_mlp_train_step: bad arguments");
}
