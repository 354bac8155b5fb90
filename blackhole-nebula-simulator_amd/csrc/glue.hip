// Element-wise work between the passes of a step, kept on the device: search lengths and
// their regrowth (sph.py:371-386), viscosity switches (viscosity.py:82-106 and the Balsara
// limiter), Navier-Stokes stress and heating (navier_stokes.py:29-57), and sums into a device
// accumulator (the radiated-energy ledger term) without a host round trip.
#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

namespace nb {

static constexpr int kBlockG = 256;

// h_search = min(min(h, h_max) * skin, box_cap)   (sph.py:371; np.minimum of finite values)
__global__ void k_search_len(const double* __restrict__ h, double h_max, double skin, double box_cap, int64_t n,
                             double* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = h[i];
    a = h_max < a ? h_max : a;
    a = a * skin;
    out[i] = box_cap < a ? box_cap : a;
}

// sph.py:382-386: rows flagged "search too small" double their search (capped); every row's
// next starting h is its result. count (device) = number of flagged rows.
__global__ void k_grow_search(const int32_t* __restrict__ flags, const double* __restrict__ h_res,
                              const int64_t* __restrict__ rows, int64_t nrows, double box_cap,
                              double* __restrict__ h_search, double* __restrict__ h_guess, int32_t* __restrict__ count)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    int64_t i = row_of(rows, q);
    if (flags[q] == 2) {
        double a = h_search[i] * 2.0;
        h_search[i] = box_cap < a ? box_cap : a;
        atomicAdd(count, 1);
    }
    h_guess[i] = h_res[q];
}

__global__ void k_fill(double* __restrict__ p, double v, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Balsara (1995) limiter on the sound-speed part of the viscosity:
// alpha = alpha_const |div v| / (|div v| + |curl v| + eps c_s / h)
__global__ void k_alpha_balsara(const double* __restrict__ div, const double* __restrict__ curl,
                                const double* __restrict__ cs, const double* __restrict__ h, double alpha_const,
                                double eps, int64_t n, double* __restrict__ alpha)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = fabs(div[i]);
    alpha[i] = alpha_const * (a / (a + fabs(curl[i]) + eps * cs[i] / h[i]));
}

// Cullen-Dehnen switch (viscosity.py:82-106): xi, alpha_loc, decay time, backward-Euler decay;
// the step dt, or dt_arr[i] when given (block time steps: each row its own step)
__global__ void k_alpha_cd(const double* __restrict__ div, const double* __restrict__ curl,
                           const double* __restrict__ ddiv, const double* __restrict__ cs,
                           const double* __restrict__ h, const double* __restrict__ alpha_in, double dt_one,
                           const double* __restrict__ dt_arr, double alpha_max, double sigma, int64_t n,
                           double* __restrict__ alpha_out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = dt_arr ? dt_arr[i] : dt_one;
    double d2 = div[i] * div[i];
    double c2 = curl[i] * curl[i];
    double xi = d2 + c2 > 0 ? d2 / (d2 + c2) : 0.0;
    double md = -ddiv[i];
    double a = xi * (md > 0.0 ? md : 0.0);
    double hi = h[i], c = cs[i];
    double loc = 10.0 * hi * hi * a / (c * c);
    loc = loc < alpha_max ? loc : alpha_max;
    double tau = hi / (sigma * c);
    double al = alpha_in[i];
    double decayed = (al + loc * dt / tau) / (1.0 + dt / tau);
    alpha_out[i] = loc > al ? loc : decayed;
}

// navier_stokes.py:29-50 with zeta_v = 0: nu = alpha_ss c_s^2 r^1.5 (if alpha_ss > 0, else the
// given constant), stress S = rho [(-2/3 nu) tr G I + nu (G + G^T)], heating
// (-2/3 nu) (tr G)^2 + nu/2 sum (G + G^T)^2.
__global__ void k_ns_glue(const double* __restrict__ g, const double* __restrict__ rho,
                          const double* __restrict__ cs, const double* __restrict__ x, double alpha_ss,
                          double nu_const, int64_t n, double* __restrict__ nu_out, double* __restrict__ stress,
                          double* __restrict__ heat)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double nu;
    if (alpha_ss > 0.0) {
        double r = sqrt(x[3 * i] * x[3 * i] + x[3 * i + 1] * x[3 * i + 1] + x[3 * i + 2] * x[3 * i + 2]);
        nu = alpha_ss * (cs[i] * cs[i]) * pow15(r);
    } else {
        nu = nu_const;
    }
    nu_out[i] = nu;
    const double* G = g + 9 * i;
    double div = G[0] + G[4] + G[8];
    double bulk = (0.0 - 2.0 / 3.0 * nu) * div;
    double ss = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double t = G[3 * a + b] + G[3 * b + a];
            stress[9 * i + 3 * a + b] = rho[i] * ((a == b ? bulk : 0.0) + nu * t);
            ss += t * t;
        }
    heat[i] = (0.0 - 2.0 / 3.0 * nu) * (div * div) + 0.5 * nu * ss;
}

// nu_av = alpha_AV c_s h / 10 (disc_viscosity.py:46-48)
__global__ void k_nu_av(const double* __restrict__ cs, const double* __restrict__ h, double alpha_av, int64_t n,
                        double* __restrict__ nu)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) nu[i] = alpha_av * cs[i] * h[i] / 10.0;
}

// out[i] = a[i] + b[i] over m doubles
__global__ void k_add(const double* __restrict__ a, const double* __restrict__ b, int64_t m, double* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) out[i] = a[i] + b[i];
}

// sum of x into acc[0] (partials per block, then one work-item adds them in block order)
__global__ void k_sum_partial(const double* __restrict__ x, int64_t n, double* __restrict__ part)
{
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += x[i];
    __shared__ double sh[kBlockG];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// acc[0] += the sum of the partials (one workgroup, fixed tree order)
__global__ void __launch_bounds__(kBlockG) k_sum_into(const double* __restrict__ part, int nparts, double* __restrict__ acc)
{
    __shared__ double sh[kBlockG];
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += sh[0];
}

static inline unsigned grid_g(int64_t n) { return (unsigned)((n + kBlockG - 1) / kBlockG); }

}  // namespace nb

using namespace nb;

static int32_t launched(nb_ctx* ctx, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        return ctx_fail(ctx, m.c_str());
    }
    return NB_OK;
}

extern "C" int32_t nb_search_lengths(nb_ctx* ctx, void* stream, const double* h, double h_max, double skin,
                                     double box_cap, int64_t n, double* out)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_search_lengths: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_search_len, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, h, h_max, skin, box_cap,
                       n, out);
    return launched(ctx, "nb_search_lengths");
}

extern "C" int32_t nb_grow_search(nb_ctx* ctx, void* stream, const int32_t* flags, const double* h_res,
                                  const int64_t* rows, int64_t nrows, double box_cap, double* h_search,
                                  double* h_guess, int64_t* grown_out)
{
    if (!ctx || nrows < 0 || !grown_out) return ctx_fail(ctx, "nb_grow_search: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    *grown_out = 0;
    if (nrows == 0) return NB_OK;
    if (ctx_reserve(ctx, (void**)&ctx->scal, &ctx->scal_bytes, sizeof(int64_t) * 4)) return NB_ERR_HIP;
    int32_t* cnt = (int32_t*)ctx->scal + 2;
    if (hipMemsetAsync(cnt, 0, sizeof(int32_t), s) != hipSuccess) return ctx_fail(ctx, "memset");
    hipLaunchKernelGGL(k_grow_search, dim3(grid_g(nrows)), dim3(kBlockG), 0, s, flags, h_res, rows, nrows, box_cap,
                       h_search, h_guess, cnt);
    int32_t host = 0;
    if (hipMemcpyAsync(&host, cnt, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "nb_grow_search: count copy failed");
    *grown_out = host;
    return NB_OK;
}

extern "C" int32_t nb_fill(nb_ctx* ctx, void* stream, double* p, double v, int64_t n)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_fill: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_fill, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, p, v, n);
    return launched(ctx, "nb_fill");
}

extern "C" int32_t nb_alpha_balsara(nb_ctx* ctx, void* stream, const double* div, const double* curl,
                                    const double* cs, const double* h, double alpha_const, double eps, int64_t n,
                                    double* alpha)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_alpha_balsara: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_alpha_balsara, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, div, curl, cs, h,
                       alpha_const, eps, n, alpha);
    return launched(ctx, "nb_alpha_balsara");
}

extern "C" int32_t nb_alpha_switch(nb_ctx* ctx, void* stream, const double* div, const double* curl,
                                   const double* ddiv, const double* cs, const double* h, const double* alpha_in,
                                   double dt, double alpha_max, double sigma, int64_t n, double* alpha_out)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_alpha_switch: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_alpha_cd, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, div, curl, ddiv, cs, h,
                       alpha_in, dt, (const double*)nullptr, alpha_max, sigma, n, alpha_out);
    return launched(ctx, "nb_alpha_switch");
}

extern "C" int32_t nb_alpha_switch_dt(nb_ctx* ctx, void* stream, const double* div, const double* curl,
                                      const double* ddiv, const double* cs, const double* h, const double* alpha_in,
                                      const double* dt, double alpha_max, double sigma, int64_t n, double* alpha_out)
{
    if (!ctx || n < 0 || (n > 0 && !dt)) return ctx_fail(ctx, "nb_alpha_switch_dt: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_alpha_cd, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, div, curl, ddiv, cs, h,
                       alpha_in, 0.0, dt, alpha_max, sigma, n, alpha_out);
    return launched(ctx, "nb_alpha_switch_dt");
}

extern "C" int32_t nb_ns_stress(nb_ctx* ctx, void* stream, const double* g, const double* rho, const double* cs,
                                const double* x, double alpha_ss, double nu_const, int64_t n, double* nu,
                                double* stress, double* heat)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_ns_stress: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_ns_glue, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, g, rho, cs, x, alpha_ss,
                       nu_const, n, nu, stress, heat);
    return launched(ctx, "nb_ns_stress");
}

extern "C" int32_t nb_nu_av(nb_ctx* ctx, void* stream, const double* cs, const double* h, double alpha_av,
                            int64_t n, double* nu)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_nu_av: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_nu_av, dim3(grid_g(n)), dim3(kBlockG), 0, (hipStream_t)stream, cs, h, alpha_av, n, nu);
    return launched(ctx, "nb_nu_av");
}

extern "C" int32_t nb_add(nb_ctx* ctx, void* stream, const double* a, const double* b, int64_t m, double* out)
{
    if (!ctx || m < 0) return ctx_fail(ctx, "nb_add: bad arguments");
    if (m == 0) return NB_OK;
    hipLaunchKernelGGL(k_add, dim3(grid_g(m)), dim3(kBlockG), 0, (hipStream_t)stream, a, b, m, out);
    return launched(ctx, "nb_add");
}

extern "C" int32_t nb_sum_into(nb_ctx* ctx, void* stream, const double* x, int64_t n, double* acc_dev)
{
    if (!ctx || n < 0 || !acc_dev) return ctx_fail(ctx, "nb_sum_into: bad arguments");
    if (n == 0) return NB_OK;
    hipStream_t s = (hipStream_t)stream;
    int nparts = (int)std::min<int64_t>(1024, (n + kBlockG - 1) / kBlockG);
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * (nparts + 8))) return NB_ERR_HIP;
    hipLaunchKernelGGL(k_sum_partial, dim3(nparts), dim3(kBlockG), 0, s, x, n, ctx->red);
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kBlockG), 0, s, ctx->red, nparts, acc_dev);
    return launched(ctx, "nb_sum_into");
}
