// Block (individual power-of-two) time steps (blocks.py, timesteps.py): the element-wise and
// list-walking work of a system step, on the device. Each particle has a level k and a step of
// 2^(K_MAX - k) ticks on an int64 timeline; a system step goes to the earliest step end, predicts
// every particle there (gravity flow done by nb_glf_step, then the kicks below), evaluates the
// active rows, closes their steps and opens new ones at levels chosen from the step criteria,
// the level rules, the neighbour limiter, landing on the stop time, and the wake-up of
// neighbours whose steps are too long. Arithmetic follows the reference expression by
// expression (NumPy's evaluation order), so that the FP64 results are the same bits.
#include <cmath>

#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

namespace nb {

static constexpr int kBlockB = 256;
static constexpr int64_t kKMax = 40;  // timesteps.py:15

static inline unsigned grid_b(int64_t n) { return (unsigned)((n + kBlockB - 1) / kBlockB); }

__device__ __forceinline__ int64_t ticks_of(int64_t level) { return (int64_t)1 << (kKMax - level); }

// |x| as numpy.linalg.norm along a row: sqrt((x0^2 + x1^2) + x2^2)
__device__ __forceinline__ double norm3(const double* x)
{
    return sqrt((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]);
}

// np.minimum: NaN if either operand is NaN
__device__ __forceinline__ double np_min(double a, double b)
{
    if (a != a) return a;
    if (b != b) return b;
    return b < a ? b : a;
}

// Heat budget of a kick u += dt du/dt split into its parts (blocks.py:434-440).
__device__ __forceinline__ void add_heat(int64_t i, double dt, const double* dudt, const double* visc,
                                         const double* cond, const double* disk, double* hv, double* hc, double* hd,
                                         double* hp)
{
    const double v = visc[i], c = cond[i], d = disk[i];
    hv[i] += dt * v;
    hc[i] += dt * c;
    hd[i] += dt * d;
    hp[i] += dt * (((dudt[i] - v) - c) - d);
}

// Elapsed time s = (n_next - n_beg) tick and half step L/2 = 0.5 (n_end - n_beg) tick per
// particle (blocks.py:505-507).
__global__ void k_block_elapsed(const int64_t* __restrict__ n_beg, const int64_t* __restrict__ n_end, int64_t n_next,
                                double tick, int64_t n, double* __restrict__ s_out, double* __restrict__ half_out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    s_out[i] = (double)(n_next - n_beg[i]) * tick;
    half_out[i] = (0.5 * (double)(n_end[i] - n_beg[i])) * tick;
}

// Predictions at tick n_next (blocks.py:502-530): s = (n_next - n_beg) tick, L/2 = 0.5 (n_end -
// n_beg) tick, p = p_flow + (s - L/2) a; with beta cooling (beta finite) u_half is cooled over s
// at r_mid (the exact law), the active particles (n_end = n_next) keep the cooled u_half and book
// the radiated m du into e_part (0 elsewhere), luminosity and heat_cool; u = u_half + (s - L/2) du/dt.
__global__ void k_block_predict(const int64_t* __restrict__ n_beg, const int64_t* __restrict__ n_end, int64_t n_next,
                                double tick, const double* __restrict__ p_flow, const double* __restrict__ a,
                                double* __restrict__ u_half, const double* __restrict__ dudt,
                                const double* __restrict__ x_beg, const double* __restrict__ x,
                                const double* __restrict__ mass, double u_floor_one, const double* __restrict__ floors,
                                double beta, int64_t n, double* __restrict__ p_out, double* __restrict__ u_out,
                                double* __restrict__ e_part, double* __restrict__ lum, double* __restrict__ heat_cool)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = (double)(n_next - n_beg[i]) * tick;
    const double half = (0.5 * (double)(n_end[i] - n_beg[i])) * tick;
    const double lag = s - half;
    for (int c = 0; c < 3; c++) p_out[3 * i + c] = p_flow[3 * i + c] + lag * a[3 * i + c];
    double uh = u_half[i];
    if (beta < INFINITY) {
        const double r_mid = 0.5 * (norm3(x_beg + 3 * i) + norm3(x + 3 * i));
        const double t_cool = beta / pow(r_mid, -1.5);
        const double fl = floors ? floors[i] : u_floor_one;
        const double uc = fl + (uh - fl) * exp(-s / t_cool);
        if (n_end[i] == n_next) {
            const double dm = mass[i] * (uh - uc);
            e_part[i] = dm;
            lum[i] = dm / (2.0 * half);
            heat_cool[i] += uh - uc;
            u_half[i] = uc;
        } else {
            e_part[i] = 0.0;
        }
        uh = uc;
    }
    u_out[i] = uh + lag * dudt[i];
}

// Closing kicks of the active rows (blocks.py:543-546): p = p_flow + (L/2) a, u = u_half + (L/2)
// du/dt, heat budget over L/2.
__global__ void k_block_close(const int64_t* __restrict__ rows, int64_t nrows, const int64_t* __restrict__ n_beg,
                              const int64_t* __restrict__ n_end, double tick, const double* __restrict__ p_flow,
                              const double* __restrict__ a, const double* __restrict__ u_half,
                              const double* __restrict__ dudt, const double* __restrict__ visc,
                              const double* __restrict__ cond, const double* __restrict__ disk, double* __restrict__ p,
                              double* __restrict__ u, double* __restrict__ hv, double* __restrict__ hc,
                              double* __restrict__ hd, double* __restrict__ hp)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    const double half = (0.5 * (double)(n_end[i] - n_beg[i])) * tick;
    for (int c = 0; c < 3; c++) p[3 * i + c] = p_flow[3 * i + c] + half * a[3 * i + c];
    u[i] = u_half[i] + half * dudt[i];
    add_heat(i, half, dudt, visc, cond, disk, hv, hc, hd, hp);
}

// New levels of the active rows (blocks.py:318-335, 367-372; timesteps.py:26-60): the smallest
// step criterion (Courant C h / v_sig, force C_force sqrt(h / |a|), viscous 0.25 h^2 / nu, orbit
// eta r^1.5), its level (the smallest k >= 0 with dt_max 2^-k <= dt, which is what the
// reference's log2 formula with its guard yields), plus the forced offset, then the rule for
// longer steps. bad[0]: a criterion not positive (or NaN); bad[1]: a level beyond K_MAX.
__global__ void k_block_levels(const int64_t* __restrict__ rows, int64_t nrows, const double* __restrict__ x,
                               const double* __restrict__ h, const double* __restrict__ vsig,
                               const double* __restrict__ a, const double* __restrict__ nu, double courant,
                               double c_force, double eta_orb, int hydro, double dt_max,
                               const int64_t* __restrict__ level_offset, int64_t n_now, int64_t* __restrict__ level,
                               int32_t* __restrict__ bad)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    double dt;
    if (!hydro) {
        dt = eta_orb * pow(norm3(x + 3 * i), 1.5);
    } else {
        const double hi = h[i];
        dt = courant * hi / vsig[i];
        const double am = norm3(a + 3 * i);
        dt = np_min(dt, am > 0 ? c_force * sqrt(hi / am) : INFINITY);
        if (nu) dt = np_min(dt, nu[i] > 0 ? 0.25 * (hi * hi) / nu[i] : INFINITY);
        if (eta_orb > 0) dt = np_min(dt, eta_orb * pow(norm3(x + 3 * i), 1.5));
    }
    if (!(dt > 0)) {
        atomicAdd(&bad[0], 1);
        return;
    }
    int64_t k = 0;
    while (k <= kKMax && ldexp(dt_max, -(int)k) > dt) k++;
    if (k > kKMax) {
        atomicAdd(&bad[1], 1);
        return;
    }
    k += level_offset[i];
    const int64_t cur = level[i];
    int64_t nk = k >= cur ? k : cur;
    if (k < cur)
        for (int64_t kk = k; kk < cur; kk++)
            if (n_now % ticks_of(kk) == 0) {
                nk = kk;
                break;
            }
    level[i] = nk;
}

// One sweep of the neighbour limiter over the rows (blocks.py:338-353): level_i rises to the
// largest neighbour level minus gap. Levels only rise, so sweeps in any order (and reads that
// see another row's rise early or late) reach the same least fixed point as the reference's
// sequential sweeps; the host repeats until nothing changes.
__global__ void k_block_limit(const int64_t* __restrict__ rows, int64_t nrows, const int64_t* __restrict__ offset,
                              const int32_t* __restrict__ index, int64_t gap, unsigned long long* __restrict__ level,
                              int32_t* __restrict__ changed)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    const long long mine = (long long)__hip_atomic_load(&level[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long long need = mine;
    for (int64_t k = offset[i]; k < offset[i + 1]; k++) {
        const long long lj =
            (long long)__hip_atomic_load(&level[index[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - gap;
        need = lj > need ? lj : need;
    }
    if (need > mine) {
        atomicMax((long long*)&level[i], need);
        *changed = 1;
    }
}

// Landing (blocks.py:383-393): while a step opened at n would end past n_stop, one level up
// (n < n_stop only). bad: a level beyond K_MAX.
__global__ void k_block_land(const int64_t* __restrict__ rows, int64_t nrows, int64_t n_now, int64_t n_stop,
                             int64_t* __restrict__ level, int32_t* __restrict__ bad)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    int64_t k = level[i];
    if (n_now < n_stop)
        while (n_now + ticks_of(k < kKMax ? k : kKMax) > n_stop) k++;
    level[i] = k;
    if (k > kKMax) atomicAdd(bad, 1);
}

// Wake-up targets (blocks.py:357-365): need_j = max over active neighbours i of level_i - gap.
__global__ void k_block_wake_targets(const int64_t* __restrict__ rows, int64_t nrows,
                                     const int64_t* __restrict__ offset, const int32_t* __restrict__ index,
                                     int64_t gap, const int64_t* __restrict__ level, int64_t* __restrict__ need)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    const long long t = (long long)(level[i] - gap);
    for (int64_t k = offset[i]; k < offset[i + 1]; k++) atomicMax((long long*)&need[index[k]], t);
}

// Cut of running steps too long for a woken particle (blocks.py:404-431): the level follows
// need; the step ends at the latest multiple of the new step <= n + T' if that is earlier, with
// the opening kick corrected to the new length (Durier & Dalla Vecchia 2012, Eq. C5).
__global__ void k_block_cut(int64_t n, int64_t n_now, double tick, const int64_t* __restrict__ need,
                            int64_t* __restrict__ level, int64_t* __restrict__ n_end, const double* __restrict__ a,
                            const double* __restrict__ dudt, const double* __restrict__ visc,
                            const double* __restrict__ cond, const double* __restrict__ disk,
                            double* __restrict__ p_half, double* __restrict__ u_half, double* __restrict__ hv,
                            double* __restrict__ hc, double* __restrict__ hd, double* __restrict__ hp,
                            int32_t* __restrict__ woken)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t nd = need[j];
    if (!(n_end[j] > n_now && nd > level[j])) return;
    const int64_t t_new = ticks_of(nd);
    const int64_t end_new = (n_now / t_new + 1) * t_new;
    level[j] = nd;
    if (!(end_new < n_end[j])) return;
    const double d_len = (double)(end_new - n_end[j]) * tick;
    const double hl = 0.5 * d_len;
    for (int c = 0; c < 3; c++) p_half[3 * j + c] += hl * a[3 * j + c];
    u_half[j] += hl * dudt[j];
    add_heat(j, hl, dudt, visc, cond, disk, hv, hc, hd, hp);
    n_end[j] = end_new;
    atomicAdd(woken, 1);
}

// Opening kicks of the active rows (blocks.py:443-452).
__global__ void k_block_open(const int64_t* __restrict__ rows, int64_t nrows, int64_t n_now, double tick,
                             const int64_t* __restrict__ level, int64_t* __restrict__ n_beg,
                             int64_t* __restrict__ n_end, const double* __restrict__ x, const double* __restrict__ p,
                             const double* __restrict__ u, const double* __restrict__ a,
                             const double* __restrict__ dudt, const double* __restrict__ visc,
                             const double* __restrict__ cond, const double* __restrict__ disk,
                             double* __restrict__ x_beg, double* __restrict__ p_half, double* __restrict__ u_half,
                             double* __restrict__ hv, double* __restrict__ hc, double* __restrict__ hd,
                             double* __restrict__ hp)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    const int64_t t = ticks_of(level[i]);
    const double hl = 0.5 * ((double)t * tick);
    n_beg[i] = n_now;
    n_end[i] = n_now + t;
    for (int c = 0; c < 3; c++) {
        x_beg[3 * i + c] = x[3 * i + c];
        p_half[3 * i + c] = p[3 * i + c] + hl * a[3 * i + c];
    }
    u_half[i] = u[i] + hl * dudt[i];
    add_heat(i, hl, dudt, visc, cond, disk, hv, hc, hd, hp);
}

// The rows and every entry of their lists (blocks.py:202-214).
__global__ void k_block_mark(const int64_t* __restrict__ rows, int64_t nrows, const int64_t* __restrict__ offset,
                             const int32_t* __restrict__ index, uint8_t* __restrict__ mark)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = rows[q];
    mark[i] = 1;
    for (int64_t k = offset[i]; k < offset[i + 1]; k++) mark[index[k]] = 1;
}

static int32_t launched_b(nb_ctx* ctx, const char* what)
{
    return hipGetLastError() == hipSuccess ? NB_OK : ctx_fail(ctx, what);
}

}  // namespace nb

using namespace nb;

extern "C" int32_t nb_block_elapsed(nb_ctx* ctx, void* stream, const int64_t* n_beg, const int64_t* n_end,
                                    int64_t n_next, double tick, int64_t n, double* s_out, double* half_out)
{
    if (!ctx || n < 0 || (n > 0 && (!n_beg || !n_end || !s_out || !half_out)))
        return ctx_fail(ctx, "nb_block_elapsed: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_elapsed, dim3(grid_b(n)), dim3(kBlockB), 0, (hipStream_t)stream, n_beg, n_end, n_next,
                       tick, n, s_out, half_out);
    return launched_b(ctx, "nb_block_elapsed: launch failed");
}

extern "C" int32_t nb_block_predict(nb_ctx* ctx, void* stream, const int64_t* n_beg, const int64_t* n_end,
                                    int64_t n_next, double tick, const double* p_flow, const double* a,
                                    double* u_half, const double* dudt, const double* x_beg, const double* x,
                                    const double* mass, double u_floor, const double* u_floors, double beta,
                                    int64_t n, double* p_out, double* u_out, double* e_part, double* lum,
                                    double* heat_cool)
{
    const bool cool = beta < INFINITY;
    if (!ctx || n < 0 || (n > 0 && (!n_beg || !n_end || !p_flow || !a || !u_half || !dudt || !p_out || !u_out)) ||
        (n > 0 && cool && (!x_beg || !x || !mass || !e_part || !lum || !heat_cool)))
        return ctx_fail(ctx, "nb_block_predict: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_predict, dim3(grid_b(n)), dim3(kBlockB), 0, (hipStream_t)stream, n_beg, n_end, n_next,
                       tick, p_flow, a, u_half, dudt, x_beg, x, mass, u_floor, u_floors, beta, n, p_out, u_out, e_part,
                       lum, heat_cool);
    return launched_b(ctx, "nb_block_predict: launch failed");
}

extern "C" int32_t nb_block_close(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, const int64_t* n_beg,
                                  const int64_t* n_end, double tick, const double* p_flow, const double* a,
                                  const double* u_half, const double* dudt, const double* visc, const double* cond,
                                  const double* disk, double* p, double* u, double* heat_visc, double* heat_cond,
                                  double* heat_disk, double* heat_pdv)
{
    if (!ctx || nrows < 0 || (nrows > 0 && (!rows || !n_beg || !n_end || !p_flow || !a || !u_half || !dudt || !visc ||
                                            !cond || !disk || !p || !u || !heat_visc || !heat_cond || !heat_disk ||
                                            !heat_pdv)))
        return ctx_fail(ctx, "nb_block_close: bad arguments");
    if (nrows == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_close, dim3(grid_b(nrows)), dim3(kBlockB), 0, (hipStream_t)stream, rows, nrows, n_beg,
                       n_end, tick, p_flow, a, u_half, dudt, visc, cond, disk, p, u, heat_visc, heat_cond, heat_disk,
                       heat_pdv);
    return launched_b(ctx, "nb_block_close: launch failed");
}

extern "C" int32_t nb_block_levels(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, const double* x,
                                   const double* h, const double* vsig, const double* a, const double* nu,
                                   double courant, double c_force, double eta_orb, int32_t hydro, double dt_max,
                                   const int64_t* level_offset, int64_t n_now, int64_t* level, int32_t* bad_dev)
{
    if (!ctx || nrows < 0 || !bad_dev || !(dt_max > 0) ||
        (nrows > 0 && (!rows || !x || !level_offset || !level || (hydro && (!h || !vsig || !a)))))
        return ctx_fail(ctx, "nb_block_levels: bad arguments");
    if (nrows == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_levels, dim3(grid_b(nrows)), dim3(kBlockB), 0, (hipStream_t)stream, rows, nrows, x, h,
                       vsig, a, nu, courant, c_force, eta_orb, (int)hydro, dt_max, level_offset, n_now, level, bad_dev);
    return launched_b(ctx, "nb_block_levels: launch failed");
}

extern "C" int32_t nb_block_limit(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, const int64_t* offset,
                                  const int32_t* index, int64_t gap, int64_t* level, int32_t* changed_dev)
{
    if (!ctx || nrows < 0 || !changed_dev || (nrows > 0 && (!rows || !offset || !index || !level)))
        return ctx_fail(ctx, "nb_block_limit: bad arguments");
    if (nrows == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_limit, dim3(grid_b(nrows)), dim3(kBlockB), 0, (hipStream_t)stream, rows, nrows, offset,
                       index, gap, (unsigned long long*)level, changed_dev);
    return launched_b(ctx, "nb_block_limit: launch failed");
}

extern "C" int32_t nb_block_land(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, int64_t n_now,
                                 int64_t n_stop, int64_t* level, int32_t* bad_dev)
{
    if (!ctx || nrows < 0 || !bad_dev || (nrows > 0 && (!rows || !level)))
        return ctx_fail(ctx, "nb_block_land: bad arguments");
    if (nrows == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_land, dim3(grid_b(nrows)), dim3(kBlockB), 0, (hipStream_t)stream, rows, nrows, n_now,
                       n_stop, level, bad_dev);
    return launched_b(ctx, "nb_block_land: launch failed");
}

extern "C" int32_t nb_block_wake(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, const int64_t* offset,
                                 const int32_t* index, int64_t gap, int64_t n, int64_t n_now, double tick,
                                 int64_t* need, int64_t* level, int64_t* n_end, const double* a, const double* dudt,
                                 const double* visc, const double* cond, const double* disk, double* p_half,
                                 double* u_half, double* heat_visc, double* heat_cond, double* heat_disk,
                                 double* heat_pdv, int32_t* woken_dev)
{
    if (!ctx || nrows < 0 || n < 0 || !woken_dev ||
        (n > 0 && (!need || !level || !n_end || !a || !dudt || !visc || !cond || !disk || !p_half || !u_half ||
                   !heat_visc || !heat_cond || !heat_disk || !heat_pdv)) ||
        (nrows > 0 && (!rows || !offset || !index)))
        return ctx_fail(ctx, "nb_block_wake: bad arguments");
    if (n == 0) return NB_OK;
    hipStream_t s = (hipStream_t)stream;
    if (nrows > 0)
        hipLaunchKernelGGL(k_block_wake_targets, dim3(grid_b(nrows)), dim3(kBlockB), 0, s, rows, nrows, offset, index,
                           gap, level, need);
    hipLaunchKernelGGL(k_block_cut, dim3(grid_b(n)), dim3(kBlockB), 0, s, n, n_now, tick, need, level, n_end, a, dudt,
                       visc, cond, disk, p_half, u_half, heat_visc, heat_cond, heat_disk, heat_pdv, woken_dev);
    return launched_b(ctx, "nb_block_wake: launch failed");
}

extern "C" int32_t nb_block_open(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, int64_t n_now,
                                 double tick, const int64_t* level, int64_t* n_beg, int64_t* n_end, const double* x,
                                 const double* p, const double* u, const double* a, const double* dudt,
                                 const double* visc, const double* cond, const double* disk, double* x_beg,
                                 double* p_half, double* u_half, double* heat_visc, double* heat_cond,
                                 double* heat_disk, double* heat_pdv)
{
    if (!ctx || nrows < 0 ||
        (nrows > 0 && (!rows || !level || !n_beg || !n_end || !x || !p || !u || !a || !dudt || !visc || !cond ||
                       !disk || !x_beg || !p_half || !u_half || !heat_visc || !heat_cond || !heat_disk || !heat_pdv)))
        return ctx_fail(ctx, "nb_block_open: bad arguments");
    if (nrows == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_open, dim3(grid_b(nrows)), dim3(kBlockB), 0, (hipStream_t)stream, rows, nrows, n_now,
                       tick, level, n_beg, n_end, x, p, u, a, dudt, visc, cond, disk, x_beg, p_half, u_half, heat_visc,
                       heat_cond, heat_disk, heat_pdv);
    return launched_b(ctx, "nb_block_open: launch failed");
}

extern "C" int32_t nb_block_mark(nb_ctx* ctx, void* stream, const int64_t* rows, int64_t nrows, const int64_t* offset,
                                 const int32_t* index, uint8_t* mark)
{
    if (!ctx || nrows < 0 || (nrows > 0 && (!rows || !offset || !index || !mark)))
        return ctx_fail(ctx, "nb_block_mark: bad arguments");
    if (nrows == 0) return NB_OK;
    hipLaunchKernelGGL(k_block_mark, dim3(grid_b(nrows)), dim3(kBlockB), 0, (hipStream_t)stream, rows, nrows, offset,
                       index, mark);
    return launched_b(ctx, "nb_block_mark: launch failed");
}
