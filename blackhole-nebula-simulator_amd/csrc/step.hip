// Per-particle work of the step: GN generalized leapfrog (integrator.py:106-198), v = dH/dp
// (fastphys.py:24-42), half kicks, equation of state, cooling, and the step's reductions
// (time-step minima, energy / angular-momentum ledger). All FP64, one work-item per particle.
#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

#include <mutex>
#include <string>

namespace nb {

static constexpr int kBlockS = 256;

// integrator.py:106-120
__device__ __forceinline__ void dh_dx(double x0, double x1, double x2, double p0, double p1, double p2,
                                      double& g0, double& g1, double& g2)
{
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double c = -2.0 * (r - 2.0) / ipow4(r);
    double da = 2.0 / (r * r);
    double dc = (6.0 * r - 16.0) / ipow5(r);
    double px = p0 * x0 + p1 * x1 + p2 * x2;
    double pp = p0 * p0 + p1 * p1 + p2 * p2;
    double radial = (0.5 * da * pp + 0.5 * dc * px * px) / r + 1.0 / ipow3(r);
    g0 = radial * x0 + c * px * p0;
    g1 = radial * x1 + c * px * p1;
    g2 = radial * x2 + c * px * p2;
}

// integrator.py:123-130
__device__ __forceinline__ void dh_dp(double x0, double x1, double x2, double p0, double p1, double p2,
                                      double& v0, double& v1, double& v2)
{
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double a = (r - 2.0) / r;
    double c = -2.0 * (r - 2.0) / ipow4(r);
    double cpx = c * (p0 * x0 + p1 * x1 + p2 * x2);
    v0 = a * p0 + cpx * x0;
    v1 = a * p1 + cpx * x1;
    v2 = a * p2 + cpx * x2;
}

// integrator.py:133-136
__device__ __forceinline__ bool converged(double n0, double n1, double n2, double o0, double o1, double o2)
{
    double e0 = n0 - o0, e1 = n1 - o1, e2 = n2 - o2;
    double d = e0 * e0 + e1 * e1 + e2 * e2;
    return d <= 1e-15 * 1e-15 * (n0 * n0 + n1 * n1 + n2 * n2);
}

// integrator.py:139-185. dt_arr (per particle) or dt_scalar; iters: count, or -1 / -2 / -3
// for stage-1 / stage-2 non-convergence / r <= 2.5 (then x, p are returned unchanged).
__global__ void __launch_bounds__(kBlockS)
k_glf64(const double* __restrict__ x, const double* __restrict__ p, const double* __restrict__ dt_arr,
        double dt_scalar, int64_t n, double* __restrict__ x_out, double* __restrict__ p_out,
        int32_t* __restrict__ iters, int32_t* __restrict__ nfail)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    const double p0 = p[3 * i], p1 = p[3 * i + 1], p2 = p[3 * i + 2];
    const double d = dt_arr ? dt_arr[i] : dt_scalar;
    double o[6] = {x0, x1, x2, p0, p1, p2};
    int32_t code;
    if (x0 * x0 + x1 * x1 + x2 * x2 <= 2.5 * 2.5) {
        code = -3;
    } else {
        const double half = 0.5 * d;
        int32_t it = 0;
        double h0 = p0, h1 = p1, h2 = p2;
        bool ok = false;
        for (int k = 0; k < 10; k++) {
            double g0, g1, g2;
            dh_dx(x0, x1, x2, h0, h1, h2, g0, g1, g2);
            double n0 = p0 - half * g0, n1 = p1 - half * g1, n2 = p2 - half * g2;
            it += 1;
            ok = converged(n0, n1, n2, h0, h1, h2);
            h0 = n0; h1 = n1; h2 = n2;
            if (ok) break;
        }
        if (!ok) {
            code = -1;
        } else {
            double v0, v1, v2;
            dh_dp(x0, x1, x2, h0, h1, h2, v0, v1, v2);
            double y0 = x0 + d * v0, y1 = x1 + d * v1, y2 = x2 + d * v2;
            ok = false;
            for (int k = 0; k < 10; k++) {
                double w0, w1, w2;
                dh_dp(y0, y1, y2, h0, h1, h2, w0, w1, w2);
                double n0 = x0 + half * (v0 + w0), n1 = x1 + half * (v1 + w1), n2 = x2 + half * (v2 + w2);
                it += 1;
                ok = converged(n0, n1, n2, y0, y1, y2);
                y0 = n0; y1 = n1; y2 = n2;
                if (ok) break;
            }
            if (!ok) {
                code = -2;
            } else {
                double g0, g1, g2;
                dh_dx(y0, y1, y2, h0, h1, h2, g0, g1, g2);
                o[0] = y0; o[1] = y1; o[2] = y2;
                o[3] = h0 - half * g0; o[4] = h1 - half * g1; o[5] = h2 - half * g2;
                code = it;
            }
        }
    }
    for (int c = 0; c < 3; c++) { x_out[3 * i + c] = o[c]; p_out[3 * i + c] = o[3 + c]; }
    iters[i] = code;
    if (code < 0 && nfail) atomicAdd(nfail, 1);
}

// fastphys.py:24-32
__global__ void k_v_from_p64(const double* __restrict__ x, const double* __restrict__ p, int64_t n,
                             double* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
    double p0 = p[3 * i], p1 = p[3 * i + 1], p2 = p[3 * i + 2];
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double a = (r - 2.0) / r;
    double c = -2.0 * (r - 2.0) / ipow4(r);
    double px = p0 * x0 + p1 * x1 + p2 * x2;
    out[3 * i] = a * p0 + c * px * x0;
    out[3 * i + 1] = a * p1 + c * px * x1;
    out[3 * i + 2] = a * p2 + c * px * x2;
}

// out = y + s * a over m doubles (coupled.py:229-235 kicks: p + (dt/2) a; u + (dt/2) du/dt)
__global__ void k_axpy(const double* __restrict__ y, const double* __restrict__ a, double s, int64_t m,
                       double* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) out[i] = y[i] + s * a[i];
}

// k_axpy with s = scale * s_dev[0] read on the device (a step whose dt the host writes between
// replays of a recorded graph); the same arithmetic as k_axpy with s formed on the host
__global__ void k_axpy_dev(const double* __restrict__ y, const double* __restrict__ a,
                           const double* __restrict__ s_dev, double scale, int64_t m, double* __restrict__ out)
{
    const double s = scale * s_dev[0];
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) out[i] = y[i] + s * a[i];
}

// k_axpy_dev over two pairs in one launch: out3 = y3 + s a3 (3 n values), out1 = y1 + s a1 (n):
// a momentum and an energy kick of the same dt
__global__ void k_kick2_dev(const double* __restrict__ y3, const double* __restrict__ a3,
                            const double* __restrict__ y1, const double* __restrict__ a1,
                            const double* __restrict__ s_dev, double scale, int64_t n, double* __restrict__ out3,
                            double* __restrict__ out1)
{
    const double s = scale * s_dev[0];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < 3 * n) out3[i] = y3[i] + s * a3[i];
    else if (i < 4 * n) out1[i - 3 * n] = y1[i - 3 * n] + s * a1[i - 3 * n];
}

// eos.py:21-28 with hydro.py:130-131: P = (gamma-1) rho u, c_s = sqrt(gamma (gamma-1) u);
// counts particles with u <= 0.
__global__ void k_eos_ideal(const double* __restrict__ rho, const double* __restrict__ u, double gm1,
                            double ggm1, int64_t n, double* __restrict__ pressure, double* __restrict__ cs,
                            int32_t* __restrict__ nbad)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ui = u[i];
    if (ui <= 0.0) atomicAdd(nbad, 1);
    pressure[i] = gm1 * rho[i] * ui;
    cs[i] = sqrt(ggm1 * ui);
}

// ----------------------------------------------------------------- gas + radiation EOS
// fastphys.py:45-111 with eos.py:76-117 (local thermodynamic equilibrium, one temperature):
//   u = 3 k T / (2 mu m_p) + a T^4 / rho,  P = rho k T / (mu m_p) + a T^4 / 3  (SI),
//   beta = P_gas / P,  Gamma_1 = beta + (4 - 3 beta)^2 (gamma - 1) / (beta + 12 (gamma - 1)(1 - beta)),
//   c_s = sqrt(Gamma_1 P / rho),
// with T from (rho, u) by Newton on log T inside the bracket [log 1, log 1e13] K, which every
// step narrows (a step leaving it, or not finite, bisects), from the smaller of the gas-only and
// radiation-only temperatures, until a step in log T is below 1e-12. SI constants are those of
// units.py (CODATA 2018).
constexpr double kKB = 1.380649e-23, kMP = 1.67262192369e-27, kMuGas = 0.6, kCLight = 299792458.0;
constexpr double kARad = 4.0 * 5.670374419e-8 / kCLight;
constexpr double kTMin = 1.0, kTMax = 1.0e13, kEosTol = 1e-12;
constexpr int kEosMaxIter = 200;

__device__ __forceinline__ double gasrad_u_si(double rho_si, double t)
{
    const double t4 = (t * t) * (t * t);
    return 1.5 * kKB * t / (kMuGas * kMP) + kARad * t4 / rho_si;
}

// P [code], c_s, T [K] of code rho and u (dens: kg/m^3 per code density). bad[0] counts u outside
// u(1 K) .. u(1e13 K) (the reference's ValueError), bad[1] solves that did not converge.
__global__ void k_eos_gasrad(const double* __restrict__ rho, const double* __restrict__ u, double dens, int64_t n,
                             double* __restrict__ pressure, double* __restrict__ cs, double* __restrict__ t_out,
                             int32_t* __restrict__ bad)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = rho[i] * dens, uu = u[i] * (kCLight * kCLight);
    if (!(uu >= gasrad_u_si(r, kTMin) && uu <= gasrad_u_si(r, kTMax))) {
        atomicAdd(&bad[0], 1);
        pressure[i] = cs[i] = NAN;
        if (t_out) t_out[i] = NAN;
        return;
    }
    const double a1 = 1.5 * kKB / (kMuGas * kMP);
    double lo = log(kTMin), hi = log(kTMax);
    const double t0 = fmin(uu / a1, pow(uu * r / kARad, 0.25));
    double lt = log(fmin(fmax(t0, kTMin), kTMax));
    bool ok = false;
    for (int it = 0; it < kEosMaxIter; it++) {
        const double t = exp(lt);
        const double t4 = (t * t) * (t * t);
        const double f = a1 * t + kARad * t4 / r - uu;
        const double dfdl = a1 * t + 4.0 * kARad * t4 / r;
        if (f < 0.0) lo = lt;
        if (f > 0.0) hi = lt;
        double nw = lt - f / dfdl;
        if (nw <= lo || nw >= hi || !isfinite(nw)) nw = 0.5 * (lo + hi);
        const double step = fabs(nw - lt);
        lt = nw;
        if (step < kEosTol) {
            ok = true;
            break;
        }
    }
    if (!ok) atomicAdd(&bad[1], 1);
    const double t = exp(lt), t4 = (t * t) * (t * t);
    const double p_gas = r * kKB * t / (kMuGas * kMP), p_rad = kARad * t4 / 3.0;
    const double beta = p_gas / (p_gas + p_rad);
    const double p_code = (p_gas + p_rad) / (dens * (kCLight * kCLight));
    const double g = 5.0 / 3.0 - 1.0;
    const double g1 = beta + (4.0 - 3.0 * beta) * (4.0 - 3.0 * beta) * g / (beta + 12.0 * g * (1.0 - beta));
    pressure[i] = p_code;
    cs[i] = sqrt(g1 * p_code / rho[i]);
    if (t_out) t_out[i] = t;
}

// hydro.py:155-163 (gas + radiation): the code u of gas at temperature t_k at each density
__global__ void k_gasrad_u(const double* __restrict__ rho, double dens, double t_k, int64_t n, double* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = gasrad_u_si(rho[i] * dens, t_k) / (kCLight * kCLight);
}

// cooling.py:33-40 with coupled.py:244-251: u_floor + (u - u_floor) exp(-dt / t_cool),
// t_cool = beta r_mid^1.5, r_mid = (|x_old| + |x_new|)/2. e_part[i] = m (u - u_new). The floor is
// one number (ideal gas) or one per particle (floors: gas + radiation, set by the density).
__global__ void k_cool_beta(const double* __restrict__ u, const double* __restrict__ x_old,
                            const double* __restrict__ x_new, const double* __restrict__ mass, double u_floor_one,
                            const double* __restrict__ floors, double dt, double beta, int64_t n,
                            double* __restrict__ u_out, double* __restrict__ e_part)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u_floor = floors ? floors[i] : u_floor_one;
    double r0 = sqrt(x_old[3 * i] * x_old[3 * i] + x_old[3 * i + 1] * x_old[3 * i + 1] + x_old[3 * i + 2] * x_old[3 * i + 2]);
    double r1 = sqrt(x_new[3 * i] * x_new[3 * i] + x_new[3 * i + 1] * x_new[3 * i + 1] + x_new[3 * i + 2] * x_new[3 * i + 2]);
    double r_mid = 0.5 * (r0 + r1);
    double t_cool = beta / pow(r_mid, -1.5);
    double decay = exp(-dt / t_cool);
    double un = u_floor + (u[i] - u_floor) * decay;
    e_part[i] = mass[i] * (u[i] - un);
    u_out[i] = un;
}

// Optically thin power-law cooling du/dt = -k rho sqrt(u) above u_floor, integrated exactly at
// fixed rho: sqrt(u) falls linearly by k rho dt / 2, stopping at the floor.
__global__ void k_cool_powerlaw(const double* __restrict__ u, const double* __restrict__ rho,
                                const double* __restrict__ mass, double u_floor, double k, double dt_host,
                                const double* __restrict__ dt_dev, int64_t n, double* __restrict__ u_out,
                                double* __restrict__ e_part)
{
    const double dt = dt_dev ? dt_dev[0] : dt_host;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ui = u[i];
    double un = ui;
    if (ui > u_floor) {
        double s = sqrt(ui) - 0.5 * k * rho[i] * dt;
        double sf = sqrt(u_floor);
        un = s > sf ? s * s : u_floor;
    }
    e_part[i] = mass[i] * (ui - un);
    u_out[i] = un;
}

// ----------------------------------------------------------------- reductions
// Time-step minima (coupled.py:213-226): out[0] = min h/v_sig, out[1] = min sqrt(h/|a|) over
// |a| > 0, out[2] = min viscous dt C h^2/nu over nu > 0 (if nu given), out[3] = min |x|.
// Minima are exact whatever the reduction order. A NaN term makes its minimum NaN, as np.min
// does (fmin would drop it): a particle that blew up then gives the reference's NaN dt instead
// of a finite one taken from the others.
__device__ __forceinline__ double nan_min(double a, double b) { return (a < b || a != a) ? a : b; }
__global__ void k_dt_partial(const double* __restrict__ h, const double* __restrict__ vsig,
                             const double* __restrict__ acc, const double* __restrict__ nu, double c_visc,
                             const double* __restrict__ x, int64_t n, double* __restrict__ part)
{
    double v[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double hi = h[i];
        v[0] = nan_min(v[0], hi / vsig[i]);
        double a0 = acc[3 * i], a1 = acc[3 * i + 1], a2 = acc[3 * i + 2];
        double am = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        if (am > 0.0) v[1] = nan_min(v[1], sqrt(hi / am));
        if (nu && nu[i] > 0.0) v[2] = nan_min(v[2], c_visc * (hi * hi) / nu[i]);
        double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        v[3] = nan_min(v[3], sqrt(x0 * x0 + x1 * x1 + x2 * x2));
    }
    __shared__ double sh[4][kBlockS];
    for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = nan_min(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[4 * blockIdx.x + k] = sh[k][0];
}

// k_dt_partial per segment of an ensemble: one workgroup per segment [off[s], off[s + 1]) writes
// its four minima to out[4 s + k].
__global__ void __launch_bounds__(kBlockS)
k_dt_seg(const double* __restrict__ h, const double* __restrict__ vsig, const double* __restrict__ acc,
         const double* __restrict__ nu, double c_visc, const double* __restrict__ x, const int64_t* __restrict__ off,
         double* __restrict__ out)
{
    double v[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    for (int64_t i = off[blockIdx.x] + threadIdx.x; i < off[blockIdx.x + 1]; i += blockDim.x) {
        const double hi = h[i];
        v[0] = nan_min(v[0], hi / vsig[i]);
        const double a0 = acc[3 * i], a1 = acc[3 * i + 1], a2 = acc[3 * i + 2];
        const double am = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        if (am > 0.0) v[1] = nan_min(v[1], sqrt(hi / am));
        if (nu && nu[i] > 0.0) v[2] = nan_min(v[2], c_visc * (hi * hi) / nu[i]);
        const double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        v[3] = nan_min(v[3], sqrt(x0 * x0 + x1 * x1 + x2 * x2));
    }
    __shared__ double sh[4][kBlockS];
    for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = nan_min(sh[k][threadIdx.x], sh[k][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) out[4 * blockIdx.x + k] = sh[k][0];
}

// out row i = y row i + s[i] a row i (rows of w values): a kick with a step per particle
__global__ void k_axpy_rows(const double* __restrict__ y, const double* __restrict__ a, const double* __restrict__ s,
                            int64_t n, int w, double* __restrict__ out)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * w) return;
    out[t] = y[t] + s[t / w] * a[t];
}

// out[i] = scale * per_seg[cloud[i]]
__global__ void k_expand_seg(const double* __restrict__ per_seg, const int32_t* __restrict__ cloud, int64_t n,
                             double scale, double* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = scale * per_seg[cloud[i]];
}

// Ledger sums (coupled.py:261-270): out = [sum m (H + u), sum m |H|, sum m (x cross p)_xyz,
// sum m (u - u_cool) extra term (e_part, may be null)]. H is the GN Hamiltonian
// (gravity.py:92-102) or |p|^2/2 without the hole.
__device__ __forceinline__ double gn_h(double x0, double x1, double x2, double p0, double p1, double p2)
{
    double r = sqrt(x0 * x0 + x1 * x1 + x2 * x2);
    double a = (r - 2.0) / r;
    double c = -2.0 * (r - 2.0) / ipow4(r);
    double pp = p0 * p0 + p1 * p1 + p2 * p2;
    double px = p0 * x0 + p1 * x1 + p2 * x2;
    return 0.5 * a * pp + 0.5 * c * (px * px) - 1.0 / r;
}

__global__ void k_ledger_partial(const double* __restrict__ x, const double* __restrict__ p,
                                 const double* __restrict__ u, const double* __restrict__ mass,
                                 const double* __restrict__ e_part, int gn, int64_t n, double* __restrict__ part)
{
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double x0 = x[3 * i], x1 = x[3 * i + 1], x2 = x[3 * i + 2];
        double p0 = p[3 * i], p1 = p[3 * i + 1], p2 = p[3 * i + 2];
        double hh = gn ? gn_h(x0, x1, x2, p0, p1, p2) : 0.5 * (p0 * p0 + p1 * p1 + p2 * p2);
        double m = mass[i];
        v[0] += m * (hh + (u ? u[i] : 0.0));
        v[1] += m * fabs(hh);
        v[2] += m * (x1 * p2 - x2 * p1);
        v[3] += m * (x2 * p0 - x0 * p2);
        v[4] += m * (x0 * p1 - x1 * p0);
        if (e_part) v[5] += e_part[i];
    }
    __shared__ double sh[6][kBlockS];
    for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 6; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[6 * blockIdx.x + k] = sh[k][0];
}

// Final pass of a reduction over `nparts` partial rows of width w (w <= 8): op 0 = min, 1 = sum.
// One block: min in any order; sums as per-thread partials combined in a fixed tree.
__global__ void k_reduce_final(const double* __restrict__ part, int nparts, int w, int op, double* __restrict__ out)
{
    __shared__ double sh[8][kBlockS];
    for (int k = 0; k < w; k++) {
        double v = op == 0 ? INFINITY : 0.0;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
            double y = part[w * b + k];
            v = op == 0 ? nan_min(v, y) : v + y;
        }
        sh[k][threadIdx.x] = v;
    }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < w; k++)
                sh[k][threadIdx.x] = op == 0 ? nan_min(sh[k][threadIdx.x], sh[k][threadIdx.x + s])
                                             : sh[k][threadIdx.x] + sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < w; k++) out[k] = sh[k][0];
}

static inline unsigned grid_s(int64_t n) { return (unsigned)((n + kBlockS - 1) / kBlockS); }

}  // namespace nb

using namespace nb;

static int32_t launch_check_s(nb_ctx* ctx, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        return ctx_fail(ctx, m.c_str());
    }
    return NB_OK;
}

extern "C" int32_t nb_glf_step(nb_ctx* ctx, void* stream, const double* x, const double* p, const double* dt_arr,
                               double dt_scalar, int64_t n, double* x_out, double* p_out, int32_t* iters,
                               int32_t* nfail_dev)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_glf_step: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_glf64, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, x, p, dt_arr, dt_scalar, n,
                       x_out, p_out, iters, nfail_dev);
    return launch_check_s(ctx, "nb_glf_step");
}

extern "C" int32_t nb_gn_v_from_p(nb_ctx* ctx, void* stream, const double* x, const double* p, int64_t n,
                                  double* out)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_gn_v_from_p: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_v_from_p64, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, x, p, n, out);
    return launch_check_s(ctx, "nb_gn_v_from_p");
}

extern "C" int32_t nb_axpy(nb_ctx* ctx, void* stream, const double* y, const double* a, double s, int64_t m,
                           double* out)
{
    if (!ctx || m < 0) return ctx_fail(ctx, "nb_axpy: bad arguments");
    if (m == 0) return NB_OK;
    hipLaunchKernelGGL(k_axpy, dim3(grid_s(m)), dim3(kBlockS), 0, (hipStream_t)stream, y, a, s, m, out);
    return launch_check_s(ctx, "nb_axpy");
}

extern "C" int32_t nb_axpy_dev(nb_ctx* ctx, void* stream, const double* y, const double* a, const double* s_dev,
                               double scale, int64_t m, double* out)
{
    if (!ctx || m < 0 || !s_dev) return ctx_fail(ctx, "nb_axpy_dev: bad arguments");
    if (m == 0) return NB_OK;
    hipLaunchKernelGGL(k_axpy_dev, dim3(grid_s(m)), dim3(kBlockS), 0, (hipStream_t)stream, y, a, s_dev, scale, m, out);
    return launch_check_s(ctx, "nb_axpy_dev");
}

extern "C" int32_t nb_kick2_dev(nb_ctx* ctx, void* stream, const double* y3, const double* a3, const double* y1,
                                const double* a1, const double* s_dev, double scale, int64_t n, double* out3,
                                double* out1)
{
    if (!ctx || n < 0 || !s_dev || !y3 || !a3 || !y1 || !a1 || !out3 || !out1)
        return ctx_fail(ctx, "nb_kick2_dev: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_kick2_dev, dim3(grid_s(4 * n)), dim3(kBlockS), 0, (hipStream_t)stream, y3, a3, y1, a1, s_dev,
                       scale, n, out3, out1);
    return launch_check_s(ctx, "nb_kick2_dev");
}

extern "C" int32_t nb_eos_ideal(nb_ctx* ctx, void* stream, const double* rho, const double* u, double gamma,
                                int64_t n, double* pressure, double* cs, int32_t* nbad_dev)
{
    if (!ctx || n < 0 || !nbad_dev) return ctx_fail(ctx, "nb_eos_ideal: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_eos_ideal, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, rho, u, gamma - 1.0,
                       gamma * (gamma - 1.0), n, pressure, cs, nbad_dev);
    return launch_check_s(ctx, "nb_eos_ideal");
}

extern "C" int32_t nb_cool_beta(nb_ctx* ctx, void* stream, const double* u, const double* x_old,
                                const double* x_new, const double* mass, double u_floor, double dt, double beta,
                                int64_t n, double* u_out, double* e_part)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_cool_beta: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_cool_beta, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, u, x_old, x_new, mass,
                       u_floor, (const double*)nullptr, dt, beta, n, u_out, e_part);
    return launch_check_s(ctx, "nb_cool_beta");
}

extern "C" int32_t nb_cool_beta_floors(nb_ctx* ctx, void* stream, const double* u, const double* x_old,
                                       const double* x_new, const double* mass, const double* u_floor, double dt,
                                       double beta, int64_t n, double* u_out, double* e_part)
{
    if (!ctx || n < 0 || (n > 0 && !u_floor)) return ctx_fail(ctx, "nb_cool_beta_floors: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_cool_beta, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, u, x_old, x_new, mass,
                       0.0, u_floor, dt, beta, n, u_out, e_part);
    return launch_check_s(ctx, "nb_cool_beta_floors");
}

extern "C" int32_t nb_eos_gasrad(nb_ctx* ctx, void* stream, const double* rho, const double* u, double dens_kg_m3,
                                 int64_t n, double* pressure, double* cs, double* t_out, int32_t* bad_dev)
{
    if (!ctx || n < 0 || !bad_dev || !(dens_kg_m3 > 0.0)) return ctx_fail(ctx, "nb_eos_gasrad: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_eos_gasrad, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, rho, u, dens_kg_m3, n,
                       pressure, cs, t_out, bad_dev);
    return launch_check_s(ctx, "nb_eos_gasrad");
}

extern "C" int32_t nb_gasrad_u(nb_ctx* ctx, void* stream, const double* rho, double dens_kg_m3, double t_k, int64_t n,
                               double* u_out)
{
    if (!ctx || n < 0 || !(dens_kg_m3 > 0.0)) return ctx_fail(ctx, "nb_gasrad_u: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_gasrad_u, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, rho, dens_kg_m3, t_k, n,
                       u_out);
    return launch_check_s(ctx, "nb_gasrad_u");
}

extern "C" int32_t nb_cool_powerlaw(nb_ctx* ctx, void* stream, const double* u, const double* rho,
                                    const double* mass, double u_floor, double k, double dt, int64_t n,
                                    double* u_out, double* e_part)
{
    if (!ctx || n < 0) return ctx_fail(ctx, "nb_cool_powerlaw: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_cool_powerlaw, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, u, rho, mass,
                       u_floor, k, dt, (const double*)nullptr, n, u_out, e_part);
    return launch_check_s(ctx, "nb_cool_powerlaw");
}

extern "C" int32_t nb_cool_powerlaw_dev(nb_ctx* ctx, void* stream, const double* u, const double* rho,
                                        const double* mass, double u_floor, double k, const double* dt_dev,
                                        int64_t n, double* u_out, double* e_part)
{
    if (!ctx || n < 0 || !dt_dev) return ctx_fail(ctx, "nb_cool_powerlaw_dev: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_cool_powerlaw, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, u, rho, mass,
                       u_floor, k, 0.0, dt_dev, n, u_out, e_part);
    return launch_check_s(ctx, "nb_cool_powerlaw_dev");
}

static int32_t reduce_two_level(nb_ctx* ctx, hipStream_t s, int w, int op, int nparts, double* host_out)
{
    double* fin = ctx->red + (size_t)w * nparts;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlockS), 0, s, ctx->red, nparts, w, op, fin);
    if (hipMemcpyAsync(host_out, fin, sizeof(double) * w, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "reduction copy failed");
    return NB_OK;
}

extern "C" int32_t nb_dt_minima(nb_ctx* ctx, void* stream, const double* h, const double* vsig, const double* acc,
                                const double* nu, double c_visc, const double* x, int64_t n, double* out4)
{
    if (!ctx || n <= 0 || !out4) return ctx_fail(ctx, "nb_dt_minima: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    int nparts = (int)std::min<int64_t>(1024, (n + kBlockS - 1) / kBlockS);
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 4 * (nparts + 1))) return NB_ERR_HIP;
    hipLaunchKernelGGL(k_dt_partial, dim3(nparts), dim3(kBlockS), 0, s, h, vsig, acc, nu, c_visc, x, n, ctx->red);
    return reduce_two_level(ctx, s, 4, 0, nparts, out4);
}

extern "C" int32_t nb_dt_minima_dev(nb_ctx* ctx, void* stream, const double* h, const double* vsig,
                                    const double* acc, const double* nu, double c_visc, const double* x, int64_t n,
                                    double* out4_dev)
{
    if (!ctx || n <= 0 || !out4_dev || !h || !vsig || !acc || !x) return ctx_fail(ctx, "nb_dt_minima_dev: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    int nparts = (int)std::min<int64_t>(1024, (n + kBlockS - 1) / kBlockS);
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 4 * (nparts + 1))) return NB_ERR_HIP;
    hipLaunchKernelGGL(k_dt_partial, dim3(nparts), dim3(kBlockS), 0, s, h, vsig, acc, nu, c_visc, x, n, ctx->red);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlockS), 0, s, ctx->red, nparts, 4, 0, out4_dev);
    return launch_check_s(ctx, "nb_dt_minima_dev");
}

// The next step's dt from the four minima of nb_dt_minima_dev, on the device: coupled.py:213-226
// (stable_dt) as the host takes it from the same minima (the Courant, force, viscous and orbit
// criteria in that order, each kept only where strictly smaller, as Python's min(dt, c)), with
// r^1.5 from pow15. dts = (dt, dt_prev) becomes (the new dt, the old dt). For a driver that
// launches the next step before reading this one back (it compares the value with the host's
// when it does).
__global__ void k_dt_advance(const double* __restrict__ m, double courant, double c_force, int viscous, int hole,
                             double eta_orb, double* __restrict__ dts)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double dt = courant * m[0];
    if (isfinite(m[1])) dt = pymin(dt, c_force * m[1]);
    if (viscous) dt = pymin(dt, m[2]);
    if (hole) dt = pymin(dt, eta_orb * pow15(m[3]));
    dts[1] = dts[0];
    dts[0] = dt;
}

extern "C" int32_t nb_dt_advance(nb_ctx* ctx, void* stream, const double* minima4_dev, double courant,
                                 double c_force, int32_t viscous, int32_t hole, double eta_orb, double* dts2_dev)
{
    if (!ctx || !minima4_dev || !dts2_dev) return ctx_fail(ctx, "nb_dt_advance: bad arguments");
    hipLaunchKernelGGL(k_dt_advance, dim3(1), dim3(64), 0, (hipStream_t)stream, minima4_dev, courant, c_force,
                       (int)(viscous != 0), (int)(hole != 0), eta_orb, dts2_dev);
    return launch_check_s(ctx, "nb_dt_advance");
}

// ------------------------------------------------------- read-back of a step launched ahead
// The counts a step driver reads back, reduced on the device: acc[0] += the per-row gather and
// extras counts (the search-list entries), acc[1] += the tight-row lengths (integral doubles),
// acc[2] += the rows that took more than one Newton step.
__global__ void __launch_bounds__(kBlockS)
k_step_sums(const int32_t* __restrict__ row_counts, const double* __restrict__ nneigh,
            const int32_t* __restrict__ iters, int64_t n, unsigned long long* __restrict__ acc)
{
    long long a = 0, b = 0, c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a += (long long)row_counts[2 * i] + (long long)row_counts[2 * i + 1];
        b += (long long)nneigh[i];
        c += iters[i] > 1 ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    __shared__ long long sh[3][kBlockS / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
        sh[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        long long t = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += sh[threadIdx.x][k];
        atomicAdd(&acc[threadIdx.x], (unsigned long long)t);
    }
}

// out = [status words (as doubles), minima (4), search entries, tight entries, dts (2), slow
// rows]; the status words and acc are left at zero for the next step
__global__ void k_step_pack(int32_t* __restrict__ status, int nstatus, const double* __restrict__ minima,
                            const double* __restrict__ dts, unsigned long long* __restrict__ acc,
                            double* __restrict__ out)
{
    const int t = threadIdx.x;
    if (t < nstatus) {
        out[t] = (double)status[t];
        status[t] = 0;
    }
    if (t < 4) out[nstatus + t] = minima[t];
    if (t == 0) {
        out[nstatus + 4] = (double)acc[0];
        out[nstatus + 5] = (double)acc[1];
        out[nstatus + 6] = dts[0];
        out[nstatus + 7] = dts[1];
        out[nstatus + 8] = (double)acc[2];
        acc[0] = 0ull;
        acc[1] = 0ull;
        acc[2] = 0ull;
    }
}

extern "C" int32_t nb_step_readback(nb_ctx* ctx, void* stream, int32_t* status, int32_t nstatus,
                                    const double* minima4, const int32_t* row_counts, const double* nneigh,
                                    const int32_t* iters, int64_t n, const double* dts2, unsigned long long* acc3,
                                    double* out)
{
    if (!ctx || !status || !minima4 || !row_counts || !nneigh || !iters || !dts2 || !acc3 || !out || n < 0 ||
        nstatus < 0 || nstatus > 64)
        return ctx_fail(ctx, "nb_step_readback: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    // 256 workgroups (a few particles per thread): three atomics each on three shared counters
    const int64_t blocks = std::min<int64_t>(256, (n + kBlockS - 1) / kBlockS);
    if (blocks > 0)
        hipLaunchKernelGGL(k_step_sums, dim3((unsigned)blocks), dim3(kBlockS), 0, s, row_counts, nneigh, iters, n,
                           acc3);
    hipLaunchKernelGGL(k_step_pack, dim3(1), dim3(64), 0, s, status, (int)nstatus, minima4, dts2, acc3, out);
    return launch_check_s(ctx, "nb_step_readback");
}

// ------------------------------------------------------- nb_dt_minima_dev without allocation
// For callers that cannot allocate (a stream being recorded into a hipGraph, a serving loop that
// forbids hipMalloc): the partial minima live in `scratch`, a fixed device array of this file,
// instead of the context's arena. k_dt_partial's grid is capped at kStaticParts workgroups, each
// striding over the particles, so the array (kStaticParts + 1 rows of 4 doubles, 32.8 KB per
// GPU, held until the process ends) fits every n. Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (concurrent streams) race; calls in stream order do not.
// The reduction reads only the rows its own launch wrote, so nothing needs clearing between
// calls.
static constexpr int kStaticParts = 1024;
static constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[4 * (kStaticParts + 1)];
static double* g_scratch_dev[kMaxDevices] = {};  // scratch's address on each device
static std::mutex g_scratch_mu;

// Called by nb_ctx_create with `device` current: looks the array's address up once per device.
// In host code `scratch` names a host-side stand-in, never the device memory: kernels and copies
// get the address found here.
int static_scratch_init(int device, std::string* err)
{
    if (device < 0 || device >= kMaxDevices) {
        if (err) *err = "static scratch: device index out of range";
        return 1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        if (err) *err = std::string("static scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e);
        return 1;
    }
    g_scratch_dev[device] = (double*)ptr;
    return 0;
}

extern "C" int32_t nb_dt_minima_dev_static(nb_ctx* ctx, void* stream, const double* h, const double* vsig,
                                           const double* acc, const double* nu, double c_visc, const double* x,
                                           int64_t n, double* out4_dev)
{
    if (!ctx || n <= 0 || !out4_dev || !h || !vsig || !acc || !x)
        return ctx_fail(ctx, "nb_dt_minima_dev_static: bad arguments");
    if (n > INT64_MAX / 3) return ctx_fail(ctx, "nb_dt_minima_dev_static: n beyond the largest (N,3) array indexable");
    double* part = nullptr;
    if (ctx->device >= 0 && ctx->device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        part = g_scratch_dev[ctx->device];
    }
    if (!part) return ctx_fail_code(ctx, NB_ERR_HIP, "nb_dt_minima_dev_static: no static scratch on this device");
    const int nparts = (int)std::min<int64_t>(kStaticParts, (n + kBlockS - 1) / kBlockS);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dt_partial, dim3(nparts), dim3(kBlockS), 0, s, h, vsig, acc, nu, c_visc, x, n, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlockS), 0, s, part, nparts, 4, 0, out4_dev);
    return launch_check_s(ctx, "nb_dt_minima_dev_static");
}

extern "C" int32_t nb_dt_minima_seg(nb_ctx* ctx, void* stream, const double* h, const double* vsig,
                                    const double* acc, const double* nu, double c_visc, const double* x,
                                    const int64_t* seg_off, int32_t nseg, double* out)
{
    if (!ctx || nseg < 1 || !seg_off || !out || !h || !vsig || !acc || !x)
        return ctx_fail(ctx, "nb_dt_minima_seg: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 4 * nseg + sizeof(int64_t) * (nseg + 1)))
        return NB_ERR_HIP;
    int64_t* off = (int64_t*)(ctx->red + 4 * nseg);
    if (hipMemcpyAsync(off, seg_off, sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "nb_dt_minima_seg: offset copy failed");
    hipLaunchKernelGGL(k_dt_seg, dim3(nseg), dim3(kBlockS), 0, s, h, vsig, acc, nu, c_visc, x, off, ctx->red);
    if (hipMemcpyAsync(out, ctx->red, sizeof(double) * 4 * nseg, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "nb_dt_minima_seg: result copy failed");
    return NB_OK;
}

extern "C" int32_t nb_dt_minima_seg_dev(nb_ctx* ctx, void* stream, const double* h, const double* vsig,
                                        const double* acc, const double* nu, double c_visc, const double* x,
                                        const int64_t* seg_off_dev, int32_t nseg, double* out_dev)
{
    if (!ctx || nseg < 1 || !h || !vsig || !acc || !x || !seg_off_dev || !out_dev)
        return ctx_fail(ctx, "nb_dt_minima_seg_dev: bad arguments");
    hipLaunchKernelGGL(k_dt_seg, dim3(nseg), dim3(kBlockS), 0, (hipStream_t)stream, h, vsig, acc, nu, c_visc, x,
                       seg_off_dev, out_dev);
    return launch_check_s(ctx, "nb_dt_minima_seg_dev");
}

extern "C" int32_t nb_axpy_rows(nb_ctx* ctx, void* stream, const double* y, const double* a, const double* s_rows,
                                int64_t n, int32_t width, double* out)
{
    if (!ctx || n < 0 || width <= 0 || !y || !a || !s_rows || !out) return ctx_fail(ctx, "nb_axpy_rows: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_axpy_rows, dim3(grid_s(n * width)), dim3(kBlockS), 0, (hipStream_t)stream, y, a, s_rows, n,
                       (int)width, out);
    return launch_check_s(ctx, "nb_axpy_rows");
}

extern "C" int32_t nb_expand_seg(nb_ctx* ctx, void* stream, const double* per_seg, const int32_t* cloud, int64_t n,
                                 double scale, double* out)
{
    if (!ctx || n < 0 || !per_seg || !cloud || !out) return ctx_fail(ctx, "nb_expand_seg: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_expand_seg, dim3(grid_s(n)), dim3(kBlockS), 0, (hipStream_t)stream, per_seg, cloud, n, scale,
                       out);
    return launch_check_s(ctx, "nb_expand_seg");
}

extern "C" int32_t nb_ledger(nb_ctx* ctx, void* stream, const double* x, const double* p, const double* u,
                             const double* mass, const double* e_part, int32_t gn, int64_t n, double* out6)
{
    if (!ctx || n <= 0 || !out6) return ctx_fail(ctx, "nb_ledger: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    int nparts = (int)std::min<int64_t>(1024, (n + kBlockS - 1) / kBlockS);
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 6 * (nparts + 1))) return NB_ERR_HIP;
    hipLaunchKernelGGL(k_ledger_partial, dim3(nparts), dim3(kBlockS), 0, s, x, p, u, mass, e_part, gn, n, ctx->red);
    return reduce_two_level(ctx, s, 6, 1, nparts, out6);
}
