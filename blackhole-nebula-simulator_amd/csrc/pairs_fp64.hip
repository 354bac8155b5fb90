// FP64 pair loops: density/h Newton solve, forces, gradients, Navier-Stokes and disc
// viscosity. One work-item per particle row; each row's sums run over its CSR list in list
// order (ascending neighbour index), so every result is the reference kernel's own sequence
// of roundings (sph.py:215-1117). Neighbour data are gathered through L2: the lists of nearby
// rows overlap, and particles stored in cell order keep those rows in the same wavefront.
#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"
#include "nb_linalg.h"

namespace nb {

static constexpr int kBlockP = 256;

// ----------------------------------------------------------------- density
// sph.py:215-228 (_sums) over the cached distances r[0:cnt] (stride 1) and masses
template <int64_t S>
__device__ __forceinline__ void sums64(const double* __restrict__ r, const double* __restrict__ m,
                                       double m_uniform, int64_t cnt, double h, int kid, double cnorm,
                                       double& rs_out, double& dws_out)
{
    double rs = 0.0, dws = 0.0;
    const double inv_h = 1.0 / h;
    const double ih3 = ipow3(inv_h);
    const double ncih4 = -cnorm * ipow4(inv_h);
    for (int64_t k = 0; k < cnt; k++) {
        double q = r[k * S] * inv_h;
        if (q < 2.0) {
            double mk = m ? m[k * S] : m_uniform;
            double fq = kernel_f(q, kid);
            rs += mk * cnorm * ih3 * fq;
            dws += mk * (ncih4 * (3.0 * fq + q * kernel_df(q, kid)));
        }
    }
    rs_out = rs;
    dws_out = dws;
}

// sph.py:251-297 (_density_kernel). Distances of the entries closer than 2 min(h_search, h_max)
// are cached once per row in the slots of the row's own list entries (so in blocked rows the
// cache is read and written 64 rows at a time, contiguously); masses too unless every particle
// has the same mass (m_uniform > 0).
template <class L>
__global__ void __launch_bounds__(kBlockP)
k_density64(const double* __restrict__ pos, const double* __restrict__ mass, double m_uniform,
            const double* __restrict__ h_start, const double* __restrict__ h_search,
            L lists, Box box,
            const int64_t* __restrict__ rows, int64_t nrows, int kid, double cnorm, double eta,
            double tol, double h_max, double* __restrict__ rcache, double* __restrict__ mcache,
            double* __restrict__ h_out, double* __restrict__ rho, double* __restrict__ omega,
            int32_t* __restrict__ iters, int32_t* __restrict__ flags)
{
    int64_t q = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = row_of(rows, q);
    const double m_i = mass[i];
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double r_max = 2.0 * pymin(h_search[i], h_max);
    const double r_max2 = r_max * r_max;
    constexpr int64_t S = L::stride;
    const int64_t kn = lists.count(i);
    double* r = rcache + lists.base(i);
    double* m = mcache ? mcache + lists.base(i) : nullptr;
    int64_t cnt = 0;
    for (int64_t k = 0; k < kn; k++) {
        const int32_t j = lists.at(i, k);
        double dx = min_image(pos[3 * j] - xi, box.l[0]);
        double dy = min_image(pos[3 * j + 1] - yi, box.l[1]);
        double dz = min_image(pos[3 * j + 2] - zi, box.l[2]);
        double d2 = 0.0;
        d2 += dx * dx;
        d2 += dy * dy;
        d2 += dz * dz;
        if (d2 < r_max2) {
            r[cnt * S] = sqrt(d2);
            if (m) m[cnt * S] = mass[j];
            cnt++;
        }
    }
    double hi = pymin(h_search[i], h_max);
    double lo = 0.0;
    double h = pymin(h_start[i], hi);
    const double h0 = h;
    int flag = 3;
    int it = 0;
    const double eta3 = ipow3(eta);
    double rs, dws;
    while (it < 50) {
        it += 1;
        sums64<S>(r, m, m_uniform, cnt, h, kid, cnorm, rs, dws);
        double f = rs - m_i * ipow3(eta / h);
        double df = dws + 3.0 * m_i * eta3 / ipow4(h);
        if (f > 0) hi = h;
        else lo = h;
        double h_new = df > 0 ? h - f / df : 0.5 * (lo + hi);
        if (!(lo <= h_new && h_new <= hi)) h_new = 0.5 * (lo + hi);
        if (fabs(h_new - h) < tol * h0) {
            h = h_new;
            flag = 0;
            break;
        }
        h = h_new;
    }
    const double h_lim = pymin(h_search[i], h_max);
    if (h >= h_lim * (1 - 1e-3)) {
        double rs_lim, dum;
        sums64<S>(r, m, m_uniform, cnt, h_lim, kid, cnorm, rs_lim, dum);
        if (rs_lim - m_i * ipow3(eta / h_lim) < 0) {
            h = h_lim;
            flag = h_lim >= h_max ? 1 : 2;
        }
    }
    sums64<S>(r, m, m_uniform, cnt, h, kid, cnorm, rs, dws);
    h_out[q] = h;
    rho[q] = rs;
    omega[q] = flag == 1 ? 1.0 : 1.0 + h / (3.0 * rs) * dws;
    iters[q] = it;
    flags[q] = flag;
}

// ----------------------------------------------------------------- forces
// sph.py:525-652 (_forces_kernel)
template <class L>
__global__ void __launch_bounds__(kBlockP)
k_forces64(const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ mass,
           const double* __restrict__ h, const double* __restrict__ rho, const double* __restrict__ omega,
           const double* __restrict__ pressure, const double* __restrict__ u, const double* __restrict__ c_s,
           const double* __restrict__ alpha, const double* __restrict__ gv,
           L lists, Box box,
           const int64_t* __restrict__ rows, int64_t nrows, nb_force_params P,
           double* __restrict__ acc, double* __restrict__ dudt, double* __restrict__ vsig_max,
           double* __restrict__ vheat, double* __restrict__ cheat)
{
    int64_t q = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t a = row_of(rows, q);
    const int kid = P.kid;
    const double cnorm = P.cnorm;
    const bool use_visc = P.use_visc, use_cond = P.use_cond, use_recon = P.use_recon;
    const double rho_a = rho[a], om_a = omega[a], h_a = h[a];
    const double inv_or_a = 1.0 / (om_a * rho_a);
    const double pa0 = pos[3 * a], pa1 = pos[3 * a + 1], pa2 = pos[3 * a + 2];
    const double va0 = vel[3 * a], va1 = vel[3 * a + 1], va2 = vel[3 * a + 2];
    const double P_a = pressure[a];
    const double acs_a = use_visc ? alpha[a] * c_s[a] : 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    double work = 0.0, heat = 0.0, visc_heat = 0.0, cond_heat = 0.0;
    double vmax = use_visc ? c_s[a] : 0.0;
    double ga_m[9];
    if (use_visc && use_recon)
        for (int k = 0; k < 9; k++) ga_m[k] = gv[9 * a + k];
    const double den_a = om_a * rho_a * rho_a;
    const int64_t kn = lists.count(a);
    for (int64_t k = 0; k < kn; k++) {
        const int32_t b = lists.at(a, k);
        if (b == a) continue;
        double dx0 = pa0 - pos[3 * b];
        double dx1 = pa1 - pos[3 * b + 1];
        double dx2 = pa2 - pos[3 * b + 2];
        if (box.l[0] > 0.0) dx0 -= box.l[0] * rint(dx0 / box.l[0]);
        if (box.l[1] > 0.0) dx1 -= box.l[1] * rint(dx1 / box.l[1]);
        if (box.l[2] > 0.0) dx2 -= box.l[2] * rint(dx2 / box.l[2]);
        const double r = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
        const double rx = dx0 / r, ry = dx1 / r, rz = dx2 / r;
        const double h_b = h[b];
        const double fa = dwdr(r, h_a, kid, cnorm);
        const double fb = dwdr(r, h_b, kid, cnorm);
        const double vdotr = (va0 - vel[3 * b]) * rx + (va1 - vel[3 * b + 1]) * ry + (va2 - vel[3 * b + 2]) * rz;
        const double rho_b = rho[b], om_b = omega[b], m_b = mass[b];
        double num_a = P_a;
        double num_b = pressure[b];
        if (use_visc && use_recon) {
            const double* gb_m = gv + 9 * (int64_t)b;
            double ga0 = ga_m[0] * dx0 + ga_m[1] * dx1 + ga_m[2] * dx2;
            double ga1 = ga_m[3] * dx0 + ga_m[4] * dx1 + ga_m[5] * dx2;
            double ga2 = ga_m[6] * dx0 + ga_m[7] * dx1 + ga_m[8] * dx2;
            double gb0 = gb_m[0] * dx0 + gb_m[1] * dx1 + gb_m[2] * dx2;
            double gb1 = gb_m[3] * dx0 + gb_m[4] * dx1 + gb_m[5] * dx2;
            double gb2 = gb_m[6] * dx0 + gb_m[7] * dx1 + gb_m[8] * dx2;
            double pa = dx0 * ga0 + dx1 * ga1 + dx2 * ga2;
            double pb = dx0 * gb0 + dx1 * gb1 + dx2 * gb2;
            double phi = 0.0;
            if (pb != 0.0) {
                double ratio = pa / pb;
                if (ratio > 0.0) phi = pymin(1.0, 4.0 * ratio / ((1.0 + ratio) * (1.0 + ratio)));
            }
            if (P.closeness) {
                double eta_ab = pymin(r / h_a, r / h_b);
                if (eta_ab < P.eta_crit) {
                    double t = (eta_ab - P.eta_crit) / P.eta_fold;
                    phi *= exp(-t * t);
                }
            }
            double vdr = ((va0 - 0.5 * phi * ga0) - (vel[3 * b] + 0.5 * phi * gb0)) * rx
                       + ((va1 - 0.5 * phi * ga1) - (vel[3 * b + 1] + 0.5 * phi * gb1)) * ry
                       + ((va2 - 0.5 * phi * ga2) - (vel[3 * b + 2] + 0.5 * phi * gb2)) * rz;
            bool on = vdr < 0.0 && vdotr < 0.0;
            if (P.gate == 1) on = vdr < 0.0;
            else if (P.gate == 2 && vdotr > vdr) vdr = vdotr;
            double vsig_a = acs_a + P.beta * fabs(vdr);
            double vsig_b = alpha[b] * c_s[b] + P.beta * fabs(vdr);
            double vsig_cfl = acs_a + P.beta * fabs(vdotr);
            if (vsig_cfl > vmax) vmax = vsig_cfl;
            if (on) {
                num_a += -0.5 * rho_a * vsig_a * vdr;
                num_b += -0.5 * rho_b * vsig_b * vdr;
                double dh = m_b * vsig_a * 0.5 * vdr * vdotr * fa * inv_or_a;
                heat -= dh;
                visc_heat -= dh;
            }
        } else if (use_visc) {
            double vsig_a = acs_a + P.beta * fabs(vdotr);
            double vsig_b = alpha[b] * c_s[b] + P.beta * fabs(vdotr);
            if (vsig_a > vmax) vmax = vsig_a;
            if (vdotr < 0.0) {
                num_a += -0.5 * rho_a * vsig_a * vdotr;
                num_b += -0.5 * rho_b * vsig_b * vdotr;
                double dh = m_b * vsig_a * 0.5 * vdotr * vdotr * fa * inv_or_a;
                heat -= dh;
                visc_heat -= dh;
            }
        }
        if (use_cond) {
            double vu;
            if (P.cond_signal == 0) vu = sqrt(fabs(P_a - pressure[b]) / (0.5 * (rho_a + rho_b)));
            else vu = fabs(vdotr);
            double dc = m_b * P.alpha_u * vu * (u[a] - u[b]) * 0.5 * (fa * inv_or_a + fb / (om_b * rho_b));
            heat += dc;
            cond_heat += dc;
        }
        const double f = m_b * (num_a / den_a * fa + num_b / (om_b * rho_b * rho_b) * fb);
        ax -= f * rx;
        ay -= f * ry;
        az -= f * rz;
        work += m_b * vdotr * fa;
    }
    acc[3 * q] = ax;
    acc[3 * q + 1] = ay;
    acc[3 * q + 2] = az;
    dudt[q] = P_a * inv_or_a / rho_a * work + heat;
    vsig_max[q] = vmax;
    if (vheat) vheat[q] = visc_heat;
    if (cheat) cheat[q] = cond_heat;
}

// ----------------------------------------------------------------- gradients
// sph.py:837-884 (_gradients_kernel); with_acc = false means acc is zero.
// sph.py:785-815 (_velocity_gradient_kernel) when gv_out is given (then div/curl/ddiv unused).
template <class L>
__global__ void __launch_bounds__(kBlockP)
k_gradients64(const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ acc,
              const double* __restrict__ mass, const double* __restrict__ h,
              L lists, Box box,
              const int64_t* __restrict__ rows, int64_t nrows, int kid, double cnorm,
              double* __restrict__ div, double* __restrict__ curl, double* __restrict__ ddiv,
              double* __restrict__ gv_out)
{
    int64_t q = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t a = row_of(rows, q);
    double rmat[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double gvm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double gam[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool want_acc = acc != nullptr && gv_out == nullptr;
    int64_t cnt = 0;
    const double h_a = h[a];
    const int64_t kn = lists.count(a);
    for (int64_t k = 0; k < kn; k++) {
        const int32_t b = lists.at(a, k);
        if (b == a) continue;
        cnt += 1;
        double dx[3];
        for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box.l[c]);
        double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double g = dwdr(r, h_a, kid, cnorm) / r;
        double mb = mass[b];
        for (int i = 0; i < 3; i++) {
            double dv = vel[3 * b + i] - vel[3 * a + i];
            double da = want_acc ? acc[3 * b + i] - acc[3 * a + i] : 0.0;
            for (int j = 0; j < 3; j++) {
                double gj = mb * g * dx[j];
                rmat[3 * i + j] -= dx[i] * gj;
                gvm[3 * i + j] += dv * gj;
                gam[3 * i + j] += da * gj;
            }
        }
    }
    double dv[9], da[9];
    if (near_singular(rmat, cnt)) {
        double scale = (rmat[0] + rmat[4] + rmat[8]) / 3.0;
        double inv_s = scale > 0.0 ? 1.0 / scale : 0.0;
        for (int k = 0; k < 9; k++) { dv[k] = gvm[k] * inv_s; da[k] = gam[k] * inv_s; }
    } else {
        double rinv[9];
        inv3(rmat, rinv);
        matmul3(gvm, rinv, dv);
        matmul3(gam, rinv, da);
    }
    if (gv_out) {
        for (int k = 0; k < 9; k++) gv_out[9 * a + k] = dv[k];
        return;
    }
    div[q] = dv[0] + dv[4] + dv[8];
    double c0 = dv[7] - dv[5], c1 = dv[2] - dv[6], c2 = dv[3] - dv[1];
    curl[q] = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    double s = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s += dv[3 * i + j] * dv[3 * j + i];
    ddiv[q] = da[0] + da[4] + da[8] - s;
}

// ----------------------------------------------------------------- Navier-Stokes
// sph.py:957-981 (_ns_gradient_kernel), out written at row a
template <class L>
__global__ void __launch_bounds__(kBlockP)
k_ns_gradients64(const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ mass,
                 const double* __restrict__ h, const double* __restrict__ rho, const double* __restrict__ omega,
                 L lists, Box box,
                 const int64_t* __restrict__ rows, int64_t nrows, int kid, double cnorm, double* __restrict__ out)
{
    int64_t q = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t a = row_of(rows, q);
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t kn = lists.count(a);
    for (int64_t k = 0; k < kn; k++) {
        const int32_t b = lists.at(a, k);
        if (b == a) continue;
        double dx[3];
        for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box.l[c]);
        double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double f = mass[b] * dwdr(r, h[a], kid, cnorm) / r;
        for (int i = 0; i < 3; i++) {
            double vab = vel[3 * a + i] - vel[3 * b + i];
            for (int j = 0; j < 3; j++) g[3 * i + j] -= vab * f * dx[j];
        }
    }
    double inv = 1.0 / (omega[a] * rho[a]);
    for (int k = 0; k < 9; k++) out[9 * a + k] = g[k] * inv;
}

// sph.py:1002-1027 (_ns_force_kernel)
template <class L>
__global__ void __launch_bounds__(kBlockP)
k_ns_forces64(const double* __restrict__ pos, const double* __restrict__ mass, const double* __restrict__ h,
              const double* __restrict__ rho, const double* __restrict__ omega, const double* __restrict__ stress,
              L lists, Box box,
              const int64_t* __restrict__ rows, int64_t nrows, int kid, double cnorm, double* __restrict__ acc)
{
    int64_t q = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t a = row_of(rows, q);
    const double ia = 1.0 / (omega[a] * rho[a] * rho[a]);
    double ax[3] = {0.0, 0.0, 0.0};
    const int64_t kn = lists.count(a);
    for (int64_t k = 0; k < kn; k++) {
        const int32_t b = lists.at(a, k);
        if (b == a) continue;
        double dx[3];
        for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box.l[c]);
        double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double fa = dwdr(r, h[a], kid, cnorm) / r * ia;
        double fb = dwdr(r, h[b], kid, cnorm) / r / (omega[b] * rho[b] * rho[b]);
        for (int i = 0; i < 3; i++) {
            double s = 0.0;
            for (int j = 0; j < 3; j++) s += (stress[9 * a + 3 * i + j] * fa + stress[9 * (int64_t)b + 3 * i + j] * fb) * dx[j];
            ax[i] += mass[b] * s;
        }
    }
    for (int i = 0; i < 3; i++) acc[3 * q + i] = ax[i];
}

// ----------------------------------------------------------------- disc viscosity
// sph.py:1084-1117 (_disc_kernel)
template <class L>
__global__ void __launch_bounds__(kBlockP)
k_disc64(const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ mass,
         const double* __restrict__ h, const double* __restrict__ rho, const double* __restrict__ omega,
         const double* __restrict__ c_s, double alpha_av, L lists, Box box, const int64_t* __restrict__ rows, int64_t nrows,
         int kid, double cnorm, double* __restrict__ acc, double* __restrict__ dudt)
{
    int64_t q = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t a = row_of(rows, q);
    const double ia = 1.0 / (omega[a] * rho[a] * rho[a]);
    double ax[3] = {0.0, 0.0, 0.0};
    double heat = 0.0;
    const int64_t kn = lists.count(a);
    for (int64_t k = 0; k < kn; k++) {
        const int32_t b = lists.at(a, k);
        if (b == a) continue;
        double dx[3];
        for (int c = 0; c < 3; c++) dx[c] = min_image(pos[3 * a + c] - pos[3 * b + c], box.l[c]);
        double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double vdr = 0.0;
        for (int c = 0; c < 3; c++) vdr += (vel[3 * a + c] - vel[3 * b + c]) * dx[c] / r;
        double qa = -rho[a] * h[a] / (2.0 * r) * alpha_av * c_s[a] * vdr * ia;
        double qb = -rho[b] * h[b] / (2.0 * r) * alpha_av * c_s[b] * vdr / (omega[b] * rho[b] * rho[b]);
        double fa = dwdr(r, h[a], kid, cnorm);
        double fb = dwdr(r, h[b], kid, cnorm);
        double f = mass[b] * (qa * fa + qb * fb);
        for (int c = 0; c < 3; c++) ax[c] -= f * dx[c] / r;
        heat += mass[b] * qa * vdr * fa;
    }
    for (int c = 0; c < 3; c++) acc[3 * q + c] = ax[c];
    dudt[q] = heat;
}

// pairs_mixed.hip: the precision-1 branch of nb_forces
int32_t forces_mixed(nb_ctx* ctx, hipStream_t s, const double* pos, const double* vel, const double* mass,
                     const double* h, const double* rho, const double* omega, const double* pressure, const double* u,
                     const double* c_s, const double* alpha, const nb_lists* lists, const double* box3,
                     const int64_t* rows, int64_t nrows, int64_t n, const nb_force_params* P, double* acc,
                     double* dudt, double* vsig_max, double* vheat, double* cheat);

static inline unsigned grid_p(int64_t n) { return (unsigned)((n + kBlockP - 1) / kBlockP); }

static inline Box make_box(const double* box3)
{
    Box b{{0.0, 0.0, 0.0}};
    if (box3) for (int c = 0; c < 3; c++) b.l[c] = box3[c];
    return b;
}

}  // namespace nb

using namespace nb;

static int32_t launch_check(nb_ctx* ctx, const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        return ctx_fail(ctx, m.c_str());
    }
    return NB_OK;
}

#define NB_DISPATCH(L, KERNEL, GRID, STREAM, ...)                                                       \
    do {                                                                                                \
        const int ts_ = ktime_begin(ctx, (STREAM), #KERNEL);                                            \
        if ((L)->length && (L)->row_major) {                                                            \
            RowView v_{(L)->index, (L)->length, (L)->offset};                                           \
            hipLaunchKernelGGL(KERNEL<RowView>, dim3(GRID), dim3(kBlockP), 0, STREAM, __VA_ARGS__);      \
        } else if ((L)->length) {                                                                       \
            BlkView v_{(L)->index, (L)->length, (L)->row_cap, (L)->offset};                                          \
            hipLaunchKernelGGL(KERNEL<BlkView>, dim3(GRID), dim3(kBlockP), 0, STREAM, __VA_ARGS__);      \
        } else {                                                                                        \
            CsrView v_{(L)->offset, (L)->index};                                                        \
            hipLaunchKernelGGL(KERNEL<CsrView>, dim3(GRID), dim3(kBlockP), 0, STREAM, __VA_ARGS__);      \
        }                                                                                               \
        ktime_end(ctx, (STREAM), ts_);                                                                  \
    } while (0)

static bool lists_ok(const nb_lists* L) { return L && L->index && (L->length || L->offset); }

extern "C" int32_t nb_density(nb_ctx* ctx, void* stream, const double* pos, const double* mass, double m_uniform,
                              const double* h_start, const double* h_search, const nb_lists* lists,
                              const double* box3, const int64_t* rows, int64_t nrows, int32_t kid, double cnorm,
                              double eta, double tol, double h_max, double* h_out, double* rho, double* omega,
                              int32_t* iters, int32_t* flags)
{
    if (!ctx || nrows < 0 || !lists_ok(lists) || lists->n_entries < 0) return ctx_fail(ctx, "nb_density: bad arguments");
    if (nrows == 0) return NB_OK;
    const int64_t ne = lists->n_entries;
    size_t need = sizeof(double) * (size_t)(ne + 1) * (m_uniform > 0.0 ? 1 : 2);
    if (ctx_reserve(ctx, (void**)&ctx->dscratch, &ctx->dscratch_bytes, need)) return NB_ERR_HIP;
    double* mcache = m_uniform > 0.0 ? nullptr : ctx->dscratch + (ne + 1);
    NB_DISPATCH(lists, k_density64, grid_p(nrows), (hipStream_t)stream, pos, mass, m_uniform, h_start, h_search, v_,
                make_box(box3), rows, nrows, kid, cnorm, eta, tol, h_max, ctx->dscratch, mcache, h_out, rho, omega,
                iters, flags);
    return launch_check(ctx, "nb_density");
}

extern "C" int32_t nb_forces(nb_ctx* ctx, void* stream, const double* pos, const double* vel, const double* mass,
                             const double* h, const double* rho, const double* omega, const double* pressure,
                             const double* u, const double* c_s, const double* alpha, const double* grad_v,
                             const nb_lists* lists, const double* box3, const int64_t* rows, int64_t nrows,
                             const nb_force_params* params, double* acc, double* dudt, double* vsig_max, double* vheat,
                             double* cheat)
{
    if (!ctx || !params || nrows < 0 || !lists_ok(lists)) return ctx_fail(ctx, "nb_forces: bad arguments");
    if (params->use_visc && (!alpha || !c_s)) return ctx_fail(ctx, "nb_forces: viscosity needs alpha and c_s");
    if (params->use_cond && !u) return ctx_fail(ctx, "nb_forces: conductivity needs u");
    if (params->use_recon && !grad_v) return ctx_fail(ctx, "nb_forces: reconstruction needs grad_v");
    if (nrows == 0) return NB_OK;
    if (params->precision == 1 && !params->use_recon) {
        const int64_t n = rows ? (int64_t)lists->n_rows : nrows;
        if (n <= 0) return ctx_fail(ctx, "nb_forces: mixed precision on a row subset needs nb_lists.n_rows");
        return forces_mixed(ctx, (hipStream_t)stream, pos, vel, mass, h, rho, omega, pressure, u, c_s, alpha, lists,
                            box3, rows, nrows, n, params, acc, dudt, vsig_max, vheat, cheat);
    }
    NB_DISPATCH(lists, k_forces64, grid_p(nrows), (hipStream_t)stream, pos, vel, mass, h, rho, omega, pressure, u, c_s,
                alpha, grad_v, v_, make_box(box3), rows, nrows, *params, acc, dudt, vsig_max, vheat, cheat);
    return launch_check(ctx, "nb_forces");
}

extern "C" int32_t nb_gradients(nb_ctx* ctx, void* stream, const double* pos, const double* vel, const double* acc,
                                const double* mass, const double* h, const nb_lists* lists, const double* box3,
                                const int64_t* rows, int64_t nrows, int32_t kid, double cnorm, double* div,
                                double* curl, double* ddiv)
{
    if (!ctx || nrows < 0 || !lists_ok(lists)) return ctx_fail(ctx, "nb_gradients: bad arguments");
    if (nrows == 0) return NB_OK;
    NB_DISPATCH(lists, k_gradients64, grid_p(nrows), (hipStream_t)stream, pos, vel, acc, mass, h, v_, make_box(box3),
                rows, nrows, kid, cnorm, div, curl, ddiv, (double*)nullptr);
    return launch_check(ctx, "nb_gradients");
}

extern "C" int32_t nb_velocity_gradients(nb_ctx* ctx, void* stream, const double* pos, const double* vel,
                                         const double* mass, const double* h, const nb_lists* lists,
                                         const double* box3, const int64_t* rows, int64_t nrows, int32_t kid,
                                         double cnorm, double* out)
{
    if (!ctx || nrows < 0 || !out || !lists_ok(lists)) return ctx_fail(ctx, "nb_velocity_gradients: bad arguments");
    if (nrows == 0) return NB_OK;
    NB_DISPATCH(lists, k_gradients64, grid_p(nrows), (hipStream_t)stream, pos, vel, (const double*)nullptr, mass, h, v_,
                make_box(box3), rows, nrows, kid, cnorm, (double*)nullptr, (double*)nullptr, (double*)nullptr, out);
    return launch_check(ctx, "nb_velocity_gradients");
}

extern "C" int32_t nb_ns_gradients(nb_ctx* ctx, void* stream, const double* pos, const double* vel, const double* mass,
                                   const double* h, const double* rho, const double* omega, const nb_lists* lists,
                                   const double* box3, const int64_t* rows, int64_t nrows, int32_t kid, double cnorm,
                                   double* out)
{
    if (!ctx || nrows < 0 || !out || !lists_ok(lists)) return ctx_fail(ctx, "nb_ns_gradients: bad arguments");
    if (nrows == 0) return NB_OK;
    NB_DISPATCH(lists, k_ns_gradients64, grid_p(nrows), (hipStream_t)stream, pos, vel, mass, h, rho, omega, v_,
                make_box(box3), rows, nrows, kid, cnorm, out);
    return launch_check(ctx, "nb_ns_gradients");
}

extern "C" int32_t nb_ns_forces(nb_ctx* ctx, void* stream, const double* pos, const double* mass, const double* h,
                                const double* rho, const double* omega, const double* stress, const nb_lists* lists,
                                const double* box3, const int64_t* rows, int64_t nrows, int32_t kid, double cnorm,
                                double* acc)
{
    if (!ctx || nrows < 0 || !lists_ok(lists)) return ctx_fail(ctx, "nb_ns_forces: bad arguments");
    if (nrows == 0) return NB_OK;
    NB_DISPATCH(lists, k_ns_forces64, grid_p(nrows), (hipStream_t)stream, pos, mass, h, rho, omega, stress, v_,
                make_box(box3), rows, nrows, kid, cnorm, acc);
    return launch_check(ctx, "nb_ns_forces");
}

extern "C" int32_t nb_disc_forces(nb_ctx* ctx, void* stream, const double* pos, const double* vel, const double* mass,
                                  const double* h, const double* rho, const double* omega, const double* c_s,
                                  double alpha_av, const nb_lists* lists, const double* box3, const int64_t* rows,
                                  int64_t nrows, int32_t kid, double cnorm, double* acc, double* dudt)
{
    if (!ctx || nrows < 0 || !lists_ok(lists)) return ctx_fail(ctx, "nb_disc_forces: bad arguments");
    if (nrows == 0) return NB_OK;
    NB_DISPATCH(lists, k_disc64, grid_p(nrows), (hipStream_t)stream, pos, vel, mass, h, rho, omega, c_s, alpha_av, v_,
                make_box(box3), rows, nrows, kid, cnorm, acc, dudt);
    return launch_check(ctx, "nb_disc_forces");
}
