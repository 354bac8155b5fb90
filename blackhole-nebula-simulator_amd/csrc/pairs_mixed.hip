// Mixed-precision pair passes: gradients (for the viscosity switches) and forces with FP32 pair
// arithmetic.
//
// A group of 8 lanes serves one row: lane l takes entries l, l + 8, ... so that the group
// reads a row-major or CSR row contiguously. Each lane accumulates its pairs in FP32 with
// compensated sums (Kahan's; TwoSum for the pair forces, whose terms must cancel exactly between
// the two rows of a pair) and the group combines the lane partials in FP64. Pair differences of
// the FP64 state, separations and velocity differences, are formed in FP64 and then rounded to
// FP32: positions and velocities carry the bulk orbital motion (r ~ 2000, v ~ 0.03 c), against
// which an FP32 position resolves only ~1e-4 of a unit length and an FP32 velocity ~2e-9 c, as
// large as the velocity differences inside a cloud moving as one (its du/dt and viscosity switch
// then came out at percent-level error). Every other per-neighbour quantity comes from a packed
// FP32 record prepared once per pass (32 bytes for forces, 16 for gradients: whole 16-byte loads).
#include <string>

#include "nb_common.h"
#include "nb_ctx.h"
#include "nb_linalg.h"
#include "nebula_hip.h"

namespace nb {

// Lanes per row of the force and gradient passes: 8 (rows of ~60 pairs give each lane 7-8; the
// per-row combination of the lane sums costs a quarter less than with 16 lanes, and the FP64
// 3 x 3 solve of the gradients runs on 8 rows per wavefront instead of 4). The div/curl pass
// takes 4: its per-row work (15 sums combined, the 3 x 3 solve on one lane per row) outweighs
// its pair arithmetic, and 16 rows per wavefront share it (0.41 -> 0.36 ms at 1M, where the
// force pass at 4 lanes went from 0.37 to 0.42 ms).
static constexpr int kPairW = 8;
static constexpr int kDivCurlW = 4;
static constexpr int kBlockM = 256;

struct __align__(16) PF {
    float4 a;  // 1/h, P / (Omega rho^2), 1 / (Omega rho), alpha c_s
    float4 b;  // m s (s = mass_scale), u, P, rho
};

// The force pass weighs each pair by the product of both FP32 masses, (m_a s)(m_b s): that
// product is the same bits in row a and in row b, so the pair force m_a m_b S_ab r_ab of one row
// is exactly minus that of the other, and the row's acceleration is that sum over the FP64 mass
// m_a s^2. Weighing by the FP32 neighbour mass alone (a = sum m_b S r) rounds m_a fl(m_b S)
// and m_b fl(m_a S) differently when masses differ, which leaks momentum and angular momentum
// at 1e-7 relative per step. s is a power of two taken from the first particle's mass, so that
// the FP32 products stay far from under- and overflow whatever the mass unit; scaling by it
// changes no mantissa.
__device__ __forceinline__ double mass_scale(const double* __restrict__ mass)
{
    int e = 0;
    (void)frexp(mass[0], &e);
    return ldexp(1.0, -e);
}

struct __align__(16) PG {
    float4 a;  // 1/h, cnorm / h^4, m, 0
};

__global__ void k_pack_forces(const double* __restrict__ mass,
                              const double* __restrict__ h, const double* __restrict__ rho,
                              const double* __restrict__ omega, const double* __restrict__ pressure,
                              const double* __restrict__ u, const double* __restrict__ c_s,
                              const double* __restrict__ alpha, double cnorm, int64_t n, PF* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i], r = rho[i], om = omega[i], p = pressure[i], ms = mass_scale(mass);
    PF v;
    v.a = make_float4((float)(1.0 / hi), (float)(p / (om * r * r)), (float)(1.0 / (om * r)),
                      (float)(alpha ? alpha[i] * c_s[i] : 0.0));
    v.b = make_float4((float)(mass[i] * ms), (float)(u ? u[i] : 0.0), (float)p, (float)r);
    out[i] = v;
}

__global__ void k_pack_grad(const double* __restrict__ mass, const double* __restrict__ h, double cnorm, int64_t n,
                            PG* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i];
    PG v;
    v.a = make_float4((float)(1.0 / hi), (float)(cnorm / ipow4(hi)), (float)mass[i], 0.f);
    out[i] = v;
}

// kernels_nb.py:31-43 in FP32
__device__ __forceinline__ float df32(float q, int kid)
{
    if (q >= 2.f) return 0.f;
    if (kid == 0) {
        if (q < 1.f) return -3.f * q + 2.25f * q * q;
        const float t = 2.f - q;
        return -0.75f * t * t;
    }
    const float x = 1.f - 0.5f * q, x2 = x * x;
    return -7.f / 3.f * q * x2 * x2 * x * (5.f * q + 2.f);
}

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v)
    {
        const float t = v - c;
        const float u = s + t;
        c = (u - s) - t;
        s = u;
    }
    // adds the product a * b, formed in one fused multiply-add with the compensation
    __device__ __forceinline__ void add_prod(float a, float b)
    {
        const float t = __builtin_fmaf(a, b, -c);
        const float u = s + t;
        c = (u - s) - t;
        s = u;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// Compensated sum whose every term enters exactly (Knuth's TwoSum, the Kahan-Babuska form): the
// rounding error of each addition is recovered exactly and gathered in c. The pair forces use it:
// their terms are rounded products fl(f r_x), which row b forms with the opposite sign, so the
// two rows hold exactly opposite terms and their sums are accurate to O(n u^2) of the terms.
// Kahan's form (above) rounds y = x - c before the addition, an error of u |x| per term that
// differs between row a and row b: momentum then leaks at about u / sqrt(N) per pass (2e-10 of
// sum m|a| at 16k particles, against 1e-16 in FP64).
struct TwoSumAcc {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float p)
    {
        const float t = s + p;
        const float z = t - s;
        c += (s - (t - z)) + (p - z);
        s = t;
    }
    // a term far below s (here the low part of a pair force, itself exactly opposite in the other
    // row) goes straight into the compensation
    __device__ __forceinline__ void add_low(float l) { c += l; }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

// butterfly over the W lanes of a group: every lane ends with the same value
template <int W>
__device__ __forceinline__ double gsum(double v)
{
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

template <int W>
__device__ __forceinline__ float gmax(float v)
{
    for (int o = W / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, W));
    return v;
}

// separation x_a - x_b (minimum image), formed in FP64, rounded to FP32
__device__ __forceinline__ void sep32(const double* pos, double pa0, double pa1, double pa2, int64_t b, const Box& box,
                                      float& d0, float& d1, float& d2)
{
    d0 = (float)min_image(pa0 - pos[3 * b], box.l[0]);
    d1 = (float)min_image(pa1 - pos[3 * b + 1], box.l[1]);
    d2 = (float)min_image(pa2 - pos[3 * b + 2], box.l[2]);
}

// the neighbour's position, velocity and packed record, loaded before the pair's arithmetic
template <class R>
struct Nb {
    R rec;
    double x0, x1, x2, v0, v1, v2;
};

template <class R>
__device__ __forceinline__ Nb<R> load_nb(const double* __restrict__ pos, const double* __restrict__ vel,
                                         const R* __restrict__ recs, int64_t b)
{
    return Nb<R>{recs[b], pos[3 * b], pos[3 * b + 1], pos[3 * b + 2], vel[3 * b], vel[3 * b + 1], vel[3 * b + 2]};
}

// sph.py:525-652 without reconstruction: pressure force with grad-h terms, Monaghan viscosity on
// approaching pairs, conductivity (pressure or velocity signal), P dV work, largest v_sig. The
// pair force is weighed by both masses (see PF) and the row's sum divided by m_a s^2 in FP64;
// the energy sums carry the neighbour's scaled mass and are divided by s (exact). kMode: bit 0
// viscosity, bit 1 conductivity (compile-time, so that a pass without them holds no register
// for them).
template <class L, int W, int kMode>
__device__ __forceinline__ void forces_mixed_row(const double* __restrict__ pos, const double* __restrict__ vel,
                                                 const double* __restrict__ mass, const PF* __restrict__ pf,
                                                 const double* __restrict__ c_s, const L& lists, const Box& box,
                                                 const int64_t* __restrict__ rows, int64_t nrows,
                                                 const nb_force_params& P, double* __restrict__ acc,
                                                 double* __restrict__ dudt, double* __restrict__ vsig_max,
                                                 double* __restrict__ vheat, double* __restrict__ cheat)
{
    const int glane = threadIdx.x & (W - 1);
    const int64_t q = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / W;
    if (q >= nrows) return;  // whole groups leave together
    const int64_t a = row_of(rows, q);
    const double pa0 = pos[3 * a], pa1 = pos[3 * a + 1], pa2 = pos[3 * a + 2];
    const double va0 = vel[3 * a], va1 = vel[3 * a + 1], va2 = vel[3 * a + 2];
    const PF A = pf[a];
    // a massless row particle keeps the neighbour-mass weighting (its own mass would zero the sum)
    const float wa = A.b.x > 0.f ? A.b.x : 1.f;
    const float cnorm = (float)P.cnorm;
    const float iha2 = A.a.x * A.a.x, cn4a = cnorm * (iha2 * iha2);  // cnorm / h_a^4
    const float beta = (float)P.beta, alpha_u = (float)P.alpha_u;
    constexpr bool use_visc = kMode & 1, use_cond = kMode & 2;
    const int kid = P.kid;
    TwoSumAcc ax, ay, az;
    Kahan work, visc, cond;
    float vmax = 0.f;
    const int64_t m = lists.count(a);
    auto pair = [&](const Nb<PF>& N) {
        const PF& B = N.rec;
        const double D0 = min_image(pa0 - N.x0, box.l[0]);
        const double D1 = min_image(pa1 - N.x1, box.l[1]);
        const double D2 = min_image(pa2 - N.x2, box.l[2]);
        const float dx0 = (float)D0, dx1 = (float)D1, dx2 = (float)D2;
        const float r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
        const float inv_r = rsqrt32(r2);
        const float r = r2 * inv_r;
        const float rx = dx0 * inv_r, ry = dx1 * inv_r, rz = dx2 * inv_r;
        const float ihb2 = B.a.x * B.a.x;
        const float fa = cn4a * df32(r * A.a.x, kid);
        const float fb = (cnorm * (ihb2 * ihb2)) * df32(r * B.a.x, kid);
        const float dv0 = (float)(va0 - N.v0), dv1 = (float)(va1 - N.v1), dv2 = (float)(va2 - N.v2);
        const float vdotr = dv0 * rx + dv1 * ry + dv2 * rz;
        float na = A.a.y, nbb = B.a.y;  // (P - rho v_sig v.r / 2) / (Omega rho^2) of either side
        if (use_visc) {
            const float bv = beta * fabsf(vdotr);
            const float vsa = A.a.w + bv;
            vmax = fmaxf(vmax, vsa);
            if (vdotr < 0.f) {
                const float vsb = B.a.w + bv;
                na += -0.5f * vsa * vdotr * A.a.z;
                nbb += -0.5f * vsb * vdotr * B.a.z;
                visc.add(-(B.b.x * vsa * 0.5f * vdotr * vdotr * fa * A.a.z));
            }
        }
        if (use_cond) {
            const float vu =
                P.cond_signal == 0 ? sqrt32(fabsf(A.b.z - B.b.z) * rcp32(0.5f * (A.b.w + B.b.w))) : fabsf(vdotr);
            cond.add(B.b.x * alpha_u * vu * (A.b.y - B.b.y) * 0.5f * (fa * A.a.z + fb * B.a.z));
        }
        // The pair force -f x_ab / r, split as hi + lo: hi = fl(g d) with g = f / r and d the FP32
        // separation; lo = (g d - hi) + g (x_ab - d), the product's rounding error (exact, by
        // FMA) and the separation's. Both are exactly opposite in row b, and hi + lo lies along
        // the FP64 separation to O(u^2), so the pair exerts no torque beyond that (hi alone
        // would tilt each force by u and leave a torque of u |x_ab| |F|).
        const float f = (wa * B.b.x) * (na * fa + nbb * fb);
        const float g = f * inv_r;
        const float h0 = g * dx0, h1 = g * dx1, h2 = g * dx2;
        ax.add(-h0);
        ay.add(-h1);
        az.add(-h2);
        ax.add_low(-__builtin_fmaf(g, (float)(D0 - (double)dx0), __builtin_fmaf(g, dx0, -h0)));
        ay.add_low(-__builtin_fmaf(g, (float)(D1 - (double)dx1), __builtin_fmaf(g, dx1, -h1)));
        az.add_low(-__builtin_fmaf(g, (float)(D2 - (double)dx2), __builtin_fmaf(g, dx2, -h2)));
        work.add_prod(B.b.x * vdotr, fa);
    };
    // two entries per pass, both neighbours' records loaded before either pair's arithmetic: the
    // same pairs in the same order per lane as one entry at a time (self skipped)
    for (int64_t k = glane; k < m; k += 2 * W) {
        const int32_t b0 = lists.at(a, k);
        const int32_t b1 = k + W < m ? lists.at(a, k + W) : (int32_t)a;
        const Nb<PF> N0 = load_nb(pos, vel, pf, b0), N1 = load_nb(pos, vel, pf, b1);
        if (b0 != a) pair(N0);
        if (b1 != a) pair(N1);
    }
    const double sx = gsum<W>(ax.value()), sy = gsum<W>(ay.value()), sz = gsum<W>(az.value());
    const double sw = gsum<W>(work.value()), sv = gsum<W>(visc.value()), sc = gsum<W>(cond.value());
    const float vm = gmax<W>(vmax);
    if (glane == 0) {
        const double ms = mass_scale(mass);
        const double da = A.b.x > 0.f ? (ms * ms) * mass[a] : ms;  // m_a s^2: acc = sum / (m_a s^2)
        acc[3 * q] = sx / da;
        acc[3 * q + 1] = sy / da;
        acc[3 * q + 2] = sz / da;
        dudt[q] = (double)A.a.y * (sw / ms) + sv / ms + sc / ms;
        vsig_max[q] = use_visc ? fmax(c_s[a], (double)vm) : 0.0;
        if (vheat) vheat[q] = sv / ms;
        if (cheat) cheat[q] = sc / ms;
    }
}

#define NB_FORCES_KERNEL(MODE)                                                                                     \
    template <class L, int W>                                                                                      \
    __global__ void __launch_bounds__(kBlockM) k_forces_mixed_##MODE(                                              \
        const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ mass,           \
        const PF* __restrict__ pf, const double* __restrict__ c_s, L lists, Box box,                               \
        const int64_t* __restrict__ rows, int64_t nrows, nb_force_params P, double* __restrict__ acc,              \
        double* __restrict__ dudt, double* __restrict__ vsig_max, double* __restrict__ vheat,                      \
        double* __restrict__ cheat)                                                                                \
    {                                                                                                              \
        forces_mixed_row<L, W, MODE>(pos, vel, mass, pf, c_s, lists, box, rows, nrows, P, acc, dudt, vsig_max,     \
                                     vheat, cheat);                                                                \
    }
NB_FORCES_KERNEL(0)  // pressure only
NB_FORCES_KERNEL(1)  // + viscosity
NB_FORCES_KERNEL(3)  // + viscosity and conductivity
#undef NB_FORCES_KERNEL

// sph.py:837-884: R, G_v and G_a of a row from FP32 pair terms, inverted in FP64 (kAcc false:
// accelerations zero, G_a not summed).
template <class L, bool kAcc, int W>
__device__ __forceinline__ void gradients_mixed_row(const double* __restrict__ pos, const double* __restrict__ vel,
                                                    const double* __restrict__ accel,
                                                    const PG* __restrict__ pg, const L& lists, const Box& box,
                                                    const int64_t* __restrict__ rows, int64_t nrows, int kid,
                                                    double* __restrict__ div, double* __restrict__ curl,
                                                    double* __restrict__ ddiv)
{
    const int glane = threadIdx.x & (W - 1);
    const int64_t q = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / W;
    if (q >= nrows) return;
    const int64_t a = row_of(rows, q);
    const double pa0 = pos[3 * a], pa1 = pos[3 * a + 1], pa2 = pos[3 * a + 2];
    const PG A = pg[a];
    const double va0 = vel[3 * a], va1 = vel[3 * a + 1], va2 = vel[3 * a + 2];
    Kahan R[6], Gv[9], Ga[kAcc ? 9 : 1];  // R is symmetric: its upper triangle (00 01 02 11 12 22)
    double aa[3] = {0.0, 0.0, 0.0};
    if (kAcc)
        for (int c = 0; c < 3; c++) aa[c] = accel[3 * a + c];
    int cnt = 0;
    const int64_t m = lists.count(a);
    auto pair = [&](const Nb<PG>& N, int32_t b) {
        cnt++;
        const PG& B = N.rec;
        float dx[3];
        dx[0] = (float)min_image(pa0 - N.x0, box.l[0]);
        dx[1] = (float)min_image(pa1 - N.x1, box.l[1]);
        dx[2] = (float)min_image(pa2 - N.x2, box.l[2]);
        const float r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        const float inv_r = rsqrt32(r2);
        const float g = A.a.y * df32(r2 * inv_r * A.a.x, kid) * inv_r;
        const float dv[3] = {(float)(N.v0 - va0), (float)(N.v1 - va1), (float)(N.v2 - va2)};
        float da[3] = {0.f, 0.f, 0.f};
        if (kAcc)
            for (int c = 0; c < 3; c++) da[c] = (float)(accel[3 * b + c] - aa[c]);
        const float mg = B.a.z * g;
        const float gx[3] = {mg * dx[0], mg * dx[1], mg * dx[2]};
        R[0].add_prod(-dx[0], gx[0]);
        R[1].add_prod(-dx[0], gx[1]);
        R[2].add_prod(-dx[0], gx[2]);
        R[3].add_prod(-dx[1], gx[1]);
        R[4].add_prod(-dx[1], gx[2]);
        R[5].add_prod(-dx[2], gx[2]);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                Gv[3 * i + j].add_prod(dv[i], gx[j]);
                if (kAcc) Ga[kAcc ? 3 * i + j : 0].add_prod(da[i], gx[j]);
            }
    };
    // two entries per pass, both neighbours' records loaded first (same pairs and order per lane)
    for (int64_t k = glane; k < m; k += 2 * W) {
        const int32_t b0 = lists.at(a, k);
        const int32_t b1 = k + W < m ? lists.at(a, k + W) : (int32_t)a;
        const Nb<PG> N0 = load_nb(pos, vel, pg, b0), N1 = load_nb(pos, vel, pg, b1);
        if (b0 != a) pair(N0, b0);
        if (b1 != a) pair(N1, b1);
    }
    double rm[9], gvm[9], gam[9];
    {
        const double r00 = gsum<W>(R[0].value()), r01 = gsum<W>(R[1].value()), r02 = gsum<W>(R[2].value());
        const double r11 = gsum<W>(R[3].value()), r12 = gsum<W>(R[4].value()), r22 = gsum<W>(R[5].value());
        rm[0] = r00, rm[1] = r01, rm[2] = r02;
        rm[3] = r01, rm[4] = r11, rm[5] = r12;
        rm[6] = r02, rm[7] = r12, rm[8] = r22;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        gvm[k] = gsum<W>(Gv[k].value());
        gam[k] = kAcc ? gsum<W>(Ga[kAcc ? k : 0].value()) : 0.0;
    }
    for (int o = W / 2; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, W);
    if (glane != 0) return;
    double dvm[9], dam[9];
    if (near_singular(rm, cnt)) {
        const double scale = (rm[0] + rm[4] + rm[8]) / 3.0;
        const double inv_s = scale > 0.0 ? 1.0 / scale : 0.0;
        for (int k = 0; k < 9; k++) {
            dvm[k] = gvm[k] * inv_s;
            dam[k] = gam[k] * inv_s;
        }
    } else {
        double rinv[9];
        inv3(rm, rinv);
        matmul3(gvm, rinv, dvm);
        matmul3(gam, rinv, dam);
    }
    div[q] = dvm[0] + dvm[4] + dvm[8];
    const double c0 = dvm[7] - dvm[5], c1 = dvm[2] - dvm[6], c2 = dvm[3] - dvm[1];
    curl[q] = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    double ss = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ss += dvm[3 * i + j] * dvm[3 * j + i];
    ddiv[q] = dam[0] + dam[4] + dam[8] - ss;
}

template <class L, int W>
__global__ void __launch_bounds__(kBlockM)
k_gradients_mixed(const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ accel,
                  const PG* __restrict__ pg, L lists, Box box, const int64_t* __restrict__ rows, int64_t nrows, int kid,
                  double* __restrict__ div, double* __restrict__ curl, double* __restrict__ ddiv)
{
    gradients_mixed_row<L, true, W>(pos, vel, accel, pg, lists, box, rows, nrows, kid, div, curl, ddiv);
}

template <class L, int W>
__global__ void __launch_bounds__(kBlockM)
k_divcurl_mixed(const double* __restrict__ pos, const double* __restrict__ vel, const double* __restrict__ accel,
                const PG* __restrict__ pg, L lists, Box box, const int64_t* __restrict__ rows, int64_t nrows, int kid,
                double* __restrict__ div, double* __restrict__ curl, double* __restrict__ ddiv)
{
    gradients_mixed_row<L, false, W>(pos, vel, accel, pg, lists, box, rows, nrows, kid, div, curl, ddiv);
}

static inline unsigned grid_m(int64_t nrows, int tile, int w) { return (unsigned)((nrows * w + tile - 1) / tile); }
static inline unsigned grid_1(int64_t n) { return (unsigned)((n + kBlockM - 1) / kBlockM); }

static inline Box box_of(const double* box3)
{
    Box b{{0.0, 0.0, 0.0}};
    if (box3)
        for (int c = 0; c < 3; c++) b.l[c] = box3[c];
    return b;
}

#define NB_DISPATCH_W(L, KERNEL, W, NROWS, STREAM, ...)                                                 \
    do {                                                                                                \
        const dim3 grid_(grid_m((NROWS), ctx->tile, (W)));                                              \
        if ((L)->length && (L)->row_major) {                                                            \
            RowView v_{(L)->index, (L)->length, (L)->offset};                                           \
            hipLaunchKernelGGL((KERNEL<RowView, W>), grid_, dim3(ctx->tile), 0, STREAM, __VA_ARGS__);    \
        } else if ((L)->length) {                                                                       \
            BlkView v_{(L)->index, (L)->length, (L)->row_cap, (L)->offset};                             \
            hipLaunchKernelGGL((KERNEL<BlkView, W>), grid_, dim3(ctx->tile), 0, STREAM, __VA_ARGS__);    \
        } else {                                                                                        \
            CsrView v_{(L)->offset, (L)->index};                                                        \
            hipLaunchKernelGGL((KERNEL<CsrView, W>), grid_, dim3(ctx->tile), 0, STREAM, __VA_ARGS__);    \
        }                                                                                               \
    } while (0)

#define NB_DISPATCH_M(L, KERNEL, W, NROWS, STREAM, ...)                                                 \
    do {                                                                                                \
        const int ts_ = ktime_begin(ctx, (STREAM), #KERNEL);                                            \
        NB_DISPATCH_W(L, KERNEL, W, NROWS, STREAM, __VA_ARGS__);                                        \
        ktime_end(ctx, (STREAM), ts_);                                                                  \
    } while (0)

static int32_t launched(nb_ctx* ctx, const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return NB_OK;
    const std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return ctx_fail(ctx, m.c_str());
}

// the mixed branch of nb_forces (precision 1, no reconstruction); arguments checked by the caller
int32_t forces_mixed(nb_ctx* ctx, hipStream_t s, const double* pos, const double* vel, const double* mass,
                     const double* h, const double* rho, const double* omega, const double* pressure, const double* u,
                     const double* c_s, const double* alpha, const nb_lists* lists, const double* box3,
                     const int64_t* rows, int64_t nrows, int64_t n, const nb_force_params* P, double* acc,
                     double* dudt, double* vsig_max, double* vheat, double* cheat)
{
    if (ctx_reserve(ctx, (void**)&ctx->pack, &ctx->pack_bytes, sizeof(PF) * (size_t)n)) return NB_ERR_HIP;
    PF* pf = (PF*)ctx->pack;
    hipLaunchKernelGGL(k_pack_forces, dim3(grid_1(n)), dim3(kBlockM), 0, s, mass, h, rho, omega, pressure,
                       P->use_cond ? u : nullptr, c_s, P->use_visc ? alpha : nullptr, P->cnorm, n, pf);
    const int mode = (P->use_visc ? 1 : 0) | (P->use_visc && P->use_cond ? 2 : 0);
    const int ts = ktime_begin(ctx, s, "k_forces_mixed");
#define NB_FORCES_MIXED(MODE)                                                                                      \
    NB_DISPATCH_W(lists, k_forces_mixed_##MODE, kPairW, nrows, s, pos, vel, mass, pf, c_s, v_, box_of(box3), rows, \
                  nrows, *P, acc, dudt, vsig_max, vheat, cheat)
    switch (mode) {
        case 0: NB_FORCES_MIXED(0); break;
        case 1: NB_FORCES_MIXED(1); break;
        default: NB_FORCES_MIXED(3); break;
    }
#undef NB_FORCES_MIXED
    ktime_end(ctx, s, ts);
    return launched(ctx, "nb_forces (mixed)");
}

}  // namespace nb

using namespace nb;

extern "C" int32_t nb_gradients_mixed(nb_ctx* ctx, void* stream, const double* pos, const double* vel,
                                      const double* acc, const double* mass, const double* h, const nb_lists* lists,
                                      const double* box3, const int64_t* rows, int64_t nrows, int32_t kid, double cnorm,
                                      double* div, double* curl, double* ddiv)
{
    if (!ctx || nrows < 0 || !lists || !lists->index || !(lists->length || lists->offset))
        return ctx_fail(ctx, "nb_gradients_mixed: bad arguments");
    if (nrows == 0) return NB_OK;
    const int64_t n = rows ? (int64_t)lists->n_rows : nrows;
    if (n <= 0) return ctx_fail(ctx, "nb_gradients_mixed: a row subset needs nb_lists.n_rows");
    if (ctx_reserve(ctx, (void**)&ctx->pack, &ctx->pack_bytes, sizeof(PG) * (size_t)n)) return NB_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    PG* pg = (PG*)ctx->pack;
    hipLaunchKernelGGL(k_pack_grad, dim3(grid_1(n)), dim3(kBlockM), 0, s, mass, h, cnorm, n, pg);
    if (acc)
        NB_DISPATCH_M(lists, k_gradients_mixed, kPairW, nrows, s, pos, vel, acc, pg, v_, box_of(box3), rows, nrows,
                      kid, div, curl, ddiv);
    else
        NB_DISPATCH_M(lists, k_divcurl_mixed, kDivCurlW, nrows, s, pos, vel, acc, pg, v_, box_of(box3), rows, nrows,
                      kid, div, curl, ddiv);
    return launched(ctx, "nb_gradients_mixed");
}
