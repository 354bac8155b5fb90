// Reuse of neighbour lists between density passes ("Verlet lists", sph.py:62-147): the exact
// completeness test of lists built at positions x_ref with search lengths h_search.
//
// A pair missing from the lists had r_build >= 2 max(hs_i, hs_j), so now
// r >= 2 max(hs_i, hs_j) - d_i - d_j with d the moves since the build. If every particle outside
// a small "fast" set F moved at most d_rest, the lists hold every pair now interacting with
// h_i <= h_eff_i = hs_i - (d_i + d_rest) / 2 between particles outside F; the rows of F are
// checked directly against the current positions. The host picks F from the largest moves and
// calls these three kernels per candidate size of F (sph.py:117-147).
#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

namespace nb {

static constexpr int kBlockR = 256;

static inline unsigned grid_r(int64_t n) { return (unsigned)((n + kBlockR - 1) / kBlockR); }

// |minimum_image(pos - x_ref)| per particle (sph.py:123-124)
__global__ void k_displacement(const double* __restrict__ pos, const double* __restrict__ x_ref, Box box, int64_t n,
                               double* __restrict__ disp)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < 3; c++) {
        double d = pos[3 * i + c] - x_ref[3 * i + c];
        if (box.l[c] > 0.0) d -= box.l[c] * rint(d / box.l[c]);
        s += d * d;
    }
    disp[i] = sqrt(s);
}

// h_eff = hs - (disp + d_rest) / 2 (sph.py:134); bad[0] counts solved rows with h_eff <= h_guess / 2,
// bad[1] other particles with h_guess > h_eff (sph.py:135-137). role: bit 0 set for the rows solved,
// bit 1 for the others (whose h is fixed); h_lim = h_guess for the others, else h_eff (sph.py:142-143).
__global__ void k_reuse_limits(const double* __restrict__ h_search, const double* __restrict__ disp, double d_rest,
                               const double* __restrict__ h_guess, const uint8_t* __restrict__ role, int64_t n,
                               double* __restrict__ h_eff, double* __restrict__ h_lim, int32_t* __restrict__ bad)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double he = h_search[i] - 0.5 * (disp[i] + d_rest);
    const double hg = h_guess[i];
    const uint8_t r = role[i];
    h_eff[i] = he;
    h_lim[i] = (r & 2) ? hg : he;
    if ((r & 1) && !(he > 0.5 * hg)) atomicAdd(&bad[0], 1);
    if ((r & 2) && !(hg <= he)) atomicAdd(&bad[1], 1);
}

// sph.py:86-111 (_fast_rows_complete): for every fast particle j and every particle k with
// r_jk <= 2 max(h_lim_j, h_lim_k) at the current positions, k must be in j's sorted CSR row;
// *missing counts the pairs that are not. One thread per (j, k).
__global__ void k_fast_rows(const double* __restrict__ pos, const int64_t* __restrict__ fast, int64_t nfast,
                            const double* __restrict__ h_lim, Box box, const int64_t* __restrict__ offset,
                            const int32_t* __restrict__ index, int64_t n, int32_t* __restrict__ missing)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nfast * n) return;
    const int64_t j = fast[t / n], k = t % n;
    double d2 = 0.0;
    for (int c = 0; c < 3; c++) {
        double dx = pos[3 * k + c] - pos[3 * j + c];
        if (box.l[c] > 0.0) dx -= box.l[c] * rint(dx / box.l[c]);
        d2 += dx * dx;
    }
    const double rr = 2.0 * pymax(h_lim[j], h_lim[k]);
    if (!(d2 <= rr * rr)) return;
    int64_t a = offset[j], b = offset[j + 1];
    const int64_t hi = b;
    while (a < b) {
        const int64_t m = (a + b) / 2;
        if ((int64_t)index[m] < k) a = m + 1;
        else b = m;
    }
    if (a >= hi || (int64_t)index[a] != k) atomicAdd(missing, 1);
}

static int32_t launched(nb_ctx* ctx, const char* what)
{
    return hipGetLastError() == hipSuccess ? NB_OK : ctx_fail(ctx, what);
}

static Box box_from(const double* box3)
{
    Box b{{0.0, 0.0, 0.0}};
    if (box3)
        for (int c = 0; c < 3; c++) b.l[c] = box3[c];
    return b;
}

}  // namespace nb

using namespace nb;

extern "C" int32_t nb_reuse_displacements(nb_ctx* ctx, void* stream, const double* pos, const double* x_ref,
                                          const double* box3, int64_t n, double* disp)
{
    if (!ctx || n < 0 || (n > 0 && (!pos || !x_ref || !disp))) return ctx_fail(ctx, "nb_reuse_displacements: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_displacement, dim3(grid_r(n)), dim3(kBlockR), 0, (hipStream_t)stream, pos, x_ref,
                       box_from(box3), n, disp);
    return launched(ctx, "nb_reuse_displacements: launch failed");
}

extern "C" int32_t nb_reuse_limits(nb_ctx* ctx, void* stream, const double* h_search, const double* disp, double d_rest,
                                   const double* h_guess, const uint8_t* role, int64_t n, double* h_eff, double* h_lim,
                                   int32_t* bad_dev)
{
    if (!ctx || n < 0 || !bad_dev || (n > 0 && (!h_search || !disp || !h_guess || !role || !h_eff || !h_lim)))
        return ctx_fail(ctx, "nb_reuse_limits: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_reuse_limits, dim3(grid_r(n)), dim3(kBlockR), 0, (hipStream_t)stream, h_search, disp, d_rest,
                       h_guess, role, n, h_eff, h_lim, bad_dev);
    return launched(ctx, "nb_reuse_limits: launch failed");
}

extern "C" int32_t nb_reuse_fast_rows(nb_ctx* ctx, void* stream, const double* pos, const int64_t* fast, int64_t nfast,
                                      const double* h_lim, const double* box3, const int64_t* offset,
                                      const int32_t* index, int64_t n, int32_t* missing_dev)
{
    if (!ctx || n < 0 || nfast < 0 || !missing_dev || (nfast > 0 && (!fast || !pos || !h_lim || !offset || !index)))
        return ctx_fail(ctx, "nb_reuse_fast_rows: bad arguments");
    if (nfast == 0 || n == 0) return NB_OK;
    if (nfast > (INT64_MAX / 2) / n) return ctx_fail(ctx, "nb_reuse_fast_rows: fast set too large");
    const int64_t total = nfast * n;
    hipLaunchKernelGGL(k_fast_rows, dim3((unsigned)((total + kBlockR - 1) / kBlockR)), dim3(kBlockR), 0,
                       (hipStream_t)stream, pos, fast, nfast, h_lim, box_from(box3), offset, index, n, missing_dev);
    return launched(ctx, "nb_reuse_fast_rows: launch failed");
}
