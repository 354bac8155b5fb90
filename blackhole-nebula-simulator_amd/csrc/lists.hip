// Neighbour lists on the GPU: cell grid, gather, symmetric union, tighten.
//
// The lists are those of the reference's find_neighbours_numba (neighbours.py:240-281):
// row i holds every j with d2 < (2 max(h_i, h_j))^2, where d2 = (dx*dx + dy*dy) + dz*dz in
// FP64 with dx = x_j - x_i (minimum image on periodic axes), sorted ascending, self included.
// The predicate is evaluated with exactly the reference's operations, so the lists are
// bit-identical whatever cell size the grid uses: the grid only bounds which candidates are
// tested, and its cell ranges are conservative (a candidate passing the predicate always lies
// in the range, by monotonicity of the rounded floor((x -+ r - lo) / cell)).
//
// Pipeline (all on one stream):
//   k_bbox/k_bbox_final    bounding box and the largest / mean search length
//   k_cell_id              dense raster cell id per particle
//   radix sort (hipCUB)    particles ordered by cell, ascending index within a cell
//   k_count_cells + scan   cell start offsets
//   k_gather               one work-item per particle: candidates from the cell range, the
//                          predicate with r = 2 h_i, written to a padded row; pairs that j's own
//                          gather will not see (d2 >= (2 h_j)^2) are flagged and counted for j
//   k_sort_rows            each padded row sorted ascending (insertion sort: rows arrive
//                          almost sorted when particles are stored in cell order)
//   k_extras_fill          the flagged pairs transposed into per-row extra lists
//   k_merge_emit           sorted gather row merged with its sorted extras -> CSR
#include "nb_common.h"
#include "nb_ctx.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nb {

static constexpr int kBlock = 256;
static constexpr uint32_t kFlag = 0x80000000u;
static constexpr uint32_t kMask = 0x7fffffffu;

// numpy's np.mod for floats (npy_divmod): result has the sign of the divisor
__device__ __forceinline__ double np_mod(double a, double b)
{
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0) != (m < 0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

__device__ __forceinline__ double wrapped(const double* pos, int64_t i, int c, const Box& box)
{
    double x = pos[3 * i + c];
    return box.l[c] > 0.0 ? np_mod(x, box.l[c]) : x;
}

// ----------------------------------------------------------------- bounding box
// out per block: minx, miny, minz, maxx, maxy, maxz, maxh, sumh
__global__ void k_bbox(const double* __restrict__ pos, const double* __restrict__ h, int64_t n,
                       Box box, double* __restrict__ part)
{
    double v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 3; c++) {
            double x = wrapped(pos, i, c, box);
            v[c] = fmin(v[c], x);
            v[3 + c] = fmax(v[3 + c], x);
        }
        v[6] = fmax(v[6], h[i]);
        v[7] += h[i];
    }
    __shared__ double sh[8][kBlock];
    for (int k = 0; k < 8; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
            for (int k = 3; k < 7; k++) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
            sh[7][threadIdx.x] += sh[7][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; k++) part[8 * blockIdx.x + k] = sh[k][0];
}

__global__ void k_bbox_final(const double* __restrict__ part, int nparts, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v[8] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0, 0.0};
    for (int b = 0; b < nparts; b++) {
        for (int k = 0; k < 3; k++) v[k] = fmin(v[k], part[8 * b + k]);
        for (int k = 3; k < 7; k++) v[k] = fmax(v[k], part[8 * b + k]);
        v[7] += part[8 * b + 7];
    }
    for (int k = 0; k < 8; k++) out[k] = v[k];
}

// ----------------------------------------------------------------- grid
struct Grid {
    double lo[3];
    double inv[3];
    int32_t nc[3];
};

__device__ __forceinline__ int32_t cell_of(double x, const Grid& g, int c)
{
    double f = floor((x - g.lo[c]) * g.inv[c]);
    int64_t k = (int64_t)f;
    if (f < 0.0) k = 0;  // also catches values far below the int range
    if (k >= g.nc[c]) k = g.nc[c] - 1;
    return (int32_t)k;
}

__global__ void k_cell_id(const double* __restrict__ pos, int64_t n, Box box, Grid g,
                          uint32_t* __restrict__ cell_id, uint32_t* __restrict__ iota,
                          uint32_t* __restrict__ cell_count)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t cx = cell_of(wrapped(pos, i, 0, box), g, 0);
    int32_t cy = cell_of(wrapped(pos, i, 1, box), g, 1);
    int32_t cz = cell_of(wrapped(pos, i, 2, box), g, 2);
    uint32_t c = ((uint32_t)cx * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[2] + (uint32_t)cz;
    cell_id[i] = c;
    iota[i] = (uint32_t)i;
    atomicAdd(&cell_count[c], 1u);
}

// neighbours.py:129-145: first cell and count along one axis; periodic searches that span the
// whole axis visit every cell once.
__device__ __forceinline__ void cell_range(double xc, double r, const Grid& g, int c, bool periodic,
                                           int64_t& a_out, int64_t& cnt_out)
{
    double fa = floor((xc - r - g.lo[c]) * g.inv[c]);
    double fb = floor((xc + r - g.lo[c]) * g.inv[c]);
    if (periodic) {
        double count = fb - fa + 1.0;
        if (count >= (double)g.nc[c]) { a_out = 0; cnt_out = g.nc[c]; return; }
        a_out = (int64_t)fa;
        cnt_out = (int64_t)count;
        return;
    }
    fa = fmax(fa, 0.0);
    fb = fmin(fb, (double)(g.nc[c] - 1));
    a_out = (int64_t)fa;
    cnt_out = fb >= fa ? (int64_t)(fb - fa) + 1 : 0;
}

__device__ __forceinline__ int32_t wrap_cell(int64_t raw, int32_t nc)
{
    int64_t m = raw % nc;
    return (int32_t)(m < 0 ? m + nc : m);
}

// One work-item per particle. Row i of the padded buffer holds the gather set
// {j : d2 < (2 h_i)^2}; entries carry kFlag when d2 >= (2 h_j)^2 (pair unseen from j's side),
// and ecount[j] counts those. rowcnt[i] is the true count even when it exceeds cap (then the
// row is truncated and the caller rebuilds with a larger cap; max_count records the largest).
__global__ void __launch_bounds__(kBlock)
k_gather(const double* __restrict__ pos, const double* __restrict__ h, int64_t n, Box box, Grid g,
         const uint32_t* __restrict__ start, const uint32_t* __restrict__ order,
         uint32_t* __restrict__ gbuf, int32_t cap, int32_t* __restrict__ rowcnt,
         unsigned long long* __restrict__ ecount, int32_t* __restrict__ max_count)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = NB_R_KERN * h[i];
    const double r2 = r * r;
    const double xi = wrapped(pos, i, 0, box), yi = wrapped(pos, i, 1, box), zi = wrapped(pos, i, 2, box);
    int64_t ax, nx, ay, ny, az, nz;
    cell_range(xi, r, g, 0, box.l[0] > 0.0, ax, nx);
    cell_range(yi, r, g, 1, box.l[1] > 0.0, ay, ny);
    cell_range(zi, r, g, 2, box.l[2] > 0.0, az, nz);
    int32_t m = 0;
    uint32_t* row = gbuf + (int64_t)i * cap;
    for (int64_t di = 0; di < nx; di++) {
        int32_t cx = wrap_cell(ax + di, g.nc[0]);
        for (int64_t dj = 0; dj < ny; dj++) {
            int32_t cy = wrap_cell(ay + dj, g.nc[1]);
            uint32_t base = ((uint32_t)cx * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[2];
            for (int64_t dk = 0; dk < nz; dk++) {
                int32_t cz = wrap_cell(az + dk, g.nc[2]);
                uint32_t c = base + (uint32_t)cz;
                uint32_t k0 = start[c], k1 = start[c + 1];
                for (uint32_t k = k0; k < k1; k++) {
                    uint32_t j = order[k];
                    double dx = min_image(wrapped(pos, j, 0, box) - xi, box.l[0]);
                    double dy = min_image(wrapped(pos, j, 1, box) - yi, box.l[1]);
                    double dz = min_image(wrapped(pos, j, 2, box) - zi, box.l[2]);
                    double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < r2) {
                        double rj = NB_R_KERN * h[j];
                        uint32_t e = j;
                        if (!(d2 < rj * rj)) {
                            e |= kFlag;
                            atomicAdd(&ecount[j], 1ull);
                        }
                        if (m < cap) row[m] = e;
                        m++;
                    }
                }
            }
        }
    }
    rowcnt[i] = m;
    atomicMax(max_count, m);
}

// Insertion sort of each padded row by index (flag bit ignored). Rows arrive nearly sorted
// when the particles are stored in cell order, which makes this linear in practice.
__global__ void k_sort_rows(uint32_t* __restrict__ gbuf, int32_t cap, const int32_t* __restrict__ rowcnt,
                            int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* row = gbuf + (int64_t)i * cap;
    int32_t m = rowcnt[i];
    for (int32_t a = 1; a < m; a++) {
        uint32_t v = row[a];
        uint32_t key = v & kMask;
        int32_t b = a - 1;
        while (b >= 0 && (row[b] & kMask) > key) {
            row[b + 1] = row[b];
            b--;
        }
        row[b + 1] = v;
    }
}

__global__ void k_extras_fill(const uint32_t* __restrict__ gbuf, int32_t cap,
                              const int32_t* __restrict__ rowcnt,
                              const unsigned long long* __restrict__ eoff,
                              unsigned int* __restrict__ efill, uint32_t* __restrict__ ebuf, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* row = gbuf + (int64_t)i * cap;
    int32_t m = rowcnt[i];
    for (int32_t a = 0; a < m; a++) {
        uint32_t v = row[a];
        if (v & kFlag) {
            uint32_t j = v & kMask;
            unsigned int s = atomicAdd(&efill[j], 1u);
            ebuf[eoff[j] + s] = (uint32_t)i;
        }
    }
}

__global__ void k_row_total(const int32_t* __restrict__ rowcnt, const unsigned long long* __restrict__ ecount,
                            unsigned long long* __restrict__ tot, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) tot[i] = (unsigned long long)rowcnt[i] + ecount[i];
    if (i == n) tot[i] = 0;
}

// The extras of a row are sorted in place, then merged with the sorted gather row. The two
// sets are disjoint by construction (an extra j of row i has d2 >= (2 h_i)^2).
__global__ void k_merge_emit(const uint32_t* __restrict__ gbuf, int32_t cap,
                             const int32_t* __restrict__ rowcnt, const unsigned long long* __restrict__ eoff,
                             uint32_t* __restrict__ ebuf, const unsigned long long* __restrict__ off,
                             int32_t* __restrict__ out_idx, int64_t out_cap, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* e = ebuf + eoff[i];
    int64_t ne = (int64_t)(eoff[i + 1] - eoff[i]);
    for (int64_t a = 1; a < ne; a++) {
        uint32_t v = e[a];
        int64_t b = a - 1;
        while (b >= 0 && e[b] > v) {
            e[b + 1] = e[b];
            b--;
        }
        e[b + 1] = v;
    }
    const uint32_t* g = gbuf + (int64_t)i * cap;
    int32_t ng = rowcnt[i];
    int64_t o = (int64_t)off[i];
    if (o + ng + ne > out_cap) return;
    int32_t p = 0;
    int64_t q = 0;
    while (p < ng || q < ne) {
        uint32_t v;
        if (q >= ne || (p < ng && (g[p] & kMask) < e[q])) v = g[p++] & kMask;
        else v = e[q++];
        out_idx[o++] = (int32_t)v;
    }
}

// ----------------------------------------------------------------- tighten
// neighbours.py:284-303: keep j with d2 < (2 max(h_i, h_j))^2 (pos as given, minimum image).
template <bool kFill>
__global__ void k_tighten(const double* __restrict__ pos, const double* __restrict__ h, Box box,
                          const int64_t* __restrict__ offset, const int32_t* __restrict__ index,
                          const int64_t* __restrict__ rows, int64_t nrows,
                          unsigned long long* __restrict__ counts, const unsigned long long* __restrict__ out_off,
                          int32_t* __restrict__ out_idx)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    int64_t i = row_of(rows, q);
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double hi = h[i];
    int64_t m = 0;
    int64_t o = kFill ? (int64_t)out_off[i] : 0;
    for (int64_t k = offset[i]; k < offset[i + 1]; k++) {
        int32_t j = index[k];
        double dx = min_image(pos[3 * j] - xi, box.l[0]);
        double dy = min_image(pos[3 * j + 1] - yi, box.l[1]);
        double dz = min_image(pos[3 * j + 2] - zi, box.l[2]);
        double d2 = 0.0;
        d2 += dx * dx;
        d2 += dy * dy;
        d2 += dz * dz;
        double rr = NB_R_KERN * pymax(hi, h[j]);
        if (d2 < rr * rr) {
            if (kFill) out_idx[o + m] = j;
            m++;
        }
    }
    if (!kFill) counts[i] = (unsigned long long)m;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static inline unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace nb

// =================================================================== host side
using namespace nb;

namespace {

// exclusive scan of n+1 values (the last input is 0), into out (n+1); returns out[n]
int scan_u64(nb_ctx* ctx, hipStream_t s, const unsigned long long* in, unsigned long long* out,
             int64_t count)
{
    size_t bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)count, s) != hipSuccess)
        return ctx_fail(ctx, "hipcub scan sizing failed");
    if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
    if (hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, bytes, in, out, (int)count, s) != hipSuccess)
        return ctx_fail(ctx, "hipcub scan failed");
    return NB_OK;
}

}  // namespace

static const int64_t kMaxAxisCells = 2048;
static const double kMaxTotalCells = 134217728.0;  // 2^27 cells (512 MiB of start offsets)

extern "C" int32_t nb_lists_build(nb_ctx* ctx, void* stream, const double* pos, const double* h,
                                  int64_t n, const double* box3, int64_t* out_offset,
                                  int32_t* out_index, int64_t capacity, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || !total_out) return ctx_fail(ctx, "nb_lists_build: bad arguments");
    ctx->lists_ready = 0;
    if (n >= (int64_t)kMask) return ctx_fail(ctx, "nb_lists_build: too many particles");
    Box box{{0.0, 0.0, 0.0}};
    if (box3) for (int c = 0; c < 3; c++) box.l[c] = box3[c];
    if (n == 0) {
        if (hipMemsetAsync(out_offset, 0, sizeof(int64_t), s) != hipSuccess) return ctx_fail(ctx, "memset");
        *total_out = 0;
        ctx->lists_n = 0;
        ctx->lists_total = 0;
        ctx->lists_ready = 1;
        return NB_OK;
    }
    // bounding box (host needs it for the grid shape)
    int nparts = (int)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 8 * (nparts + 1)))
        return NB_ERR_HIP;
    hipLaunchKernelGGL(k_bbox, dim3(nparts), dim3(kBlock), 0, s, pos, h, n, box, ctx->red);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(1), 0, s, ctx->red, nparts, ctx->red + 8 * nparts);
    double bb[8];
    if (hipMemcpyAsync(bb, ctx->red + 8 * nparts, sizeof(bb), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "nb_lists_build: bounding box copy failed");
    for (int c = 0; c < 3; c++)
        if (box.l[c] > 0.0 && NB_R_KERN * bb[6] >= 0.5 * box.l[c])
            return ctx_fail_code(ctx, NB_ERR_BOX,
                                 "kernel support exceeds half the periodic box: neighbours ambiguous");
    if (!(bb[6] == bb[6]) || !std::isfinite(bb[0] + bb[1] + bb[2] + bb[3] + bb[4] + bb[5]))
        return ctx_fail(ctx, "nb_lists_build: non-finite positions or smoothing lengths");
    // grid: cell about the mean search length, bounded in count
    Grid g;
    double span[3];
    for (int c = 0; c < 3; c++) {
        g.lo[c] = box.l[c] > 0.0 ? 0.0 : bb[c];
        span[c] = box.l[c] > 0.0 ? box.l[c] : bb[3 + c] - bb[c];
    }
    double maxspan = std::max(span[0], std::max(span[1], span[2]));
    double cell = std::max(std::max(bb[7] / (double)n, maxspan / (double)kMaxAxisCells), 1e-300);
    for (;;) {
        double total = 1.0;
        for (int c = 0; c < 3; c++) {
            if (box.l[c] > 0.0) {
                double k = std::max(std::floor(box.l[c] / cell), 1.0);
                g.nc[c] = (int32_t)k;
                g.inv[c] = k / box.l[c];
            } else {
                g.nc[c] = (int32_t)std::max(std::floor(span[c] / cell) + 1.0, 1.0);
                g.inv[c] = 1.0 / cell;
            }
            total *= (double)g.nc[c];
        }
        if (total <= kMaxTotalCells) break;
        cell *= 1.25;
    }
    const int64_t ncells = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
    // buffers
    if (ctx_reserve(ctx, (void**)&ctx->cell_id, &ctx->cell_id_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->order, &ctx->order_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->keys_out, &ctx->keys_out_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->iota, &ctx->iota_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->cell_count, &ctx->cell_count_bytes, sizeof(uint32_t) * (ncells + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->cell_start, &ctx->cell_start_bytes, sizeof(uint32_t) * (ncells + 1)))
        return NB_ERR_HIP;
    if (hipMemsetAsync(ctx->cell_count, 0, sizeof(uint32_t) * (ncells + 1), s) != hipSuccess)
        return ctx_fail(ctx, "memset cell_count");
    hipLaunchKernelGGL(k_cell_id, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, n, box, g,
                       ctx->cell_id, ctx->iota, ctx->cell_count);
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < ncells && end_bit < 32) end_bit++;
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ctx->cell_id, ctx->keys_out, ctx->iota,
                                           ctx->order, (int)n, 0, end_bit, s) != hipSuccess)
        return ctx_fail(ctx, "hipcub sort sizing failed");
    if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
    if (hipcub::DeviceRadixSort::SortPairs(ctx->cub_tmp, bytes, ctx->cell_id, ctx->keys_out, ctx->iota,
                                           ctx->order, (int)n, 0, end_bit, s) != hipSuccess)
        return ctx_fail(ctx, "hipcub sort failed");
    bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, ctx->cell_count, ctx->cell_start,
                                         (int)(ncells + 1), s) != hipSuccess)
        return ctx_fail(ctx, "hipcub scan sizing failed");
    if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
    if (hipcub::DeviceScan::ExclusiveSum(ctx->cub_tmp, bytes, ctx->cell_count, ctx->cell_start,
                                         (int)(ncells + 1), s) != hipSuccess)
        return ctx_fail(ctx, "hipcub scan failed");
    // gather into padded rows; grow the row capacity until every row fits
    if (ctx_reserve(ctx, (void**)&ctx->rowcnt, &ctx->rowcnt_bytes, sizeof(int32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->ecount, &ctx->ecount_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->eoff, &ctx->eoff_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->tot, &ctx->tot_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->efill, &ctx->efill_bytes, sizeof(unsigned int) * n) ||
        ctx_reserve(ctx, (void**)&ctx->offs, &ctx->offs_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->scal, &ctx->scal_bytes, sizeof(int64_t) * 4))
        return NB_ERR_HIP;
    int32_t cap = ctx->row_cap > 0 ? ctx->row_cap : 256;
    for (int attempt = 0; attempt < 8; attempt++) {
        if (ctx_reserve(ctx, (void**)&ctx->gbuf, &ctx->gbuf_bytes, sizeof(uint32_t) * (size_t)n * cap))
            return NB_ERR_HIP;
        if (hipMemsetAsync(ctx->ecount, 0, sizeof(unsigned long long) * (n + 1), s) != hipSuccess ||
            hipMemsetAsync(ctx->scal, 0, sizeof(int64_t) * 4, s) != hipSuccess)
            return ctx_fail(ctx, "memset");
        hipLaunchKernelGGL(k_gather, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h, n, box, g,
                           ctx->cell_start, ctx->order, ctx->gbuf, cap, ctx->rowcnt, ctx->ecount,
                           (int32_t*)ctx->scal);
        int32_t maxc = 0;
        if (hipMemcpyAsync(&maxc, ctx->scal, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return ctx_fail(ctx, "nb_lists_build: gather failed");
        if (maxc <= cap) break;
        cap = ((maxc + maxc / 4 + 63) / 64) * 64;
        if (attempt == 7) return ctx_fail(ctx, "nb_lists_build: row capacity did not settle");
    }
    ctx->row_cap = cap;
    hipLaunchKernelGGL(k_sort_rows, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, cap, ctx->rowcnt, n);
    int rc = scan_u64(ctx, s, ctx->ecount, ctx->eoff, n + 1);
    if (rc) return rc;
    unsigned long long ne_total = 0;
    if (hipMemcpyAsync(&ne_total, ctx->eoff + n, sizeof(ne_total), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "nb_lists_build: extras count copy failed");
    if (ctx_reserve(ctx, (void**)&ctx->ebuf, &ctx->ebuf_bytes, sizeof(uint32_t) * (ne_total + 1)))
        return NB_ERR_HIP;
    if (hipMemsetAsync(ctx->efill, 0, sizeof(unsigned int) * n, s) != hipSuccess) return ctx_fail(ctx, "memset");
    hipLaunchKernelGGL(k_extras_fill, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, cap, ctx->rowcnt,
                       ctx->eoff, ctx->efill, ctx->ebuf, n);
    hipLaunchKernelGGL(k_row_total, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->rowcnt, ctx->ecount,
                       ctx->tot, n);
    rc = scan_u64(ctx, s, ctx->tot, ctx->offs, n + 1);
    if (rc) return rc;
    unsigned long long total = 0;
    if (hipMemcpyAsync(&total, ctx->offs + n, sizeof(total), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "nb_lists_build: total copy failed");
    if (hipMemcpyAsync(out_offset, ctx->offs, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "offset copy failed");
    *total_out = (int64_t)total;
    ctx->lists_n = n;
    ctx->lists_total = (int64_t)total;
    ctx->lists_ready = 1;
    ctx->last_grid_cells = ncells;
    ctx->last_cell = cell;
    if ((int64_t)total > capacity || !out_index) return NB_RETRY_CAPACITY;
    hipLaunchKernelGGL(k_merge_emit, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, cap, ctx->rowcnt,
                       ctx->eoff, ctx->ebuf, ctx->offs, out_index, capacity, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_lists_build: kernel launch failed");
    ctx->lists_ready = 2;
    return NB_OK;
}

extern "C" int32_t nb_lists_emit(nb_ctx* ctx, void* stream, int32_t* out_index, int64_t capacity)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->lists_ready) return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_lists_emit: no lists built");
    if (ctx->lists_total > capacity) return ctx_fail(ctx, "nb_lists_emit: capacity too small");
    int64_t n = ctx->lists_n;
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_merge_emit, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, ctx->row_cap,
                       ctx->rowcnt, ctx->eoff, ctx->ebuf, ctx->offs, out_index, capacity, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_lists_emit: kernel launch failed");
    return NB_OK;
}

extern "C" int32_t nb_tighten(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                              const double* box3, const int64_t* offset, const int32_t* index,
                              const int64_t* rows, int64_t nrows, int64_t* out_offset, int32_t* out_index,
                              int64_t capacity, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || nrows < 0 || !total_out) return ctx_fail(ctx, "nb_tighten: bad arguments");
    Box box{{0.0, 0.0, 0.0}};
    if (box3) for (int c = 0; c < 3; c++) box.l[c] = box3[c];
    if (ctx_reserve(ctx, (void**)&ctx->tcount, &ctx->tcount_bytes, sizeof(unsigned long long) * (n + 1)))
        return NB_ERR_HIP;
    // rows not in ``rows`` stay empty
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->tcount, n + 1, 0ull);
    if (nrows > 0)
        hipLaunchKernelGGL((k_tighten<false>), dim3(grid_of(nrows)), dim3(kBlock), 0, s, pos, h, box, offset,
                           index, rows, nrows, ctx->tcount, nullptr, nullptr);
    int rc = scan_u64(ctx, s, ctx->tcount, (unsigned long long*)out_offset, n + 1);
    if (rc) return rc;
    int64_t total = 0;
    if (hipMemcpyAsync(&total, out_offset + n, sizeof(total), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "nb_tighten: total copy failed");
    *total_out = total;
    if (total > capacity) return NB_RETRY_CAPACITY;
    if (nrows > 0)
        hipLaunchKernelGGL((k_tighten<true>), dim3(grid_of(nrows)), dim3(kBlock), 0, s, pos, h, box, offset,
                           index, rows, nrows, nullptr, (const unsigned long long*)out_offset, out_index);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_tighten: kernel launch failed");
    return NB_OK;
}
