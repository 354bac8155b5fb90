// Neighbour lists on the GPU: cell grid, gather, symmetric union, tighten, grid ordering.
//
// The lists are those of the reference's find_neighbours_numba (neighbours.py:240-281):
// row i holds every j with d2 < (2 max(h_i, h_j))^2, where d2 = (dx*dx + dy*dy) + dz*dz in
// FP64 with dx = x_j - x_i (minimum image on periodic axes), sorted ascending, self included.
// The predicate is evaluated with exactly the reference's operations, so the lists are
// bit-identical whatever cell size the grid uses: the grid only bounds which candidates are
// tested, and its cell ranges are conservative (a candidate passing the predicate always lies
// in the range, by monotonicity of the rounded floor((x -+ r - lo) / cell)).
//
// Row layout. Gathered rows, and the lists handed to the pair passes, are stored in 64-row
// blocks: entry k of row i sits at ((i / 64) * cap + k) * 64 + i % 64, so when the 64 lanes of
// a wavefront walk their own rows in step, every load and store of entry k is one contiguous
// 256-byte transaction. The CSR form (int64 offsets) is produced for callers that want the
// reference's Neighbours arrays.
//
// Grid. Cells are cubes of about the mean search length (at least 1/2048 of the widest span,
// at most 2^27 cells), numbered in raster order (x slowest). The mean is an order-independent
// fixed-point sum (each h quantised to h / h_max * 2^32), so the grid, and the particle order it
// induces, do not depend on the reduction order. Candidates are visited cell by cell in raster
// order and, inside a cell, in ascending index: when the particles are stored in this grid's
// order (nb_grid_order), every gathered row comes out sorted and the sort pass does no work.
//
// Pipeline (one stream): k_bbox / k_bbox_final / k_hsum (grid shape), k_cell_id, radix sort
// (rocPRIM radix sort: by cell, ascending index inside a cell), k_count via atomics + scan (cell starts),
// k_gather (one work-item per particle; pairs that j's own gather will not see, d2 >= (2 h_j)^2,
// are flagged and counted for j), k_sort_rows (only rows that came out unsorted), k_extras_fill
// (the flagged pairs transposed into per-row extra lists), k_merge_* (sorted gather row merged
// with its sorted extras, into CSR or blocked rows).
#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nb {

static constexpr int kBlock = 256;
static constexpr uint32_t kFlag = 0x80000000u;
static constexpr uint32_t kMask = 0x7fffffffu;

__device__ __forceinline__ int64_t blk(int64_t i, int64_t k, int64_t cap) { return ((i >> 6) * cap + k) * 64 + (i & 63); }

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

// ----------------------------------------------------------------- grid shape
// partial rows per block: minx, miny, minz, maxx, maxy, maxz, maxh, (unused)
__global__ void k_bbox(const double* __restrict__ pos, const double* __restrict__ h, int64_t n, Box box,
                       double* __restrict__ part)
{
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < 3; c++) {
            double x = wrapped(pos, i, c, box);
            v[c] = fmin(v[c], x);
            v[3 + c] = fmax(v[3 + c], x);
        }
        v[6] = fmax(v[6], h[i]);
    }
    __shared__ double sh[7][kBlock];
    for (int k = 0; k < 7; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
            for (int k = 3; k < 7; k++) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 7; k++) part[8 * blockIdx.x + k] = sh[k][0];
}

// out[0..6] = bbox + max h; out[7] = 0 (the fixed-point sum lands in the u64 after it).
// One block; min / max are exact in any order.
__global__ void k_bbox_final(const double* __restrict__ part, int nparts, double* __restrict__ out,
                             unsigned long long* __restrict__ hsum)
{
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        for (int k = 0; k < 3; k++) v[k] = fmin(v[k], part[8 * b + k]);
        for (int k = 3; k < 7; k++) v[k] = fmax(v[k], part[8 * b + k]);
    }
    __shared__ double sh[7][kBlock];
    for (int k = 0; k < 7; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
            for (int k = 3; k < 7; k++) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 7; k++) out[k] = sh[k][0];
        out[7] = 0.0;
        *hsum = 0ull;
    }
}

// order-independent sum of floor(h_i * (2^32 / h_max)) (integer additions commute exactly)
__global__ void k_hsum(const double* __restrict__ h, int64_t n, const double* __restrict__ bb,
                       unsigned long long* __restrict__ hsum)
{
    const double hmax = bb[6];
    const double scale = hmax > 0.0 ? 4294967296.0 / hmax : 0.0;
    unsigned long long s = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += (unsigned long long)(h[i] * scale);
    __shared__ unsigned long long sh[kBlock];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(hsum, sh[0]);
}

struct Grid {
    double lo[3];
    double inv[3];
    int32_t nc[3];
};

// The segmented grids (the mixed-precision passes: an ensemble, or one cloud as one segment) cut
// each x-y cell's column into kSegZSub cells along z. The search walks columns and takes one
// contiguous run of cells per column, so shorter z cells trim each run closer to the search
// sphere's chord at no extra cost per column.
static constexpr double kSegZSub = 4.0;

// One particle system of an ensemble in the segmented grid: its own cells, numbered from base.
struct GridSeg {
    Grid g;
    uint32_t base;
    uint32_t pad;
};

__device__ __forceinline__ int32_t cell_of(double x, const Grid& g, int c)
{
    double f = floor((x - g.lo[c]) * g.inv[c]);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)g.nc[c]) return g.nc[c] - 1;
    return (int32_t)f;
}

__device__ __forceinline__ uint32_t cell_key(const double* pos, int64_t i, const Box& box, const Grid& g)
{
    const int32_t cx = cell_of(wrapped(pos, i, 0, box), g, 0);
    const int32_t cy = cell_of(wrapped(pos, i, 1, box), g, 1);
    const int32_t cz = cell_of(wrapped(pos, i, 2, box), g, 2);
    return ((uint32_t)cx * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[2] + (uint32_t)cz;
}

// Bounding box, largest h and the fixed-point sum of h (as k_bbox / k_hsum) of each segment
// [off[s], off[s + 1]), in `parts` workgroups per segment (grid nseg x parts): k_seg_bbox_part
// writes each workgroup's box and h_max (8 doubles per workgroup), k_seg_bbox_final combines them
// into out[9 s + 0..6]; k_seg_hsum_part / _final do the same for the u64 sum (out[9 s + 8]).
// Minima, maxima and integer sums do not depend on the split.
__device__ __forceinline__ void bbox_block_reduce(double* v, double (*sh)[kBlock])
{
    for (int k = 0; k < 7; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + st]);
            for (int k = 3; k < 7; k++) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + st]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ unsigned long long u64_block_sum(unsigned long long v, unsigned long long* su)
{
    su[threadIdx.x] = v;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) su[threadIdx.x] += su[threadIdx.x + st];
        __syncthreads();
    }
    return su[0];
}

__global__ void __launch_bounds__(kBlock)
k_seg_bbox_part(const double* __restrict__ pos, const double* __restrict__ h, const int64_t* __restrict__ off,
                int parts, double* __restrict__ part)
{
    const int64_t a = off[blockIdx.x], b = off[blockIdx.x + 1];
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int64_t i = a + (int64_t)blockIdx.y * blockDim.x + threadIdx.x; i < b; i += (int64_t)parts * blockDim.x) {
        for (int c = 0; c < 3; c++) {
            v[c] = fmin(v[c], pos[3 * i + c]);
            v[3 + c] = fmax(v[3 + c], pos[3 * i + c]);
        }
        v[6] = fmax(v[6], h[i]);
    }
    __shared__ double sh[7][kBlock];
    bbox_block_reduce(v, sh);
    if (threadIdx.x == 0) {
        double* o = part + 8 * ((int64_t)blockIdx.x * parts + blockIdx.y);
        for (int k = 0; k < 7; k++) o[k] = sh[k][0];
    }
}

__global__ void __launch_bounds__(kBlock) k_seg_bbox_final(const double* __restrict__ part, int parts,
                                                           double* __restrict__ out)
{
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int q = threadIdx.x; q < parts; q += blockDim.x) {
        const double* o = part + 8 * ((int64_t)blockIdx.x * parts + q);
        for (int k = 0; k < 3; k++) v[k] = fmin(v[k], o[k]);
        for (int k = 3; k < 7; k++) v[k] = fmax(v[k], o[k]);
    }
    __shared__ double sh[7][kBlock];
    bbox_block_reduce(v, sh);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 7; k++) out[9 * blockIdx.x + k] = sh[k][0];
        out[9 * blockIdx.x + 7] = 0.0;
    }
}

__global__ void __launch_bounds__(kBlock)
k_seg_hsum_part(const double* __restrict__ h, const int64_t* __restrict__ off, const double* __restrict__ out,
                int parts, unsigned long long* __restrict__ hpart)
{
    const int64_t a = off[blockIdx.x], b = off[blockIdx.x + 1];
    const double hmax = out[9 * blockIdx.x + 6];
    const double scale = hmax > 0.0 ? 4294967296.0 / hmax : 0.0;
    unsigned long long hs = 0;
    for (int64_t i = a + (int64_t)blockIdx.y * blockDim.x + threadIdx.x; i < b; i += (int64_t)parts * blockDim.x)
        hs += (unsigned long long)(h[i] * scale);
    __shared__ unsigned long long su[kBlock];
    const unsigned long long t = u64_block_sum(hs, su);
    if (threadIdx.x == 0) hpart[(int64_t)blockIdx.x * parts + blockIdx.y] = t;
}

__global__ void __launch_bounds__(kBlock) k_seg_hsum_final(const unsigned long long* __restrict__ hpart, int parts,
                                                           double* __restrict__ out)
{
    unsigned long long hs = 0;
    for (int q = threadIdx.x; q < parts; q += blockDim.x) hs += hpart[(int64_t)blockIdx.x * parts + q];
    __shared__ unsigned long long su[kBlock];
    const unsigned long long t = u64_block_sum(hs, su);
    if (threadIdx.x == 0) ((unsigned long long*)out)[9 * blockIdx.x + 8] = t;
}

__device__ __forceinline__ uint32_t seg_key(const double* pos, int64_t i, const GridSeg* segs, const int32_t* cloud)
{
    const GridSeg& G = segs[cloud[i]];
    const Box open{{0.0, 0.0, 0.0}};
    return G.base + cell_key(pos, i, open, G.g);
}

// segmented cell ids (global numbering), as k_cell_id
__global__ void __launch_bounds__(kBlock)
k_cell_id_seg(const double* __restrict__ pos, int64_t n, const GridSeg* __restrict__ segs,
              const int32_t* __restrict__ cloud, uint32_t* __restrict__ cell_id, uint32_t* __restrict__ iota,
              uint32_t* __restrict__ cell_count, long long* __restrict__ part)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    long long bad = 0;
    if (i < n) {
        const uint32_t c = seg_key(pos, i, segs, cloud);
        cell_id[i] = c;
        iota[i] = (uint32_t)i;
        if (cell_count) atomicAdd(&cell_count[c], 1u);  // NULL: the starts come from the sorted ids
        if (part && i + 1 < n && seg_key(pos, i + 1, segs, cloud) < c) bad = 1;
    }
    if (part) block_partial<false>(bad, part);
}

// The grid of the last order pass, reused by the density pass that follows it (positions and
// search lengths the same, permuted into cell order): every particle must lie inside its
// segment's grid (unclamped cell coordinates) and inside its cell's range of the sorted order
// (start[c] <= i < start[c + 1]); that makes the ranges exactly this arrangement's cells.
// part: the particles that fail.
__global__ void __launch_bounds__(kBlock)
k_check_grid_seg(const double* __restrict__ pos, int64_t n, const GridSeg* __restrict__ segs,
                 const int32_t* __restrict__ cloud, const uint32_t* __restrict__ start, long long* __restrict__ part)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    long long bad = 0;
    if (i < n) {
        const GridSeg& G = segs[cloud[i]];
        int32_t cc[3];
        bool inside = true;
        for (int a = 0; a < 3; a++) {
            const double f = floor((pos[3 * i + a] - G.g.lo[a]) * G.g.inv[a]);
            inside = inside && f >= 0.0 && f < (double)G.g.nc[a];
            cc[a] = inside ? (int32_t)f : 0;
        }
        if (!inside) {
            bad = 1;
        } else {
            const uint32_t c =
                G.base + ((uint32_t)cc[0] * (uint32_t)G.g.nc[1] + (uint32_t)cc[1]) * (uint32_t)G.g.nc[2] + (uint32_t)cc[2];
            bad = start[c] <= (uint32_t)i && (uint32_t)i < start[c + 1] ? 0 : 1;
        }
    }
    block_partial<false>(bad, part);
}

// cell of each particle, particles per cell, and (part, one partial per workgroup) the number of
// neighbouring pairs i, i + 1 out of cell order: zero when the particles are stored sorted.
__global__ void __launch_bounds__(kBlock)
k_cell_id(const double* __restrict__ pos, int64_t n, Box box, Grid g, uint32_t* __restrict__ cell_id,
          uint32_t* __restrict__ iota, uint32_t* __restrict__ cell_count, long long* __restrict__ part)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    long long bad = 0;
    if (i < n) {
        const uint32_t c = cell_key(pos, i, box, g);
        cell_id[i] = c;
        iota[i] = (uint32_t)i;
        atomicAdd(&cell_count[c], 1u);
        if (part && i + 1 < n && cell_key(pos, i + 1, box, g) < c) bad = 1;
    }
    if (part) block_partial<false>(bad, part);
}

// Sum (bit s of maxmask clear) or max (set) of the per-workgroup partials of nslots slots
// (slot s at part[s * nb + b]) into out[s] (zeroed beforehand): each workgroup reduces a strided
// share in LDS and adds its result with one atomic per slot (a few hundred at most).
// status words that a reduced slot's nonzero total sets (-1: none), for the fixed passes
struct SlotStatus {
    int word[5];
};

__global__ void __launch_bounds__(kBlock)
k_partials_final(const long long* __restrict__ part, int64_t nb, int nslots, unsigned maxmask, long long* out,
                 int32_t* __restrict__ status, SlotStatus ss)
{
    __shared__ long long sh[kBlock];
    for (int sl = 0; sl < nslots; sl++) {
        const bool mx = (maxmask >> sl) & 1u;
        long long v = 0;
        for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
            const long long t = part[sl * nb + b];
            v = mx ? (t > v ? t : v) : v + t;
        }
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int k = blockDim.x / 2; k > 0; k >>= 1) {
            if ((int)threadIdx.x < k) {
                const long long t = sh[threadIdx.x + k];
                sh[threadIdx.x] = mx ? (t > sh[threadIdx.x] ? t : sh[threadIdx.x]) : sh[threadIdx.x] + t;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (mx) atomicMax(&out[sl], sh[0]);
            else atomicAdd((unsigned long long*)&out[sl], (unsigned long long)sh[0]);
            // the slots hold counts and maxima (never negative): the total is nonzero exactly
            // when some workgroup's part is, and every such workgroup stores the same 1
            if (status && sl < 5 && ss.word[sl] >= 0 && sh[0] != 0) status[ss.word[sl]] = 1;
        }
        __syncthreads();
    }
}

// out[k] = src[k] for a few device scalars (collects values for one host copy)
__global__ void k_collect(long long* out, const unsigned long long* a, const unsigned long long* b)
{
    if (threadIdx.x == 0) {
        if (a) out[0] = (long long)*a;
        if (b) out[1] = (long long)*b;
    }
}

// neighbours.py:129-145: first cell and count along one axis; periodic searches that span the
// whole axis visit every cell once.
__device__ __forceinline__ void cell_range(double xc, double r, const Grid& g, int c, bool periodic, int64_t& a_out,
                                           int64_t& cnt_out)
{
    double fa = floor((xc - r - g.lo[c]) * g.inv[c]);
    double fb = floor((xc + r - g.lo[c]) * g.inv[c]);
    if (periodic) {
        double count = fb - fa + 1.0;
        if (count >= (double)g.nc[c]) {
            a_out = 0;
            cnt_out = g.nc[c];
            return;
        }
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

// ----------------------------------------------------------------- gather
// One work-item per particle; row i (blocked, capacity cap) holds {j : d2 < (2 h_i)^2} in
// traversal order; entries carry kFlag when d2 >= (2 h_j)^2 (pair unseen from j's side) and
// ecount[j] counts those. rowcnt[i] is the true count even past cap (then the caller rebuilds
// with a larger cap; the longest row goes to the workgroup partials). Rows that come out
// unsorted are marked and counted. order NULL: particles stored in cell order.
__global__ void __launch_bounds__(kBlock)
k_gather(const double* __restrict__ pos, const double* __restrict__ h, int64_t n, Box box, Grid g,
         const uint32_t* __restrict__ start, const uint32_t* __restrict__ order, uint32_t* __restrict__ gbuf,
         int64_t cap, int32_t* __restrict__ rowcnt, uint8_t* __restrict__ unsorted,
         unsigned long long* __restrict__ ecount, long long* __restrict__ part)
{
    // lanes past the end redo the last particle without writing (full blocks for the reductions)
    int64_t i = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    const bool live = i < n;
    if (!live) i = n - 1;
    const double r = NB_R_KERN * h[i];
    const double r2 = r * r;
    const double xi = wrapped(pos, i, 0, box), yi = wrapped(pos, i, 1, box), zi = wrapped(pos, i, 2, box);
    int64_t ax, nx, ay, ny, az, nz;
    cell_range(xi, r, g, 0, box.l[0] > 0.0, ax, nx);
    cell_range(yi, r, g, 1, box.l[1] > 0.0, ay, ny);
    cell_range(zi, r, g, 2, box.l[2] > 0.0, az, nz);
    int32_t m = 0;
    int64_t last = -1;
    bool sorted = true;
    uint32_t* row = gbuf + blk(i, 0, cap);
    for (int64_t di = 0; di < nx; di++) {
        const int32_t cx = wrap_cell(ax + di, g.nc[0]);
        for (int64_t dj = 0; dj < ny; dj++) {
            const int32_t cy = wrap_cell(ay + dj, g.nc[1]);
            const uint32_t base = ((uint32_t)cx * (uint32_t)g.nc[1] + (uint32_t)cy) * (uint32_t)g.nc[2];
            for (int64_t dk = 0; dk < nz; dk++) {
                const uint32_t c = base + (uint32_t)wrap_cell(az + dk, g.nc[2]);
                const uint32_t k1 = start[c + 1];
                for (uint32_t k = start[c]; k < k1; k++) {
                    const uint32_t j = order ? order[k] : k;
                    const double dx = min_image(wrapped(pos, j, 0, box) - xi, box.l[0]);
                    const double dy = min_image(wrapped(pos, j, 1, box) - yi, box.l[1]);
                    const double dz = min_image(wrapped(pos, j, 2, box) - zi, box.l[2]);
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < r2) {
                        const double rj = NB_R_KERN * h[j];
                        uint32_t e = j;
                        if (!(d2 < rj * rj)) {
                            e |= kFlag;
                            if (live) atomicAdd(&ecount[j], 1ull);
                        }
                        if ((int64_t)j < last) sorted = false;
                        last = j;
                        if (live && m < cap) row[(int64_t)m * 64] = e;
                        m++;
                    }
                }
            }
        }
    }
    if (live) {
        rowcnt[i] = m;
        unsorted[i] = sorted ? 0 : 1;
    }
    int32_t pv[2] = {(int32_t)m, live && !sorted ? 1 : 0};
    block_partials<2>(pv, 1u, part);
}

// Insertion sort of the rows marked unsorted (flag bit ignored in the key).
__global__ void k_sort_rows(uint32_t* __restrict__ gbuf, int64_t cap, const int32_t* __restrict__ rowcnt,
                            const uint8_t* __restrict__ unsorted, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || !unsorted[i]) return;
    uint32_t* row = gbuf + blk(i, 0, cap);
    const int32_t m = rowcnt[i];
    for (int32_t a = 1; a < m; a++) {
        const uint32_t v = row[(int64_t)a * 64];
        const uint32_t key = v & kMask;
        int32_t b = a - 1;
        while (b >= 0 && (row[(int64_t)b * 64] & kMask) > key) {
            row[(int64_t)(b + 1) * 64] = row[(int64_t)b * 64];
            b--;
        }
        row[(int64_t)(b + 1) * 64] = v;
    }
}

__global__ void k_extras_fill(const uint32_t* __restrict__ gbuf, int64_t cap, const int32_t* __restrict__ rowcnt,
                              const unsigned long long* __restrict__ eoff, unsigned int* __restrict__ efill,
                              uint32_t* __restrict__ ebuf, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* row = gbuf + blk(i, 0, cap);
    const int32_t m = rowcnt[i];
    for (int32_t a = 0; a < m; a++) {
        const uint32_t v = row[(int64_t)a * 64];
        if (v & kFlag) {
            const uint32_t j = v & kMask;
            const unsigned int s = atomicAdd(&efill[j], 1u);
            ebuf[eoff[j] + s] = (uint32_t)i;
        }
    }
}

// final row lengths (gather + extras), their largest value (workgroup partials) and, for the CSR scan,
// tot[n] = 0.
__global__ void __launch_bounds__(kBlock)
k_row_total(const int32_t* __restrict__ rowcnt, const unsigned long long* __restrict__ ecount,
            unsigned long long* __restrict__ tot, long long* __restrict__ part, int64_t n,
            int32_t* __restrict__ row_counts)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    long long t = 0;
    if (i < n) {
        t = (long long)rowcnt[i] + (long long)ecount[i];
        tot[i] = (unsigned long long)t;
        if (row_counts) {  // (gather row, extras) of the particle, the next pass's capacity hint
            row_counts[2 * i] = rowcnt[i];
            row_counts[2 * i + 1] = (int32_t)ecount[i];
        }
    }
    if (i == n) tot[i] = 0;
    block_partial<true>(t, part);
}

// Sorted extras of row i (sorted here, in place), merged with the sorted gather row. The sets are
// disjoint by construction (an extra j of row i has d2 >= (2 h_i)^2). kBlocked: output rows in
// the blocked layout (out_len, capacity cap_out); else CSR at off[i].
template <bool kBlocked>
__global__ void k_merge(const uint32_t* __restrict__ gbuf, int64_t cap, const int32_t* __restrict__ rowcnt,
                        const unsigned long long* __restrict__ eoff, uint32_t* __restrict__ ebuf,
                        const unsigned long long* __restrict__ off, int32_t* __restrict__ out_idx,
                        int32_t* __restrict__ out_len, int64_t cap_out, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* e = ebuf + eoff[i];
    const int64_t ne = (int64_t)(eoff[i + 1] - eoff[i]);
    for (int64_t a = 1; a < ne; a++) {
        const uint32_t v = e[a];
        int64_t b = a - 1;
        while (b >= 0 && e[b] > v) {
            e[b + 1] = e[b];
            b--;
        }
        e[b + 1] = v;
    }
    const uint32_t* g = gbuf + blk(i, 0, cap);
    const int32_t ng = rowcnt[i];
    if (ng + ne > cap_out) return;  // the caller asked with too small a capacity
    int32_t* out = kBlocked ? out_idx + blk(i, 0, cap_out) : out_idx + (int64_t)off[i];
    const int64_t stride = kBlocked ? 64 : 1;
    int32_t p = 0;
    int64_t q = 0, o = 0;
    while (p < ng || q < ne) {
        uint32_t v;
        if (q >= ne || (p < ng && (g[(int64_t)p * 64] & kMask) < e[q])) v = g[(int64_t)(p++) * 64] & kMask;
        else v = e[q++];
        out[o * stride] = (int32_t)v;
        o++;
    }
    if (kBlocked) out_len[i] = (int32_t)o;
}

// ----------------------------------------------------------------- tighten
// neighbours.py:284-303: keep j with d2 < (2 max(h_i, h_j))^2 (positions as given, minimum image);
// kIncl keeps d2 <= (2 max(h_i, h_j))^2, the pair set of the reference's cKDTree twin
// (neighbours.py:53-71: query_ball_point includes the boundary). CSR in -> CSR out (count pass,
// then fill pass).
template <bool kFill, bool kIncl>
__global__ void k_tighten_csr(const double* __restrict__ pos, const double* __restrict__ h, Box box,
                              const int64_t* __restrict__ offset, const int32_t* __restrict__ index,
                              const int64_t* __restrict__ rows, int64_t nrows, unsigned long long* __restrict__ counts,
                              const unsigned long long* __restrict__ out_off, int32_t* __restrict__ out_idx)
{
    int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t i = row_of(rows, q);
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const double hi = h[i];
    int64_t m = 0;
    const int64_t o = kFill ? (int64_t)out_off[i] : 0;
    for (int64_t k = offset[i]; k < offset[i + 1]; k++) {
        const int32_t j = index[k];
        const double dx = min_image(pos[3 * j] - xi, box.l[0]);
        const double dy = min_image(pos[3 * j + 1] - yi, box.l[1]);
        const double dz = min_image(pos[3 * j + 2] - zi, box.l[2]);
        double d2 = 0.0;
        d2 += dx * dx;
        d2 += dy * dy;
        d2 += dz * dz;
        const double rr = NB_R_KERN * pymax(hi, h[j]);
        if (kIncl ? d2 <= rr * rr : d2 < rr * rr) {
            if (kFill) out_idx[o + m] = j;
            m++;
        }
    }
    if (!kFill) counts[i] = (unsigned long long)m;
}

// Blocked in -> blocked out, one pass (the tightened row never outgrows the search row).
// total (device) accumulates the entries kept.
__global__ void __launch_bounds__(kBlock)
k_tighten_blk(const double* __restrict__ pos, const double* __restrict__ h, Box box, const int32_t* __restrict__ in_len,
              const int32_t* __restrict__ in_idx, int64_t cap, int64_t n, int32_t* __restrict__ out_len,
              int32_t* __restrict__ out_idx, long long* __restrict__ part)
{
    int64_t i = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    int32_t m = 0;
    if (i < n) {
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        const double hi = h[i];
        const int32_t* src = in_idx + blk(i, 0, cap);
        int32_t* dst = out_idx + blk(i, 0, cap);
        const int32_t len = in_len[i];
        for (int32_t k = 0; k < len; k++) {
            const int32_t j = src[(int64_t)k * 64];
            const double dx = min_image(pos[3 * j] - xi, box.l[0]);
            const double dy = min_image(pos[3 * j + 1] - yi, box.l[1]);
            const double dz = min_image(pos[3 * j + 2] - zi, box.l[2]);
            double d2 = 0.0;
            d2 += dx * dx;
            d2 += dy * dy;
            d2 += dz * dz;
            const double rr = NB_R_KERN * pymax(hi, h[j]);
            if (d2 < rr * rr) {
                dst[(int64_t)m * 64] = j;
                m++;
            }
        }
        out_len[i] = m;
    }
    block_partial<false>(m, part);
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// dst row i = src row order[i] (rows of w doubles)
__global__ void k_permute(const double* __restrict__ src, const int32_t* __restrict__ order, int64_t n, int w,
                          double* __restrict__ dst)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * w) return;
    int64_t i = t / w;
    int c = (int)(t - i * w);
    dst[t] = src[(int64_t)order[i] * w + c];
}

__global__ void k_permute_i64(const int64_t* __restrict__ src, const int32_t* __restrict__ order, int64_t n,
                              int64_t* __restrict__ dst)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[order[i]];
}

// several arrays of 8-byte values at once (bit copies: doubles or int64)
static constexpr int kMaxPermute = 16;
struct PermSet {
    const uint64_t* src[kMaxPermute];
    const uint64_t* src2[kMaxPermute];  // NULL: a copy; else a kick, dst = src + s src2 (doubles)
    uint64_t* dst[kMaxPermute];
    int32_t width[kMaxPermute];
    int32_t count;
    int32_t total;       // sum of widths
    const double* s_dev;  // the kicks' s = scale * s_dev[0] (a device dt)
    double scale;
};

// Work-item t writes element t of the concatenated destinations (array after array), so each
// wavefront stores one contiguous run of one array; the reads follow the order, which is close to
// the identity from one step to the next.
__global__ void k_permute_many(PermSet P, const int32_t* __restrict__ order, int64_t n)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * P.total) return;
    int k = 0;
    while (t >= n * P.width[k]) {
        t -= n * P.width[k];
        k++;
    }
    const int w = P.width[k];
    const int64_t i = w == 1 ? t : t / w;
    const int c = (int)(t - i * w);
    const int64_t from = (int64_t)order[i] * w + c;
    if (!P.src2[k]) {
        P.dst[k][t] = P.src[k][from];
    } else {  // the kick k_axpy_dev forms (y + s a, s = scale s_dev[0]) of the moved element
        const double s = P.scale * P.s_dev[0];
        const double y = __longlong_as_double((long long)P.src[k][from]);
        const double a = __longlong_as_double((long long)P.src2[k][from]);
        P.dst[k][t] = (uint64_t)__double_as_longlong(y + s * a);
    }
}

// ----------------------------------------------------------------- fused step passes
// The coupled step needs, per density pass, the search lists only for the density solve and,
// afterwards, only the tightened lists. So the gather also keeps each accepted pair's d2 (and
// the neighbour's mass when masses differ) next to the index, and the same work-item solves its
// row's h-rho right after gathering it: the solve sums over the gathered row filtered by
// d2 < (2 min(h_search, h_max))^2, which is exactly the reference's distance pass over the
// final (sorted) row, because extras always have d2 >= (2 h_search_i)^2 and a gather row in
// grid order is already sorted. Rows that came out unsorted (or overflowed) are solved after
// the sort pass. The tightened rows are then emitted straight from the gather rows and the
// extras (which carry their d2 too), with the final smoothing lengths.
struct DensArgs {
    const double* mass;
    double m_uniform;  // > 0: every mass equals it
    const double* h_start;
    double h_max;
    int kid;
    double cnorm, eta, tol;
    double* h_out;
    double* rho;
    double* omega;
    int32_t* iters;
    int32_t* flags;
    uint8_t* pending;  // 1: row not solved yet
    long long* part;   // per-workgroup partials: longest row, unsorted rows, rows left pending
};

// sph.py:215-228 on the cached d2 of a blocked row: FP64 (exact order) or FP32 with
// compensated (Kahan) accumulation (mixed precision).
template <bool kMixed>
__device__ __forceinline__ void row_sums(const double* __restrict__ d2row, const double* __restrict__ mrow,
                                         double m_uniform, int32_t m, double r_max2, double h, int kid, double cnorm,
                                         double& rs_out, double& dws_out)
{
    if (!kMixed) {
        double rs = 0.0, dws = 0.0;
        const double inv_h = 1.0 / h;
        const double ih3 = ipow3(inv_h);
        const double ncih4 = -cnorm * ipow4(inv_h);
        for (int32_t k = 0; k < m; k++) {
            const double d2 = d2row[(int64_t)k * 64];
            if (d2 < r_max2) {
                const double q = sqrt(d2) * inv_h;
                if (q < 2.0) {
                    const double mk = mrow ? mrow[(int64_t)k * 64] : m_uniform;
                    const double fq = kernel_f(q, kid);
                    rs += mk * cnorm * ih3 * fq;
                    dws += mk * (ncih4 * (3.0 * fq + q * kernel_df(q, kid)));
                }
            }
        }
        rs_out = rs;
        dws_out = dws;
    } else {
        // FP32 terms relative to the row's own scale: sum of m_k f(q_k) and m_k (3 f + q f')
        float s1 = 0.f, c1 = 0.f, s2 = 0.f, c2 = 0.f;
        const float inv_hf = (float)(1.0 / h);
        const float rmax2f = (float)r_max2;
        const float mu = (float)(m_uniform > 0.0 ? m_uniform : 1.0);
        for (int32_t k = 0; k < m; k++) {
            const float d2 = (float)d2row[(int64_t)k * 64];
            if (d2 < rmax2f) {
                const float q = __fsqrt_rn(d2) * inv_hf;
                if (q < 2.f) {
                    const float mk = mrow ? (float)mrow[(int64_t)k * 64] : mu;
                    float fq, dfq;
                    if (kid == 0) {
                        if (q < 1.f) {
                            fq = 1.f - 1.5f * q * q + 0.75f * q * q * q;
                            dfq = -3.f * q + 2.25f * q * q;
                        } else {
                            const float t = 2.f - q;
                            fq = 0.25f * t * t * t;
                            dfq = -0.75f * t * t;
                        }
                    } else {
                        const float x = 1.f - 0.5f * q, x2 = x * x;
                        fq = x2 * x2 * x2 * (35.f / 12.f * q * q + 3.f * q + 1.f);
                        dfq = -7.f / 3.f * q * x2 * x2 * x * (5.f * q + 2.f);
                    }
                    const float t1 = mk * fq - c1;
                    const float u1 = s1 + t1;
                    c1 = (u1 - s1) - t1;
                    s1 = u1;
                    const float t2 = mk * (3.f * fq + q * dfq) - c2;
                    const float u2 = s2 + t2;
                    c2 = (u2 - s2) - t2;
                    s2 = u2;
                }
            }
        }
        const double inv_h = 1.0 / h;
        const double scale_m = mrow ? 1.0 : (m_uniform > 0.0 ? m_uniform / (double)mu : 1.0);
        rs_out = (double)s1 * scale_m * cnorm * ipow3(inv_h);
        dws_out = -(double)s2 * scale_m * cnorm * ipow4(inv_h);
    }
}

// sph.py:251-297 for one row whose gathered d2 (and masses) sit in blocked buffers.
template <bool kMixed>
__device__ void solve_row(int64_t i, int32_t m, const double* __restrict__ dbuf, const double* __restrict__ mbuf,
                          int64_t base, double h_search_i, const DensArgs& A)
{
    const double m_i = A.mass[i];
    const double r_max = 2.0 * pymin(h_search_i, A.h_max);
    const double r_max2 = r_max * r_max;
    const double* d2row = dbuf + base;
    const double* mrow = mbuf ? mbuf + base : nullptr;
    double hi = pymin(h_search_i, A.h_max);
    double lo = 0.0;
    double h = pymin(A.h_start[i], hi);
    const double h0 = h;
    int flag = 3;
    int it = 0;
    const double eta3 = ipow3(A.eta);
    double rs, dws;
    while (it < 50) {
        it += 1;
        row_sums<kMixed>(d2row, mrow, A.m_uniform, m, r_max2, h, A.kid, A.cnorm, rs, dws);
        const double f = rs - m_i * ipow3(A.eta / h);
        const double df = dws + 3.0 * m_i * eta3 / ipow4(h);
        if (f > 0) hi = h;
        else lo = h;
        double h_new = df > 0 ? h - f / df : 0.5 * (lo + hi);
        if (!(lo <= h_new && h_new <= hi)) h_new = 0.5 * (lo + hi);
        if (fabs(h_new - h) < A.tol * h0) {
            h = h_new;
            flag = 0;
            break;
        }
        h = h_new;
    }
    const double h_lim = pymin(h_search_i, A.h_max);
    if (h >= h_lim * (1 - 1e-3)) {
        double rs_lim, dum;
        row_sums<kMixed>(d2row, mrow, A.m_uniform, m, r_max2, h_lim, A.kid, A.cnorm, rs_lim, dum);
        if (rs_lim - m_i * ipow3(A.eta / h_lim) < 0) {
            h = h_lim;
            flag = h_lim >= A.h_max ? 1 : 2;
        }
    }
    row_sums<kMixed>(d2row, mrow, A.m_uniform, m, r_max2, h, A.kid, A.cnorm, rs, dws);
    A.h_out[i] = h;
    A.rho[i] = rs;
    A.omega[i] = flag == 1 ? 1.0 : 1.0 + h / (3.0 * rs) * dws;
    A.iters[i] = it;
    A.flags[i] = flag;
}

// Capacities of 64-row blocks with capacities of their own: bsize[b] = 64 x the largest row need
// of block b (row i of the block at boff[b] + i % 64, entry k 64 further per k; boff = exclusive
// scan of bsize). Row need: mode 0, cap0; mode 1, a + a / 4 + 16 rounded up to a multiple of 4
// (a = cnt[i * cstride], the previous pass's count of the same particle: headroom for the drift
// between steps); mode 2, cnt[i * cstride] (+ add[i] when given), exact. bsize[nblk] = 0.
__global__ void k_block_caps(const int32_t* __restrict__ cnt, int cstride, const unsigned long long* __restrict__ add,
                             int64_t n, int mode, int64_t cap0, unsigned long long* __restrict__ bsize, int64_t nblk)
{
    const int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);  // one wavefront per block
    if (blk > nblk) return;
    const int64_t i = blk * 64 + (threadIdx.x & 63);
    long long need = 0;
    if (blk < nblk && i < n) {
        if (mode == 0) {
            need = cap0;
        } else {
            const long long a = cnt[i * cstride];
            need = mode == 1 ? ((a + a / 4 + 16 + 3) / 4) * 4 : a + (add ? (long long)add[i] : 0);
        }
        need = need < 1 ? 1 : need;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const long long t = __shfl_xor(need, o, 64);
        need = t > need ? t : need;
    }
    if ((threadIdx.x & 63) == 0) bsize[blk] = (unsigned long long)need * 64ull;
}

// Gather (as k_gather, open boxes) keeping d2 (and masses) in 64-row blocks of capacities of
// their own (boff), then the row's density solve. Rows longer than their block's capacity are
// counted in full (rowcnt) and stored up to the capacity; the caller sizes the blocks from those
// counts and gathers again.
template <bool kMixed>
__global__ void __launch_bounds__(kBlock)
k_gather_solve(const double* __restrict__ pos, const double* __restrict__ h, int64_t n, Grid g,
               const uint32_t* __restrict__ start, const uint32_t* __restrict__ order, uint32_t* __restrict__ gbuf,
               double* __restrict__ dbuf, double* __restrict__ mbuf, const unsigned long long* __restrict__ boff,
               int32_t* __restrict__ rowcnt, uint8_t* __restrict__ unsorted, unsigned long long* __restrict__ ecount,
               DensArgs A)
{
    // lanes past the end redo the last particle without writing (full blocks for the reductions)
    int64_t i = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    const bool live = i < n;
    if (!live) i = n - 1;
    const double hs = h[i];
    const double r = NB_R_KERN * hs;
    const double r2 = r * r;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    int64_t ax, nx, ay, ny, az, nz;
    cell_range(xi, r, g, 0, false, ax, nx);
    cell_range(yi, r, g, 1, false, ay, ny);
    cell_range(zi, r, g, 2, false, az, nz);
    int32_t m = 0;
    int64_t last = -1;
    bool sorted = true;
    const int64_t base = (int64_t)boff[i >> 6] + (i & 63);
    const int32_t cap = (int32_t)((boff[(i >> 6) + 1] - boff[i >> 6]) >> 6);
    for (int64_t di = 0; di < nx; di++) {
        const uint32_t cx = (uint32_t)(ax + di);
        for (int64_t dj = 0; dj < ny; dj++) {
            const uint32_t cy = (uint32_t)(ay + dj);
            const uint32_t c0 = (cx * (uint32_t)g.nc[1] + cy) * (uint32_t)g.nc[2] + (uint32_t)az;
            // the z-run of cells is one contiguous range of the cell-sorted particles
            const uint32_t k0 = start[c0], k1 = start[c0 + (uint32_t)nz];
            for (uint32_t k = k0; k < k1; k++) {
                const uint32_t j = order ? order[k] : k;
                const double dx = pos[3 * j] - xi;
                const double dy = pos[3 * j + 1] - yi;
                const double dz = pos[3 * j + 2] - zi;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < r2) {
                    const double rj = NB_R_KERN * h[j];
                    uint32_t e = j;
                    if (!(d2 < rj * rj)) {
                        e |= kFlag;
                        if (live) atomicAdd(&ecount[j], 1ull);
                    }
                    if ((int64_t)j < last) sorted = false;
                    last = j;
                    if (live && m < cap) {
                        gbuf[base + (int64_t)m * 64] = e;
                        dbuf[base + (int64_t)m * 64] = d2;
                        if (mbuf) mbuf[base + (int64_t)m * 64] = A.mass[j];
                    }
                    m++;
                }
            }
        }
    }
    const bool done_here = sorted && m <= cap;
    int32_t pv[4] = {(int32_t)m, live && !sorted ? 1 : 0, live && !done_here ? 1 : 0, live && m > cap ? 1 : 0};
    block_partials<4>(pv, 1u, A.part);
    if (!live) return;
    rowcnt[i] = m;
    unsorted[i] = sorted ? 0 : 1;
    A.pending[i] = done_here ? 0 : 1;
    if (done_here) solve_row<kMixed>(i, m, dbuf, mbuf, base, hs, A);
}

// Insertion sort of unsorted rows (blocks at boff), carrying d2 and mass along.
__global__ void k_sort_rows_payload(uint32_t* __restrict__ gbuf, double* __restrict__ dbuf, double* __restrict__ mbuf,
                                    const unsigned long long* __restrict__ boff, const int32_t* __restrict__ rowcnt,
                                    const uint8_t* __restrict__ unsorted, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || !unsorted[i]) return;
    const int64_t base = (int64_t)boff[i >> 6] + (i & 63);
    const int32_t m = rowcnt[i];
    for (int32_t a = 1; a < m; a++) {
        const uint32_t v = gbuf[base + (int64_t)a * 64];
        const double dv = dbuf[base + (int64_t)a * 64];
        const double mv = mbuf ? mbuf[base + (int64_t)a * 64] : 0.0;
        const uint32_t key = v & kMask;
        int32_t b = a - 1;
        while (b >= 0 && (gbuf[base + (int64_t)b * 64] & kMask) > key) {
            gbuf[base + (int64_t)(b + 1) * 64] = gbuf[base + (int64_t)b * 64];
            dbuf[base + (int64_t)(b + 1) * 64] = dbuf[base + (int64_t)b * 64];
            if (mbuf) mbuf[base + (int64_t)(b + 1) * 64] = mbuf[base + (int64_t)b * 64];
            b--;
        }
        gbuf[base + (int64_t)(b + 1) * 64] = v;
        dbuf[base + (int64_t)(b + 1) * 64] = dv;
        if (mbuf) mbuf[base + (int64_t)(b + 1) * 64] = mv;
    }
}

template <bool kMixed>
__global__ void k_solve_pending(const double* __restrict__ h, int64_t n, const double* __restrict__ dbuf,
                                const double* __restrict__ mbuf, const unsigned long long* __restrict__ boff,
                                const int32_t* __restrict__ rowcnt, DensArgs A)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || !A.pending[i]) return;
    solve_row<kMixed>(i, rowcnt[i], dbuf, mbuf, (int64_t)boff[i >> 6] + (i & 63), h[i], A);
}

// extras with their d2
__global__ void k_extras_fill_d2(const uint32_t* __restrict__ gbuf, const double* __restrict__ dbuf,
                                 const unsigned long long* __restrict__ boff, const int32_t* __restrict__ rowcnt,
                                 const unsigned long long* __restrict__ eoff, unsigned int* __restrict__ efill,
                                 uint32_t* __restrict__ ebuf, double* __restrict__ edbuf, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t base = (int64_t)boff[i >> 6] + (i & 63);
    const int32_t m = rowcnt[i];
    auto put = [&](uint32_t v, int32_t a) {
        if (v & kFlag) {
            const uint32_t j = v & kMask;
            const unsigned int s = atomicAdd(&efill[j], 1u);
            ebuf[eoff[j] + s] = (uint32_t)i;
            edbuf[eoff[j] + s] = dbuf[base + (int64_t)a * 64];
        }
    };
    // four entries' flags loaded before any is handled (same entries, same order per row)
    int32_t a = 0;
    for (; a + 4 <= m; a += 4) {
        uint32_t v[4];
#pragma unroll
        for (int t = 0; t < 4; t++) v[t] = gbuf[base + (int64_t)(a + t) * 64];
#pragma unroll
        for (int t = 0; t < 4; t++) put(v[t], a + t);
    }
    for (; a < m; a++) put(gbuf[base + (int64_t)a * 64], a);
}

// neighbours.py:284-303 on the merged row (gather row + sorted extras, both with d2), with the
// final smoothing lengths hf: keep d2 < (2 max(hf_i, hf_j))^2, into 64-row blocks at toff (each
// block's capacity the longest merged row of its rows, so every kept entry has a slot).
__global__ void __launch_bounds__(kBlock)
k_tight_emit(const uint32_t* __restrict__ gbuf, const double* __restrict__ dbuf,
             const unsigned long long* __restrict__ boff, const int32_t* __restrict__ rowcnt,
             const unsigned long long* __restrict__ eoff, uint32_t* __restrict__ ebuf, double* __restrict__ edbuf,
             const double* __restrict__ hf, int64_t n, const unsigned long long* __restrict__ toff,
             int32_t* __restrict__ out_len, int32_t* __restrict__ out_idx, long long* __restrict__ part)
{
    int64_t i = xcd_tile() * (int64_t)blockDim.x + threadIdx.x;
    int32_t o = 0;
    if (i < n) {
        uint32_t* e = ebuf + eoff[i];
        double* ed = edbuf + eoff[i];
        const int64_t ne = (int64_t)(eoff[i + 1] - eoff[i]);
        for (int64_t a = 1; a < ne; a++) {
            const uint32_t v = e[a];
            const double dv = ed[a];
            int64_t b = a - 1;
            while (b >= 0 && e[b] > v) {
                e[b + 1] = e[b];
                ed[b + 1] = ed[b];
                b--;
            }
            e[b + 1] = v;
            ed[b + 1] = dv;
        }
        const int64_t base = (int64_t)boff[i >> 6] + (i & 63);
        const int64_t obase = (int64_t)toff[i >> 6] + (i & 63);
        const int32_t cap_out = (int32_t)((toff[(i >> 6) + 1] - toff[i >> 6]) >> 6);
        const int32_t ng = rowcnt[i];
        const double hi = hf[i];
        int32_t p = 0;
        int64_t q = 0;
        // the head of each input held in registers; the one consumed is refilled while the
        // neighbour's h is loaded (the same merge order and entries)
        uint32_t gj = 0, ej = 0;
        double gd = 0.0, edv = 0.0;
        if (ng > 0) {
            gj = gbuf[base] & kMask;
            gd = dbuf[base];
        }
        if (ne > 0) {
            ej = e[0];
            edv = ed[0];
        }
        while (p < ng || q < ne) {
            const bool take_g = q >= ne || (p < ng && gj < ej);
            const uint32_t j = take_g ? gj : ej;
            const double d2 = take_g ? gd : edv;
            const double hj = hf[j];
            if (take_g) {
                p++;
                if (p < ng) {
                    gj = gbuf[base + (int64_t)p * 64] & kMask;
                    gd = dbuf[base + (int64_t)p * 64];
                }
            } else {
                q++;
                if (q < ne) {
                    ej = e[q];
                    edv = ed[q];
                }
            }
            const double rr = NB_R_KERN * pymax(hi, hj);
            if (d2 < rr * rr && o < cap_out) {
                out_idx[obase + (int64_t)o * 64] = (int32_t)j;
                o++;
            }
        }
        out_len[i] = o;
    }
    block_partial<false>(o, part);
}

// ----------------------------------------------------------------- mixed-precision step passes
// Lane groups: kG consecutive lanes serve one particle. The search sphere is cut into grid
// columns (x-y cells); in each column only the z-cells the sphere reaches at that column's x-y
// distance are visited, and that z-run is one contiguous range of the cell-sorted particles.
// The group lays the ranges of a chunk of kCols columns end to end and walks the concatenation
// kG candidates at a time (lane l takes candidate l of each batch; a binary search over the
// chunk's prefix sums, held in the lanes' registers, finds its column), so every lane tests a
// candidate on every pass whatever the column lengths. Candidate positions and search lengths
// come in one 32-byte record. The FP64 list predicate keeps the lists exact; accepted entries
// are compacted with a ballot (lane l's slot is the number of accepting lanes below it) into
// rows of their own capacity (row i at goff[i] .. goff[i + 1]) of neighbour indices.
// The first evaluation of the h-rho Newton solve (sph.py:251-297, at the starting h) is summed
// on the fly, in FP32 with compensated accumulation per lane, combined in FP64 by the group,
// and the group takes the Newton step in FP64. Rows converged after that step (the common case)
// get rho and Omega at their final h from the emission pass, which reads the row anyway; the
// others are solved in full by k_solve_slow_mixed.
static constexpr int kG = 16;
static constexpr int kCols = 2 * kG;  // columns per chunk: two per lane
static constexpr int kTab = 384;      // candidate slots per lane group in LDS (multiple of kG)
static constexpr int kQueue = 32;     // extras queued per row in LDS (more go out at once)

// Status words (int32, device) of a step recorded for replay: what did not fit, or what the
// caller must redo eagerly. Words 6 and 7 are written by nb_glf_step (failures) and
// nb_eos_ideal (u <= 0) when given status + 6 / status + 7.
enum : int { kStGather = 0, kStExtras = 1, kStIndex = 2, kStCells = 3, kStGrow = 4, kStUnsorted = 5 };

// Orders the LDS accesses of one wavefront (lanes read what other lanes of the wave wrote).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// f(q) and f'(q) of kernels_nb.py:12-43 in FP32, without branches: the cubic spline as
// (2 - q)_+^3 / 4 - (1 - q)_+^3 (the same polynomial on each piece), so that the lanes of a
// wavefront never split on q; both forms are exactly zero from q = 2 on.
__device__ __forceinline__ void kernel_fdf32(float q, int kid, float& fq, float& dfq)
{
    if (kid == 0) {
        const float t2 = fmaxf(2.f - q, 0.f), t1 = fmaxf(1.f - q, 0.f);
        const float s2 = t2 * t2, s1 = t1 * t1;
        fq = __builtin_fmaf(0.25f * s2, t2, -(s1 * t1));
        dfq = __builtin_fmaf(3.f, s1, -0.75f * s2);
    } else {
        const float x = fmaxf(1.f - 0.5f * q, 0.f), x2 = x * x;
        fq = x2 * x2 * x2 * (35.f / 12.f * q * q + 3.f * q + 1.f);
        dfq = -7.f / 3.f * q * x2 * x2 * x * (5.f * q + 2.f);
    }
}

template <int W = kG>
__device__ __forceinline__ double group_sum(double v)
{
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// compensated FP32 sums of m f(q) and m (3 f + q f') (the two density-solve sums)
struct DensSums {
    float a1 = 0.f, c1 = 0.f, a2 = 0.f, c2 = 0.f;
    __device__ __forceinline__ void add(float q, float mk, int kid)
    {
        float fq, dfq;
        kernel_fdf32(q, kid, fq, dfq);  // zero terms past the support
        const float t1 = __builtin_fmaf(mk, fq, -c1);  // the products fused with the compensation
        const float u1 = a1 + t1;
        c1 = (u1 - a1) - t1;
        a1 = u1;
        const float t2 = __builtin_fmaf(mk, __builtin_fmaf(q, dfq, 3.f * fq), -c2);
        const float u2 = a2 + t2;
        c2 = (u2 - a2) - t2;
        a2 = u2;
    }
    // group totals (FP64) over W lanes, identical on every lane of the group
    template <int W = kG>
    __device__ __forceinline__ void total(double& s1, double& s2) const
    {
        s1 = group_sum<W>((double)a1 - (double)c1);
        s2 = group_sum<W>((double)a2 - (double)c2);
    }
};

struct MixArgs {
    const double* mass;
    double m_uniform;
    const double* h_start;
    double h_max;
    int kid;
    double cnorm, eta, tol;
    double* h_out;
    double* rho;
    double* omega;
    int32_t* iters;
    int32_t* flags;
    uint8_t* slow;    // 1: row solved by k_solve_slow_mixed (not converged in one Newton step)
    long long* part;  // per-workgroup partials: longest row, slow rows
};

// Row capacities of the mixed lists (gather row gc, extras segment ec): uniform (mode 0), from
// the previous pass's counts of the same particles with headroom (mode 1; hint[2 i] gather row,
// hint[2 i + 1] extras), or from this pass's exact counts after an overflow (mode 2). Entry n is
// 0, so the exclusive scans end in the totals. Rows are kept 16-byte aligned. The headroom of
// mode 1 (a quarter plus 32 entries; extras doubled plus 16) is sized so that a single row of a
// million outgrowing it, which repeats the whole pass, stays rare.
__global__ void k_caps_mixed(const int32_t* __restrict__ hint, const int32_t* __restrict__ rowcnt,
                             const unsigned int* __restrict__ efill, int64_t n, int mode, int64_t cap_u,
                             int64_t ecap_u, unsigned long long* __restrict__ gc, unsigned long long* __restrict__ ec)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > n) return;
    unsigned long long g = 0, e = 0;
    if (i < n) {
        if (mode == 0) {
            g = (unsigned long long)cap_u;
            e = (unsigned long long)ecap_u;
        } else if (mode == 1) {
            const unsigned long long a = hint[2 * i] > 0 ? (unsigned long long)hint[2 * i] : 0ull;
            const unsigned long long b = hint[2 * i + 1] > 0 ? (unsigned long long)hint[2 * i + 1] : 0ull;
            g = a + a / 4 + 32;
            e = 2 * b + 16;
        } else {
            const unsigned long long a = (unsigned long long)rowcnt[i], b = efill[i];
            g = a + a / 16 + 4;
            e = b + b / 8 + 2;
        }
        g = (g + 3) & ~3ull;
        e = (e + 3) & ~3ull;
    }
    gc[i] = g;
    ec[i] = e;
}

// (x, y, z, h_search) records of the candidates
__global__ void k_pack_posh(const double* __restrict__ pos, const double* __restrict__ h, int64_t n,
                            double4* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], h[i]);
}

// The z cells [z0, z1] that a search sphere reaches in grid column (cx, cy), in FP32 cell units
// (centre ux, uy, uz; radius rc in x-y cells; zr z cells per x-y cell); false when it misses the
// column. The caller widens the radius to rc (1 + 1e-5) + 1e-3 and the x-y distances here are
// shortened by 1e-3 cells, far beyond the FP32 rounding of these values (below 1e-3 cells for
// grids of up to 2048 cells a side, relative 1e-7 in rc^2), so the range only grows: extra
// candidates fail the exact FP64 test.
__device__ __forceinline__ bool column_zrange32(float ux, float uy, float uz, float rc, float zr, int32_t cx, int32_t cy,
                                                int32_t nz, int32_t& z0, int32_t& z1)
{
    const float ex = fmaxf(fmaxf((float)cx - ux, ux - (float)(cx + 1)) - 1e-3f, 0.f);
    const float ey = fmaxf(fmaxf((float)cy - uy, uy - (float)(cy + 1)) - 1e-3f, 0.f);
    const float rest = rc * rc - (ex * ex + ey * ey);
    if (!(rest > 0.f)) return false;
    const float hz = (sqrt32(rest) + 2e-3f) * zr + 4e-3f;
    const float fa = fmaxf(floorf(uz - hz), 0.f);
    const float fb = fminf(floorf(uz + hz), (float)(nz - 1));
    if (fb < fa) return false;
    z0 = (int32_t)fa;
    z1 = (int32_t)fb;
    return true;
}

// Column c of an nx x ny block of columns as (c / ny, c % ny), for c < 2^22 (the float quotient
// is off by at most one, which the remainder corrects)
__device__ __forceinline__ void column_of(uint32_t c, uint32_t ny, float inv_ny, int32_t& qx, int32_t& qy)
{
    qx = (int32_t)((float)c * inv_ny);
    qy = (int32_t)c - qx * (int32_t)ny;
    if (qy < 0) {
        qx--;
        qy += (int32_t)ny;
    } else if (qy >= (int32_t)ny) {
        qx++;
        qy -= (int32_t)ny;
    }
}

// f(q) and f'(q) of one kernel (kKid 0 cubic spline, 1 Wendland C4), as kernel_fdf32
template <int kKid>
__device__ __forceinline__ void kernel_fdf32_k(float q, float& fq, float& dfq)
{
    if (kKid == 0) {
        const float t2 = fmaxf(2.f - q, 0.f), t1 = fmaxf(1.f - q, 0.f);
        const float s2 = t2 * t2, s1 = t1 * t1;
        fq = __builtin_fmaf(0.25f * s2, t2, -(s1 * t1));
        dfq = __builtin_fmaf(3.f, s1, -0.75f * s2);
    } else {
        const float x = fmaxf(1.f - 0.5f * q, 0.f), x2 = x * x;
        fq = x2 * x2 * x2 * (35.f / 12.f * q * q + 3.f * q + 1.f);
        dfq = -7.f / 3.f * q * x2 * x2 * x * (5.f * q + 2.f);
    }
}

// Extras (pairs seen only from the other side, d2 >= (2 h_search_j)^2) go straight to row j's
// extras segment (eoff[j] .. eoff[j + 1], filled with an atomic counter efill[j]). Rows and
// segments that outgrow their capacity keep counting (the pass is repeated with capacities from
// the exact counts). kSeg: an ensemble of
// independent systems; particle i searches only the cells of its own system cloud[i], whose grid
// and first cell come from segs. Rows hold neighbour indices only (4 bytes per entry): the passes
// that read them form d2 again from the positions, in the same FP64 operations, so they see the
// same d2 bits. kMasses (masses differ): row i stores its particle's FP32 mass in rmass[i], which
// those passes read by index. The first-step sums take every candidate's FP32 term (zero off the
// row) into FP64 per lane, branch-free. kSums false (the FP64 pass, whose solve sums every row
// itself): the lists and extras only, no first step, masses or solve outputs.
template <bool kMasses, bool kSeg, int kKid, bool kSums = true>
__global__ void __launch_bounds__(256, 5)
k_gather_mixed(const double4* __restrict__ posh, int64_t n, Grid g_one, const GridSeg* __restrict__ segs,
               const int32_t* __restrict__ cloud, const uint32_t* __restrict__ start,
               const uint32_t* __restrict__ order, uint32_t* __restrict__ ridx, float* __restrict__ rmass,
               const unsigned long long* __restrict__ goff, int32_t* __restrict__ rowcnt,
               unsigned int* __restrict__ efill, uint32_t* __restrict__ ebuf,
               const unsigned long long* __restrict__ eoff, MixArgs A)
{
    const int glane = threadIdx.x & (kG - 1);
    const int gshift = (threadIdx.x & 63) & ~(kG - 1);
    const int64_t i = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / kG;
    const bool live = i < n;
    const int64_t ii = live ? i : n - 1;
    const Grid g = kSeg ? segs[cloud[ii]].g : g_one;
    const uint32_t cbase = kSeg ? segs[cloud[ii]].base : 0u;
    const double4 pi = posh[ii];
    const double xi = pi.x, yi = pi.y, zi = pi.z, hs = pi.w;
    const double r = NB_R_KERN * hs;
    const double r2 = r * r;
    int64_t ax, nx, ay, ny;
    cell_range(xi, r, g, 0, false, ax, nx);
    cell_range(yi, r, g, 1, false, ay, ny);
    const uint32_t ncol = live ? (uint32_t)(nx * ny) : 0u;
    const uint32_t ny32 = (uint32_t)ny;
    const float inv_ny = 1.f / (float)(ny > 0 ? ny : 1);
    const float ux = (float)((xi - g.lo[0]) * g.inv[0]), uy = (float)((yi - g.lo[1]) * g.inv[1]);
    const float uz = (float)((zi - g.lo[2]) * g.inv[2]);
    const float rc = (float)(r * g.inv[0]) * (1.f + 1e-5f) + 1e-3f;
    const float zr = (float)(g.inv[2] / g.inv[0]);
    const uint32_t nc1 = (uint32_t)g.nc[1], nc2 = (uint32_t)g.nc[2];
    const double r_lim = pymin(hs, A.h_max);
    const double h0 = pymin(A.h_start[ii], r_lim);
    const float inv_h0 = (float)(1.0 / h0);
    const float mu = (float)(A.m_uniform > 0.0 ? A.m_uniform : 1.0);
    double s1 = 0.0, s2 = 0.0;  // per-lane first-step sums m f(q) and m (3 f + q f')
    int32_t m = 0;
    unsigned int emax = 0, eover = 0;
    const unsigned long long gb0 = goff[ii];
    const int64_t cap = (int64_t)(goff[ii + 1] - gb0);
    uint32_t* row = ridx + gb0;
    __shared__ uint32_t ctab[256 / kG][kTab];     // one candidate table per lane group
    __shared__ uint32_t cqueue[256 / kG][kQueue];  // its row's extras, handed out after the walk
    uint32_t* tab = ctab[threadIdx.x / kG];
    uint32_t* queue = cqueue[threadIdx.x / kG];
    const unsigned below = (1u << glane) - 1u;
    int qm = 0;
    // every loop below runs the same number of times on all lanes of the wave (ballot bounds),
    // so the shuffles always see their whole group active
    for (uint32_t c0 = 0;; c0 += kCols) {
        if (__ballot(c0 < ncol) == 0) break;
        const int32_t m0 = m;  // the row's entries before this chunk
        uint32_t k0a = 0, lena = 0, k0b = 0, lenb = 0;
        int32_t z0, z1, qx, qy;
        const uint32_t ca = c0 + glane, cb = c0 + kG + glane;
        if (ca < ncol) {
            column_of(ca, ny32, inv_ny, qx, qy);
            const int32_t cx = (int32_t)ax + qx, cy = (int32_t)ay + qy;
            if (column_zrange32(ux, uy, uz, rc, zr, cx, cy, (int32_t)nc2, z0, z1)) {
                const uint32_t base = cbase + ((uint32_t)cx * nc1 + (uint32_t)cy) * nc2;
                k0a = start[base + (uint32_t)z0];
                lena = start[base + (uint32_t)z1 + 1u] - k0a;
            }
        }
        if (cb < ncol) {
            column_of(cb, ny32, inv_ny, qx, qy);
            const int32_t cx = (int32_t)ax + qx, cy = (int32_t)ay + qy;
            if (column_zrange32(ux, uy, uz, rc, zr, cx, cy, (int32_t)nc2, z0, z1)) {
                const uint32_t base = cbase + ((uint32_t)cx * nc1 + (uint32_t)cy) * nc2;
                k0b = start[base + (uint32_t)z0];
                lenb = start[base + (uint32_t)z1 + 1u] - k0b;
            }
        }
        uint32_t ia = lena, ib = lenb;  // inclusive prefix sums over the group
        for (int o = 1; o < kG; o <<= 1) {
            const uint32_t ta = __shfl_up(ia, o, kG), tb = __shfl_up(ib, o, kG);
            if (glane >= o) {
                ia += ta;
                ib += tb;
            }
        }
        const uint32_t tota = __shfl(ia, kG - 1, kG);
        const uint32_t total = tota + __shfl(ib, kG - 1, kG);
        const uint32_t pa = ia - lena, pb = tota + ib - lenb;  // exclusive prefixes of both columns
        // the chunk's first kTab candidates by position: each lane lists its two columns' runs
        wave_sync();  // the previous chunk's reads are done
        for (uint32_t t = 0; t < lena && pa + t < (uint32_t)kTab; t++) tab[pa + t] = k0a + t;
        for (uint32_t t = 0; t < lenb && pb + t < (uint32_t)kTab; t++) tab[pb + t] = k0b + t;
        wave_sync();
        for (uint32_t vb = 0;; vb += kG) {
            if (__ballot(vb < total) == 0) break;
            const uint32_t v = vb + (uint32_t)glane;
            uint32_t kc = 0;
            if (vb < (uint32_t)kTab) {  // the same on every lane of the group
                kc = v < total ? tab[v] : 0u;
            } else {
                const uint32_t vs = v < total ? v : total - 1;
                // the chunk column holding candidate vs: the last one whose prefix is <= vs
                int lo = 0;
                for (int step = kCols / 2; step > 0; step >>= 1) {
                    const int mid = lo + step;
                    const uint32_t qa = __shfl(pa, mid & (kG - 1), kG), qb = __shfl(pb, mid & (kG - 1), kG);
                    if ((mid < kG ? qa : qb) <= vs) lo = mid;
                }
                const uint32_t qa = __shfl(pa, lo & (kG - 1), kG), qb = __shfl(pb, lo & (kG - 1), kG);
                const uint32_t ka = __shfl(k0a, lo & (kG - 1), kG), kb = __shfl(k0b, lo & (kG - 1), kG);
                kc = (lo < kG ? ka : kb) + (vs - (lo < kG ? qa : qb));
            }
            bool pred = false, flag = false;
            uint32_t j = 0;
            double d2 = 0.0;
            double mjd = 0.0;
            if (v < total) {
                const uint32_t k = kc;
                j = order ? order[k] : k;
                const double4 pj = posh[j];
                if (kSums && kMasses) mjd = A.mass[j];
                const double dx = pj.x - xi;
                const double dy = pj.y - yi;
                const double dz = pj.z - zi;
                d2 = dx * dx + dy * dy + dz * dz;
                pred = d2 < r2;
                // evaluated on every lane (no branch on pred), so the record's search length
                // arrives with its position rather than in a load issued after the test
                const double rj = NB_R_KERN * pj.w;
                flag = pred & !(d2 < rj * rj);
            }
            const unsigned long long bal = __ballot(pred);
            const unsigned gb = (unsigned)((bal >> gshift) & ((1ull << kG) - 1));
            if (pred) {
                // the chunk's entries go to the table slots already read (slot - m0 < vb + kG),
                // and out to the row after the walk: no store inside the walk for its loads to
                // wait behind
                const int slot = m + __popc(gb & below);
                if (slot - m0 < kTab) tab[slot - m0] = j;
                else if (slot < cap) row[slot] = j;
            }
            const unsigned long long fbal = __ballot(flag);
            if (fbal) {
                const unsigned fg = (unsigned)((fbal >> gshift) & ((1ull << kG) - 1));
                if (flag) {
                    const int qs = qm + __popc(fg & below);
                    if (qs < kQueue) {
                        queue[qs] = j;
                    } else {
                        const unsigned int es = atomicAdd(&efill[j], 1u);
                        const unsigned long long eb = eoff[j];
                        if (es < eoff[j + 1] - eb) ebuf[eb + es] = (uint32_t)i;
                        else eover = 1;
                        emax = es + 1 > emax ? es + 1 : emax;
                    }
                }
                qm += __popc(fg);
            }
            if (kSums) {
                // the first-step sums on every lane: terms off the row add an exact zero
                const float q = sqrt32((float)d2) * inv_h0;
                float fq, dfq;
                kernel_fdf32_k<kKid>(q, fq, dfq);
                const float mj = kMasses ? (float)mjd : mu;
                s1 += pred ? (double)(mj * fq) : 0.0;
                s2 += pred ? (double)(mj * __builtin_fmaf(q, dfq, 3.f * fq)) : 0.0;
            }
            m += __popc(gb);
        }
        // the chunk's entries, from the table to the row (contiguous per group)
        wave_sync();
        const int32_t mc = m - m0 < kTab ? m - m0 : kTab;
        for (int32_t t = glane; t < mc; t += kG)
            if (m0 + t < cap) row[m0 + t] = tab[t];
        wave_sync();  // read before the next chunk's candidates overwrite the table
    }
    // the queued extras to their segments
    const int qn = qm < kQueue ? qm : kQueue;
    for (int t = glane; t < qn; t += kG) {
        const uint32_t j = queue[t];
        const unsigned int es = atomicAdd(&efill[j], 1u);
        const unsigned long long eb = eoff[j];
        if (es < eoff[j + 1] - eb) ebuf[eb + es] = (uint32_t)i;
        else eover = 1;
        emax = es + 1 > emax ? es + 1 : emax;
    }
    if (live && glane == 0) {
        rowcnt[i] = m;
        if (kSums && kMasses) rmass[i] = (float)A.mass[i];  // the FP32 mass the sums of later passes take
    }
    if (!kSums) {
        // [0] longest gather row, [2] an extras segment overflowed, [3] gather rows over
        // capacity, [4] longest extras segment ([1] no rows left slow: none solved here)
        int32_t pv[5] = {live ? m : 0, 0, (int32_t)eover, live && glane == 0 && m > cap ? 1 : 0, (int32_t)emax};
        block_partials<5>(pv, 0b10101u, A.part);
        return;
    }
    // the loop's first pass at h0 (sph.py:260-290), the Newton update and convergence test in FP64
    const double S1 = group_sum<kG>(s1), S2 = group_sum<kG>(s2);
    const double m_i = A.mass[ii];
    const double mscale = kMasses ? 1.0 : (A.m_uniform > 0.0 ? A.m_uniform / (double)mu : 1.0);
    const double inv_h = 1.0 / h0;
    const double rs = S1 * mscale * A.cnorm * ipow3(inv_h);
    const double dws = -S2 * mscale * A.cnorm * ipow4(inv_h);
    const double f = rs - m_i * ipow3(A.eta / h0);
    const double df = dws + 3.0 * m_i * ipow3(A.eta) / ipow4(h0);
    double lo_b = 0.0, hi_b = r_lim;
    if (f > 0) hi_b = h0;
    else lo_b = h0;
    double h_new = df > 0 ? h0 - f / df : 0.5 * (lo_b + hi_b);
    if (!(lo_b <= h_new && h_new <= hi_b)) h_new = 0.5 * (lo_b + hi_b);
    const bool fast = m <= cap && fabs(h_new - h0) < A.tol * h0 && !(h_new >= r_lim * (1 - 1e-3));
    // [0] longest gather row, [1] rows left slow, [2] an extras segment overflowed, [3] gather rows
    // over capacity, [4] longest extras segment
    int32_t pv[5] = {live ? m : 0, live && glane == 0 && !fast ? 1 : 0, (int32_t)eover,
                     live && glane == 0 && m > cap ? 1 : 0, (int32_t)emax};
    block_partials<5>(pv, 0b10101u, A.part);
    if (live && glane == 0) {
        A.slow[i] = fast ? 0 : 1;
        if (fast) {
            A.h_out[i] = h_new;
            A.iters[i] = 1;
            A.flags[i] = 0;
        }
    }
}

// the two density sums of row i at inverse length inv_h, over the stored row (group-wide); d2 as
// the gather formed it
__device__ __forceinline__ void row_dens_sums(const double4* posh, const double4& pi, const uint32_t* irow,
                                              const float* mf, int32_t m, int glane, float inv_h, int kid, float mu,
                                              double& s1, double& s2)
{
    DensSums sums;
    for (int32_t k = glane; k < m; k += kG) {
        const uint32_t j = irow[k];
        const double4 pj = posh[j];
        const double dx = pj.x - pi.x;
        const double dy = pj.y - pi.y;
        const double dz = pj.z - pi.z;
        const float d2f = (float)(dx * dx + dy * dy + dz * dz);
        sums.add(sqrt32(d2f) * inv_h, mf ? mf[j] : mu, kid);
    }
    sums.total(s1, s2);
}

// position of the t-th set bit (t counted from 0) of a mask with more than t bits set: a binary
// search on the popcounts of halves
__device__ __forceinline__ int nth_set_bit(unsigned long long mask, int t)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long lo = mask & ((1ull << w) - 1ull);
        const int c = __popcll(lo);
        if (t >= c) {
            t -= c;
            mask >>= w;
            pos += w;
        } else {
            mask = lo;
        }
    }
    return pos;
}

// Rows not converged in one Newton step: the whole solve of sph.py:251-297 over the stored row
// (sums FP32 compensated, Newton steps and tests FP64), rho and Omega at the final h.
template <bool kMasses>
__global__ void __launch_bounds__(256)
k_solve_slow_mixed(const double* __restrict__ h_search, int64_t n, const double4* __restrict__ posh,
                   const uint32_t* __restrict__ ridx, const float* __restrict__ rmass, const unsigned long long* __restrict__ goff,
                   const int32_t* __restrict__ rowcnt, MixArgs A, const long long* __restrict__ slow_total,
                   int per_group)
{
    const int glane = threadIdx.x & (kG - 1);
    // Each wavefront takes 64 rows (one a lane), reads their slow flags into one ballot mask, and
    // its 64 / kG lane groups walk the set bits in order, one row per group at a time (whole
    // groups take the same row, so the group shuffles see all their lanes). No counters or
    // atomics; a wavefront with no slow row exits. Many slow rows (a start far from the root):
    // 64 consecutive rows a wavefront, in grid order, so the rows in flight at once share their
    // neighbour records in L2. Few (slow rows come in clusters of neighbours): rows strided by
    // the launch's wavefront count, so a cluster spreads over many wavefronts instead of
    // queueing in one. per_group (a launch of kG lanes per row, chosen by a host that knows
    // the count is large): 64 / kG consecutive rows a wavefront, the first lane of each group
    // voting, which keeps small systems with many slow rows parallel.
    static_assert(kG == 16, "per-group rows: one vote each 16 lanes");
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int lane = threadIdx.x & 63;
    const bool spread = !per_group && *slow_total * 16 < n;
    const int sh = per_group ? 4 : 0;  // lanes per row: 1 << sh
    const int64_t w = spread ? (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6 : 0;
    const int64_t r0 = spread ? w : (xcd_tile() * (int64_t)blockDim.x + (threadIdx.x & ~63)) >> sh;
    const int64_t step = spread ? waves : 1;
    const int64_t mine = r0 + (int64_t)(lane >> sh) * step;
    const bool vote = (lane & ((1 << sh) - 1)) == 0 && mine < n && A.slow[mine] != 0;
    const unsigned long long mask = __ballot(vote);
    const int nslow = __popcll(mask);
    for (int t = lane / kG; t < nslow; t += 64 / kG) {
        const int64_t i = r0 + (int64_t)(nth_set_bit(mask, t) >> sh) * step;
        // the stored row (shorter than rowcnt only after an overflow the caller is told about)
        const int32_t m = (int32_t)min((unsigned long long)rowcnt[i], goff[i + 1] - goff[i]);
        const double4 pi = posh[i];
        const uint32_t* irow = ridx + goff[i];
        const float* mf = kMasses ? rmass : nullptr;
        const float mu = (float)(A.m_uniform > 0.0 ? A.m_uniform : 1.0);
        const double mscale = kMasses ? 1.0 : (A.m_uniform > 0.0 ? A.m_uniform / (double)mu : 1.0);
        const double m_i = A.mass[i];
        const double r_lim = pymin(h_search[i], A.h_max);
        double hi = r_lim, lo = 0.0;
        double hh = pymin(A.h_start[i], hi);
        const double h0 = hh;
        int flag = 3, it = 0;
        const double eta3 = ipow3(A.eta);
        double s1, s2;
        while (it < 50) {
            it++;
            row_dens_sums(posh, pi, irow, mf, m, glane, (float)(1.0 / hh), A.kid, mu, s1, s2);
            const double inv_h = 1.0 / hh;
            const double rs = s1 * mscale * A.cnorm * ipow3(inv_h);
            const double dws = -s2 * mscale * A.cnorm * ipow4(inv_h);
            const double f = rs - m_i * ipow3(A.eta / hh);
            const double df = dws + 3.0 * m_i * eta3 / ipow4(hh);
            if (f > 0) hi = hh;
            else lo = hh;
            double h_new = df > 0 ? hh - f / df : 0.5 * (lo + hi);
            if (!(lo <= h_new && h_new <= hi)) h_new = 0.5 * (lo + hi);
            if (fabs(h_new - hh) < A.tol * h0) {
                hh = h_new;
                flag = 0;
                break;
            }
            hh = h_new;
        }
        if (hh >= r_lim * (1 - 1e-3)) {
            row_dens_sums(posh, pi, irow, mf, m, glane, (float)(1.0 / r_lim), A.kid, mu, s1, s2);
            const double rs_lim = s1 * mscale * A.cnorm * ipow3(1.0 / r_lim);
            if (rs_lim - m_i * ipow3(A.eta / r_lim) < 0) {
                hh = r_lim;
                flag = r_lim >= A.h_max ? 1 : 2;
            }
        }
        row_dens_sums(posh, pi, irow, mf, m, glane, (float)(1.0 / hh), A.kid, mu, s1, s2);
        const double inv_h = 1.0 / hh;
        const double rs = s1 * mscale * A.cnorm * ipow3(inv_h);
        const double dws = -s2 * mscale * A.cnorm * ipow4(inv_h);
        if (glane == 0) {
            A.h_out[i] = hh;
            A.rho[i] = rs;
            A.omega[i] = flag == 1 ? 1.0 : 1.0 + hh / (3.0 * rs) * dws;
            A.iters[i] = it;
            A.flags[i] = flag;
        }
    }
}

// final row lengths of the mixed lists (gather row + extras) into mc (entry n: 0, for the scan
// of the tight rows' offsets), the two counts per row for the next pass's capacities
// (row_counts, may be NULL), and partials of their largest value and of their sum (the
// search-list entries). Runs after a pass without overflow: the counts are the stored entries.
__global__ void __launch_bounds__(kBlock)
k_row_total_rm(const int32_t* __restrict__ rowcnt, const unsigned int* __restrict__ efill,
               const unsigned long long* __restrict__ goff, const unsigned long long* __restrict__ eoff,
               unsigned long long* __restrict__ mc, int32_t* __restrict__ row_counts, long long* __restrict__ part,
               int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    long long t = 0;
    if (i < n) {
        // the stored entries (the counts exceed them only after an overflow the caller is told about)
        t = (long long)min((unsigned long long)rowcnt[i], goff[i + 1] - goff[i]) +
            (long long)min((unsigned long long)efill[i], eoff[i + 1] - eoff[i]);
        mc[i] = (unsigned long long)t;
        if (row_counts) {
            row_counts[2 * i] = rowcnt[i];
            row_counts[2 * i + 1] = (int32_t)efill[i];
        }
    } else if (i == n) {
        mc[n] = 0;
    }
    long long pv[2] = {t, t};
    block_partials<2, long long>(pv, 1u, part);
}

// Extras arrive in atomic order: each row's segment is put in index order (so that the tightened
// rows, and every sum over them, are the same from run to run), out of place into es. A
// 16-lane group per row ranks up to 64 entries held in registers against each other (every entry
// goes to its rank, the number of smaller entries; entries of a segment are distinct). Longer
// segments, few and clustered where large search spheres meet a dense region, are sorted by the
// whole workgroup.
static constexpr int kSortRegs = 4;
static constexpr int kSortWave = 512;  // longest extras segment sorted by one wavefront (bitonic, LDS)
static constexpr int kSortLds = 2048;  // longest extras segment sorted by the workgroup (bitonic, LDS)

// the stored extras of row i (its count exceeds the segment only after a reported overflow)
__device__ __forceinline__ int64_t extras_of(const unsigned int* efill, const unsigned long long* eoff, int64_t i,
                                             unsigned long long& c0)
{
    c0 = eoff[i];
    return (int64_t)min((unsigned long long)efill[i], eoff[i + 1] - c0);
}

// Ranks of one group's segment of at most kG * R entries, R per lane in registers: every entry
// goes to its rank, the number of smaller entries (entries of a segment are distinct; pads
// compare above all of them).
template <int R>
__device__ __forceinline__ void rank_segment(const uint32_t* __restrict__ e, int64_t ne, int glane,
                                             uint32_t* __restrict__ out)
{
    uint32_t v[R];
    int rank[R];
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int64_t k = glane + kG * t;
        v[t] = k < ne ? e[k] : 0xffffffffu;
        rank[t] = 0;
    }
#pragma unroll
    for (int t2 = 0; t2 < R; t2++)
        for (int l = 0; l < kG; l++) {
            const uint32_t w = __shfl(v[t2], l, kG);
#pragma unroll
            for (int t = 0; t < R; t++) rank[t] += w < v[t] ? 1 : 0;
        }
#pragma unroll
    for (int t = 0; t < R; t++) {
        const int64_t k = glane + kG * t;
        if (k < ne) out[rank[t]] = v[t];
    }
}

// Short segments, ranked with as few registers per lane as they need (most hold at most kG
// entries); longer ones go to the work lists of k_sort_extras_mid (rows up to kSortWave
// entries, list[0 ..], count cnt[0]) and k_sort_extras_long (list[n ..], count cnt[1]).
__global__ void __launch_bounds__(256)
k_sort_extras(const uint32_t* __restrict__ ebuf, const unsigned int* __restrict__ efill,
              const unsigned long long* __restrict__ eoff, int64_t n, uint32_t* __restrict__ es,
              uint32_t* __restrict__ list, unsigned int* __restrict__ cnt)
{
    const int glane = threadIdx.x & (kG - 1);
    const int64_t i = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / kG;
    if (i >= n) return;  // whole groups leave together
    unsigned long long b0;
    const int64_t ne = extras_of(efill, eoff, i, b0);
    if (ne == 0) return;
    if (ne > kG * kSortRegs) {
        if (glane == 0) {
            const int which = ne > kSortWave ? 1 : 0;
            list[which * n + atomicAdd(&cnt[which], 1u)] = (uint32_t)i;
        }
        return;
    }
    if (ne <= kG) rank_segment<1>(ebuf + b0, ne, glane, es + b0);
    else if (ne <= 2 * kG) rank_segment<2>(ebuf + b0, ne, glane, es + b0);
    else rank_segment<kSortRegs>(ebuf + b0, ne, glane, es + b0);
}

// Bitonic sort of a segment's indices in LDS (p2 a power of two >= m, pads last), then the
// segment written in index order. sync(): the barrier of the threads taking part (tid of nt).
template <class Sync>
__device__ __forceinline__ void bitonic_segment(const uint32_t* e, int m, uint32_t* sk, int tid, int nt, uint32_t* eo,
                                                Sync sync)
{
    int p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (int t = tid; t < p2; t += nt) sk[t] = t < m ? e[t] : 0xffffffffu;
    sync();
    for (int k = 2; k <= p2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < p2; t += nt) {
                const int u = t ^ j;
                if (u > t) {
                    const bool up = (t & k) == 0;
                    const uint32_t a = sk[t], b = sk[u];
                    if ((a > b) == up) {
                        sk[t] = b;
                        sk[u] = a;
                    }
                }
            }
            sync();
        }
    for (int t = tid; t < m; t += nt) eo[t] = sk[t];
    sync();  // the LDS slice is free again
}

// Segments of kG * kSortRegs .. kSortWave entries (a particle seen by many others with larger
// search lengths; few, and clustered where large search spheres meet a dense region): one
// wavefront per row, the rows of the list spread over the whole launch.
__global__ void __launch_bounds__(256)
k_sort_extras_mid(const uint32_t* __restrict__ ebuf, const unsigned int* __restrict__ efill,
                  const unsigned long long* __restrict__ eoff, uint32_t* __restrict__ es,
                  const uint32_t* __restrict__ list, const unsigned int* __restrict__ cnt)
{
    __shared__ uint32_t sk[4][kSortWave];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned int count = cnt[0];
    const unsigned int stride = gridDim.x * (blockDim.x >> 6);
    for (unsigned int r = blockIdx.x * (blockDim.x >> 6) + wave; r < count; r += stride) {
        const int64_t row = list[r];
        unsigned long long c0;
        const int m = (int)extras_of(efill, eoff, row, c0);
        bitonic_segment(ebuf + c0, m, sk[wave], lane, 64, es + c0, [] { wave_sync(); });
    }
}

// Segments longer than kSortWave: one workgroup per row (bitonic in LDS up to kSortLds entries,
// else ranks counted against the segment read through LDS in tiles).
__global__ void __launch_bounds__(256)
k_sort_extras_long(const uint32_t* __restrict__ ebuf, const unsigned int* __restrict__ efill,
                   const unsigned long long* __restrict__ eoff, int64_t n, uint32_t* __restrict__ es,
                   const uint32_t* __restrict__ list, const unsigned int* __restrict__ cnt)
{
    __shared__ uint32_t sk[kSortLds];
    const int tid = threadIdx.x, bs = blockDim.x;
    const unsigned int count = cnt[1];
    for (unsigned int r = blockIdx.x; r < count; r += gridDim.x) {
        const int64_t row = list[n + r];
        unsigned long long c0;
        const int64_t m = extras_of(efill, eoff, row, c0);
        const uint32_t* e = ebuf + c0;
        if (m <= kSortLds) {
            bitonic_segment(e, (int)m, sk, tid, bs, es + c0, [] { __syncthreads(); });
            continue;
        }
        for (int64_t kb = 0; kb < m; kb += bs) {
            const int64_t k = kb + tid;
            const uint32_t v = k < m ? e[k] : 0u;
            int64_t rank = 0;
            for (int64_t tb = 0; tb < m; tb += kSortLds) {
                const int tl = (int)min((int64_t)kSortLds, m - tb);
                __syncthreads();
                for (int t = tid; t < tl; t += bs) sk[t] = e[tb + t];
                __syncthreads();
                if (k < m)
                    for (int q = 0; q < tl; q++) rank += sk[q] < v ? 1 : 0;
            }
            if (k < m) es[c0 + rank] = v;
        }
        __syncthreads();
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n);

// The three passes of the extras sort (work-list counters cleared by a kernel: graph-safe).
static void sort_extras(nb_ctx* ctx, hipStream_t s, int64_t n, unsigned blocks, int tile)
{
    uint32_t* list = ctx->xlist;
    unsigned int* cnt = (unsigned int*)(ctx->xlist + 2 * n);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)cnt, 0u, (int64_t)2);
    NB_TIMED(ctx, s, "k_sort_extras", k_sort_extras, dim3(blocks), dim3(tile), 0, s, ctx->ebuf, ctx->efill,
             ctx->rm_eo, n, ctx->ebuf_s, list, cnt);
    NB_TIMED(ctx, s, "k_sort_extras_mid", k_sort_extras_mid, dim3(1024), dim3(256), 0, s, ctx->ebuf, ctx->efill,
             ctx->rm_eo, ctx->ebuf_s, (const uint32_t*)list, (const unsigned int*)cnt);
    NB_TIMED(ctx, s, "k_sort_extras_long", k_sort_extras_long, dim3(256), dim3(256), 0, s, ctx->ebuf, ctx->efill,
             ctx->rm_eo, n, ctx->ebuf_s, (const uint32_t*)list, (const unsigned int*)cnt);
}

// what the emission pass needs to finish the rows converged in one Newton step
struct FinArgs {
    const float* rmass;  // FP32 particle masses, by particle index (NULL: uniform masses)
    double m_uniform;
    int kid;
    double cnorm;
    const uint8_t* slow;
    double* rho;
    double* omega;
};

// Tightened rows (row i at toff[i], capacity its merged length): gather row entries, then the
// row's extras (sorted), kept when d2 < (2 max(h_i, h_j))^2 at the final h, with d2 formed in
// FP64 from the positions as the search formed it, so the rows are exactly tighten() of the
// search lists at h_final. The same pass sums rho and Omega at the final h of the rows the gather
// finished in one Newton step, over FP32 d2 (extras never count: their d2 >= (2 h_search_i)^2 >=
// (2 h_i)^2).
// Lanes per row of the emission: 16. The pass waits on its loads (index, then the neighbour's
// record), and more rows per wavefront with fewer lanes each issue fewer of them at once: at 1M,
// 4 lanes took 1.17 ms, 8 took 0.80, 16 take 0.69 and 32 took 0.72.
static constexpr int kEmitW = 16;
static constexpr int kEmitBuf = 128;  // kept entries per row held in LDS until the walk ends

template <int kE>
__global__ void __launch_bounds__(256)
k_tight_emit_rm(const double4* __restrict__ posh, const uint32_t* __restrict__ ridx,
                const unsigned long long* __restrict__ goff, const int32_t* __restrict__ rowcnt,
                const unsigned int* __restrict__ efill, const unsigned long long* __restrict__ eoff,
                const uint32_t* __restrict__ ebuf, int64_t n, const unsigned long long* __restrict__ toff, int64_t* __restrict__ out_off,
                int32_t* __restrict__ out_len, int32_t* __restrict__ out_idx, unsigned long long idx_cap,
                int32_t* __restrict__ status, FinArgs F, long long* __restrict__ part)
{
    const int glane = threadIdx.x & (kE - 1);
    const int gshift = (threadIdx.x & 63) & ~(kE - 1);
    const int64_t i = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / kE;
    const bool live = i < n;
    const int64_t ii = live ? i : n - 1;
    const double4 pi = posh[ii];
    const double hi = pi.w;
    const unsigned long long g0 = goff[ii], e0 = eoff[ii], t0 = toff[ii];
    // the stored entries (the counts exceed them only after an overflow the caller is told about)
    const int32_t ng = live ? (int32_t)min((unsigned long long)rowcnt[ii], goff[ii + 1] - g0) : 0;
    const int64_t ne = live ? (int64_t)min((unsigned long long)efill[ii], eoff[ii + 1] - e0) : 0;
    const int64_t len = ng + ne;
    const bool fin = live && !F.slow[ii];
    const float inv_hf = (float)(1.0 / hi);
    const float mu = (float)(F.m_uniform > 0.0 ? F.m_uniform : 1.0);
    DensSums sums;
    int32_t o = 0;
    // index of entry k (< len) of the merged row: gather row entries, then the sorted extras
    auto index_at = [&](int64_t k) -> uint32_t {
        return k < ng ? ridx[g0 + k] : ebuf[e0 + (k - ng)];
    };
    auto dist2 = [&](const double4& pj) -> double {
        const double dx = pj.x - pi.x;
        const double dy = pj.y - pi.y;
        const double dz = pj.z - pi.z;
        return dx * dx + dy * dy + dz * dz;
    };
    // kept at the final h
    auto keep = [&](double d2, double hj) -> bool {
        const double rr = NB_R_KERN * pymax(hi, hj);
        return d2 < rr * rr;
    };
    // the row's first kEmitBuf kept entries wait in LDS and go out after the walk (contiguous per
    // group), so that no store sits in the walk for its loads to wait behind; later ones go at once
    __shared__ int32_t obuf[256 / kE][kEmitBuf];
    int32_t* ob = obuf[threadIdx.x / kE];
    auto put = [&](int32_t k, uint32_t j) {
        const unsigned long long at = t0 + (unsigned long long)k;
        if (at < idx_cap) out_idx[at] = (int32_t)j;
        else if (status) status[kStIndex] = 1;
    };
    auto emit = [&](bool pred, uint32_t j) {
        const unsigned long long bal = __ballot(pred);
        const unsigned gb = (unsigned)((bal >> gshift) & ((1ull << kE) - 1));
        if (pred) {
            const int32_t k = o + __popc(gb & ((1u << glane) - 1));
            if (k < kEmitBuf) ob[k] = (int32_t)j;
            else put(k, j);
        }
        o += __popc(gb);
    };
    // Two batches of kE entries per pass. The indices of the pass after next are loaded while this
    // pass's records arrive (the records depend on the indices): the same entries, sums and output
    // order as one batch at a time. Lanes past the row's end read the row's own record.
    const uint32_t self = (uint32_t)ii;
    auto index_or_self = [&](int64_t k) -> uint32_t { return k < len ? index_at(k) : self; };
    uint32_t na = index_or_self(glane), nc = index_or_self(kE + glane);
    uint32_t na2 = index_or_self(2 * kE + glane), nc2 = index_or_self(3 * kE + glane);
    for (int64_t kb = 0;; kb += 2 * kE) {
        if (__ballot(kb < len) == 0) break;
        const int64_t ka = kb + glane, kc = kb + kE + glane;
        const bool va = ka < len, vc = kc < len;
        const uint32_t ja = na, jc = nc;
        const double4 qa = posh[ja], qc = posh[jc];
        const float ma = F.rmass ? F.rmass[ja] : mu, mc = F.rmass ? F.rmass[jc] : mu;
        na = na2;
        nc = nc2;
        na2 = index_or_self(ka + 4 * kE);
        nc2 = index_or_self(kc + 4 * kE);
        const double da = dist2(qa), dc = dist2(qc);
        if (fin && va && ka < ng) sums.add(sqrt32((float)da) * inv_hf, ma, F.kid);
        if (fin && vc && kc < ng) sums.add(sqrt32((float)dc) * inv_hf, mc, F.kid);
        const bool pa = va && keep(da, qa.w);
        const bool pc = vc && keep(dc, qc.w);
        emit(pa, ja);
        emit(pc, jc);
    }
    wave_sync();
    for (int32_t k = glane; k < (o < kEmitBuf ? o : kEmitBuf); k += kE) put(k, (uint32_t)ob[k]);
    double s1, s2;
    sums.total<kE>(s1, s2);
    if (live && glane == 0) {
        // a row cut by the index capacity (reported in status) keeps what was written
        out_len[i] = t0 + (unsigned long long)o <= idx_cap ? o : (t0 < idx_cap ? (int32_t)(idx_cap - t0) : 0);
        out_off[i] = (int64_t)t0;
        if (i == n - 1) out_off[n] = (int64_t)toff[n];
        if (fin) {
            const double mscale = F.rmass ? 1.0 : (F.m_uniform > 0.0 ? F.m_uniform / (double)mu : 1.0);
            const double inv_h = 1.0 / hi;
            const double rs = s1 * mscale * F.cnorm * ipow3(inv_h);
            const double dws = -s2 * mscale * F.cnorm * ipow4(inv_h);
            F.rho[i] = rs;
            F.omega[i] = 1.0 + hi / (3.0 * rs) * dws;
        }
    }
    block_partial<false>(live && glane == 0 ? o : 0, part);
}

// ----------------------------------------------------------------- FP64 step on the lane-group lists
// The FP64 step's density pass on the lists of k_gather_mixed (rows of neighbour indices in the
// search order, which for particles stored in grid order is ascending index order, and sorted
// extras segments), with the reference's FP64 arithmetic bit for bit.

// sph.py:215-228 for one row, 16 lanes. Each lane forms the terms of entries glane, glane + 16,
// ... (the terms of entries outside 2 min(h_search, h_max) or past q = 2 exact zeros); the
// group writes each batch of 16 to LDS, and lane 0 adds the density terms, lane 1 the dW/dh
// terms, in row order: the reference's left-to-right sums, bit for bit (adding +0.0 changes
// neither sum: neither can be -0.0). cache_row forms r = sqrt(d2) (+inf outside 2 min(h_search,
// h_max)) and takes the mass of each entry once: the first kRowCache batches into registers, the
// next kRowLds batches into the group's LDS (each lane reads back only what it wrote), so the
// Newton evaluations of rows up to 256 entries make no global load (q = r / h is the same
// product either way); entries past them are formed from the records each evaluation.
static constexpr int kRowCache = 8;  // batches of 16 entries held in each lane's registers
static constexpr int kRowLds = 8;    // batches of 16 entries held in LDS (16 B each)

template <bool kMasses>
__device__ __forceinline__ void cache_row(const double4* __restrict__ posh, const double4& pi,
                                          const uint32_t* __restrict__ row, int32_t m, const double* mass,
                                          double r_max2, int glane, double (&rc)[kRowCache],
                                          double (&mc)[kRowCache], double2* __restrict__ tc)
{
    auto entry = [&](int32_t k, double& r, double& mk) {
        r = __builtin_inf();
        mk = 0.0;
        if (k < m) {
            const uint32_t j = row[k];
            const double4 pj = posh[j];
            const double dx = pj.x - pi.x;
            const double dy = pj.y - pi.y;
            const double dz = pj.z - pi.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < r_max2) r = sqrt(d2);
            if (kMasses) mk = mass[j];
        }
    };
#pragma unroll
    for (int b = 0; b < kRowCache; b++) {
        double r, mk;
        entry(b * kG + glane, r, mk);
        rc[b] = r;
        if (kMasses) mc[b] = mk;
        // four batches' records in flight at a time (all eight would hold registers the
        // evaluations need)
        if ((b & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    for (int b = 0; b < kRowLds && (kRowCache + b) * kG < m; b++) {
        double r, mk;
        entry((kRowCache + b) * kG + glane, r, mk);
        tc[b * kG + glane] = make_double2(r, mk);
    }
}

// adds the batch's terms (cnt of them) in order: lane 0 of the group the density sum, lane 1
// dW/dh (the other lanes skip the reads)
__device__ __forceinline__ void add_batch(double2* terms, int glane, double trs, double tdws, int32_t cnt, double& acc)
{
    terms[glane] = make_double2(trs, tdws);
    wave_sync();
    if (glane < 2) {
        const double* t = reinterpret_cast<const double*>(terms) + glane;
        if (cnt == kG) {
#pragma unroll
            for (int l = 0; l < kG; l++) acc += t[2 * l];
        } else {
            for (int32_t l = 0; l < cnt; l++) acc += t[2 * l];
        }
    }
    wave_sync();  // every lane has read the terms before the next batch overwrites them
    // the next batch's terms are formed after this one is added (scheduling them earlier only
    // holds more registers: the adds are a dependent chain either way)
    __builtin_amdgcn_sched_barrier(0);
}

// kernel_f and kernel_df of nb_common.h without branches: both pieces of the cubic spline are
// formed by the same expressions and one is selected (the same bits for q < 2; the caller
// discards q >= 2)
template <int kKid>
__device__ __forceinline__ void kernel_fdf(double q, double& f, double& df)
{
    if (kKid == 0) {
        const double f_in = 1.0 - 1.5 * q * q + 0.75 * q * q * q;
        const double df_in = -3.0 * q + 2.25 * q * q;
        const double t = 2.0 - q;
        const double f_out = 0.25 * t * t * t;
        const double df_out = -0.75 * t * t;
        const bool in = q < 1.0;
        f = in ? f_in : f_out;
        df = in ? df_in : df_out;
    } else {
        const double x = 1.0 - 0.5 * q;
        const double x2 = x * x;
        f = x2 * x2 * x2 * (35.0 / 12.0 * q * q + 3.0 * q + 1.0);
        df = -7.0 / 3.0 * q * x2 * x2 * x * (5.0 * q + 2.0);
    }
}

template <bool kMasses, int kKid>
__device__ __forceinline__ void row_sums_group(const double4* __restrict__ posh, const double4& pi,
                                               const uint32_t* __restrict__ row, int32_t m, const double* mass,
                                               double m_uniform, double r_max2, double h, double cnorm, int glane,
                                               const double (&rc)[kRowCache], const double (&mc)[kRowCache],
                                               const double2* __restrict__ tc, double2* terms, double& rs_out,
                                               double& dws_out)
{
    double acc = 0.0;
    const double inv_h = 1.0 / h;
    const double ih3 = ipow3(inv_h);
    const double ncih4 = -cnorm * ipow4(inv_h);
    // the row length as if redefined here (an empty statement): the batch tests and counts
    // derived from it are formed in each evaluation instead of being hoisted out of the Newton
    // loop, where they would hold registers across it
    asm volatile("" : "+v"(m));
#pragma unroll
    for (int b = 0; b < kRowCache; b++) {
        const int32_t k0 = b * kG;
        if (k0 >= m) break;
        const double q = rc[b] * inv_h;
        const double mk = kMasses ? mc[b] : m_uniform;
        double fq, dfq;
        kernel_fdf<kKid>(q, fq, dfq);
        const double trs_q = mk * cnorm * ih3 * fq;
        const double tdws_q = mk * (ncih4 * (3.0 * fq + q * dfq));
        const bool inside = q < 2.0;  // also false for entries past the row or the search (r = inf)
        add_batch(terms, glane, inside ? trs_q : 0.0, inside ? tdws_q : 0.0, m - k0 < kG ? m - k0 : kG, acc);
    }
    for (int b = 0; b < kRowLds; b++) {
        const int32_t k0 = (kRowCache + b) * kG;
        if (k0 >= m) break;
        const double2 e = tc[b * kG + glane];
        const double q = e.x * inv_h;
        const double mk = kMasses ? e.y : m_uniform;
        double fq, dfq;
        kernel_fdf<kKid>(q, fq, dfq);
        const double trs_q = mk * cnorm * ih3 * fq;
        const double tdws_q = mk * (ncih4 * (3.0 * fq + q * dfq));
        const bool inside = q < 2.0;
        add_batch(terms, glane, inside ? trs_q : 0.0, inside ? tdws_q : 0.0, m - k0 < kG ? m - k0 : kG, acc);
    }
    for (int32_t k0 = (kRowCache + kRowLds) * kG; k0 < m; k0 += kG) {
        const int32_t k = k0 + glane;
        double trs = 0.0, tdws = 0.0;
        if (k < m) {
            const uint32_t j = row[k];
            const double4 pj = posh[j];
            const double dx = pj.x - pi.x;
            const double dy = pj.y - pi.y;
            const double dz = pj.z - pi.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < r_max2) {
                const double q = sqrt(d2) * inv_h;
                if (q < 2.0) {
                    const double mk = kMasses ? mass[j] : m_uniform;
                    const double fq = kernel_f(q, kKid);
                    trs = mk * cnorm * ih3 * fq;
                    tdws = mk * (ncih4 * (3.0 * fq + q * kernel_df(q, kKid)));
                }
            }
        }
        add_batch(terms, glane, trs, tdws, m - k0 < kG ? m - k0 : kG, acc);
    }
    // lane 0 of the group holds the density sum, lane 1 dW/dh
    rs_out = __shfl(acc, 0, kG);
    dws_out = __shfl(acc, 1, kG);
}

// sph.py:251-297 for every row (as solve_row over a blocked row), 16 lanes per row.
template <bool kMasses, int kKid>
__global__ void __launch_bounds__(256, 4)
k_solve64(const double4* __restrict__ posh, const uint32_t* __restrict__ ridx,
          const unsigned long long* __restrict__ goff, const int32_t* __restrict__ rowcnt, int64_t n, DensArgs A)
{
    __shared__ double2 terms[256 / kG][kG];
    __shared__ double2 tcache[256 / kG][kRowLds * kG];  // (r, mass) of row entries 128 .. 255
    const int glane = threadIdx.x & (kG - 1);
    const int64_t i = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / kG;
    if (i >= n) return;  // whole groups leave together
    const double4 pi = posh[i];
    const double h_search_i = pi.w;
    const int32_t m = (int32_t)min((unsigned long long)rowcnt[i], goff[i + 1] - goff[i]);
    const uint32_t* row = ridx + goff[i];
    double2* tg = terms[threadIdx.x / kG];
    const double* mass = A.mass;
    const double m_i = A.mass[i];
    const double r_max = 2.0 * pymin(h_search_i, A.h_max);
    const double r_max2 = r_max * r_max;
    double rc[kRowCache], mc[kRowCache];
    double2* tc = tcache[threadIdx.x / kG];
    cache_row<kMasses>(posh, pi, row, m, mass, r_max2, glane, rc, mc, tc);
    double hi = pymin(h_search_i, A.h_max);
    double lo = 0.0;
    double h = pymin(A.h_start[i], hi);
    const double h0 = h;
    const double h_lim = hi;
    int flag = 3;
    int it = 0;
    const double eta3 = ipow3(A.eta);
    double rs = 0.0, dws = 0.0;
    // solve_row's evaluations from one call site: Newton steps (phase 0), the check at h_lim
    // (phase 1, when h ends within 1e-3 of it), the sums at the final h (phase 2)
    int phase = 0;
    double h_eval = h;
    for (;;) {
        double s1, s2;
        row_sums_group<kMasses, kKid>(posh, pi, row, m, mass, A.m_uniform, r_max2, h_eval, A.cnorm, glane, rc, mc, tc,
                                      tg, s1, s2);
        if (phase == 2) {
            rs = s1;
            dws = s2;
            break;
        }
        if (phase == 0) {
            it += 1;
            const double f = s1 - m_i * ipow3(A.eta / h);
            const double df = s2 + 3.0 * m_i * eta3 / ipow4(h);
            if (f > 0) hi = h;
            else lo = h;
            double h_new = df > 0 ? h - f / df : 0.5 * (lo + hi);
            if (!(lo <= h_new && h_new <= hi)) h_new = 0.5 * (lo + hi);
            const bool done = fabs(h_new - h) < A.tol * h0;
            h = h_new;
            if (done) flag = 0;
            if (!done && it < 50) {
                h_eval = h;
                continue;
            }
            if (h >= h_lim * (1 - 1e-3)) {
                phase = 1;
                h_eval = h_lim;
                continue;
            }
        } else if (s1 - m_i * ipow3(A.eta / h_lim) < 0) {  // phase 1
            h = h_lim;
            flag = h_lim >= A.h_max ? 1 : 2;
        }
        phase = 2;
        h_eval = h;
    }
    if (glane == 0) {
        A.h_out[i] = h;
        A.rho[i] = rs;
        A.omega[i] = flag == 1 ? 1.0 : 1.0 + h / (3.0 * rs) * dws;
        A.iters[i] = it;
        A.flags[i] = flag;
    }
}

// Tight-row blocks of the FP64 step: capacity of each 64-row block the longest merged row (gather
// row + extras) of its rows, times 64 (bsize[nblk] = 0, for the exclusive scan).
__global__ void k_block_caps64(const int32_t* __restrict__ rowcnt, const unsigned int* __restrict__ efill,
                               const unsigned long long* __restrict__ goff, const unsigned long long* __restrict__ eoff,
                               int64_t n, unsigned long long* __restrict__ bsize, int64_t nblk)
{
    const int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);  // one wavefront per block
    if (blk > nblk) return;
    const int64_t i = blk * 64 + (threadIdx.x & 63);
    long long need = 0;
    if (blk < nblk && i < n)
        need = (long long)min((unsigned long long)rowcnt[i], goff[i + 1] - goff[i]) +
               (long long)min((unsigned long long)efill[i], eoff[i + 1] - eoff[i]);
    need = need < 1 ? 1 : need;
    for (int o = 32; o > 0; o >>= 1) {
        const long long t = __shfl_xor(need, o, 64);
        need = t > need ? t : need;
    }
    if ((threadIdx.x & 63) == 0) bsize[blk] = blk < nblk ? (unsigned long long)need * 64ull : 0ull;
}

// The FP64 step's tight rows (neighbours.py:284-323 on the search lists at the final h) in 64-row
// blocks (entry k of row i at boff[i / 64] + 64 k + i % 64): the gather row (ascending) merged
// with the row's sorted extras, each entry kept when d2 < (2 max(h_i, h_j))^2, so that each row
// comes out in ascending order as tighten() leaves it. 16 lanes per row: the kept extras (up to
// kMergeLds) are compacted into LDS first; the gather row is then tightened 16 entries at a time,
// each kept entry going to its merged position (the kept entries before it plus the kept extras
// below it, by binary search), and each kept extra counts the kept gather entries below it (a
// binary search over the batch); the extras go out last. Rows with more extras merge on one lane.
static constexpr int kMergeLds = 128;

__device__ __forceinline__ int lower_bound_u32(const uint32_t* a, int n, uint32_t v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256)
k_emit64(const double4* __restrict__ posh, const uint32_t* __restrict__ ridx,
         const unsigned long long* __restrict__ goff, const int32_t* __restrict__ rowcnt,
         const unsigned int* __restrict__ efill, const unsigned long long* __restrict__ eoff,
         const uint32_t* __restrict__ es, int64_t n, const unsigned long long* __restrict__ boff,
         int32_t* __restrict__ out_len, int32_t* __restrict__ out_idx, long long* __restrict__ part)
{
    __shared__ uint32_t ek[256 / kG][kMergeLds];  // kept extras, ascending
    __shared__ int32_t ecnt[256 / kG][kMergeLds];  // kept gather entries below each
    __shared__ uint32_t jb[256 / kG][kG];         // the batch's indices (ascending)
    const int glane = threadIdx.x & (kG - 1);
    const int grp = threadIdx.x / kG;
    const int gshift = (threadIdx.x & 63) & ~(kG - 1);
    const unsigned below = (1u << glane) - 1u;
    const int64_t i = (xcd_tile() * (int64_t)blockDim.x + threadIdx.x) / kG;
    const bool live = i < n;
    const int64_t ii = live ? i : n - 1;
    const double4 pi = posh[ii];
    const double hi = pi.w;
    const int32_t ng = live ? (int32_t)min((unsigned long long)rowcnt[ii], goff[ii + 1] - goff[ii]) : 0;
    const unsigned long long e0 = eoff[ii];
    const int32_t ne = live ? (int32_t)min((unsigned long long)efill[ii], eoff[ii + 1] - e0) : 0;
    const uint32_t* row = ridx + goff[ii];
    const uint32_t* ex = es + e0;
    int32_t* orow = out_idx + (boff[ii >> 6] + (ii & 63));
    auto keep = [&](uint32_t j, double& d2) -> bool {
        const double4 pj = posh[j];
        const double dx = pj.x - pi.x;
        const double dy = pj.y - pi.y;
        const double dz = pj.z - pi.z;
        d2 = dx * dx + dy * dy + dz * dz;
        const double rr = NB_R_KERN * pymax(hi, pj.w);
        return d2 < rr * rr;
    };
    int32_t total = 0;
    if (ne > kMergeLds) {
        // many extras: one lane merges the two ascending lists
        if (glane == 0) {
            int32_t a = 0, b = 0, o = 0;
            while (a < ng || b < ne) {
                const bool take_g = b >= ne || (a < ng && row[a] < ex[b]);
                const uint32_t j = take_g ? row[a++] : ex[b++];
                double d2;
                if (keep(j, d2)) orow[(int64_t)(o++) * 64] = (int32_t)j;
            }
            total = o;
        }
    } else {
        // the kept extras, compacted into LDS (order kept)
        int32_t nek = 0;
        for (int32_t t0 = 0; t0 < ne; t0 += kG) {
            const int32_t t = t0 + glane;
            double d2 = 0.0;
            const uint32_t e = t < ne ? ex[t] : 0u;
            const bool kp = t < ne && keep(e, d2);
            const unsigned kb = (unsigned)((__ballot(kp) >> gshift) & 0xffffull);
            if (kp) {
                const int32_t q = nek + __popc(kb & below);
                ek[grp][q] = e;
                ecnt[grp][q] = 0;
            }
            nek += __popc(kb);
        }
        wave_sync();
        int32_t g = 0;  // kept gather entries so far
        for (int32_t k0 = 0; k0 < ng; k0 += kG) {
            const int32_t k = k0 + glane;
            double d2 = 0.0;
            const uint32_t j = k < ng ? row[k] : 0xffffffffu;
            const bool kp = k < ng && keep(j, d2);
            const unsigned kb = (unsigned)((__ballot(kp) >> gshift) & 0xffffull);
            if (kp) {
                const int32_t pos = g + __popc(kb & below) + lower_bound_u32(ek[grp], nek, j);
                orow[(int64_t)pos * 64] = (int32_t)j;
            }
            if (nek > 0) {
                jb[grp][glane] = j;  // ascending; past the row's end above every index
                wave_sync();
                for (int32_t q = glane; q < nek; q += kG) {
                    // the kept gather entries of this batch below extra q
                    const int p = lower_bound_u32(jb[grp], kG, ek[grp][q]);
                    ecnt[grp][q] += __popc(kb & ((1u << p) - 1u));
                }
                wave_sync();
            }
            g += __popc(kb);
        }
        for (int32_t q = glane; q < nek; q += kG) orow[(int64_t)(q + ecnt[grp][q]) * 64] = (int32_t)ek[grp][q];
        total = g + nek;
    }
    if (live && glane == 0) out_len[i] = total;
    block_partial<false>(live && glane == 0 ? total : 0, part);
}

static inline unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// workgroups of k_solve_slow_mixed: one lane per row
static inline unsigned slow_grid(int64_t n, int tile) { return (unsigned)((n + tile - 1) / tile); }

// ------------------------------------------------------------------ fixed-capacity passes
__global__ void k_fill_i64(long long* __restrict__ p, long long v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Graph-safe device-wide exclusive scan and stable radix sort (kernels only). Library sorts and
// scans of large inputs clear their state with memsets, which misbehave in replayed graphs.
static constexpr int kScanThreads = 256;
static constexpr int kScanItems = 4;
static constexpr int kScanTile = kScanThreads * kScanItems;

// exclusive scan of v over the workgroup (kScanThreads lanes); total returned to every lane
template <class T>
__device__ __forceinline__ T block_exclusive(T v, T& total)
{
    __shared__ T wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const T t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    __syncthreads();
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    T before = 0, all = 0;
    for (int k = 0; k < kScanThreads / 64; k++) {
        if (k < w) before += wsum[k];
        all += wsum[k];
    }
    total = all;
    return before + x - v;
}

template <class T>
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const T* __restrict__ in, int64_t count, T* __restrict__ sums)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    T v = 0;
    for (int k = 0; k < kScanItems; k++)
        if (base + k < count) v += in[base + k];
    T total;
    block_exclusive(v, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of the tile sums in place, one workgroup
template <class T>
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(T* __restrict__ sums, int64_t ntiles)
{
    T carry = 0;
    for (int64_t c = 0; c < ntiles; c += kScanThreads) {
        const int64_t i = c + threadIdx.x;
        const T v = i < ntiles ? sums[i] : T(0);
        T total;
        const T ex = block_exclusive(v, total);
        if (i < ntiles) sums[i] = carry + ex;
        carry += total;
    }
}

template <class T>
__global__ void __launch_bounds__(kScanThreads)
k_scan_apply(const T* __restrict__ in, int64_t count, const T* __restrict__ sums, T* __restrict__ out)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T local = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < count ? in[base + k] : T(0);
        local += v[k];
    }
    T total;
    T run = sums[blockIdx.x] + block_exclusive(local, total);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < count) out[base + k] = run;
        run += v[k];
    }
}

// Two exclusive scans of count values each in one set of launches (blockIdx.y or, for the tile
// sums, blockIdx.x picks the array), each result clamped to its capacity: out_a[i] = min(scan,
// cap_a), and status[word_a] = 1 when the total (entry count - 1) exceeds cap_a (rows past the
// capacity then get none: the pass overflows and says so).
struct Scan2 {
    const unsigned long long* in[2];
    unsigned long long* out[2];
    unsigned long long cap[2];
    int word[2];
};

__global__ void __launch_bounds__(kScanThreads) k_scan2_tiles(Scan2 S, int64_t count, unsigned long long* __restrict__ sums,
                                                              int64_t ntiles)
{
    const unsigned long long* in = S.in[blockIdx.y];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    unsigned long long v = 0;
    for (int k = 0; k < kScanItems; k++)
        if (base + k < count) v += in[base + k];
    unsigned long long total;
    block_exclusive(v, total);
    if (threadIdx.x == 0) sums[blockIdx.y * ntiles + blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan2_sums(unsigned long long* __restrict__ sums, int64_t ntiles)
{
    unsigned long long* sm = sums + blockIdx.x * ntiles;
    unsigned long long carry = 0;
    for (int64_t c = 0; c < ntiles; c += kScanThreads) {
        const int64_t i = c + threadIdx.x;
        const unsigned long long v = i < ntiles ? sm[i] : 0ull;
        unsigned long long total;
        const unsigned long long ex = block_exclusive(v, total);
        if (i < ntiles) sm[i] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan2_apply(Scan2 S, int64_t count,
                                                              const unsigned long long* __restrict__ sums,
                                                              int64_t ntiles, int32_t* __restrict__ status)
{
    const int a = blockIdx.y;
    const unsigned long long* in = S.in[a];
    unsigned long long* out = S.out[a];
    const unsigned long long cap = S.cap[a];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    unsigned long long v[kScanItems];
    unsigned long long local = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < count ? in[base + k] : 0ull;
        local += v[k];
    }
    unsigned long long total;
    unsigned long long run = sums[a * ntiles + blockIdx.x] + block_exclusive(local, total);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < count) {
            out[base + k] = run < cap ? run : cap;
            if (base + k == count - 1 && run > cap) status[S.word[a]] = 1;
        }
        run += v[k];
    }
}

static constexpr int kSortThreads = 256;
static constexpr int kRadixBits = 4;
static constexpr int kRadix = 1 << kRadixBits;

// digit counts of each workgroup's keys: hist[d * nblocks + block]
__global__ void __launch_bounds__(kSortThreads)
k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t cnt[kRadix];
    if (threadIdx.x < kRadix) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();
    if (threadIdx.x < kRadix) hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// stable scatter by digit: position = start of (digit, block) + rank among the workgroup's
// earlier keys with the same digit (waves in order, lanes in order by ballot)
__global__ void __launch_bounds__(kSortThreads)
k_radix_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
                const uint32_t* __restrict__ offs, uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out)
{
    __shared__ uint32_t wcnt[kSortThreads / 64][kRadix];
    const int64_t i = (int64_t)blockIdx.x * kSortThreads + threadIdx.x;
    const bool live = i < n;
    const uint32_t key = live ? keys[i] : 0u;
    const int d = live ? (int)((key >> shift) & (kRadix - 1)) : -1;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t rank = 0;
    for (int b = 0; b < kRadix; b++) {
        const unsigned long long m = __ballot(d == b);
        if (d == b) rank = (uint32_t)__popcll(m & lt);
        if (lane == 0) wcnt[w][b] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (live) {
        uint32_t before = 0;
        for (int k = 0; k < w; k++) before += wcnt[k][d];
        const uint32_t pos = offs[(int64_t)d * gridDim.x + blockIdx.x] + before + rank;
        keys_out[pos] = key;
        vals_out[pos] = vals[i];
    }
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// The segmented grid plan of build_grid_seg on the device (same arithmetic, so the same grids),
// from k_seg_bbox's boxes. When the cells of all segments exceed cmax, every segment's starting
// cell edge grows by 1.25 until they fit (any cell edge gives the same lists; only the visit
// order, hence FP32 sums, change) and status[kStCells] is set.
__global__ void k_plan_seg(const double* __restrict__ bb, const int64_t* __restrict__ off, int nseg,
                           double cell_factor, double cmax, double max_axis, double max_cells,
                           GridSeg* __restrict__ segs, int32_t* __restrict__ status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double grow = 1.0;
    for (int attempt = 0; attempt < 200; attempt++) {
        double total = 0.0;
        for (int k = 0; k < nseg; k++) {
            const double* b = bb + 9 * k;
            const unsigned long long hsum = ((const unsigned long long*)bb)[9 * k + 8];
            const double hmax = b[6];
            double span[3];
            for (int c = 0; c < 3; c++) {
                segs[k].g.lo[c] = b[c];
                span[c] = b[3 + c] - b[c];
            }
            const double scale = hmax > 0.0 ? 4294967296.0 / hmax : 0.0;
            const double hmean = scale > 0.0 ? ((double)hsum / scale) / (double)(off[k + 1] - off[k]) : 0.0;
            const double maxspan = fmax(span[0], fmax(span[1], span[2]));
            double cell = fmax(fmax(cell_factor * hmean, maxspan / max_axis), 1e-300) * grow;
            double cells;
            for (;;) {
                cells = 1.0;
                for (int c = 0; c < 3; c++) {
                    const double cc = c == 2 ? cell / kSegZSub : cell;
                    segs[k].g.nc[c] = (int32_t)fmax(floor(span[c] / cc) + 1.0, 1.0);
                    segs[k].g.inv[c] = 1.0 / cc;
                    cells *= (double)segs[k].g.nc[c];
                }
                if (cells <= max_cells) break;
                cell *= 1.25;
            }
            segs[k].base = (uint32_t)total;
            segs[k].pad = 0;
            total += cells;
        }
        if (total <= cmax) {
            if (attempt > 0) status[kStCells] = 1;
            return;
        }
        grow *= 1.25;
    }
    status[kStCells] = 1;
}

// rows whose h-rho root lies beyond their search (flag 2): the caller redoes the step eagerly
__global__ void k_status_grow(const int32_t* __restrict__ flags, int64_t n, int32_t* __restrict__ status)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && flags[i] == 2) atomicAdd(&status[kStGrow], 1);
}

}  // namespace nb

// =================================================================== host side
// rocPRIM's device-wide primitives, called directly: the stable LSD radix sort of (key, value)
// pairs by the key's low `bits` bits, and the exclusive prefix sum. Each is called twice, first
// with temp = nullptr for its scratch size, then to run.
template <class T>
static hipError_t excl_sum(void* temp, size_t& bytes, const T* in, T* out, int64_t count, hipStream_t s)
{
    return rocprim::exclusive_scan(temp, bytes, in, out, T(0), (size_t)count, rocprim::plus<T>(), s);
}

// rocPRIM's default takes its block-sort + merge path up to 2^20 items (21 launches, 125 us for
// 2^20 cell ids of 24 bits on MI355X); above 512K items its onesweep radix passes are faster
// (107 us at 2^20, the same stable order), below it the merge path is (80 against 94 us at 256K)
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 512 * 1024>;

static hipError_t sort_pairs(void* temp, size_t& bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                             uint32_t* vout, int64_t n, int bits, hipStream_t s)
{
    return rocprim::radix_sort_pairs<SortConfig>(temp, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s);
}

using namespace nb;

namespace {

int scan_u64(nb_ctx* ctx, hipStream_t s, const unsigned long long* in, unsigned long long* out, int64_t count)
{
    size_t bytes = 0;
    if (excl_sum(nullptr, bytes, in, out, (int64_t)(count), s) != hipSuccess)
        return ctx_fail(ctx, "scan sizing failed");
    if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
    if (excl_sum(ctx->cub_tmp, bytes, in, out, (int64_t)(count), s) != hipSuccess)
        return ctx_fail(ctx, "scan failed");
    return NB_OK;
}

int sync_copy(nb_ctx* ctx, hipStream_t s, void* host, const void* dev, size_t bytes, const char* what)
{
    if (hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, what);
    return NB_OK;
}

const int64_t kMaxAxisCells = 2048;
const double kMaxTotalCells = 134217728.0;  // 2^27

Box make_box(const double* box3)
{
    Box b{{0.0, 0.0, 0.0}};
    if (box3)
        for (int c = 0; c < 3; c++) b.l[c] = box3[c];
    return b;
}

// Grid of the given positions and search lengths, particles sorted by cell into ctx->order and
// cell starts into ctx->cell_start.
// Scratch of the per-workgroup partials (4 slots for launches of up to n / 16 + 1 workgroups) and
// of the collected scalars (ctx->scal, 16 values).
int reserve_reductions(nb_ctx* ctx, int64_t n)
{
    // 5 slots for launches of up to n / 4 + 1 workgroups (lane groups of 16 in 64-thread workgroups)
    if (ctx_reserve(ctx, (void**)&ctx->part, &ctx->part_bytes, sizeof(long long) * 5 * (size_t)(n / 4 + 1024)) ||
        ctx_reserve(ctx, (void**)&ctx->scal, &ctx->scal_bytes, sizeof(long long) * 16))
        return NB_ERR_HIP;
    return NB_OK;
}

inline void finish_partials(nb_ctx* ctx, hipStream_t s, int64_t nb, int nslots, unsigned maxmask, int at,
                            int32_t* status = nullptr, SlotStatus ss = SlotStatus{{-1, -1, -1, -1, -1}})
{
    long long* out = (long long*)ctx->scal + at;
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, s, out, 0ll, (int64_t)nslots);  // a kernel: graph-safe
    const int64_t blocks = std::min<int64_t>(256, (nb + 4 * kBlock - 1) / (4 * kBlock));
    hipLaunchKernelGGL(k_partials_final, dim3((unsigned)blocks), dim3(kBlock), 0, s, ctx->part, nb, nslots, maxmask,
                       out, status, ss);
}

// The particle order of the last grid for the gather kernels: NULL when the particles were stored
// in cell order already (identity).
inline const uint32_t* order_arg(const nb_ctx* ctx) { return ctx->order_ident ? nullptr : ctx->order; }

// Grid of the given positions and search lengths and the particles' cell order (ctx->order,
// cell_start). check_sorted: when the particles are already stored in cell order (the step
// keeps them so) the sort is skipped and ctx->order_ident set.
int build_grid(nb_ctx* ctx, hipStream_t s, const double* pos, const double* h, int64_t n, const Box& box, Grid& g,
               int64_t& ncells, bool check_sorted, double cell_factor = 1.0)
{
    // every list builder starts here and reuses the context's list buffers: lists of an
    // earlier build can no longer be emitted
    ctx->lists_ready = 0;
    ctx->fx_grid_from_order = 0;
    ctx->hydro_ready = 0;
    ctx->hydro_mixed_ready = 0;
    ctx->hydro64_ready = 0;
    ctx->order_ident = 0;
    int nparts = (int)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
    if (ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 8 * (nparts + 2))) return NB_ERR_HIP;
    if (reserve_reductions(ctx, n)) return NB_ERR_HIP;
    double* fin = ctx->red + 8 * nparts;
    unsigned long long* hsum = (unsigned long long*)(fin + 8);
    hipLaunchKernelGGL(k_bbox, dim3(nparts), dim3(kBlock), 0, s, pos, h, n, box, ctx->red);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, s, ctx->red, nparts, fin, hsum);
    hipLaunchKernelGGL(k_hsum, dim3(nparts), dim3(kBlock), 0, s, h, n, fin, hsum);
    double bb[9];
    int rc = sync_copy(ctx, s, bb, fin, sizeof(double) * 9, "grid: bounding box copy failed");
    if (rc) return rc;
    unsigned long long hs;
    std::memcpy(&hs, &bb[8], sizeof(hs));
    const double hmax = bb[6];
    for (int c = 0; c < 3; c++)
        if (box.l[c] > 0.0 && NB_R_KERN * hmax >= 0.5 * box.l[c])
            return ctx_fail_code(ctx, NB_ERR_BOX, "kernel support exceeds half the periodic box: neighbours ambiguous");
    if (!std::isfinite(bb[0] + bb[1] + bb[2] + bb[3] + bb[4] + bb[5]) || !std::isfinite(hmax))
        return ctx_fail(ctx, "grid: non-finite positions or smoothing lengths");
    double span[3];
    for (int c = 0; c < 3; c++) {
        g.lo[c] = box.l[c] > 0.0 ? 0.0 : bb[c];
        span[c] = box.l[c] > 0.0 ? box.l[c] : bb[3 + c] - bb[c];
    }
    const double scale = hmax > 0.0 ? 4294967296.0 / hmax : 0.0;
    const double hmean = scale > 0.0 ? ((double)hs / scale) / (double)n : 0.0;
    const double maxspan = std::max(span[0], std::max(span[1], span[2]));
    double cell = std::max(std::max(cell_factor * hmean, maxspan / (double)kMaxAxisCells), 1e-300);
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
    ncells = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
    ctx->last_grid_cells = ncells;
    ctx->last_cell = cell;
    if (ctx_reserve(ctx, (void**)&ctx->cell_id, &ctx->cell_id_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->order, &ctx->order_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->keys_out, &ctx->keys_out_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->iota, &ctx->iota_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->cell_count, &ctx->cell_count_bytes, sizeof(uint32_t) * (ncells + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->cell_start, &ctx->cell_start_bytes, sizeof(uint32_t) * (ncells + 1)))
        return NB_ERR_HIP;
    if (hipMemsetAsync(ctx->cell_count, 0, sizeof(uint32_t) * (ncells + 1), s) != hipSuccess)
        return ctx_fail(ctx, "memset cell_count");
    NB_TIMED(ctx, s, "k_cell_id", k_cell_id, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, n, box, g, ctx->cell_id, ctx->iota,
                       ctx->cell_count, check_sorted ? ctx->part : nullptr);
    bool sorted = false;
    if (check_sorted) {
        finish_partials(ctx, s, grid_of(n), 1, 0u, 0);
        long long bad = 0;
        rc = sync_copy(ctx, s, &bad, ctx->scal, sizeof(bad), "grid: order check failed");
        if (rc) return rc;
        sorted = bad == 0;
    }
    if (!sorted) {
        int end_bit = 1;
        while ((int64_t(1) << end_bit) < ncells && end_bit < 32) end_bit++;
        size_t bytes = 0;
        if (sort_pairs(nullptr, bytes, ctx->cell_id, ctx->keys_out, ctx->iota, ctx->order, n, end_bit, s) != hipSuccess)
            return ctx_fail(ctx, "sort sizing failed");
        if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
        if (sort_pairs(ctx->cub_tmp, bytes, ctx->cell_id, ctx->keys_out, ctx->iota, ctx->order, n, end_bit, s) != hipSuccess)
            return ctx_fail(ctx, "sort failed");
    }
    ctx->order_ident = sorted ? 1 : 0;
    size_t bytes = 0;
    if (excl_sum(nullptr, bytes, ctx->cell_count, ctx->cell_start, (int64_t)(ncells + 1), s) !=
        hipSuccess)
        return ctx_fail(ctx, "scan sizing failed");
    if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
    if (excl_sum(ctx->cub_tmp, bytes, ctx->cell_count, ctx->cell_start, (int64_t)(ncells + 1), s) !=
        hipSuccess)
        return ctx_fail(ctx, "scan failed");
    return NB_OK;
}

// Grids of the segments of an ensemble (open boxes), each planned as build_grid plans one grid
// from its own box and search lengths, numbered one after the other (GridSeg.base), and the
// particles' order by global cell id (ctx->order, cell_start, order_ident as build_grid). cloud[i]:
// Workgroups per segment of the segment box passes, and the bytes of their partials.
static int seg_parts(int64_t n, int64_t nseg)
{
    const int64_t p = (n + nseg * 8192 - 1) / (nseg * 8192);
    return (int)std::max<int64_t>(1, std::min<int64_t>(128, p));
}

static size_t seg_part_bytes(int64_t n, int64_t nseg)
{
    return (size_t)nseg * (size_t)seg_parts(n, nseg) * (8 * sizeof(double) + sizeof(unsigned long long));
}

// Each segment's box, h_max and fixed-point h sum into out (9 values per segment), from the
// partials in ctx->segpart (seg_part_bytes reserved).
static void launch_seg_bbox(nb_ctx* ctx, hipStream_t s, const double* pos, const double* h, int64_t n,
                            const int64_t* off_dev, int32_t nseg, double* out)
{
    const int parts = seg_parts(n, nseg);
    double* part = (double*)ctx->segpart;
    unsigned long long* hpart = (unsigned long long*)(part + 8 * (size_t)nseg * parts);
    hipLaunchKernelGGL(k_seg_bbox_part, dim3((unsigned)nseg, (unsigned)parts), dim3(kBlock), 0, s, pos, h, off_dev,
                       parts, part);
    hipLaunchKernelGGL(k_seg_bbox_final, dim3((unsigned)nseg), dim3(kBlock), 0, s, (const double*)part, parts, out);
    hipLaunchKernelGGL(k_seg_hsum_part, dim3((unsigned)nseg, (unsigned)parts), dim3(kBlock), 0, s, h, off_dev,
                       (const double*)out, parts, hpart);
    hipLaunchKernelGGL(k_seg_hsum_final, dim3((unsigned)nseg), dim3(kBlock), 0, s,
                       (const unsigned long long*)hpart, parts, out);
}

// the segment of particle i; off (host, nseg + 1): the segments' particle ranges.
int build_grid_seg(nb_ctx* ctx, hipStream_t s, const double* pos, const double* h, int64_t n, const int32_t* cloud,
                   const int64_t* off, int32_t nseg, double cell_factor, bool check_sorted, int64_t& ncells)
{
    ctx->lists_ready = 0;
    ctx->hydro_ready = 0;
    ctx->hydro_mixed_ready = 0;
    ctx->hydro64_ready = 0;
    ctx->order_ident = 0;
    ctx->fx_grid_from_order = 0;
    if (nseg < 1 || off[0] != 0 || off[nseg] != n) return ctx_fail(ctx, "segmented grid: bad segment offsets");
    for (int k = 0; k < nseg; k++)
        if (off[k + 1] <= off[k]) return ctx_fail(ctx, "segmented grid: empty segment");
    if (reserve_reductions(ctx, n) ||
        ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 9 * nseg + sizeof(int64_t) * (nseg + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->segs, &ctx->segs_bytes, sizeof(GridSeg) * nseg) ||
        ctx_reserve(ctx, &ctx->segpart, &ctx->segpart_bytes, seg_part_bytes(n, nseg)))
        return NB_ERR_HIP;
    int64_t* off_dev = (int64_t*)(ctx->red + 9 * nseg);
    if (hipMemcpyAsync(off_dev, off, sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "segmented grid: offset copy failed");
    launch_seg_bbox(ctx, s, pos, h, n, off_dev, nseg, ctx->red);
    std::vector<double> bb((size_t)9 * nseg);
    int rc = sync_copy(ctx, s, bb.data(), ctx->red, sizeof(double) * 9 * nseg, "segmented grid: box copy failed");
    if (rc) return rc;
    ctx->host_segs.resize((size_t)nseg * sizeof(GridSeg));
    GridSeg* hs = (GridSeg*)ctx->host_segs.data();
    double total_cells = 0.0;
    for (int k = 0; k < nseg; k++) {
        const double* b = &bb[(size_t)9 * k];
        unsigned long long hsum;
        std::memcpy(&hsum, &b[8], sizeof(hsum));
        const double hmax = b[6];
        if (!std::isfinite(b[0] + b[1] + b[2] + b[3] + b[4] + b[5]) || !std::isfinite(hmax))
            return ctx_fail(ctx, "segmented grid: non-finite positions or smoothing lengths");
        double span[3];
        for (int c = 0; c < 3; c++) {
            hs[k].g.lo[c] = b[c];
            span[c] = b[3 + c] - b[c];
        }
        const double scale = hmax > 0.0 ? 4294967296.0 / hmax : 0.0;
        const double hmean = scale > 0.0 ? ((double)hsum / scale) / (double)(off[k + 1] - off[k]) : 0.0;
        const double maxspan = std::max(span[0], std::max(span[1], span[2]));
        double cell = std::max(std::max(cell_factor * hmean, maxspan / (double)kMaxAxisCells), 1e-300);
        double cells;
        for (;;) {
            cells = 1.0;
            for (int c = 0; c < 3; c++) {
                const double cc = c == 2 ? cell / kSegZSub : cell;
                hs[k].g.nc[c] = (int32_t)std::max(std::floor(span[c] / cc) + 1.0, 1.0);
                hs[k].g.inv[c] = 1.0 / cc;
                cells *= (double)hs[k].g.nc[c];
            }
            if (cells <= kMaxTotalCells) break;
            cell *= 1.25;
        }
        hs[k].base = (uint32_t)total_cells;
        hs[k].pad = 0;
        total_cells += cells;
        if (total_cells > 2147483648.0) return ctx_fail(ctx, "segmented grid: more than 2^31 cells in total");
    }
    ncells = (int64_t)total_cells;
    ctx->last_grid_cells = ncells;
    if (hipMemcpyAsync(ctx->segs, hs, sizeof(GridSeg) * nseg, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ctx_fail(ctx, "segmented grid: table copy failed");
    if (ctx_reserve(ctx, (void**)&ctx->cell_id, &ctx->cell_id_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->order, &ctx->order_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->keys_out, &ctx->keys_out_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->iota, &ctx->iota_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->cell_count, &ctx->cell_count_bytes, sizeof(uint32_t) * (ncells + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->cell_start, &ctx->cell_start_bytes, sizeof(uint32_t) * (ncells + 1)))
        return NB_ERR_HIP;
    if (hipMemsetAsync(ctx->cell_count, 0, sizeof(uint32_t) * (ncells + 1), s) != hipSuccess)
        return ctx_fail(ctx, "memset cell_count");
    NB_TIMED(ctx, s, "k_cell_id", k_cell_id_seg, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, n, (const GridSeg*)ctx->segs,
             cloud, ctx->cell_id, ctx->iota, ctx->cell_count, check_sorted ? ctx->part : nullptr);
    bool sorted = false;
    if (check_sorted) {
        finish_partials(ctx, s, grid_of(n), 1, 0u, 0);
        long long bad = 0;
        rc = sync_copy(ctx, s, &bad, ctx->scal, sizeof(bad), "segmented grid: order check failed");
        if (rc) return rc;
        sorted = bad == 0;
    }
    if (!sorted) {
        int end_bit = 1;
        while ((int64_t(1) << end_bit) < ncells && end_bit < 32) end_bit++;
        size_t bytes = 0;
        if (sort_pairs(nullptr, bytes, ctx->cell_id, ctx->keys_out, ctx->iota, ctx->order, n, end_bit, s) != hipSuccess)
            return ctx_fail(ctx, "sort sizing failed");
        if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
        if (sort_pairs(ctx->cub_tmp, bytes, ctx->cell_id, ctx->keys_out, ctx->iota, ctx->order, n, end_bit, s) != hipSuccess)
            return ctx_fail(ctx, "sort failed");
    }
    ctx->order_ident = sorted ? 1 : 0;
    size_t bytes = 0;
    if (excl_sum(nullptr, bytes, ctx->cell_count, ctx->cell_start, (int64_t)(ncells + 1), s) !=
        hipSuccess)
        return ctx_fail(ctx, "scan sizing failed");
    if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
    if (excl_sum(ctx->cub_tmp, bytes, ctx->cell_count, ctx->cell_start, (int64_t)(ncells + 1), s) !=
        hipSuccess)
        return ctx_fail(ctx, "scan failed");
    return NB_OK;
}

// Grid, gather, row sort, extras and final row lengths: everything up to the emission.
int gather_stage(nb_ctx* ctx, hipStream_t s, const double* pos, const double* h, int64_t n, const Box& box)
{
    Grid g;
    int64_t ncells = 0;
    int rc = build_grid(ctx, s, pos, h, n, box, g, ncells, true);
    if (rc) return rc;
    if (ctx_reserve(ctx, (void**)&ctx->rowcnt, &ctx->rowcnt_bytes, sizeof(int32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->unsorted, &ctx->unsorted_bytes, n) ||
        ctx_reserve(ctx, (void**)&ctx->ecount, &ctx->ecount_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->eoff, &ctx->eoff_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->tot, &ctx->tot_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->efill, &ctx->efill_bytes, sizeof(unsigned int) * n) ||
        ctx_reserve(ctx, (void**)&ctx->offs, &ctx->offs_bytes, sizeof(unsigned long long) * (n + 1)))
        return NB_ERR_HIP;
    const int64_t nblk = (n + 63) / 64;
    const int64_t nb = grid_of(n);
    int64_t cap = ctx->row_cap > 0 ? ctx->row_cap : 192;
    long long sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int attempt = 0; attempt < 8; attempt++) {
        if (ctx_reserve(ctx, (void**)&ctx->gbuf, &ctx->gbuf_bytes, sizeof(uint32_t) * (size_t)(nblk * cap * 64)))
            return NB_ERR_HIP;
        if (hipMemsetAsync(ctx->ecount, 0, sizeof(unsigned long long) * (n + 1), s) != hipSuccess)
            return ctx_fail(ctx, "memset");
        NB_TIMED(ctx, s, "k_gather", k_gather, dim3(nb), dim3(kBlock), 0, s, pos, h, n, box, g, ctx->cell_start, order_arg(ctx),
                           ctx->gbuf, cap, ctx->rowcnt, ctx->unsorted, ctx->ecount, ctx->part);
        finish_partials(ctx, s, nb, 2, 1u, 0);  // [0] longest row, [1] unsorted rows
        rc = sync_copy(ctx, s, sc, ctx->scal, sizeof(long long) * 2, "gather failed");
        if (rc) return rc;
        if (sc[0] <= cap) break;
        cap = ((sc[0] + sc[0] / 8 + 15) / 16) * 16;
        if (attempt == 7) return ctx_fail(ctx, "row capacity did not settle");
    }
    ctx->row_cap = cap;
    ctx->last_unsorted = sc[1];
    if (sc[1] > 0)
        NB_TIMED(ctx, s, "k_sort_rows", k_sort_rows, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, cap, ctx->rowcnt, ctx->unsorted, n);
    rc = scan_u64(ctx, s, ctx->ecount, ctx->eoff, n + 1);
    if (rc) return rc;
    NB_TIMED(ctx, s, "k_row_total", k_row_total, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->rowcnt, ctx->ecount, ctx->tot,
                       ctx->part, n, (int32_t*)nullptr);
    finish_partials(ctx, s, grid_of(n + 1), 1, 1u, 2);  // [2] longest merged row
    rc = scan_u64(ctx, s, ctx->tot, ctx->offs, n + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, s, (long long*)ctx->scal + 3, ctx->eoff + n, ctx->offs + n);
    rc = sync_copy(ctx, s, sc, ctx->scal, sizeof(long long) * 5, "list totals copy failed");
    if (rc) return rc;
    const unsigned long long ne_total = (unsigned long long)sc[3];
    if (ctx_reserve(ctx, (void**)&ctx->ebuf, &ctx->ebuf_bytes, sizeof(uint32_t) * (ne_total + 1))) return NB_ERR_HIP;
    if (hipMemsetAsync(ctx->efill, 0, sizeof(unsigned int) * n, s) != hipSuccess) return ctx_fail(ctx, "memset");
    NB_TIMED(ctx, s, "k_extras_fill", k_extras_fill, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, cap, ctx->rowcnt, ctx->eoff,
                       ctx->efill, ctx->ebuf, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "gather: kernel launch failed");
    ctx->lists_n = n;
    ctx->lists_total = sc[4];
    ctx->lists_maxrow = sc[2];
    ctx->lists_ready = 1;
    return NB_OK;
}

}  // namespace

extern "C" int32_t nb_grid_order(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                 const double* box3, double cell_factor, int32_t* order_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || !order_out || !(cell_factor > 0.0)) return ctx_fail(ctx, "nb_grid_order: bad arguments");
    if (n == 0) return NB_OK;
    if (n >= (int64_t)kMask) return ctx_fail(ctx, "nb_grid_order: too many particles");
    Grid g;
    int64_t ncells;
    int rc = build_grid(ctx, s, pos, h, n, make_box(box3), g, ncells, false, cell_factor);
    if (rc) return rc;
    if (hipMemcpyAsync(order_out, ctx->order, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "nb_grid_order: copy failed");
    return NB_OK;
}

extern "C" int32_t nb_grid_order_seg(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                     const int32_t* cloud, const int64_t* seg_off, int32_t nseg, double cell_factor,
                                     int32_t* order_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n <= 0 || !order_out || !cloud || !seg_off || !(cell_factor > 0.0))
        return ctx_fail(ctx, "nb_grid_order_seg: bad arguments");
    if (n >= (int64_t)kMask) return ctx_fail(ctx, "nb_grid_order_seg: too many particles");
    int64_t ncells;
    int rc = build_grid_seg(ctx, s, pos, h, n, cloud, seg_off, nseg, cell_factor, false, ncells);
    if (rc) return rc;
    if (hipMemcpyAsync(order_out, ctx->order, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "nb_grid_order_seg: copy failed");
    return NB_OK;
}

extern "C" int32_t nb_lists_build(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                  const double* box3, int64_t* out_offset, int32_t* out_index, int64_t capacity,
                                  int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || !total_out || !out_offset) return ctx_fail(ctx, "nb_lists_build: bad arguments");
    ctx->lists_ready = 0;
    if (n >= (int64_t)kMask) return ctx_fail(ctx, "nb_lists_build: too many particles");
    if (n == 0) {
        if (hipMemsetAsync(out_offset, 0, sizeof(int64_t), s) != hipSuccess) return ctx_fail(ctx, "memset");
        *total_out = 0;
        ctx->lists_n = 0;
        ctx->lists_total = 0;
        ctx->lists_ready = 1;
        return NB_OK;
    }
    int rc = gather_stage(ctx, s, pos, h, n, make_box(box3));
    if (rc) return rc;
    if (hipMemcpyAsync(out_offset, ctx->offs, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "offset copy failed");
    *total_out = ctx->lists_total;
    if (ctx->lists_total > capacity || !out_index) return NB_RETRY_CAPACITY;
    return nb_lists_emit(ctx, stream, out_index, capacity);
}

extern "C" int32_t nb_lists_emit(nb_ctx* ctx, void* stream, int32_t* out_index, int64_t capacity)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->lists_ready) return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_lists_emit: no lists built");
    if (ctx->lists_total > capacity) return ctx_fail(ctx, "nb_lists_emit: capacity too small");
    const int64_t n = ctx->lists_n;
    if (n == 0) return NB_OK;
    NB_TIMED(ctx, s, "k_merge", k_merge<false>, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, (int64_t)ctx->row_cap,
                       ctx->rowcnt, ctx->eoff, ctx->ebuf, ctx->offs, out_index, (int32_t*)nullptr, (int64_t)1 << 40, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_lists_emit: kernel launch failed");
    ctx->lists_ready = 2;
    return NB_OK;
}

extern "C" int32_t nb_lists_build_blocked(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                          const double* box3, int32_t* out_len, int32_t* out_index, int64_t row_cap,
                                          int64_t* need_cap_out, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || !need_cap_out || !total_out) return ctx_fail(ctx, "nb_lists_build_blocked: bad arguments");
    ctx->lists_ready = 0;
    if (n >= (int64_t)kMask) return ctx_fail(ctx, "nb_lists_build_blocked: too many particles");
    *need_cap_out = 0;
    *total_out = 0;
    if (n == 0) {
        ctx->lists_n = 0;
        ctx->lists_total = 0;
        ctx->lists_ready = 1;
        return NB_OK;
    }
    int rc = gather_stage(ctx, s, pos, h, n, make_box(box3));
    if (rc) return rc;
    *need_cap_out = ctx->lists_maxrow;
    *total_out = ctx->lists_total;
    if (ctx->lists_maxrow > row_cap || !out_index || !out_len) return NB_RETRY_CAPACITY;
    return nb_lists_emit_blocked(ctx, stream, out_len, out_index, row_cap);
}

extern "C" int32_t nb_lists_emit_blocked(nb_ctx* ctx, void* stream, int32_t* out_len, int32_t* out_index,
                                         int64_t row_cap)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->lists_ready) return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_lists_emit_blocked: no lists built");
    if (ctx->lists_maxrow > row_cap) return ctx_fail(ctx, "nb_lists_emit_blocked: row capacity too small");
    const int64_t n = ctx->lists_n;
    if (n == 0) return NB_OK;
    NB_TIMED(ctx, s, "k_merge", k_merge<true>, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, (int64_t)ctx->row_cap,
                       ctx->rowcnt, ctx->eoff, ctx->ebuf, ctx->offs, out_index, out_len, row_cap, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_lists_emit_blocked: kernel launch failed");
    ctx->lists_ready = 2;
    return NB_OK;
}

extern "C" int32_t nb_lists_stats(const nb_ctx* ctx, int64_t* out4)
{
    if (!ctx || !out4) return NB_ERR_ARG;
    out4[0] = ctx->lists_total;
    out4[1] = ctx->lists_maxrow;
    out4[2] = ctx->last_unsorted;
    out4[3] = ctx->last_grid_cells;
    return NB_OK;
}

template <bool kIncl>
static int32_t tighten_csr(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                           const double* box3, const int64_t* offset, const int32_t* index, const int64_t* rows,
                           int64_t nrows, int64_t* out_offset, int32_t* out_index, int64_t capacity, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || nrows < 0 || !total_out) return ctx_fail(ctx, "nb_tighten: bad arguments");
    const Box box = make_box(box3);
    if (ctx_reserve(ctx, (void**)&ctx->tcount, &ctx->tcount_bytes, sizeof(unsigned long long) * (n + 1)))
        return NB_ERR_HIP;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->tcount, n + 1, 0ull);
    if (nrows > 0)
        hipLaunchKernelGGL((k_tighten_csr<false, kIncl>), dim3(grid_of(nrows)), dim3(kBlock), 0, s, pos, h, box, offset,
                           index, rows, nrows, ctx->tcount, nullptr, nullptr);
    int rc = scan_u64(ctx, s, ctx->tcount, (unsigned long long*)out_offset, n + 1);
    if (rc) return rc;
    int64_t total = 0;
    rc = sync_copy(ctx, s, &total, out_offset + n, sizeof(total), "nb_tighten: total copy failed");
    if (rc) return rc;
    *total_out = total;
    if (total > capacity) return NB_RETRY_CAPACITY;
    if (nrows > 0)
        hipLaunchKernelGGL((k_tighten_csr<true, kIncl>), dim3(grid_of(nrows)), dim3(kBlock), 0, s, pos, h, box, offset,
                           index, rows, nrows, nullptr, (const unsigned long long*)out_offset, out_index);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_tighten: kernel launch failed");
    return NB_OK;
}

extern "C" int32_t nb_tighten(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                              const double* box3, const int64_t* offset, const int32_t* index, const int64_t* rows,
                              int64_t nrows, int64_t* out_offset, int32_t* out_index, int64_t capacity,
                              int64_t* total_out)
{
    return tighten_csr<false>(ctx, stream, pos, h, n, box3, offset, index, rows, nrows, out_offset, out_index,
                              capacity, total_out);
}

extern "C" int32_t nb_tighten_inclusive(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                        const double* box3, const int64_t* offset, const int32_t* index,
                                        const int64_t* rows, int64_t nrows, int64_t* out_offset, int32_t* out_index,
                                        int64_t capacity, int64_t* total_out)
{
    return tighten_csr<true>(ctx, stream, pos, h, n, box3, offset, index, rows, nrows, out_offset, out_index,
                             capacity, total_out);
}

extern "C" int32_t nb_tighten_blocked(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                      const double* box3, const int32_t* in_len, const int32_t* in_index,
                                      int64_t row_cap, int32_t* out_len, int32_t* out_index, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n < 0 || !total_out) return ctx_fail(ctx, "nb_tighten_blocked: bad arguments");
    *total_out = 0;
    if (n == 0) return NB_OK;
    if (reserve_reductions(ctx, n)) return NB_ERR_HIP;
    NB_TIMED(ctx, s, "k_tighten_blk", k_tighten_blk, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h, make_box(box3), in_len, in_index,
                       row_cap, n, out_len, out_index, ctx->part);
    finish_partials(ctx, s, grid_of(n), 1, 0u, 8);
    long long t = 0;
    int rc = sync_copy(ctx, s, &t, (long long*)ctx->scal + 8, sizeof(t), "nb_tighten_blocked: total copy failed");
    if (rc) return rc;
    *total_out = (int64_t)t;
    return NB_OK;
}

extern "C" int32_t nb_permute(nb_ctx* ctx, void* stream, const double* src, const int32_t* order, int64_t n,
                              int32_t width, double* dst)
{
    if (!ctx || n < 0 || width <= 0 || src == dst) return ctx_fail(ctx, "nb_permute: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_permute, dim3(grid_of(n * width)), dim3(kBlock), 0, (hipStream_t)stream, src, order, n,
                       (int)width, dst);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_permute: kernel launch failed");
    return NB_OK;
}

extern "C" int32_t nb_permute_many(nb_ctx* ctx, void* stream, int32_t count, const void* const* src, void* const* dst,
                                   const int32_t* width, const int32_t* order, int64_t n)
{
    if (!ctx || count < 0 || count > kMaxPermute || n < 0 || !order || (count && (!src || !dst || !width)))
        return ctx_fail(ctx, "nb_permute_many: bad arguments");
    PermSet P{};
    P.count = count;
    P.total = 0;
    for (int k = 0; k < count; k++) {
        if (!src[k] || !dst[k] || src[k] == dst[k] || width[k] <= 0) return ctx_fail(ctx, "nb_permute_many: bad array");
        P.src[k] = (const uint64_t*)src[k];
        P.dst[k] = (uint64_t*)dst[k];
        P.width[k] = width[k];
        P.total += width[k];
    }
    if (n == 0 || count == 0) return NB_OK;
    hipLaunchKernelGGL(k_permute_many, dim3(grid_of(n * P.total)), dim3(kBlock), 0, (hipStream_t)stream, P, order, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_permute_many: kernel launch failed");
    return NB_OK;
}

// nb_permute_many where some destinations are kicks of moved elements: src2[k] non-NULL makes
// dst[k] = src[k] + s src2[k] (doubles, s = scale * s_dev[0]), in the new order
extern "C" int32_t nb_permute_kick_many(nb_ctx* ctx, void* stream, int32_t count, const void* const* src,
                                        const void* const* src2, void* const* dst, const int32_t* width,
                                        const double* s_dev, double scale, const int32_t* order, int64_t n)
{
    if (!ctx || count < 0 || count > kMaxPermute || n < 0 || !order || !s_dev ||
        (count && (!src || !src2 || !dst || !width)))
        return ctx_fail(ctx, "nb_permute_kick_many: bad arguments");
    PermSet P{};
    P.count = count;
    P.total = 0;
    P.s_dev = s_dev;
    P.scale = scale;
    for (int k = 0; k < count; k++) {
        if (!src[k] || !dst[k] || src[k] == dst[k] || src2[k] == dst[k] || width[k] <= 0)
            return ctx_fail(ctx, "nb_permute_kick_many: bad array");
        P.src[k] = (const uint64_t*)src[k];
        P.src2[k] = (const uint64_t*)src2[k];
        P.dst[k] = (uint64_t*)dst[k];
        P.width[k] = width[k];
        P.total += width[k];
    }
    if (n == 0 || count == 0) return NB_OK;
    hipLaunchKernelGGL(k_permute_many, dim3(grid_of(n * P.total)), dim3(kBlock), 0, (hipStream_t)stream, P, order, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_permute_kick_many: kernel launch failed");
    return NB_OK;
}

extern "C" int32_t nb_permute_i64(nb_ctx* ctx, void* stream, const int64_t* src, const int32_t* order, int64_t n,
                                  int64_t* dst)
{
    if (!ctx || n < 0 || src == dst) return ctx_fail(ctx, "nb_permute_i64: bad arguments");
    if (n == 0) return NB_OK;
    hipLaunchKernelGGL(k_permute_i64, dim3(grid_of(n)), dim3(kBlock), 0, (hipStream_t)stream, src, order, n, dst);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_permute_i64: kernel launch failed");
    return NB_OK;
}


// ------------------------------------------------------------------ fused step passes (host)
// Exclusive offsets ctx->fb_* (nblk + 1 entries) of 64-row blocks sized by k_block_caps; their
// total (entries) in *total. Synchronises with the stream.
static int32_t block_offsets(nb_ctx* ctx, hipStream_t s, const int32_t* cnt, int cstride,
                             const unsigned long long* add, int64_t n, int mode, int64_t cap0,
                             unsigned long long* off, int64_t& total, const char* what)
{
    const int64_t nblk = (n + 63) / 64;
    const int64_t wg = (nblk + 1 + 3) / 4;  // four blocks (wavefronts) per workgroup
    hipLaunchKernelGGL(k_block_caps, dim3((unsigned)wg), dim3(kBlock), 0, s, cnt, cstride, add, n, mode, cap0,
                       ctx->fb_sz, nblk);
    int rc = scan_u64(ctx, s, ctx->fb_sz, off, nblk + 1);
    if (rc) return rc;
    unsigned long long t = 0;
    rc = sync_copy(ctx, s, &t, off + nblk, sizeof(t), what);
    total = (int64_t)t;
    return rc;
}

extern "C" int32_t nb_hydro_density(nb_ctx* ctx, void* stream, const double* pos, const double* mass, double m_uniform,
                                    const double* h_start, const double* h_search, int64_t n,
                                    const int32_t* row_hint, int32_t* row_counts, int32_t kid, double cnorm,
                                    double eta, double tol, double h_max, int32_t precision, double* h_out,
                                    double* rho, double* omega, int32_t* iters, int32_t* flags, int64_t* stats4)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n <= 0 || !stats4 || n >= (int64_t)kMask) return ctx_fail(ctx, "nb_hydro_density: bad arguments");
    const Box open{{0.0, 0.0, 0.0}};
    Grid g;
    int64_t ncells = 0;
    int rc = build_grid(ctx, s, pos, h_search, n, open, g, ncells, true);
    if (rc) return rc;
    const bool mixed = precision != 0;
    const bool masses = !(m_uniform > 0.0);
    const int64_t nblk = (n + 63) / 64;
    if (ctx_reserve(ctx, (void**)&ctx->rowcnt, &ctx->rowcnt_bytes, sizeof(int32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->unsorted, &ctx->unsorted_bytes, n) ||
        ctx_reserve(ctx, (void**)&ctx->pending, &ctx->pending_bytes, n) ||
        ctx_reserve(ctx, (void**)&ctx->ecount, &ctx->ecount_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->eoff, &ctx->eoff_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->efill, &ctx->efill_bytes, sizeof(unsigned int) * n) ||
        ctx_reserve(ctx, (void**)&ctx->tot, &ctx->tot_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->offs, &ctx->offs_bytes, sizeof(unsigned long long) * (n + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->fb_sz, &ctx->fb_sz_bytes, sizeof(unsigned long long) * (nblk + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->fb_go, &ctx->fb_go_bytes, sizeof(unsigned long long) * (nblk + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->fb_to, &ctx->fb_to_bytes, sizeof(unsigned long long) * (nblk + 1)))
        return NB_ERR_HIP;
    const int64_t nb = grid_of(n);
    long long sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    DensArgs A{mass, m_uniform, h_start, h_max, kid, cnorm, eta, tol, h_out, rho, omega, iters, flags, ctx->pending,
               ctx->part};
    // block capacities: from the hint (the previous pass's counts of the same particles), else a
    // uniform first guess; a pass with rows over their block's capacity is repeated with exact ones
    int64_t cells = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
        const int mode = attempt ? 2 : (row_hint ? 1 : 0);
        rc = block_offsets(ctx, s, attempt ? ctx->rowcnt : row_hint, attempt ? 1 : 2, nullptr, n, mode, 192,
                           ctx->fb_go, cells, "nb_hydro_density: block capacities");
        if (rc) return rc;
        if (ctx_reserve(ctx, (void**)&ctx->gbuf, &ctx->gbuf_bytes, sizeof(uint32_t) * (size_t)cells) ||
            ctx_reserve(ctx, (void**)&ctx->dscratch, &ctx->dscratch_bytes,
                        sizeof(double) * (size_t)cells * (masses ? 2 : 1)))
            return NB_ERR_HIP;
        hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)grid_of(n + 1)), dim3(kBlock), 0, s, (long long*)ctx->ecount, 0ll,
                           n + 1);
        double* mbuf = masses ? ctx->dscratch + cells : nullptr;
        if (mixed)
            NB_TIMED(ctx, s, "k_gather_solve", k_gather_solve<true>, dim3(nb), dim3(kBlock), 0, s, pos, h_search, n, g,
                               ctx->cell_start, order_arg(ctx), ctx->gbuf, ctx->dscratch, mbuf, ctx->fb_go,
                               ctx->rowcnt, ctx->unsorted, ctx->ecount, A);
        else
            NB_TIMED(ctx, s, "k_gather_solve", k_gather_solve<false>, dim3(nb), dim3(kBlock), 0, s, pos, h_search, n, g,
                               ctx->cell_start, order_arg(ctx), ctx->gbuf, ctx->dscratch, mbuf, ctx->fb_go,
                               ctx->rowcnt, ctx->unsorted, ctx->ecount, A);
        // [0] longest gather row, [1] unsorted rows, [2] rows pending, [3] rows over capacity
        finish_partials(ctx, s, nb, 4, 1u, 0);
        rc = sync_copy(ctx, s, sc, ctx->scal, sizeof(long long) * 4, "nb_hydro_density: gather failed");
        if (rc) return rc;
        if (sc[3] == 0) break;
        if (attempt == 1) return ctx_fail(ctx, "nb_hydro_density: rows exceed exact capacities");
    }
    ctx->row_cap = (int32_t)sc[0];
    ctx->fb_gtotal = cells;
    ctx->last_unsorted = sc[1];
    double* mbuf = masses ? ctx->dscratch + cells : nullptr;
    if (sc[1] > 0)
        NB_TIMED(ctx, s, "k_sort_rows_payload", k_sort_rows_payload, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf,
                           ctx->dscratch, mbuf, ctx->fb_go, ctx->rowcnt, ctx->unsorted, n);
    if (sc[2] > 0) {
        if (mixed)
            NB_TIMED(ctx, s, "k_solve_pending", k_solve_pending<true>, dim3(grid_of(n)), dim3(kBlock), 0, s, h_search, n,
                               ctx->dscratch, mbuf, ctx->fb_go, ctx->rowcnt, A);
        else
            NB_TIMED(ctx, s, "k_solve_pending", k_solve_pending<false>, dim3(grid_of(n)), dim3(kBlock), 0, s, h_search, n,
                               ctx->dscratch, mbuf, ctx->fb_go, ctx->rowcnt, A);
    }
    // extras (with d2) for the tightened rows; the longest merged row bounds the tight rows
    NB_TIMED(ctx, s, "k_row_total", k_row_total, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->rowcnt, ctx->ecount, ctx->tot,
                       ctx->part, n, row_counts);
    finish_partials(ctx, s, grid_of(n + 1), 1, 1u, 3);  // [3] longest merged row
    rc = scan_u64(ctx, s, ctx->ecount, ctx->eoff, n + 1);
    if (rc) return rc;
    rc = scan_u64(ctx, s, ctx->tot, ctx->offs, n + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, s, (long long*)ctx->scal + 4, ctx->eoff + n, ctx->offs + n);
    // tight-row blocks: capacity of each the longest merged row of its rows
    const int64_t wg = (nblk + 1 + 3) / 4;
    hipLaunchKernelGGL(k_block_caps, dim3((unsigned)wg), dim3(kBlock), 0, s, ctx->rowcnt, 1, ctx->ecount, n, 2,
                       (int64_t)0, ctx->fb_sz, nblk);
    rc = scan_u64(ctx, s, ctx->fb_sz, ctx->fb_to, nblk + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, s, (long long*)ctx->scal + 6, ctx->fb_to + nblk,
                       ctx->fb_to + nblk);
    rc = sync_copy(ctx, s, sc, ctx->scal, sizeof(long long) * 7, "nb_hydro_density: totals copy failed");
    if (rc) return rc;
    const unsigned long long ne_total = (unsigned long long)sc[4];
    if (ctx_reserve(ctx, (void**)&ctx->ebuf, &ctx->ebuf_bytes, sizeof(uint32_t) * (ne_total + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->edbuf, &ctx->edbuf_bytes, sizeof(double) * (ne_total + 1)))
        return NB_ERR_HIP;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)grid_of(n)), dim3(kBlock), 0, s, (uint32_t*)ctx->efill, 0u, n);
    NB_TIMED(ctx, s, "k_extras_fill_d2", k_extras_fill_d2, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, ctx->dscratch,
                       ctx->fb_go, ctx->rowcnt, ctx->eoff, ctx->efill, ctx->ebuf, ctx->edbuf, n);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_hydro_density: kernel launch failed");
    ctx->lists_n = n;
    ctx->hydro_ready = 1;
    ctx->lists_maxrow = sc[3];
    ctx->fb_ttotal = sc[6];
    stats4[0] = sc[0];   // longest gather row
    stats4[1] = sc[1];   // rows gathered unsorted
    stats4[2] = sc[5];   // entries of the search lists (incl. self)
    stats4[3] = sc[6];   // entries of the tight rows' blocks (each block as long as its longest merged row)
    return NB_OK;
}

extern "C" int32_t nb_hydro_tight(nb_ctx* ctx, void* stream, const double* h_final, int64_t* out_off,
                                  int32_t* out_len, int32_t* out_index, int64_t index_cap, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->hydro_ready || !total_out) return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_hydro_tight: no density pass");
    if (index_cap < ctx->fb_ttotal) return ctx_fail(ctx, "nb_hydro_tight: index capacity below the tight rows' blocks");
    if (!out_off || !out_len || !out_index) return ctx_fail(ctx, "nb_hydro_tight: bad arguments");
    const int64_t n = ctx->lists_n;
    const int64_t nblk = (n + 63) / 64;
    if (hipMemcpyAsync(out_off, ctx->fb_to, sizeof(int64_t) * (nblk + 1), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "nb_hydro_tight: offset copy failed");
    NB_TIMED(ctx, s, "k_tight_emit", k_tight_emit, dim3(grid_of(n)), dim3(kBlock), 0, s, ctx->gbuf, ctx->dscratch,
                       ctx->fb_go, ctx->rowcnt, ctx->eoff, ctx->ebuf, ctx->edbuf, h_final, n, ctx->fb_to, out_len,
                       out_index, ctx->part);
    finish_partials(ctx, s, grid_of(n), 1, 0u, 8);
    long long t = 0;
    int rc = sync_copy(ctx, s, &t, (long long*)ctx->scal + 8, sizeof(t), "nb_hydro_tight: total copy failed");
    if (rc) return rc;
    *total_out = (int64_t)t;
    return NB_OK;
}

extern "C" int32_t nb_hydro_density_mixed(nb_ctx* ctx, void* stream, const double* pos, const double* mass,
                                          double m_uniform, const double* h_start, const double* h_search, int64_t n,
                                          const int32_t* cloud, const int64_t* seg_off, int32_t nseg,
                                          double cell_factor, const int32_t* row_hint, int32_t* row_counts,
                                          int32_t kid, double cnorm, double eta, double tol, double h_max,
                                          double* h_out, double* rho, double* omega, int32_t* iters, int32_t* flags,
                                          int64_t* stats4)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n <= 0 || !stats4 || n >= (int64_t)kMask || !(cell_factor > 0.0) || (cloud && !seg_off))
        return ctx_fail(ctx, "nb_hydro_density_mixed: bad arguments");
    const Box open{{0.0, 0.0, 0.0}};
    Grid g{};
    int64_t ncells = 0;
    const bool segmented = cloud != nullptr;
    int rc = segmented ? build_grid_seg(ctx, s, pos, h_search, n, cloud, seg_off, nseg, cell_factor, true, ncells)
                       : build_grid(ctx, s, pos, h_search, n, open, g, ncells, true, cell_factor);
    if (rc) return rc;
    const bool masses = !(m_uniform > 0.0);
    const size_t u64n = sizeof(unsigned long long) * (size_t)(n + 1);
    if (ctx_reserve(ctx, (void**)&ctx->rowcnt, &ctx->rowcnt_bytes, sizeof(int32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->pending, &ctx->pending_bytes, n) ||
        ctx_reserve(ctx, (void**)&ctx->posh, &ctx->posh_bytes, sizeof(double4) * n) ||
        ctx_reserve(ctx, (void**)&ctx->efill, &ctx->efill_bytes, sizeof(unsigned int) * n) ||
        ctx_reserve(ctx, (void**)&ctx->xlist, &ctx->xlist_bytes, sizeof(uint32_t) * (size_t)(2 * n + 2)) ||
        ctx_reserve(ctx, (void**)&ctx->rm_gc, &ctx->rm_gc_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_go, &ctx->rm_go_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_ec, &ctx->rm_ec_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_eo, &ctx->rm_eo_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_to, &ctx->rm_to_bytes, u64n))
        return NB_ERR_HIP;
    NB_TIMED(ctx, s, "k_pack_posh", k_pack_posh, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h_search, n,
             (double4*)ctx->posh);
    const int64_t cap_u = ctx->row_cap_rm > 0 ? ctx->row_cap_rm : 192;
    const int64_t ecap_u = ctx->ecap_rm > 0 ? ctx->ecap_rm : 32;
    long long sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    MixArgs A{mass, m_uniform, h_start, h_max, kid, cnorm, eta, tol, h_out, rho, omega, iters, flags, ctx->pending,
              ctx->part};
    // the slow rows' work list: the extras sort's list buffer, free until that sort
    const int tile = ctx->tile;  // workgroup size of the lane-group kernels (autotuned: 64, 128 or 256)
    const int64_t blocks = (n * kG + tile - 1) / tile;
    const GridSeg* segs = (const GridSeg*)ctx->segs;
    // capacities: the previous pass's counts (row_hint) or uniform; after an overflow, this pass's
    // exact counts, which the repeated pass cannot exceed
    int mode = row_hint ? 1 : 0;
    unsigned long long tot[2] = {0, 0};
    for (int attempt = 0;; attempt++) {
        hipLaunchKernelGGL(k_caps_mixed, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, row_hint, ctx->rowcnt, ctx->efill, n,
                           mode, cap_u, ecap_u, ctx->rm_gc, ctx->rm_ec);
        rc = scan_u64(ctx, s, ctx->rm_gc, ctx->rm_go, n + 1);
        if (rc) return rc;
        rc = scan_u64(ctx, s, ctx->rm_ec, ctx->rm_eo, n + 1);
        if (rc) return rc;
        if (hipMemcpyAsync(&tot[0], ctx->rm_go + n, sizeof(unsigned long long), hipMemcpyDeviceToHost, s) != hipSuccess)
            return ctx_fail(ctx, "nb_hydro_density_mixed: capacity copy failed");
        rc = sync_copy(ctx, s, &tot[1], ctx->rm_eo + n, sizeof(unsigned long long),
                       "nb_hydro_density_mixed: capacity copy failed");
        if (rc) return rc;
        const size_t cells = (size_t)tot[0], ecells = (size_t)tot[1];
        if (ctx_reserve(ctx, (void**)&ctx->gbuf, &ctx->gbuf_bytes, sizeof(uint32_t) * cells) ||
            ctx_reserve(ctx, (void**)&ctx->fscratch, &ctx->fscratch_bytes, sizeof(float) * (masses ? (size_t)n : 1)) ||
            ctx_reserve(ctx, (void**)&ctx->ebuf, &ctx->ebuf_bytes, sizeof(uint32_t) * ecells))
            return NB_ERR_HIP;
        if (hipMemsetAsync(ctx->efill, 0, sizeof(unsigned int) * n, s) != hipSuccess) return ctx_fail(ctx, "memset");
        float* rmass = masses ? ctx->fscratch : nullptr;
#define NB_GATHER_MIXED(M, S, K)                                                                                     \
    NB_TIMED(ctx, s, "k_gather_mixed", (k_gather_mixed<M, S, K>), dim3(blocks), dim3(tile), 0, s,                   \
             (const double4*)ctx->posh, n, g, segs, cloud, ctx->cell_start, order_arg(ctx), ctx->gbuf, rmass,       \
             ctx->rm_go, ctx->rowcnt, ctx->efill, ctx->ebuf, ctx->rm_eo, A)
        if (kid == 0) {
            if (masses && segmented) NB_GATHER_MIXED(true, true, 0);
            else if (masses) NB_GATHER_MIXED(true, false, 0);
            else if (segmented) NB_GATHER_MIXED(false, true, 0);
            else NB_GATHER_MIXED(false, false, 0);
        } else {
            if (masses && segmented) NB_GATHER_MIXED(true, true, 1);
            else if (masses) NB_GATHER_MIXED(true, false, 1);
            else if (segmented) NB_GATHER_MIXED(false, true, 1);
            else NB_GATHER_MIXED(false, false, 1);
        }
#undef NB_GATHER_MIXED
        // [0] longest gather row, [1] rows left slow, [2] extras overflow, [3] rows over capacity,
        // [4] longest extras segment
        finish_partials(ctx, s, blocks, 5, 1u | 4u | 16u, 0);
        rc = sync_copy(ctx, s, sc, ctx->scal, sizeof(long long) * 5, "nb_hydro_density_mixed: gather failed");
        if (rc) return rc;
        if (sc[2] == 0 && sc[3] == 0) break;
        if (attempt == 2) return ctx_fail(ctx, "row capacity did not settle");
        mode = 2;
    }
    // uniform capacities for a later pass without counts
    ctx->row_cap_rm = (int32_t)(sc[0] + sc[0] / 8 + 8);
    ctx->ecap_rm = sc[4] + sc[4] / 4 + 4;
    ctx->rm_gtotal = (int64_t)tot[0];
    ctx->rm_etotal = (int64_t)tot[1];
    const long long slow_rows = sc[1];
    // many slow rows: one lane group per row (the kernel's per_group rows), else one lane per row
    const int per_group = slow_rows * 16 >= n ? 1 : 0;
    const unsigned slow_wgs = per_group ? (unsigned)blocks : slow_grid(n, tile);
    float* rmass = masses ? ctx->fscratch : nullptr;
    if (slow_rows > 0) {
        if (masses)
            NB_TIMED(ctx, s, "k_solve_slow_mixed", k_solve_slow_mixed<true>, dim3(slow_wgs), dim3(tile), 0, s, h_search, n,
                     (const double4*)ctx->posh, ctx->gbuf, rmass, ctx->rm_go, ctx->rowcnt, A,
                     (const long long*)ctx->scal + 1, per_group);  // scal[1]: the partials' rows left slow
        else
            NB_TIMED(ctx, s, "k_solve_slow_mixed", k_solve_slow_mixed<false>, dim3(slow_wgs), dim3(tile), 0, s, h_search,
                     n, (const double4*)ctx->posh, ctx->gbuf, rmass, ctx->rm_go, ctx->rowcnt, A,
                     (const long long*)ctx->scal + 1, per_group);  // scal[1]: the partials' rows left slow
    }
    if (ctx_reserve(ctx, (void**)&ctx->ebuf_s, &ctx->ebuf_s_bytes, sizeof(uint32_t) * (size_t)tot[1]))
        return NB_ERR_HIP;
    sort_extras(ctx, s, n, (unsigned)blocks, tile);
    // merged row lengths (reusing the gather capacities' array), their scan: the tight rows' offsets
    NB_TIMED(ctx, s, "k_row_total_rm", k_row_total_rm, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->rowcnt,
             ctx->efill, ctx->rm_go, ctx->rm_eo, ctx->rm_gc, row_counts, ctx->part, n);
    finish_partials(ctx, s, grid_of(n + 1), 2, 1u, 5);  // [5] longest merged row, [6] search-list entries
    rc = scan_u64(ctx, s, ctx->rm_gc, ctx->rm_to, n + 1);
    if (rc) return rc;
    rc = sync_copy(ctx, s, sc + 5, (long long*)ctx->scal + 5, sizeof(long long) * 2,
                   "nb_hydro_density_mixed: totals copy failed");
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_hydro_density_mixed: kernel launch failed");
    ctx->lists_n = n;
    ctx->lists_maxrow = sc[5];
    ctx->lists_total = sc[6];
    ctx->hydro_mixed_ready = 1;
    ctx->mix_kid = kid;
    ctx->mix_cnorm = cnorm;
    ctx->mix_m_uniform = m_uniform;
    stats4[0] = (int64_t)tot[0];  // gather-row entries allocated
    stats4[1] = slow_rows;        // rows not converged in one Newton step (solved in full from the stored row)
    stats4[2] = sc[6];            // search-list entries (incl. self): the capacity the tight rows need
    stats4[3] = sc[5];            // longest merged row
    return NB_OK;
}

extern "C" int32_t nb_hydro_tight_mixed(nb_ctx* ctx, void* stream, const double* pos, const double* h_final,
                                        double* rho, double* omega, int64_t* out_off, int32_t* out_len,
                                        int32_t* out_index, int64_t index_cap, int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->hydro_mixed_ready || !total_out)
        return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_hydro_tight_mixed: no density pass");
    if (!pos || !h_final || !rho || !omega || !out_off || !out_len || !out_index)
        return ctx_fail(ctx, "nb_hydro_tight_mixed: bad arguments");
    if (index_cap < ctx->lists_total)
        return ctx_fail(ctx, "nb_hydro_tight_mixed: index capacity below the search-list entries");
    const int64_t n = ctx->lists_n;
    if (ctx->posh_bytes < sizeof(double4) * (size_t)n) return ctx_fail(ctx, "nb_hydro_tight_mixed: records not reserved");
    const bool masses = !(ctx->mix_m_uniform > 0.0);
    const FinArgs F{masses ? ctx->fscratch : nullptr, ctx->mix_m_uniform, ctx->mix_kid,
                    ctx->mix_cnorm, ctx->pending, rho, omega};
    const int tile = ctx->tile;
    const unsigned eblocks = (unsigned)((n * kEmitW + tile - 1) / tile);
    if (ctx->part_bytes < sizeof(long long) * (size_t)eblocks)
        return ctx_fail(ctx, "nb_hydro_tight_mixed: reduction buffer smaller than the launch");
    // (x, y, z, h_final) records: the emission reads one per row entry
    NB_TIMED(ctx, s, "k_pack_posh", k_pack_posh, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h_final, n,
             (double4*)ctx->posh);
    NB_TIMED(ctx, s, "k_tight_emit_rm", k_tight_emit_rm<kEmitW>, dim3(eblocks), dim3(tile), 0, s,
             (const double4*)ctx->posh, ctx->gbuf, ctx->rm_go, ctx->rowcnt, ctx->efill, ctx->rm_eo, ctx->ebuf_s, n, ctx->rm_to, out_off, out_len,
             out_index, (unsigned long long)index_cap, (int32_t*)nullptr, F, ctx->part);
    finish_partials(ctx, s, eblocks, 1, 0u, 8);
    long long t = 0;
    int rc = sync_copy(ctx, s, &t, (long long*)ctx->scal + 8, sizeof(t), "nb_hydro_tight_mixed: total copy failed");
    if (rc) return rc;
    *total_out = (int64_t)t;
    return NB_OK;
}

// The FP64 step's density pass on the lane-group lists (k_gather_mixed over the FP64 grid,
// extras sorted, k_solve64 on every row), with nb_hydro_density's outputs, statistics and tight
// blocks (nb_hydro_tight64). The rows come out in ascending index order only when the particles
// are stored in this grid's order (the step's own permutation puts them so); otherwise nothing is
// computed and NB_NOT_GRID_ORDER is returned, for the caller to take nb_hydro_density.
extern "C" int32_t nb_hydro_density64(nb_ctx* ctx, void* stream, const double* pos, const double* mass,
                                      double m_uniform, const double* h_start, const double* h_search, int64_t n,
                                      const int32_t* row_hint, int32_t* row_counts, int32_t kid, double cnorm,
                                      double eta, double tol, double h_max, double* h_out, double* rho,
                                      double* omega, int32_t* iters, int32_t* flags, int64_t* stats4)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || n <= 0 || !stats4 || n >= (int64_t)kMask) return ctx_fail(ctx, "nb_hydro_density64: bad arguments");
    const Box open{{0.0, 0.0, 0.0}};
    Grid g{};
    int64_t ncells = 0;
    int rc = build_grid(ctx, s, pos, h_search, n, open, g, ncells, true);
    if (rc) return rc;
    if (!ctx->order_ident) return NB_NOT_GRID_ORDER;
    const bool masses = !(m_uniform > 0.0);
    const int64_t nblk = (n + 63) / 64;
    const size_t u64n = sizeof(unsigned long long) * (size_t)(n + 1);
    if (ctx_reserve(ctx, (void**)&ctx->rowcnt, &ctx->rowcnt_bytes, sizeof(int32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->pending, &ctx->pending_bytes, n) ||
        ctx_reserve(ctx, (void**)&ctx->posh, &ctx->posh_bytes, sizeof(double4) * n) ||
        ctx_reserve(ctx, (void**)&ctx->efill, &ctx->efill_bytes, sizeof(unsigned int) * n) ||
        ctx_reserve(ctx, (void**)&ctx->xlist, &ctx->xlist_bytes, sizeof(uint32_t) * (size_t)(2 * n + 2)) ||
        ctx_reserve(ctx, (void**)&ctx->rm_gc, &ctx->rm_gc_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_go, &ctx->rm_go_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_ec, &ctx->rm_ec_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_eo, &ctx->rm_eo_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->fb_sz, &ctx->fb_sz_bytes, sizeof(unsigned long long) * (nblk + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->fb_to, &ctx->fb_to_bytes, sizeof(unsigned long long) * (nblk + 1)))
        return NB_ERR_HIP;
    NB_TIMED(ctx, s, "k_pack_posh", k_pack_posh, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h_search, n,
             (double4*)ctx->posh);
    const int64_t cap_u = ctx->row_cap_rm > 0 ? ctx->row_cap_rm : 192;
    const int64_t ecap_u = ctx->ecap_rm > 0 ? ctx->ecap_rm : 32;
    long long sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // the gather builds the lists only (k_solve64 solves every row); its arguments carry the
    // partials and the solve's inputs
    MixArgs M{mass, m_uniform, h_start, h_max, kid, cnorm, eta, tol, h_out, rho, omega, iters, flags, ctx->pending,
              ctx->part};
    const int tile = 256;
    const int64_t blocks = (n * kG + tile - 1) / tile;
    int mode = row_hint ? 1 : 0;
    unsigned long long tot[2] = {0, 0};
    for (int attempt = 0;; attempt++) {
        hipLaunchKernelGGL(k_caps_mixed, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, row_hint, ctx->rowcnt, ctx->efill, n,
                           mode, cap_u, ecap_u, ctx->rm_gc, ctx->rm_ec);
        rc = scan_u64(ctx, s, ctx->rm_gc, ctx->rm_go, n + 1);
        if (rc) return rc;
        rc = scan_u64(ctx, s, ctx->rm_ec, ctx->rm_eo, n + 1);
        if (rc) return rc;
        if (hipMemcpyAsync(&tot[0], ctx->rm_go + n, sizeof(unsigned long long), hipMemcpyDeviceToHost, s) != hipSuccess)
            return ctx_fail(ctx, "nb_hydro_density64: capacity copy failed");
        rc = sync_copy(ctx, s, &tot[1], ctx->rm_eo + n, sizeof(unsigned long long),
                       "nb_hydro_density64: capacity copy failed");
        if (rc) return rc;
        if (ctx_reserve(ctx, (void**)&ctx->gbuf, &ctx->gbuf_bytes, sizeof(uint32_t) * (size_t)tot[0]) ||
            ctx_reserve(ctx, (void**)&ctx->ebuf, &ctx->ebuf_bytes, sizeof(uint32_t) * (size_t)tot[1]))
            return NB_ERR_HIP;
        if (hipMemsetAsync(ctx->efill, 0, sizeof(unsigned int) * n, s) != hipSuccess) return ctx_fail(ctx, "memset");
        NB_TIMED(ctx, s, "k_gather_mixed", (k_gather_mixed<false, false, 0, false>), dim3(blocks), dim3(tile), 0, s,
                 (const double4*)ctx->posh, n, g, (const GridSeg*)nullptr, (const int32_t*)nullptr, ctx->cell_start,
                 (const uint32_t*)nullptr, ctx->gbuf, (float*)nullptr, ctx->rm_go, ctx->rowcnt, ctx->efill,
                 ctx->ebuf, ctx->rm_eo, M);
        finish_partials(ctx, s, blocks, 5, 1u | 4u | 16u, 0);
        rc = sync_copy(ctx, s, sc, ctx->scal, sizeof(long long) * 5, "nb_hydro_density64: gather failed");
        if (rc) return rc;
        if (sc[2] == 0 && sc[3] == 0) break;
        if (attempt == 2) return ctx_fail(ctx, "nb_hydro_density64: row capacity did not settle");
        mode = 2;
    }
    ctx->row_cap_rm = (int32_t)(sc[0] + sc[0] / 8 + 8);
    ctx->ecap_rm = sc[4] + sc[4] / 4 + 4;
    const long long longest = sc[0];
    if (ctx_reserve(ctx, (void**)&ctx->ebuf_s, &ctx->ebuf_s_bytes, sizeof(uint32_t) * (size_t)tot[1]))
        return NB_ERR_HIP;
    sort_extras(ctx, s, n, (unsigned)blocks, tile);
    DensArgs A{mass, m_uniform, h_start, h_max, kid, cnorm, eta, tol, h_out, rho, omega, iters, flags, ctx->pending,
               ctx->part};
#define NB_SOLVE64(M_, K_)                                                                                      \
    NB_TIMED(ctx, s, "k_solve64", (k_solve64<M_, K_>), dim3((unsigned)blocks), dim3(tile), 0, s,                   \
             (const double4*)ctx->posh, ctx->gbuf, ctx->rm_go, ctx->rowcnt, n, A)
    if (masses && kid == 0) NB_SOLVE64(true, 0);
    else if (masses) NB_SOLVE64(true, 1);
    else if (kid == 0) NB_SOLVE64(false, 0);
    else NB_SOLVE64(false, 1);
#undef NB_SOLVE64
    // merged row lengths (the search-list entries) and the tight blocks' capacities
    NB_TIMED(ctx, s, "k_row_total_rm", k_row_total_rm, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->rowcnt,
             ctx->efill, ctx->rm_go, ctx->rm_eo, ctx->rm_gc, row_counts, ctx->part, n);
    finish_partials(ctx, s, grid_of(n + 1), 2, 1u, 5);  // [5] longest merged row, [6] search-list entries
    const int64_t wg = (nblk + 1 + 3) / 4;
    hipLaunchKernelGGL(k_block_caps64, dim3((unsigned)wg), dim3(kBlock), 0, s, ctx->rowcnt, ctx->efill, ctx->rm_go,
                       ctx->rm_eo, n, ctx->fb_sz, nblk);
    rc = scan_u64(ctx, s, ctx->fb_sz, ctx->fb_to, nblk + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(64), 0, s, (long long*)ctx->scal + 7, ctx->fb_to + nblk,
                       ctx->fb_to + nblk);
    rc = sync_copy(ctx, s, sc + 5, (long long*)ctx->scal + 5, sizeof(long long) * 3,
                   "nb_hydro_density64: totals copy failed");
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_hydro_density64: kernel launch failed");
    ctx->lists_n = n;
    ctx->lists_maxrow = sc[5];
    ctx->lists_total = sc[6];
    ctx->fb_ttotal = sc[7];
    ctx->hydro64_ready = 1;
    stats4[0] = longest;  // longest gather row
    stats4[1] = 0;        // rows gathered unsorted (none: the particles are in grid order)
    stats4[2] = sc[6];    // entries of the search lists (incl. self)
    stats4[3] = sc[7];    // entries of the tight rows' blocks (each block as long as its longest merged row)
    return NB_OK;
}

// The tight rows of the last nb_hydro_density64 pass at the final h (nb_hydro_tight's outputs:
// rows in 64-row blocks, block starts in out_off, lengths in out_len), from the positions.
extern "C" int32_t nb_hydro_tight64(nb_ctx* ctx, void* stream, const double* pos, const double* h_final,
                                    int64_t* out_off, int32_t* out_len, int32_t* out_index, int64_t index_cap,
                                    int64_t* total_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->hydro64_ready || !total_out)
        return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_hydro_tight64: no nb_hydro_density64 pass");
    if (index_cap < ctx->fb_ttotal) return ctx_fail(ctx, "nb_hydro_tight64: index capacity below the tight rows' blocks");
    if (!pos || !h_final || !out_off || !out_len || !out_index) return ctx_fail(ctx, "nb_hydro_tight64: bad arguments");
    const int64_t n = ctx->lists_n;
    const int64_t nblk = (n + 63) / 64;
    if (hipMemcpyAsync(out_off, ctx->fb_to, sizeof(int64_t) * (nblk + 1), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ctx_fail(ctx, "nb_hydro_tight64: offset copy failed");
    NB_TIMED(ctx, s, "k_pack_posh", k_pack_posh, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h_final, n,
             (double4*)ctx->posh);
    const unsigned eblocks = (unsigned)((n * kG + 255) / 256);
    if (ctx->part_bytes < sizeof(long long) * (size_t)eblocks)
        return ctx_fail(ctx, "nb_hydro_tight64: reduction buffer smaller than the launch");
    NB_TIMED(ctx, s, "k_emit64", k_emit64, dim3(eblocks), dim3(256), 0, s, (const double4*)ctx->posh, ctx->gbuf,
             ctx->rm_go, ctx->rowcnt, ctx->efill, ctx->rm_eo, ctx->ebuf_s, n, ctx->fb_to, out_len, out_index,
             ctx->part);
    finish_partials(ctx, s, eblocks, 1, 0u, 8);
    long long t = 0;
    int rc = sync_copy(ctx, s, &t, (long long*)ctx->scal + 8, sizeof(t), "nb_hydro_tight64: total copy failed");
    if (rc) return rc;
    *total_out = (int64_t)t;
    return NB_OK;
}

// ===================================================================== fixed-capacity passes
// The mixed step's grid and density passes with capacities fixed beforehand (nb_fixed_reserve):
// no host synchronisation, no allocation and no memset, so that a caller can record a whole step
// in a hipGraph and replay it. What does not fit is reported in the int32 status words (see
// kStGather ...); the caller then redoes the step eagerly.

namespace {

int fixed_check(nb_ctx* ctx, int64_t n, int32_t nseg, const char* what)
{
    if (!ctx->fx_ready || n > ctx->fx_n || nseg > ctx->fx_nseg || n <= 0 || nseg < 1)
        return ctx_fail(ctx, what);
    return NB_OK;
}

// scratch of the fixed passes' scans and sort histograms, in bytes
size_t fixed_tmp_bytes(int64_t n, int64_t cells)
{
    const int64_t m = std::max(std::max(n + 1, cells + 1), (int64_t)kRadix * ((n + kSortThreads - 1) / kSortThreads));
    const int64_t tiles = (m + kScanTile - 1) / kScanTile;
    return sizeof(unsigned long long) * (size_t)(tiles + 1) + 2 * sizeof(uint32_t) * (size_t)(m + 1) + 256;
}

// exclusive scan with kernels only (count entries of in into out; in may not alias out)
// The fixed-capacity passes run either eagerly (the single-synchronisation driver) or while a
// stream is being recorded into a graph. Only the recording needs the kernel-only sort and scan:
// the library's record memsets, which have misbehaved on replays.
static bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

template <class T>
int scan_fixed(nb_ctx* ctx, hipStream_t s, const T* in, T* out, int64_t count)
{
    if (!capturing(s)) {
        // not being recorded: the library scan (fewer launches; its memsets are harmless here)
        size_t bytes = 0;
        if (excl_sum(nullptr, bytes, in, out, (int64_t)(count), s) != hipSuccess)
            return ctx_fail(ctx, "scan sizing failed");
        if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
        if (excl_sum(ctx->cub_tmp, bytes, in, out, (int64_t)(count), s) != hipSuccess)
            return ctx_fail(ctx, "scan failed");
        return NB_OK;
    }
    const int64_t tiles = (count + kScanTile - 1) / kScanTile;
    if (sizeof(T) * (size_t)(tiles + 1) > ctx->fx_tmp_bytes) return ctx_fail(ctx, "fixed pass: scan scratch not reserved");
    T* sums = (T*)ctx->fx_tmp;
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, count, sums);
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(kScanThreads), 0, s, sums, tiles);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, count, sums, out);
    return NB_OK;
}

// the gather and extras offsets of the fixed density pass: both scans and their clamps in three
// launches (against two library scans and two clamp passes)
int scan2_fixed(nb_ctx* ctx, hipStream_t s, const Scan2& S, int64_t count, int32_t* status)
{
    const int64_t tiles = (count + kScanTile - 1) / kScanTile;
    if (sizeof(unsigned long long) * (size_t)(2 * tiles) > ctx->fx_tmp_bytes)
        return ctx_fail(ctx, "fixed pass: scan scratch not reserved");
    unsigned long long* sums = (unsigned long long*)ctx->fx_tmp;
    hipLaunchKernelGGL(k_scan2_tiles, dim3((unsigned)tiles, 2), dim3(kScanThreads), 0, s, S, count, sums, tiles);
    hipLaunchKernelGGL(k_scan2_sums, dim3(2), dim3(kScanThreads), 0, s, sums, tiles);
    hipLaunchKernelGGL(k_scan2_apply, dim3((unsigned)tiles, 2), dim3(kScanThreads), 0, s, S, count,
                       (const unsigned long long*)sums, tiles, status);
    return NB_OK;
}

// stable LSD radix sort of (ctx->cell_id, ctx->iota) by the key's low fx_bits bits, values into
// vout (ctx->order, or the caller's order array; keys and ping-pong buffers clobbered);
// *sorted_keys: where the sorted keys are
int sort_fixed(nb_ctx* ctx, hipStream_t s, int64_t n, const uint32_t** sorted_keys, uint32_t* vout)
{
    if (!capturing(s)) {
        // not being recorded: the library's (also stable) radix sort over the same bits
        size_t bytes = 0;
        if (sort_pairs(nullptr, bytes, ctx->cell_id, ctx->keys_out, ctx->iota, ctx->order, n, ctx->fx_bits, s) != hipSuccess)
            return ctx_fail(ctx, "sort sizing failed");
        if (ctx_reserve(ctx, (void**)&ctx->cub_tmp, &ctx->cub_tmp_bytes, bytes)) return NB_ERR_HIP;
        if (sort_pairs(ctx->cub_tmp, bytes, ctx->cell_id, ctx->keys_out, ctx->iota, vout, n, ctx->fx_bits, s) != hipSuccess)
            return ctx_fail(ctx, "sort failed");
        *sorted_keys = ctx->keys_out;
        return NB_OK;
    }
    const int64_t nblocks = (n + kSortThreads - 1) / kSortThreads;
    const int64_t hn = (int64_t)kRadix * nblocks;
    const int64_t tiles = (hn + kScanTile - 1) / kScanTile;
    const size_t need = sizeof(uint32_t) * (size_t)(tiles + 1) + 64 + 2 * sizeof(uint32_t) * (size_t)(hn + 1);
    if (need > ctx->fx_tmp_bytes) return ctx_fail(ctx, "fixed pass: sort scratch not reserved");
    uint32_t* hist = (uint32_t*)((char*)ctx->fx_tmp + sizeof(uint32_t) * (size_t)(tiles + 1) + 64);
    uint32_t* offs = hist + hn + 1;
    uint32_t *ka = ctx->cell_id, *kb = ctx->keys_out, *va = ctx->iota, *vb = ctx->order;
    const int passes = (ctx->fx_bits + kRadixBits - 1) / kRadixBits;
    for (int p = 0; p < passes; p++) {
        const int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nblocks), dim3(kSortThreads), 0, s, ka, n, shift, hist);
        // scan_fixed's partials live at the start of fx_tmp, before hist
        const int64_t t = (hn + kScanTile - 1) / kScanTile;
        uint32_t* sums = (uint32_t*)ctx->fx_tmp;
        hipLaunchKernelGGL(k_scan_tiles<uint32_t>, dim3((unsigned)t), dim3(kScanThreads), 0, s, hist, hn, sums);
        hipLaunchKernelGGL(k_scan_sums<uint32_t>, dim3(1), dim3(kScanThreads), 0, s, sums, t);
        hipLaunchKernelGGL(k_scan_apply<uint32_t>, dim3((unsigned)t), dim3(kScanThreads), 0, s, hist, hn, sums, offs);
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nblocks), dim3(kSortThreads), 0, s, ka, va, n, shift, offs,
                           kb, vb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (va != vout)  // the passes end in the iota or the order buffer
        hipLaunchKernelGGL(k_copy_u32, dim3(grid_of(n)), dim3(kBlock), 0, s, va, vout, n);
    *sorted_keys = ka;
    return NB_OK;
}

// Cell starts from the sorted cell ids of a stable sort: start[c] = the first sorted position
// whose id is >= c (the exclusive scan of the cells' counts), for c = 0 .. the grid's cell count
// (no further: later entries of the ncap + 1 are never read). Thread k of 0 .. n owns the run of
// cells keys[k - 1] < c <= keys[k] (thread n: up to the cell count) and k is written
// to each: runs of up to kStartLane cells by their thread, up to kStartRun by the whole
// wavefront (one run at a time), longer ones as chunks of kStartChunk cells through a work list
// that k_starts_fill hands out to whole workgroups. work holds 3 words per chunk; a run of
// L > kStartRun cells makes ceil(L / kStartChunk) chunks, so ncap + 1 cells make at most
// (ncap + 1) (1 / kStartChunk + 1 / (kStartRun + 1)) of them.
static constexpr int kStartLane = 8;
static constexpr int kStartRun = 256;
static constexpr int kStartChunk = 4096;

__global__ void __launch_bounds__(kBlock)
k_starts_from_keys(const uint32_t* __restrict__ keys, int64_t n, int64_t ncap, const GridSeg* __restrict__ segs,
                   int nseg, uint32_t* __restrict__ start, uint32_t* __restrict__ work,
                   unsigned int* __restrict__ nwork, int64_t work_cap)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t lo = 0, hi = -1;  // the thread's run (empty past n)
    if (k <= n) {
        lo = k == 0 ? 0 : (int64_t)keys[k - 1] + 1;
        int64_t top = ncap;
        if (k == n) {  // the last run ends at the grid's cell count
            const GridSeg& last = segs[nseg - 1];
            const int64_t cells = (int64_t)last.base + (int64_t)last.g.nc[0] * last.g.nc[1] * last.g.nc[2];
            top = cells < ncap ? cells : ncap;
        }
        const int64_t hk = k == n ? top : (int64_t)keys[k];
        hi = hk < ncap ? hk : ncap;  // inclusive (ids are below ncap + 1)
    }
    const int64_t len = hi - lo + 1;
    if (len > 0 && len <= kStartLane)
        for (int64_t c = lo; c <= hi; c++) start[c] = (uint32_t)k;
    // every lane takes part (no early exit above): the loop bound is the same on all of them
    unsigned long long mask = __ballot(len > kStartLane && len <= kStartRun);
    const int lane = threadIdx.x & 63;
    while (mask) {
        const int src = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1ull;
        const uint32_t a = __shfl((uint32_t)lo, src, 64), b = __shfl((uint32_t)hi, src, 64);
        const uint32_t v = __shfl((uint32_t)k, src, 64);
        for (int64_t c = (int64_t)a + lane; c <= (int64_t)b; c += 64) start[c] = v;
    }
    // longer runs: their chunks' slots in the work list, one atomic per wavefront
    const int chunks = len > kStartRun ? (int)((len + kStartChunk - 1) / kStartChunk) : 0;
    int incl = chunks;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    const int wave_total = __shfl(incl, 63, 64);
    unsigned int base = 0;
    if (lane == 63 && wave_total > 0) base = atomicAdd(nwork, (unsigned int)wave_total);
    base = __shfl(base, 63, 64);
    if (chunks > 0) {
        const int64_t w = (int64_t)base + (incl - chunks);
        for (int64_t q = 0; q < chunks; q++) {
            const int64_t idx = w + q;
            if (idx >= work_cap) break;  // cannot happen for ids <= ncap (the bound above)
            const int64_t c0 = lo + q * kStartChunk;
            work[3 * idx] = (uint32_t)c0;
            work[3 * idx + 1] = (uint32_t)(c0 + kStartChunk - 1 < hi ? c0 + kStartChunk - 1 : hi);
            work[3 * idx + 2] = (uint32_t)k;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_starts_fill(uint32_t* __restrict__ start, const uint32_t* __restrict__ work, const unsigned int* __restrict__ nwork,
              int64_t work_cap)
{
    const int64_t nw = (int64_t)*nwork < work_cap ? (int64_t)*nwork : work_cap;
    for (int64_t item = blockIdx.x; item < nw; item += gridDim.x) {
        const int64_t a = work[3 * item], b = work[3 * item + 1];
        const uint32_t v = work[3 * item + 2];
        for (int64_t c = a + threadIdx.x; c <= b; c += blockDim.x) start[c] = v;
    }
}

// whether the work list of the starts of cells 0 .. ncap fits in ncap + 1 words
inline bool starts_fit(int64_t ncap)
{
    return (ncap + 1) / kStartChunk + (ncap + 1) / (kStartRun + 1) + 1 <= ncap / 3;
}

// the starts of cells 0 .. the grid's cell count (<= ncap) from n sorted ids; scratch: ncap + 1 words (the
// cells' count buffer, not needed by this path) for the work list and its counter
int starts_from_keys(nb_ctx* ctx, hipStream_t s, const uint32_t* keys, int64_t n, int64_t ncap, const GridSeg* segs,
                     int nseg, uint32_t* start, uint32_t* scratch)
{
    unsigned int* nwork = (unsigned int*)(scratch + ncap);  // the last word
    const int64_t work_cap = ncap / 3;
    if (!starts_fit(ncap)) return ctx_fail(ctx, "cell starts: work list does not fit");
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t*)nwork, 0u, (int64_t)1);
    hipLaunchKernelGGL(k_starts_from_keys, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, keys, n, ncap, segs, nseg, start,
                       scratch, nwork, work_cap);
    hipLaunchKernelGGL(k_starts_fill, dim3(1024), dim3(kBlock), 0, s, start, (const uint32_t*)scratch,
                       (const unsigned int*)nwork, work_cap);
    return NB_OK;
}

// build_grid_seg with the plan made on the device. sort: order the particles by cell (the grid
// order pass); else they must be stored in cell order already (the density pass after the step's
// permutation), which is checked (status[kStUnsorted]).
int build_grid_seg_fixed(nb_ctx* ctx, hipStream_t s, const double* pos, const double* h, int64_t n,
                         const int32_t* cloud, const int64_t* off_dev, int32_t nseg, double cell_factor, bool sort,
                         int32_t* status, uint32_t* order_out = nullptr)
{
    ctx->lists_ready = 0;
    ctx->hydro_ready = 0;
    ctx->hydro_mixed_ready = 0;
    ctx->hydro64_ready = 0;
    if (!sort && ctx->fx_grid_from_order) {
        // the order pass's grid, checked against the particles as they now lie (a failure is
        // reported as not in cell order: the caller redoes the step eagerly)
        ctx->fx_grid_from_order = 0;
        hipLaunchKernelGGL(k_check_grid_seg, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, n, (const GridSeg*)ctx->segs,
                           cloud, (const uint32_t*)ctx->cell_start, ctx->part);
        finish_partials(ctx, s, grid_of(n), 1, 0u, 0, status, SlotStatus{{(int)kStUnsorted, -1, -1, -1, -1}});
        ctx->order_ident = 1;
        return NB_OK;
    }
    ctx->fx_grid_from_order = 0;
    if (ctx->segpart_bytes < seg_part_bytes(n, nseg)) return ctx_fail(ctx, "fixed pass: box partials not reserved");
    launch_seg_bbox(ctx, s, pos, h, n, off_dev, nseg, ctx->red);
    hipLaunchKernelGGL(k_plan_seg, dim3(1), dim3(64), 0, s, ctx->red, off_dev, nseg, cell_factor, (double)ctx->fx_cells,
                       (double)kMaxAxisCells, kMaxTotalCells, (GridSeg*)ctx->segs, status);
    const int64_t nc1 = ctx->fx_cells + 1;
    // sorting: the starts from the sorted ids (no histogram, no scan over the cells)
    const bool from_keys = sort && starts_fit(nc1 - 1);
    if (!from_keys)
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<int64_t>(4096, (nc1 + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, ctx->cell_count, 0u, nc1);
    NB_TIMED(ctx, s, "k_cell_id", k_cell_id_seg, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, n, (const GridSeg*)ctx->segs,
             cloud, ctx->cell_id, ctx->iota, from_keys ? nullptr : ctx->cell_count, sort ? nullptr : ctx->part);
    if (sort) {
        const uint32_t* skeys = nullptr;
        int rc = sort_fixed(ctx, s, n, &skeys, order_out ? order_out : ctx->order);
        if (rc) return rc;
        ctx->order_ident = 0;
        ctx->fx_grid_from_order = 1;  // the density pass after the permutation may reuse this grid
        if (from_keys)
            return starts_from_keys(ctx, s, skeys, n, nc1 - 1, (const GridSeg*)ctx->segs, nseg, ctx->cell_start,
                                    ctx->cell_count);
    } else {
        finish_partials(ctx, s, grid_of(n), 1, 0u, 0, status, SlotStatus{{(int)kStUnsorted, -1, -1, -1, -1}});
        ctx->order_ident = 1;
    }
    return scan_fixed(ctx, s, (const uint32_t*)ctx->cell_count, ctx->cell_start, nc1);
}

}  // namespace

extern "C" int32_t nb_fixed_reserve(nb_ctx* ctx, int64_t n, int32_t nseg, int64_t gather_cap, int64_t extras_cap,
                                    int64_t index_cap, int64_t cells_cap, int32_t masses)
{
    if (!ctx || n <= 0 || n >= (int64_t)kMask || nseg < 1 || gather_cap <= 0 || extras_cap <= 0 || index_cap <= 0 ||
        cells_cap <= 0 || cells_cap >= 2147483647ll)
        return ctx_fail(ctx, "nb_fixed_reserve: bad arguments");
    ctx->fx_ready = 0;
    ctx->fx_grid_from_order = 0;
    const size_t u64n = sizeof(unsigned long long) * (size_t)(n + 1);
    if (reserve_reductions(ctx, n) ||
        ctx_reserve(ctx, (void**)&ctx->red, &ctx->red_bytes, sizeof(double) * 9 * (size_t)nseg + 8 * (size_t)(nseg + 1) + 4096) ||
        ctx_reserve(ctx, (void**)&ctx->segs, &ctx->segs_bytes, sizeof(GridSeg) * (size_t)nseg) ||
        ctx_reserve(ctx, &ctx->segpart, &ctx->segpart_bytes, seg_part_bytes(n, nseg)) ||
        ctx_reserve(ctx, (void**)&ctx->cell_id, &ctx->cell_id_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->order, &ctx->order_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->keys_out, &ctx->keys_out_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->iota, &ctx->iota_bytes, sizeof(uint32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->cell_count, &ctx->cell_count_bytes, sizeof(uint32_t) * (size_t)(cells_cap + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->cell_start, &ctx->cell_start_bytes, sizeof(uint32_t) * (size_t)(cells_cap + 1)) ||
        ctx_reserve(ctx, (void**)&ctx->rowcnt, &ctx->rowcnt_bytes, sizeof(int32_t) * n) ||
        ctx_reserve(ctx, (void**)&ctx->pending, &ctx->pending_bytes, n) ||
        ctx_reserve(ctx, (void**)&ctx->posh, &ctx->posh_bytes, sizeof(double4) * n) ||
        ctx_reserve(ctx, (void**)&ctx->efill, &ctx->efill_bytes, sizeof(unsigned int) * n) ||
        ctx_reserve(ctx, (void**)&ctx->xlist, &ctx->xlist_bytes, sizeof(uint32_t) * (size_t)(2 * n + 2)) ||
        ctx_reserve(ctx, (void**)&ctx->rm_gc, &ctx->rm_gc_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_go, &ctx->rm_go_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_ec, &ctx->rm_ec_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_eo, &ctx->rm_eo_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->rm_to, &ctx->rm_to_bytes, u64n) ||
        ctx_reserve(ctx, (void**)&ctx->gbuf, &ctx->gbuf_bytes, sizeof(uint32_t) * (size_t)gather_cap) ||
        ctx_reserve(ctx, (void**)&ctx->fscratch, &ctx->fscratch_bytes,
                    sizeof(float) * (masses ? (size_t)n : 1)) ||
        ctx_reserve(ctx, (void**)&ctx->ebuf, &ctx->ebuf_bytes, sizeof(uint32_t) * (size_t)extras_cap) ||
        ctx_reserve(ctx, (void**)&ctx->ebuf_s, &ctx->ebuf_s_bytes, sizeof(uint32_t) * (size_t)extras_cap) ||
        ctx_reserve(ctx, (void**)&ctx->pack, &ctx->pack_bytes, 48 * (size_t)n))
        return NB_ERR_HIP;
    int bits = 1;
    while ((int64_t(1) << bits) < cells_cap + 1 && bits < 32) bits++;
    if (ctx_reserve(ctx, (void**)&ctx->fx_tmp, &ctx->fx_tmp_bytes, fixed_tmp_bytes(n, cells_cap))) return NB_ERR_HIP;
    ctx->fx_n = n;
    ctx->fx_nseg = nseg;
    ctx->fx_gather = gather_cap;
    ctx->fx_extras = extras_cap;
    ctx->fx_index = index_cap;
    ctx->fx_cells = cells_cap;
    ctx->fx_bits = bits;
    ctx->fx_ready = 1;
    return NB_OK;
}

extern "C" int32_t nb_ctx_sizes(const nb_ctx* ctx, int64_t* out6)
{
    if (!ctx || !out6) return -1;
    out6[0] = ctx->rm_gtotal;        // gather-row entries allocated by the last mixed density pass
    out6[1] = ctx->rm_etotal;        // extras entries allocated by it
    out6[2] = ctx->lists_total;      // its search-list entries (the tight index capacity it needs)
    out6[3] = ctx->last_grid_cells;  // cells of the last grid
    out6[4] = ctx->lists_n;
    out6[5] = ctx->fx_ready;
    return NB_OK;
}

extern "C" int32_t nb_grid_order_seg_fixed(nb_ctx* ctx, void* stream, const double* pos, const double* h, int64_t n,
                                           const int32_t* cloud, const int64_t* seg_off_dev, int32_t nseg,
                                           double cell_factor, int32_t* status, int32_t* order_out)
{
    hipStream_t s = (hipStream_t)stream;
    if (!cloud || !seg_off_dev || !status || !order_out || !(cell_factor > 0.0))
        return ctx_fail(ctx, "nb_grid_order_seg_fixed: bad arguments");
    int rc = fixed_check(ctx, n, nseg, "nb_grid_order_seg_fixed: capacities not reserved for this size");
    if (rc) return rc;
    // the sort writes the order straight into the caller's array
    return build_grid_seg_fixed(ctx, s, pos, h, n, cloud, seg_off_dev, nseg, cell_factor, true, status,
                                (uint32_t*)order_out);
}

extern "C" int32_t nb_hydro_density_mixed_fixed(nb_ctx* ctx, void* stream, const double* pos, const double* mass,
                                                double m_uniform, const double* h_start, const double* h_search,
                                                int64_t n, const int32_t* cloud, const int64_t* seg_off_dev,
                                                int32_t nseg, double cell_factor, const int32_t* row_hint,
                                                int32_t* row_counts, int32_t kid, double cnorm, double eta,
                                                double tol, double h_max, double* h_out, double* rho, double* omega,
                                                int32_t* iters, int32_t* flags, int32_t* status)
{
    hipStream_t s = (hipStream_t)stream;
    if (!cloud || !seg_off_dev || !row_hint || !row_counts || !status || !(cell_factor > 0.0))
        return ctx_fail(ctx, "nb_hydro_density_mixed_fixed: bad arguments");
    int rc = fixed_check(ctx, n, nseg, "nb_hydro_density_mixed_fixed: capacities not reserved for this size");
    if (rc) return rc;
    const bool masses = !(m_uniform > 0.0);
    if (masses && ctx->fscratch_bytes < sizeof(float) * (size_t)n)
        return ctx_fail(ctx, "nb_hydro_density_mixed_fixed: particle masses not reserved");
    rc = build_grid_seg_fixed(ctx, s, pos, h_search, n, cloud, seg_off_dev, nseg, cell_factor, false, status);
    if (rc) return rc;
    NB_TIMED(ctx, s, "k_pack_posh", k_pack_posh, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h_search, n,
             (double4*)ctx->posh);
    hipLaunchKernelGGL(k_caps_mixed, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, row_hint, ctx->rowcnt, ctx->efill, n, 1,
                       (int64_t)0, (int64_t)0, ctx->rm_gc, ctx->rm_ec);
    {
        const Scan2 S{{(const unsigned long long*)ctx->rm_gc, (const unsigned long long*)ctx->rm_ec},
                      {ctx->rm_go, ctx->rm_eo},
                      {(unsigned long long)ctx->fx_gather, (unsigned long long)ctx->fx_extras},
                      {(int)kStGather, (int)kStExtras}};
        rc = scan2_fixed(ctx, s, S, n + 1, status);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n)), dim3(kBlock), 0, s, (uint32_t*)ctx->efill, 0u, n);
    MixArgs A{mass, m_uniform, h_start, h_max, kid, cnorm, eta, tol, h_out, rho, omega, iters, flags, ctx->pending,
              ctx->part};
    const int tile = ctx->tile;
    const int64_t blocks = (n * kG + tile - 1) / tile;
    const GridSeg* segs = (const GridSeg*)ctx->segs;
    float* rmass = masses ? ctx->fscratch : nullptr;
    const Grid g_none{};
#define NB_GATHER_FIXED(M, K)                                                                                        \
    NB_TIMED(ctx, s, "k_gather_mixed", (k_gather_mixed<M, true, K>), dim3(blocks), dim3(tile), 0, s,                \
             (const double4*)ctx->posh, n, g_none, segs, cloud, ctx->cell_start, (const uint32_t*)nullptr, ctx->gbuf, \
             rmass, ctx->rm_go, ctx->rowcnt, ctx->efill, ctx->ebuf, ctx->rm_eo, A)
    if (masses && kid == 0) NB_GATHER_FIXED(true, 0);
    else if (masses) NB_GATHER_FIXED(true, 1);
    else if (kid == 0) NB_GATHER_FIXED(false, 0);
    else NB_GATHER_FIXED(false, 1);
#undef NB_GATHER_FIXED
    // [2] extras overflow, [3] rows over capacity
    finish_partials(ctx, s, blocks, 5, 1u | 4u | 16u, 0, status,
                    SlotStatus{{-1, -1, (int)kStExtras, (int)kStGather, -1}});
    if (masses)
        NB_TIMED(ctx, s, "k_solve_slow_mixed", k_solve_slow_mixed<true>, dim3(slow_grid(n, tile)), dim3(tile), 0, s, h_search, n,
                 (const double4*)ctx->posh, ctx->gbuf, rmass, ctx->rm_go, ctx->rowcnt, A,
                 (const long long*)ctx->scal + 1, 0);  // scal[1]: the partials' rows left slow
    else
        NB_TIMED(ctx, s, "k_solve_slow_mixed", k_solve_slow_mixed<false>, dim3(slow_grid(n, tile)), dim3(tile), 0, s, h_search, n,
                 (const double4*)ctx->posh, ctx->gbuf, rmass, ctx->rm_go, ctx->rowcnt, A,
                 (const long long*)ctx->scal + 1, 0);  // scal[1]: the partials' rows left slow
    hipLaunchKernelGGL(k_status_grow, dim3(grid_of(n)), dim3(kBlock), 0, s, flags, n, status);
    sort_extras(ctx, s, n, (unsigned)blocks, tile);
    NB_TIMED(ctx, s, "k_row_total_rm", k_row_total_rm, dim3(grid_of(n + 1)), dim3(kBlock), 0, s, ctx->rowcnt,
             ctx->efill, ctx->rm_go, ctx->rm_eo, ctx->rm_gc, row_counts, ctx->part, n);
    rc = scan_fixed(ctx, s, (const unsigned long long*)ctx->rm_gc, ctx->rm_to, n + 1);
    if (rc) return rc;
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_hydro_density_mixed_fixed: kernel launch failed");
    ctx->lists_n = n;
    ctx->lists_total = 0;
    ctx->lists_maxrow = 0;
    ctx->rm_gtotal = ctx->fx_gather;
    ctx->hydro_mixed_ready = 1;
    ctx->mix_kid = kid;
    ctx->mix_cnorm = cnorm;
    ctx->mix_m_uniform = m_uniform;
    return NB_OK;
}

extern "C" int32_t nb_hydro_tight_mixed_fixed(nb_ctx* ctx, void* stream, const double* pos, const double* h_final,
                                              double* rho, double* omega, int64_t* out_off, int32_t* out_len,
                                              int32_t* out_index, int32_t* status)
{
    hipStream_t s = (hipStream_t)stream;
    if (!ctx || !ctx->hydro_mixed_ready || !ctx->fx_ready)
        return ctx_fail_code(ctx, NB_ERR_NOT_BUILT, "nb_hydro_tight_mixed_fixed: no fixed density pass");
    if (!pos || !h_final || !rho || !omega || !out_off || !out_len || !out_index || !status)
        return ctx_fail(ctx, "nb_hydro_tight_mixed_fixed: bad arguments");
    const int64_t n = ctx->lists_n;
    if (ctx->posh_bytes < sizeof(double4) * (size_t)n)
        return ctx_fail(ctx, "nb_hydro_tight_mixed_fixed: records not reserved");
    const bool masses = !(ctx->mix_m_uniform > 0.0);
    const FinArgs F{masses ? ctx->fscratch : nullptr, ctx->mix_m_uniform, ctx->mix_kid,
                    ctx->mix_cnorm, ctx->pending, rho, omega};
    const int tile = ctx->tile;
    const unsigned eblocks = (unsigned)((n * kEmitW + tile - 1) / tile);
    NB_TIMED(ctx, s, "k_pack_posh", k_pack_posh, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, h_final, n,
             (double4*)ctx->posh);
    NB_TIMED(ctx, s, "k_tight_emit_rm", k_tight_emit_rm<kEmitW>, dim3(eblocks), dim3(tile), 0, s,
             (const double4*)ctx->posh, ctx->gbuf, ctx->rm_go, ctx->rowcnt, ctx->efill, ctx->rm_eo, ctx->ebuf_s, n, ctx->rm_to, out_off, out_len,
             out_index, (unsigned long long)ctx->fx_index, status, F, ctx->part);
    if (hipGetLastError() != hipSuccess) return ctx_fail(ctx, "nb_hydro_tight_mixed_fixed: kernel launch failed");
    return NB_OK;
}

// Node counts of a recorded graph by type (hipGraphNodeType order, 16 entries): a caller checks
// that a captured step holds no memset nodes.
extern "C" int32_t nb_graph_node_counts(void* graph, int32_t* counts16)
{
    if (!graph || !counts16) return -1;
    for (int k = 0; k < 16; k++) counts16[k] = 0;
    size_t num = 0;
    if (hipGraphGetNodes((hipGraph_t)graph, nullptr, &num) != hipSuccess) return -1;
    std::vector<hipGraphNode_t> nodes(num);
    if (num && hipGraphGetNodes((hipGraph_t)graph, nodes.data(), &num) != hipSuccess) return -1;
    for (size_t k = 0; k < num; k++) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[k], &t) != hipSuccess) return -1;
        const int it = (int)t;
        counts16[it >= 0 && it < 16 ? it : 15]++;
    }
    return NB_OK;
}
