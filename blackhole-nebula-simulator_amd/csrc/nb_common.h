// Shared device helpers for the SPH kernels (gfx950).
//
// FP64 arithmetic in this library is written so that each product and sum is rounded on
// its own (the library is compiled with -ffp-contract=off): products are evaluated left to
// right and integer powers as x**3 = x*(x*x), x**4 = (x*x)*(x*x), x**5 = x*((x*x)*(x*x)),
// which is the order the reference's numba kernels use. Division and square root are the
// IEEE correctly rounded operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NB_R_KERN 2.0

namespace nb {

__device__ __forceinline__ double ipow3(double x) { return x * (x * x); }
__device__ __forceinline__ double ipow4(double x) { double x2 = x * x; return x2 * x2; }
// r^1.5 (r >= 0) rounded once, as a correctly rounded pow(r, 1.5) gives it: sqrt(r) = s + sl and
// r s = p + e carried exactly with FMA residuals, then p + (e + r sl) rounds the ~106-bit value.
__device__ __forceinline__ double pow15(double r)
{
    if (!(r > 0.0)) return r == 0.0 ? 0.0 : sqrt(r);  // 0, or NaN for r < 0 / NaN
    const double s = sqrt(r);
    const double sl = fma(-s, s, r) / (2.0 * s);
    const double p = r * s;
    const double e = fma(r, s, -p);
    return p + (e + r * sl);
}
__device__ __forceinline__ double ipow5(double x) { double x2 = x * x; return x * (x2 * x2); }

// Python's builtin min/max of two floats (second argument only if strictly smaller/larger)
__device__ __forceinline__ double pymin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double pymax(double a, double b) { return b > a ? b : a; }

// kernels_nb.py:16-28 (kid 0 cubic spline, kid 1 Wendland C4)
__device__ __forceinline__ double kernel_f(double q, int kid)
{
    if (q >= 2.0) return 0.0;
    if (kid == 0) {
        if (q < 1.0) return 1.0 - 1.5 * q * q + 0.75 * q * q * q;
        double t = 2.0 - q;
        return 0.25 * t * t * t;
    }
    double x = 1.0 - 0.5 * q;
    double x2 = x * x;
    return x2 * x2 * x2 * (35.0 / 12.0 * q * q + 3.0 * q + 1.0);
}

// kernels_nb.py:31-43
__device__ __forceinline__ double kernel_df(double q, int kid)
{
    if (q >= 2.0) return 0.0;
    if (kid == 0) {
        if (q < 1.0) return -3.0 * q + 2.25 * q * q;
        double t = 2.0 - q;
        return -0.75 * t * t;
    }
    double x = 1.0 - 0.5 * q;
    double x2 = x * x;
    return -7.0 / 3.0 * q * x2 * x2 * x * (5.0 * q + 2.0);
}

// sph.py:519-522: dW/dr = C / h^4 f'(r/h)
__device__ __forceinline__ double dwdr(double r, double h, int kid, double cnorm)
{
    return cnorm / ipow4(h) * kernel_df(r / h, kid);
}

// neighbours.py:39-43: minimum image of one separation component (box_c <= 0: open)
__device__ __forceinline__ double min_image(double d, double box_c)
{
    if (box_c > 0.0) d -= box_c * rint(d / box_c);
    return d;
}

// FP32 square root, reciprocal square root and reciprocal of the hardware (about 1 ulp): the
// pair terms of the mixed-precision passes carry FP32 rounding anyway, and the correctly rounded
// forms take several instructions each.
__device__ __forceinline__ float sqrt32(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsqrt32(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

struct Box {
    double l[3];  // periodic lengths, 0 = open axis
};

// Neighbour-list views of the pair passes. CSR rows (int64 offsets, the reference's layout),
// 64-row blocks: entry k of row i at ((i / 64) cap + k) 64 + i % 64 (or, for blocks with
// capacities of their own, at boff[i / 64] + 64 k + i % 64), so that the 64 lanes of a
// wavefront reading entry k of their own rows touch one contiguous 256-byte segment, or
// row-major rows (entry k of row i at i cap + k) for passes where a lane group shares a row.
struct CsrView {
    const int64_t* off;
    const int32_t* idx;
    static constexpr int64_t stride = 1;
    __device__ __forceinline__ int64_t count(int64_t i) const { return off[i + 1] - off[i]; }
    __device__ __forceinline__ int64_t base(int64_t i) const { return off[i]; }
    __device__ __forceinline__ int32_t at(int64_t i, int64_t k) const { return idx[off[i] + k]; }
};

struct BlkView {
    const int32_t* idx;
    const int32_t* len;
    int64_t cap;
    const int64_t* boff;  // non-null: block b starts at boff[b] (blocks with capacities of their own)
    static constexpr int64_t stride = 64;
    __device__ __forceinline__ int64_t count(int64_t i) const { return len[i]; }
    __device__ __forceinline__ int64_t base(int64_t i) const
    {
        return boff ? boff[i >> 6] + (i & 63) : ((i >> 6) * cap) * 64 + (i & 63);
    }
    __device__ __forceinline__ int32_t at(int64_t i, int64_t k) const { return idx[base(i) + k * 64]; }
};

// rows of their own capacity: row i has len[i] entries from off[i]
struct RowView {
    const int32_t* idx;
    const int32_t* len;
    const int64_t* off;
    static constexpr int64_t stride = 1;
    __device__ __forceinline__ int64_t count(int64_t i) const { return len[i]; }
    __device__ __forceinline__ int64_t base(int64_t i) const { return off[i]; }
    __device__ __forceinline__ int32_t at(int64_t i, int64_t k) const { return idx[off[i] + k]; }
};

// Wave-wide reductions (every lane of the wavefront must call them); the result is valid in
// all lanes. Used so that one lane per wavefront, not one per row, touches a global counter.
__device__ __forceinline__ int wave_max_i32(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroups of a launch are dispatched to the 8 XCDs round-robin (workgroup b on XCD b % 8),
// and each XCD has its own L2. xcd_tile() renumbers the workgroups so that each XCD works on one
// contiguous range of tiles: consecutive tiles hold neighbouring particles (grid order), whose
// neighbour data then stays in one L2 instead of being fetched by all eight. A bijection of
// [0, gridDim.x) for any grid size (1-D grids).
__device__ __forceinline__ int64_t xcd_tile()
{
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Block-wide sum or max of one value per thread, written by thread 0 to part[blockIdx.x] (a
// per-workgroup partial, combined afterwards by one small kernel: hot kernels make no
// same-address atomics). Every thread of the block must call it; blockDim.x a multiple of 64.
template <bool kMax>
__device__ __forceinline__ void block_partial(long long v, long long* part)
{
    __shared__ long long sh_part[16];
    for (int o = 32; o > 0; o >>= 1) {
        const long long t = __shfl_xor(v, o, 64);
        v = kMax ? (t > v ? t : v) : v + t;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long r = sh_part[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = kMax ? (sh_part[w] > r ? sh_part[w] : r) : r + sh_part[w];
        part[blockIdx.x] = r;
    }
}

// K per-thread counts reduced over the workgroup in one pass (one pair of barriers for all K):
// value k of workgroup b goes to part[k * gridDim.x + b], the maximum over the workgroup where
// bit k of max_mask is set, else the sum (the sums must fit in T). Every thread of the block must
// call it; blockDim.x a multiple of 64, at most 1024.
template <int K, class T = int32_t>
__device__ __forceinline__ void block_partials(T (&v)[K], unsigned max_mask, long long* part)
{
    __shared__ T sh_parts[16][K];
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const T t = __shfl_xor(v[k], o, 64);
            v[k] = (max_mask >> k) & 1u ? (t > v[k] ? t : v[k]) : v[k] + t;
        }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) sh_parts[threadIdx.x >> 6][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)K) {
        const int k = threadIdx.x;
        const bool mx = (max_mask >> k) & 1u;
        long long r = sh_parts[0][k];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            const long long x = sh_parts[w][k];
            r = mx ? (x > r ? x : r) : r + x;
        }
        part[(int64_t)k * gridDim.x + blockIdx.x] = r;
    }
}

// Particle rows of a pass: rows[q] if a row list is given, else q.
__device__ __forceinline__ int64_t row_of(const int64_t* rows, int64_t q)
{
    return rows ? rows[q] : q;
}

}  // namespace nb

// status codes of the C ABI (include/nebula_hip.h)
#define NB_OK 0
#define NB_RETRY_CAPACITY 1
#define NB_NOT_GRID_ORDER 2  // nb_hydro_density64: particles not stored in grid order (nothing computed)
#define NB_ERR_ARG -1
#define NB_ERR_HIP -2
#define NB_ERR_BOX -3
#define NB_ERR_NOT_BUILT -4
