// 3x3 linear algebra of the gradient passes (device): Gaussian elimination with partial
// pivoting (first largest |pivot|, multipliers by division), determinant from the pivots,
// inverse by solving A X = I column by column, and the near-singular test of sph.py:776-782.
#pragma once

#include "nb_common.h"

namespace nb {

// Gaussian elimination with partial pivoting (first largest |pivot|), multipliers by
// division; determinant from the pivots; inverse by solving A X = I column by column.
__device__ __forceinline__ void lu3(const double* a_in, double* lu, int* perm, int& sign)
{
    for (int k = 0; k < 9; k++) lu[k] = a_in[k];
    for (int i = 0; i < 3; i++) perm[i] = i;
    sign = 1;
    for (int j = 0; j < 3; j++) {
        int p = j;
        double best = fabs(lu[3 * j + j]);
        for (int i = j + 1; i < 3; i++) {
            double v = fabs(lu[3 * i + j]);
            if (v > best) { best = v; p = i; }
        }
        if (p != j) {
            for (int k = 0; k < 3; k++) {
                double t = lu[3 * j + k]; lu[3 * j + k] = lu[3 * p + k]; lu[3 * p + k] = t;
            }
            int t = perm[j]; perm[j] = perm[p]; perm[p] = t;
            sign = -sign;
        }
        double piv = lu[3 * j + j];
        if (piv != 0.0) {
            for (int i = j + 1; i < 3; i++) {
                double l = lu[3 * i + j] / piv;
                lu[3 * i + j] = l;
                for (int k = j + 1; k < 3; k++) lu[3 * i + k] -= l * lu[3 * j + k];
            }
        }
    }
}

__device__ __forceinline__ double det3(const double* a)
{
    double lu[9]; int perm[3], sign;
    lu3(a, lu, perm, sign);
    double d = lu[0] * lu[4] * lu[8];
    return sign < 0 ? -d : d;
}

__device__ __forceinline__ void inv3(const double* a, double* out)
{
    double lu[9]; int perm[3], sign;
    lu3(a, lu, perm, sign);
    for (int col = 0; col < 3; col++) {
        double y[3];
        for (int i = 0; i < 3; i++) {
            double s = perm[i] == col ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) s -= lu[3 * i + k] * y[k];
            y[i] = s;
        }
        for (int i = 2; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < 3; k++) s -= lu[3 * i + k] * y[k];
            y[i] = s / lu[3 * i + i];
        }
        for (int i = 0; i < 3; i++) out[3 * i + col] = y[i];
    }
}

__device__ __forceinline__ void matmul3(const double* a, const double* b, double* c)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            c[3 * i + j] = s;
        }
}

// sph.py:776-782
__device__ __forceinline__ bool near_singular(const double* rmat, int64_t cnt)
{
    double scale = (rmat[0] + rmat[4] + rmat[8]) / 3.0;
    if (cnt < 4 || !(scale > 0.0)) return true;
    return !(fabs(det3(rmat)) >= 1e-6 * ipow3(scale));
}


}  // namespace nb
