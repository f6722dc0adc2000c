// Register-resident LU factorisation and substitution of one small dense
// block (Doolittle, unit lower triangle, no pivoting), in the operation order
// of the reference's ludcmp / lusub (src/krylov/utils/inverse.c:7-71 and the
// identical src/lusgs/inverse.h:13-80).  A 1x1 "factor" is the reciprocal.
// ludcmp returns false on a zero pivot (the reference does not check).
#pragma once

#include <hip/hip_runtime.h>

namespace ucfd {

template <int B>
__device__ __forceinline__ bool ludcmp(double *A)
{
    if (B == 1) {
        if (A[0] == 0.0) return false;
        A[0] = 1.0/A[0];
        return true;
    }
    if (A[0] == 0.0) return false;
#pragma unroll
    for (int row = 1; row < B; ++row) {
        const int nrow = B*row;
        A[nrow] /= A[0];
#pragma unroll
        for (int col = 1; col < B; ++col) {
            if (row > col) {
                double val = 0.0;
#pragma unroll
                for (int k = 0; k < col; ++k) val += A[nrow + k]*A[col + B*k];
                A[nrow + col] = (A[nrow + col] - val)/A[(B + 1)*col];
            } else {
                double val = 0.0;
#pragma unroll
                for (int k = 0; k < row; ++k) val += A[nrow + k]*A[B*k + col];
                A[nrow + col] -= val;
            }
        }
        if (A[(B + 1)*row] == 0.0) return false;
    }
    return true;
}

template <int B>
__device__ __forceinline__ void lusub(const double *LU, double *b)
{
    if (B == 1) {
        b[0] *= LU[0];
        return;
    }
#pragma unroll
    for (int row = 1; row < B; ++row) {
        double val = 0.0;
#pragma unroll
        for (int col = 0; col < row; ++col) val += LU[row*B + col]*b[col];
        b[row] -= val;
    }
    b[B - 1] /= LU[B*B - 1];
#pragma unroll
    for (int row = B - 2; row > -1; --row) {
        double val = 0.0;
#pragma unroll
        for (int col = row + 1; col < B; ++col) val += LU[row*B + col]*b[col];
        b[row] = (b[row] - val)/LU[row*B + row];
    }
}

}  // namespace ucfd
