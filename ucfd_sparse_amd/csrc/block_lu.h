// Register-resident LU factorisation and substitution of one small dense
// block (Doolittle, unit lower triangle, no pivoting), in the operation order
// of the reference's ludcmp / lusub (src/krylov/utils/inverse.c:7-71 and the
// identical src/lusgs/inverse.h:13-80), and the right division by a factored
// block of the block ILU(0) factorisation (lusubmattrans, inverse.c:77-107).
// A 1x1 "factor" is the reciprocal.  ludcmp returns false on a zero pivot
// (the reference does not check).
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

// M := M (L U)^-1 for a row-major block M, with LU as left by ludcmp: first
// Y U = M column by column (Y_{:,c} = (M_{:,c} - sum_{k<c} Y_{:,k} U_kc) / U_cc),
// then X L = Y from the last column back (unit diagonal).
template <int B>
__device__ __forceinline__ void lusubmattrans(const double *LU, double *M)
{
    if (B == 1) {
        M[0] *= LU[0];
        return;
    }
#pragma unroll
    for (int s = 0; s < B; ++s) M[s*B] /= LU[0];
#pragma unroll
    for (int c = 1; c < B; ++c)
#pragma unroll
        for (int s = 0; s < B; ++s) {
            double val = 0.0;
#pragma unroll
            for (int k = 0; k < c; ++k) val += M[s*B + k]*LU[k*B + c];
            M[s*B + c] = (M[s*B + c] - val)/LU[c*B + c];
        }
#pragma unroll
    for (int c = B - 2; c > -1; --c)
#pragma unroll
        for (int s = 0; s < B; ++s) {
            double val = 0.0;
#pragma unroll
            for (int k = c + 1; k < B; ++k) val += M[s*B + k]*LU[k*B + c];
            M[s*B + c] -= val;
        }
}

}  // namespace ucfd
