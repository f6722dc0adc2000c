// Internal interface between the Krylov host code (krylov_api.cpp) and its
// kernels (krylov_kernels.hip).  Not installed.
#pragma once

#include <stdint.h>

#include "ucfd_internal.h"

// Deterministic reductions: a kernel writes one partial per block (and per
// reduced quantity), a second kernel sums them in a fixed order.
constexpr int KRY_BLOCK = 256;

// y = alpha*A*x + beta*y for a BSR matrix (row-major blk x blk blocks,
// vectors stored x[row*blk + r]); one thread per block row, the operation
// order of src/krylov/matrix/sparsemat.c:57-101.
UCFD_HIDDEN int kry_spmv_bsr(hipStream_t s, int bn, int blk, const int *rowptr, const int *colidx,
                             const double *values, double alpha, const double *x, double beta,
                             double *y, int *flag);
// Scalar CSR form (sparsemat.c:10-29).  Both raise UCFD_KFLAG_INDEX for a
// column index outside the matrix (the entry is skipped).
UCFD_HIDDEN int kry_spmv_csr(hipStream_t s, int n, const int *rowptr, const int *colidx,
                             const double *values, double alpha, const double *x, double beta,
                             double *y, int *flag);

// diagvalues[i] = LU(values[diagslots[i]]) (FP64, Doolittle without pivoting);
// raises UCFD_KFLAG_PIVOT on a zero pivot.  vals32 (nullable): FP32 copy of
// every block for the mixed-precision apply.
UCFD_HIDDEN int kry_bsgs_prepare(hipStream_t s, int bn, int blk, int64_t nnz, const int *diagslots,
                                 const double *values, double *diagvalues, float *vals32,
                                 int *flag);

// One level of the in-place ILU(0) factorisation (rows base..base+cnt-1, or
// order[base..]): scalar != 0 is ILU(0) of a CSR matrix (blk 1, ilu.c:4-48),
// else block ILU(0) (bilu.c:5-65).  diagvalues[i] gets row i's factored
// diagonal block (scalar: its pivot); raises UCFD_KFLAG_PIVOT on a zero pivot.
// rows != 0 (block form): no lower slot of a row refers to the row itself,
// which lets the lanes of a wave share a row's block operations.
UCFD_HIDDEN int kry_ilu_factor(hipStream_t s, int scalar, int rows, int blk, int base, int cnt,
                               const int *order, const int *rowptr, const int *colidx,
                               const int *diagslots, double *values, double *diagvalues,
                               int *flag);

// Sweep kinds of kry_bsgs_sweep64
enum { KRY_SW_LUSGS = 0, KRY_SW_BILU = 1, KRY_SW_ILU = 2 };

// One sweep over rows base..base+cnt-1 (or order[base..base+cnt-1] when
// order != NULL), in place on b.  KRY_SW_LUSGS: lower x_i = D_i^-1 (b_i -
// sum_{j<diag} A_ij x_j), upper x_i -= D_i^-1 sum_{j>diag} A_ij x_j.
// KRY_SW_BILU (factored values): lower x_i = b_i - sum_{j<diag} L_ij x_j,
// upper x_i = U_ii^-1 (x_i - sum_{j>diag} U_ij x_j) with diagvalues the
// factored U_ii.  KRY_SW_ILU: the scalar form (blk 1), dividing by the pivot.
// maxw: the most blocks the rows of one wave (64/blk consecutive rows) use
// in this direction (UcfdLevel::maxw; -1 = unknown): sizes the LDS staging.
UCFD_HIDDEN int kry_bsgs_sweep64(hipStream_t s, int kind, int upper, int blk, int base, int cnt,
                                 const int *order, const int *rowptr, const int *colidx,
                                 const int *diagslots, const double *values,
                                 const double *diagvalues, double *b, int maxw);
// Mixed form: block products in FP32 on FP32 values and an FP32 work vector
// x32, the diagonal solve in FP64.  lower reads the right-hand side b (FP64)
// and writes x32; upper updates x32 and writes the final value to b (FP64).
UCFD_HIDDEN int kry_bsgs_sweep32(hipStream_t s, int upper, int blk, int base, int cnt,
                                 const int *order, const int *rowptr, const int *colidx,
                                 const int *diagslots, const float *vals32,
                                 const double *diagvalues, float *x32, double *b, int maxw);

// BLAS-1/2 on device vectors.  Reductions write to out (device) in a fixed
// order; part must hold kry_parts(n) doubles per reduced quantity.
UCFD_HIDDEN int kry_parts(int64_t n);
UCFD_HIDDEN int kry_dot(hipStream_t s, int64_t n, const double *x, const double *y, double *part,
                        double *out);
UCFD_HIDDEN int kry_axpy(hipStream_t s, int64_t n, double a, const double *x, double *y);
UCFD_HIDDEN int kry_axpby(hipStream_t s, int64_t n, double a, const double *x, double b,
                          double *y);  // y = a x + b y
// dst = src * (1/sqrt(*normsq)); the same values also into dst2 when given
UCFD_HIDDEN int kry_scale_by_inv_sqrt(hipStream_t s, int64_t n, const double *src,
                                      const double *normsq, double *dst,
                                      double *dst2 = nullptr);
// h[j] = sum_i V[j*ldv + i] w[i], j < k (k <= 64): out[j] (device)
UCFD_HIDDEN int kry_gemvt(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv,
                          const double *w, double *part, double *out);
// w[i] += alpha * sum_j V[j*ldv + i] h[j], h on the device; with normsq != NULL
// also writes sum_i w_new[i]^2 to *normsq
UCFD_HIDDEN int kry_gemv(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv,
                         double alpha, const double *h, double *w, double *part,
                         double *normsq);
// kry_gemv (no norm) followed by kry_gemvt on the updated w, in one pass over
// V when k <= 32: w[i] += alpha * sum_j V[j*ldv + i] h[j]; out[j] = V_j . w
UCFD_HIDDEN int kry_gemv_gemvt(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv,
                               double alpha, const double *h, double *w, double *part,
                               double *out);
// BiCGSTAB fused reduction (bicgstab.c:5-22): out[0..4] = (rt.s, rt.t, t.s, t.t, s.s)
UCFD_HIDDEN int kry_dot5(hipStream_t s, int64_t n, const double *rt, const double *sv,
                         const double *t, double *part, double *out);
// p = r + beta*(p - omega*v); pt = p
UCFD_HIDDEN int kry_bicg_p(hipStream_t s, int64_t n, double beta, double omega, const double *r,
                           const double *v, double *p, double *pt);
// s = r - alpha*v; shat = s, with alpha = rho / *pi (pi on the device)
UCFD_HIDDEN int kry_bicg_s(hipStream_t s, int64_t n, double rho, const double *pi,
                           const double *r, const double *v, double *sv, double *shat);
// b = (float)a
UCFD_HIDDEN int kry_cast32(hipStream_t s, int64_t n, const double *a, float *b);
// x += alpha*pt + omega*shat; r = s - omega*t
UCFD_HIDDEN int kry_bicg_xr(hipStream_t s, int64_t n, double alpha, const double *pt,
                            double omega, const double *shat, const double *sv, const double *t,
                            double *x, double *r);
UCFD_HIDDEN int kry_copy(hipStream_t s, int64_t n, const double *x, double *y);
UCFD_HIDDEN int kry_fill(hipStream_t s, int64_t n, double v, double *y);
