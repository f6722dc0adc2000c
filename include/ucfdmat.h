/*
 * Sparse matrices of the Krylov API (replaces src/krylov/include/ucfdmat.h).
 *
 * A matrix handle keeps the caller's arrays (no copy), as in the reference
 * (sparsemat.c:32-50, 104-125): values changed between calls are used by the
 * next call.  Arrays may be host memory (copied to the GPU for each call) or
 * device memory (used in place, the fast path).  rowptr must not change
 * after the first use of the handle.  Errors print one line to stderr and
 * return UCFD_FAILED.
 */
#pragma once

#include "ucfdtypes.h"

typedef struct _SpMat *SpMat;
typedef const char *SpMatType;

#define CSR "csr"
#define BSR "bsr"
#define CSRCUDA "cudacsr"
#define BSRCUDA "cudabsr"

#if defined(__cplusplus)
extern "C" {
#endif

/* matrix.c:4-14 */
UCFD_EXTERN ucfd_status_t UCFDMatInit(SpMat *mat);
/* matrix.c:23-33 */
UCFD_EXTERN ucfd_status_t UCFDMatDestroy(SpMat *mat);
/* matrix.c:16-21, sparsemat.c:10-29 and 57-101: y = alpha*A*x + beta*y */
UCFD_EXTERN ucfd_status_t UCFDMatMult(UCFDReal alpha, SpMat mat, UCFDReal *x, UCFDReal beta,
                                      UCFDReal *y);
/* sparsemat.c:32-50: scalar CSR, n rows */
UCFD_EXTERN ucfd_status_t UCFDMatCreateCSR(SpMat *mat, UCFDInt n, UCFDInt *rowptr,
                                           UCFDInt *colidx, UCFDReal *values);
/* sparsemat.c:104-125: BSR, bn block rows of blk x blk row-major blocks
 * (1 <= blk <= 7), vectors stored x[row*blk + r] */
UCFD_EXTERN ucfd_status_t UCFDMatCreateBSR(SpMat *mat, UCFDInt bn, UCFDInt blk, UCFDInt *rowptr,
                                           UCFDInt *colidx, UCFDReal *values);
/* ucfdmat.h:39-40 (CUDA builds): the same GPU matrices under the CUDA names */
UCFD_EXTERN ucfd_status_t UCFDMatCreateCUDACSR(SpMat *mat, UCFDInt n, UCFDInt *rowptr,
                                               UCFDInt *colidx, UCFDReal *values);
UCFD_EXTERN ucfd_status_t UCFDMatCreateCUDABSR(SpMat *mat, UCFDInt bn, UCFDInt blk,
                                               UCFDInt *rowptr, UCFDInt *colidx, UCFDReal *values);

#if defined(__cplusplus)
}
#endif
