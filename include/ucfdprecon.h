/*
 * Preconditioners of the Krylov API (replaces src/krylov/include/ucfdprecon.h).
 *
 * Block LU-SGS, sequential (BLUSGS, precon/blusgs.c) and coloured (PBLUSGS,
 * precon/pblusgs.c), and the ILU(0) family: scalar ILU(0) of a CSR matrix
 * (ILU, precon/ilu.c), block ILU(0) (BILU, precon/bilu.c) and its coloured
 * form (PBILU, precon/pbilu.c), all on the GPU.  The sequential forms give
 * the result of the reference's row-by-row factorisations and sweeps exactly
 * (rows are grouped into dependency levels at Prepare).  The coloured forms
 * expect the matrix numbered colour by colour with icolors[0..ncolors] the
 * colour offsets; each colour range is one kernel launch, and the result is
 * that of the reference run on one thread.
 */
#pragma once

#include "ucfdmat.h"
#include "ucfdtypes.h"

typedef struct _Precon *Precon;
typedef const char *PreconType;

#define NONE "none"
#define ILU "ilu"
#define BILU "bilu"
#define BLUSGS "blu-sgs"
#define PBILU "pbilu"
#define PBLUSGS "pblu-sgs"
#define CUDABILU "cudabilu"
#define CUDABLUSGS "cudablusgs"

#if defined(__cplusplus)
extern "C" {
#endif

/* precon.c:13-27: keeps the caller's block-sparse arrays (no copy) */
UCFD_EXTERN ucfd_status_t UCFDPreconCreatefromArrays(Precon *precon, UCFDInt *rowptr,
                                                     UCFDInt *colidx, UCFDInt *diagslots,
                                                     UCFDReal *values);
/* none.c:14-30 */
UCFD_EXTERN ucfd_status_t UCFDPreconCreateNone(Precon *precon);
/* precon.c:29-34.  LU-SGS (blusgs.c:7-26): LU of every diagonal block into
 * library memory (FP64, no pivoting).  ILU family (ilu.c:4-48, bilu.c:5-65,
 * pbilu.c:5-76): the ILU(0) factorisation, in place on the caller's values
 * array, as the reference does (calling Prepare twice factors twice).  A
 * zero pivot returns UCFD_FAILED.  The first Prepare after a UCFDPreconSet*
 * call also validates the structure and builds the sweep schedule; later
 * Prepares (new values, e.g. one per Newton step) reuse them and read only
 * the values, so the structure must not change afterwards. */
UCFD_EXTERN ucfd_status_t UCFDPreconPrepare(Precon precon);
/* precon.c:36-45 */
UCFD_EXTERN ucfd_status_t UCFDPreconDestroy(Precon *precon);

/* blusgs.c:111-128 */
UCFD_EXTERN ucfd_status_t UCFDPreconSetBLUSGS(Precon *precon, UCFDInt bn, UCFDInt block);
/* pblusgs.c:107-127 (the colour offsets are read at Prepare) */
UCFD_EXTERN ucfd_status_t UCFDPreconSetPBLUSGS(Precon *precon, UCFDInt bn, UCFDInt block,
                                               UCFDInt ncolors, UCFDInt *icolors);
/* ucfdprecon.h:36 (CUDA builds): the coloured form under the CUDA name */
UCFD_EXTERN ucfd_status_t UCFDPreconSetCUDABLUSGS(Precon *precon, UCFDInt bn, UCFDInt block,
                                                  UCFDInt ncolors, UCFDInt *icolors);
/* ilu.c:99-117: ILU(0) of an n x n CSR matrix.  The solve's backward sweep
 * runs over the strictly upper entries, x_i = (y_i - sum_{j>i} U_ij x_j) /
 * U_ii; the reference's loop (ilu.c:81) also subtracts the diagonal term
 * U_ii y_i first, which this library does not replicate. */
UCFD_EXTERN ucfd_status_t UCFDPreconSetILU(Precon *precon, UCFDInt n);
/* bilu.c:138-158: block ILU(0), 1 <= block <= 7 */
UCFD_EXTERN ucfd_status_t UCFDPreconSetBILU(Precon *precon, UCFDInt bn, UCFDInt block);
/* pbilu.c:184-207: coloured block ILU(0) (the colour offsets are read at
 * Prepare) */
UCFD_EXTERN ucfd_status_t UCFDPreconSetPBILU(Precon *precon, UCFDInt bn, UCFDInt block,
                                             UCFDInt ncolors, UCFDInt *icolors);
/* ucfdprecon.h:35, cubilu.cu:321-343 (CUDA builds): the coloured form under
 * the CUDA name */
UCFD_EXTERN ucfd_status_t UCFDPreconSetCUDABILU(Precon *precon, UCFDInt bn, UCFDInt block,
                                                UCFDInt ncolors, UCFDInt *icolors);

/* Extension (the reference's UCFDPreconApply is static inline,
 * ucfdsolverimpl.h:63-67): b := M^-1 b in place. */
UCFD_EXTERN ucfd_status_t UCFDPreconApply(Precon precon, UCFDReal *b);
/* Extension: fp32 != 0 selects the mixed apply (block products in FP32 on an
 * FP32 copy of the blocks taken at Prepare, the diagonal LU and the
 * right-hand side in FP64).  Call before UCFDPreconPrepare.  Block LU-SGS
 * only (UCFD_FAILED for other types). */
UCFD_EXTERN ucfd_status_t UCFDPreconSetPrecision(Precon precon, UCFDInt fp32);

#if defined(__cplusplus)
}
#endif
