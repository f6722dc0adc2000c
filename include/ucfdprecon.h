/*
 * Preconditioners of the Krylov API (replaces src/krylov/include/ucfdprecon.h).
 *
 * Block LU-SGS, sequential (BLUSGS, precon/blusgs.c) and coloured (PBLUSGS,
 * precon/pblusgs.c), on the GPU.  The sequential form gives the result of
 * the reference's row-by-row sweeps exactly (rows are grouped into
 * dependency levels at Prepare).  The coloured form expects the matrix
 * numbered colour by colour with icolors[0..ncolors] the colour offsets;
 * each colour range is one kernel launch.  ILU-type preconditioners are not
 * provided (their setters return UCFD_FAILED).
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
/* precon.c:29-34, blusgs.c:7-26: LU of every diagonal block (FP64, no
 * pivoting; a zero pivot returns UCFD_FAILED).  Also validates the structure
 * and builds the sweep schedule; the structure must not change afterwards. */
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
/* ucfdprecon.h:30-33: not provided, return UCFD_FAILED */
UCFD_EXTERN ucfd_status_t UCFDPreconSetILU(Precon *precon, UCFDInt n);
UCFD_EXTERN ucfd_status_t UCFDPreconSetBILU(Precon *precon, UCFDInt bn, UCFDInt block);
UCFD_EXTERN ucfd_status_t UCFDPreconSetPBILU(Precon *precon, UCFDInt bn, UCFDInt block,
                                             UCFDInt ncolors, UCFDInt *icolors);

/* Extension (the reference's UCFDPreconApply is static inline,
 * ucfdsolverimpl.h:63-67): b := M^-1 b in place. */
UCFD_EXTERN ucfd_status_t UCFDPreconApply(Precon precon, UCFDReal *b);
/* Extension: fp32 != 0 selects the mixed apply (block products in FP32 on an
 * FP32 copy of the blocks taken at Prepare, the diagonal LU and the
 * right-hand side in FP64).  Call before UCFDPreconPrepare. */
UCFD_EXTERN ucfd_status_t UCFDPreconSetPrecision(Precon precon, UCFDInt fp32);

#if defined(__cplusplus)
}
#endif
