/*
 * FlowSys API (replaces src/lusgs/include/ucfdlusgs.h): a face-slot CSR
 * system with per-element arrays and the block LU-SGS method on the GPU.
 *
 * Pointers are kept, not copied (flowsys.c:18-54, 120-140); they may be host
 * memory (copied to the GPU for each call) or device memory (used in place).
 * The working arrays rhs, du, dup [nlocal, nvars] and the LU-factored
 * diagonal blocks are library-owned device memory.  Structure arrays
 * (rowptr, colidx, slots, sides) must not change after the first Prepare;
 * jmat / tjmat are read at Prepare (call Prepare again after changing them).
 * The sweeps give the reference's sequential Gauss-Seidel result exactly
 * (cells grouped into dependency levels; one launch per colour on a
 * colour-ordered numbering).  Errors print one line and return UCFD_FAILED.
 */
#pragma once

#include "ucfdtypes.h"

typedef struct _FlowSys *FlowSys;

#if defined(__cplusplus)
extern "C" {
#endif

/* flowsys.c:18-54: nlocal cells, nvars = nfvars + nturbvars, nfaces face
 * slots, nelem element types; rowptr[nlocal+1], colidx / slots / sides
 * [rowptr[nlocal]] (sides +1 / -1), face_area[nfaces],
 * face_normal[ndims*nfaces], rcp_vol[nlocal] */
UCFD_EXTERN ucfd_status_t UCFDFlowSysCreate(FlowSys *sys, UCFDInt nlocal, UCFDInt nvars,
                                            UCFDInt nfvars, UCFDInt ndims, UCFDInt nfaces,
                                            UCFDInt nelem, UCFDInt *rowptr, UCFDInt *colidx,
                                            UCFDInt *slots, UCFDInt8 *sides, UCFDReal *face_area,
                                            UCFDReal *face_normal, UCFDReal *rcp_vol);
/* flowsys.c:120-140: element arrays [nvars, neles] (uptsb, rhs, dsrc,
 * resid_out), [neles] (dt, vol); cell_ids[neles] = rank cell of each element */
UCFD_EXTERN ucfd_status_t UCFDFlowSysSetElement(FlowSys *sys, UCFDInt eidx, UCFDInt neles,
                                                UCFDInt nface, UCFDInt *cell_ids, UCFDReal *uptsb,
                                                UCFDReal *rhs, UCFDReal *dt, UCFDReal *dsrc,
                                                UCFDReal *vol, UCFDReal *resid_out);
/* flowsys.c:6-15 */
UCFD_EXTERN ucfd_status_t UCFDFlowSysDestroy(FlowSys *sys);

/* ucfdlusgs.h:17-25 and flowsys.c:64-81: the scalar LU-SGS method of the
 * FlowSys.  fspr / tfspr: spectral radius per face slot (the pointers are
 * kept); rank arrays u / du / diag are library-owned [nvars, nlocal].  Each
 * Prepare also stores per-position face records (side*normal,
 * face_area*rcp_vol and the spectral radius of its rows, 40 B per position
 * plus 8 B for the turbulence radius): NS (RANS) sweeps use the fspr
 * (tfspr) that the last NS (RANS) Prepare read and the geometry that the
 * last NS Prepare read (or a RANS Prepare with none before it), bit for bit
 * the same results as reading them again; a caller that changes them calls
 * Prepare again (as the diagonal needs anyway).  Sweeps with no Prepare
 * before them read the caller's arrays.  Pack (lusgs.c:8-52): per element, u = uptsb,
 * du = rhs, diag = 1/(dt*factor) + a0 + dsrc with factor = turb_factor on the
 * turbulence rows (cell_ids mapping two elements to one cell returns
 * UCFD_FAILED; Pack and Update keep a work buffer of
 * nlocal * (3*nvars rounded up to even) doubles).  Prepare (lusgs.c:87-140): diag += sum over the cell's
 * positions of 0.5*kappa*fspr*face_area*rcp_vol (NS rows, or turbulence rows
 * with tfspr).  Sweeps (lusgs.c:157-321): in-place Gauss-Seidel over the lower
 * (upper) neighbours, NS rows with the Euler flux difference, turbulence rows
 * with the RANS convective flux (flux.c:77-90; 1 to 4 turbulence rows).
 * Update (lusgs.c:54-84): uptsb += du.  ndims 2 or 3, nfvars = ndims + 2. */
UCFD_EXTERN ucfd_status_t UCFDFlowSysSetLUSGS(FlowSys *sys, UCFDReal *fspr, UCFDReal *tfspr);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_Pack(FlowSys sys, UCFDInt eidx, UCFDReal factor, UCFDReal a0);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_Update(FlowSys sys, UCFDInt eidx);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_NSPrepare(FlowSys sys, UCFDReal kappa);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_RANSPrepare(FlowSys sys, UCFDReal kappa);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_NSLowerSweep(FlowSys sys, UCFDReal kappa);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_NSUpperSweep(FlowSys sys, UCFDReal kappa);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_RANSLowerSweep(FlowSys sys, UCFDReal kappa);
UCFD_EXTERN ucfd_status_t UCFDLUSGS_RANSUpperSweep(FlowSys sys, UCFDReal kappa);
/* Extension: copy a library-owned scalar LU-SGS array to out (host or
 * device): which 0 = u, 1 = du, 2 = diag (each [nvars, nlocal]). */
UCFD_EXTERN ucfd_status_t UCFDLUSGS_Export(FlowSys sys, UCFDInt which, UCFDReal *out);

/* flowsys.c:91-117: jmat [2][nfvars][nfvars][nfaces] (A+, A- per face),
 * tjmat [2][nturbvars][nturbvars][nfaces] or NULL */
UCFD_EXTERN ucfd_status_t UCFDFlowSysSetBLUSGS(FlowSys *sys, UCFDReal *jmat, UCFDReal *tjmat);
/* blusgs.c:47-57: rank rhs from the element rhs, diag = (1/dt + a0) I
 * (through one record per cell in the scalar method's work buffer) */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_Pack(FlowSys sys, UCFDInt eidx, UCFDReal a0);
/* blusgs.c:96-123: turbulence diagonal with the k-omega SST / SA source
 * Jacobian plus 1/(dt turb_factor) + a0 */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_KWSST_Pack(FlowSys sys, UCFDInt eidx, UCFDReal turb_factor,
                                                UCFDReal a0);
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_SA_Pack(FlowSys sys, UCFDInt eidx, UCFDReal turb_factor,
                                             UCFDReal a0);
/* blusgs.c:145-155: uptsb += du */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_Update(FlowSys sys, UCFDInt eidx);
/* blusgs.c:187-197: resid_out = (du - dup)^2 vol, dup = du */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_SubResidual(FlowSys sys, UCFDInt eidx);
/* blusgs.c:248-270: diagonal blocks += sum of face Jacobians * fv, then LU
 * (FP64; a zero pivot returns UCFD_FAILED) */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_NSPrepare(FlowSys sys);
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_RANSPrepare(FlowSys sys);
/* blusgs.c:375-428: block Gauss-Seidel over every face neighbour, cells
 * ascending (lower) or descending (upper), in place on du */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_NSLowerSweep(FlowSys sys);
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_RANSLowerSweep(FlowSys sys);
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_NSUpperSweep(FlowSys sys);
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_RANSUpperSweep(FlowSys sys);
/* blusgs.c:431-440: du = dup = 0 */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_Reset(FlowSys sys);

/* Extension: fp32 != 0 runs the sweeps' block products in FP32 (FP32 copies
 * of the packed blocks and of du) with the FP64 diagonal solve and FP64 du;
 * takes effect at the next Prepare. */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_SetPrecision(FlowSys sys, UCFDInt fp32);
/* Extension: copy a library-owned array to out (host or device):
 * which 0 = rhs, 1 = du, 2 = dup ([nlocal, nvars]), 3 = diag
 * ([nlocal, nfvars, nfvars], LU-factored after Prepare), 4 = tdiag. */
UCFD_EXTERN ucfd_status_t UCFDBLUSGS_Export(FlowSys sys, UCFDInt which, UCFDReal *out);

#if defined(__cplusplus)
}
#endif
