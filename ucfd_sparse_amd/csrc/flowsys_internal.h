// Internal interface between the FlowSys host code (flowsys_api.cpp) and its
// kernels (flowsys_kernels.hip).  Not installed.
#pragma once

#include <stdint.h>

#include "ucfd_internal.h"

// Element-level kernels; cell_ids entries outside [0, nlocal) raise
// UCFD_KFLAG_INDEX and are skipped.
// inv [nlocal] and rec [nlocal * ((nvars + 2) & ~1)] (both optional): the
// record form of the pack
UCFD_HIDDEN int fs_launch_pack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars,
                               double a0, const int *cell_ids, const double *rhs, const double *dt,
                               double *rank_rhs, double *diag, int *flag, int *inv = nullptr,
                               double *rec = nullptr);
// kind 1 = Spalart-Allmaras, 2 = k-omega SST source Jacobian
UCFD_HIDDEN int fs_launch_tpack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars,
                                double a0, double factor, int kind, const int *cell_ids,
                                const double *uptsb, const double *dsrc, const double *dt,
                                double *tdiag, int *flag);
UCFD_HIDDEN int fs_launch_update(hipStream_t s, int neles, int nlocal, int nvars,
                                 const int *cell_ids, const double *du, double *upts, int *flag);
UCFD_HIDDEN int fs_launch_subres(hipStream_t s, int neles, int nlocal, int nvars,
                                 const int *cell_ids, const double *vol, const double *du,
                                 double *dup, double *res, int *flag);
// diag += sum over positions of the signed face Jacobian times fv, then LU;
// also writes every position's neighbour block jp[p][nv][nv] and fv (and
// their FP32 copies when jp32 != NULL).  sorted = the positions in
// ascending face slot (stable) with their cells, as (position, cell)
// pairs, slot s's being sorted[slot_ptr[s] .. slot_ptr[s+1]); partner = the
// other position of the same face slot (-1 when the slot has not exactly
// two).  A zero pivot raises UCFD_KFLAG_PIVOT.
UCFD_HIDDEN int fs_launch_prepare(hipStream_t s, int nv, int nlocal, int64_t nfaces, int64_t npos,
                                  const int *rowptr, const int *slot_ptr, const int2 *sorted,
                                  const int *partner, const int *slots, const signed char *sides,
                                  const double *face_area, const double *rcp_vol, double *diag,
                                  const double *jmat, double *jp, double *fvp, float *jp32,
                                  float *fv32, int *flag);
// One level of a block Gauss-Seidel sweep: rows base..base+cnt-1 (or
// order[base..base+cnt)), variables var0..var0+nv-1 of the [nlocal, nvars]
// arrays.  f32: FP32 products on jp32 / fv32 / du32, FP64 diagonal solve,
// du and du32 both written.  maxw: the most positions that 64/nv consecutive
// rows of the level hold (-1 = unknown); sizes the LDS of a contiguous level.
UCFD_HIDDEN int fs_launch_sweep(hipStream_t s, int nv, bool f32, int base, int cnt,
                                const int *order, int var0, int nvars, const int *rowptr,
                                const int *colidx, const double *jp, const double *fvp,
                                const float *jp32, const float *fv32, const double *rhs,
                                const double *diag, double *du, float *du32, int maxw);
// a = b = 0 (and c = 0 when not NULL), n entries
UCFD_HIDDEN int fs_launch_zero(hipStream_t s, int64_t n, double *a, double *b, float *c);

// Scalar LU-SGS (src/lusgs/lusgs.c); rank arrays u / du / diag are
// [nvars, nlocal].  lusgs_pack / rank_lusgs_update / pre_lusgs.  lpack needs
// inv (an nlocal int workspace: the inverse of cell_ids, built here).
UCFD_HIDDEN int fs_launch_lpack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars,
                                double turb_factor, double a0, const int *cell_ids,
                                const double *upts, const double *rhs, const double *dt,
                                const double *dsrc, double *u, double *du, double *diag,
                                int *flag, int *inv, double *rec);
// rec (optional, nlocal * fs_rec_words(nvars) doubles): record forms of Pack
// and Update for nvars 4..9 (one contiguous record per cell instead of a
// gather or scatter per row); other nvars, or rec NULL, use the row forms.
UCFD_HIDDEN int fs_launch_lupdate(hipStream_t s, int neles, int nlocal, int nvars,
                                  const int *cell_ids, const double *du, double *upts, int *flag,
                                  double *rec);
inline size_t fs_rec_words(int nvars) { return (size_t)((3*nvars + 1) & ~1); }
// Scalar Prepare (lusgs.c:87-113): diag rows nv0..nv1-1 += the cell's face
// sum, and the per-position face records the sweeps read: geo[p] =
// {side*normal (ndims), face_area*rcp_vol} as 4 doubles (geo may be NULL:
// only lam), lam[p] = lambdaf[slot of p].
UCFD_HIDDEN int fs_launch_lprep_pos(hipStream_t s, int ndims, int nlocal, int64_t nfaces, int nv0,
                                    int nv1, double kappa, const int *rowptr, const int *slots,
                                    const signed char *sides, const double *face_area,
                                    const double *face_normal, const double *rcp_vol,
                                    const double *lambdaf, double *diag, void *geo, double *lam);
// A sweep level reading those records (same results as fs_launch_lsweep).
// The first level of a scalar sweep schedule (no face taken): rows
// nv0..nv1-1 of du updated with a zero face sum (same bits as the sweeps).
UCFD_HIDDEN int fs_launch_lsweep_free(hipStream_t s, bool upper, int base, int cnt,
                                      const int *order, int nv0, int nv1, int nlocal,
                                      const double *diag, double *du);
UCFD_HIDDEN int fs_launch_lsweep_packed(hipStream_t s, bool upper, int ndims, int nt, int base,
                                        int cnt, const int *order, int nfvars, double kappa,
                                        int nlocal, const int *rowptr, const int *colidx,
                                        const void *geo, const double *lam, const double *diag,
                                        const double *u, double *du);
// One level of the lower (upper = false) or upper sweep: NS rows (nt = 0,
// nfvars = ndims + 2) or the nt turbulence rows (1 <= nt <= 4).
UCFD_HIDDEN int fs_launch_lsweep(hipStream_t s, bool upper, int ndims, int nt, int base, int cnt,
                                 const int *order, int nvars, int nfvars, double kappa,
                                 const double *lambdaf, int nlocal, int64_t nfaces,
                                 const int *rowptr, const int *colidx, const signed char *sides,
                                 const int *slots, const double *face_area,
                                 const double *face_normal, const double *rcp_vol,
                                 const double *diag, const double *u, double *du);
