/*
 * Colored LU-SGS, legacy array API (MI355X implementation).
 *
 * The reference's coloredlusgs.h / coloredlusgs.c are absent from its tree
 * (named at examples/omp3d.c:27, doc/1_introduction.md:71); each signature
 * below is recovered from the reference call sites it must satisfy.
 *
 * Layouts (C order, int = 32 bit, double = FP64):
 *   nei_ele[f*neles + i]  fnorm_vol[f*neles + i]  fspr[f*neles + i]
 *   vec_fnorm[(f*NDIMS + d)*neles + i]  upts/rhsb/dub[k*neles + i]
 *   dt[i]  diag[i]  icolor[p] -> cell  lcolor[i] -> colour (1-based)
 * A call with n0, ne processes the cells icolor[n0..ne-1] (one colour); a face
 * neighbour nb is "lower" iff lcolor[nb] < lcolor[i] and "upper" iff
 * lcolor[nb] > lcolor[i]; boundary faces point to the cell itself and count
 * as neither.
 *
 * Pointers may be device (hipMalloc / managed) or host memory: device arrays
 * are used in place, host arrays are staged through the GPU (correct but
 * PCIe bound).  Every function returns 0 on success and a nonzero status on
 * error (the reference prints and exits; callers that ignore the return
 * value keep working).
 */
#ifndef UCFD_COLOREDLUSGS_H
#define UCFD_COLOREDLUSGS_H

#ifdef __cplusplus
extern "C" {
#endif

/* diag[i] = 1/(dt[i]*factor) + sum_f 0.5*fspr[f,i]*fnorm_vol[f,i]  (assigned)
 * replaces: parallel_pre_lusgs, called at examples/omp3d.c:180 and
 * examples/omp3d.py:76-77; arithmetic of src/lusgs/lusgs.c:33,100-110. */
int parallel_pre_lusgs(int neles, int nface, double factor, double *fnorm_vol,
                       double *dt, double *diag, double *fspr);

/* C in-place form (13 args): dub holds the right-hand side on entry of the
 * lower sweep, dQ* after it and dQ after the upper sweep.
 * lower: dub_i = (dub_i - 0.5*df_i)/diag_i,  df over lower neighbours
 * upper: dub_i = dub_i - 0.5*df_i/diag_i,    df over upper neighbours
 * df_i = sum_f (F(Q_nb + dub_nb; n_f) - F(Q_nb; n_f) - fspr_f*dub_nb)*fnorm_vol_f
 * replaces: parallel_ns_lower_sweep / parallel_ns_upper_sweep, called at
 * examples/omp3d.c:181-188; arithmetic of src/lusgs/lusgs.c:157-257. */
int parallel_ns_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                            int *icolor, int *lcolor, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *dub,
                            double *diag, double *fspr);
int parallel_ns_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                            int *icolor, int *lcolor, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *dub,
                            double *diag, double *fspr);

/* Python two-array form (14 args): lower reads rhsb and writes dub = dQ*;
 * upper reads dub and writes rhsb = dQ (README.md:44-50).
 * replaces: ns_parallel_lower_sweep / ns_parallel_upper_sweep, bound at
 * examples/omp3d.py:24-35 and called at examples/omp3d.py:78-93. */
int ns_parallel_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                            int *icolor, int *lcolor, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *rhsb,
                            double *dub, double *diag, double *fspr);
int ns_parallel_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                            int *icolor, int *lcolor, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *rhsb,
                            double *dub, double *diag, double *fspr);

/* RANS turbulence rows (k = NFVARS .. NVARS-1, NVARS - NFVARS in {1, 2}),
 * the {rans}_{parallel}_{lower,upper}_sweep family of doc/1_introduction.md:
 * 19-22,40-41, which no reference file implements or calls.  Signatures: the
 * NS ones above (decided here) with tdiag[i] the turbulence diagonal (from
 * parallel_pre_lusgs with the turbulence spectral radius and factor =
 * turb_factor) and tfspr[f*neles + i] the turbulence spectral radius; upts
 * holds all NVARS rows (the neighbour's density and momentum give the
 * contravariant velocity).  Per taken face:
 *   df_k += ((dub_nb,k + u_nb,k)*theta - u_nb,k*theta - tfspr_f*dub_nb,k)*fnorm_vol_f,
 *   theta = (sum_d m_nb,d n_f,d)/rho_nb
 * (src/lusgs/lusgs.c:274-321 with rans_flux_container, flux.c:77-90); update
 * and the in-place / two-array conventions as for the NS rows. */
int parallel_rans_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                              int *icolor, int *lcolor, double *fnorm_vol,
                              double *vec_fnorm, double *upts, double *dub,
                              double *tdiag, double *tfspr);
int parallel_rans_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                              int *icolor, int *lcolor, double *fnorm_vol,
                              double *vec_fnorm, double *upts, double *dub,
                              double *tdiag, double *tfspr);
int rans_parallel_lower_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                              int *icolor, int *lcolor, double *fnorm_vol,
                              double *vec_fnorm, double *upts, double *rhsb,
                              double *dub, double *tdiag, double *tfspr);
int rans_parallel_upper_sweep(int n0, int ne, int neles, int nface, int *nei_ele,
                              int *icolor, int *lcolor, double *fnorm_vol,
                              double *vec_fnorm, double *upts, double *rhsb,
                              double *dub, double *tdiag, double *tfspr);

/* upts[k,i] += x[k,i] for k < NVARS (x = dub in C, rhsb in Python).
 * replaces: lusgs_parallel_ns_update, called at examples/omp3d.c:189 and
 * examples/omp3d.py:94; arithmetic of src/lusgs/lusgs.c:62-68.
 * parallel_ns_lusgs_update is the same function under the serial family's
 * word order. */
int lusgs_parallel_ns_update(int neles, double *upts, double *x);
int parallel_ns_lusgs_update(int neles, double *upts, double *x);

#ifdef __cplusplus
}
#endif

#endif /* UCFD_COLOREDLUSGS_H */
