/*
 * Sequential LU-SGS, legacy array API (MI355X implementation).
 *
 * The reference's lusgs.h is absent from its tree (named at
 * examples/mpi3d.c:26); each signature below is recovered from the
 * reference call sites.  Semantics are the sequential sweeps of
 * src/lusgs/lusgs.c:157-257: the lower sweep visits cells in ascending id and
 * uses neighbours nb < i, the upper sweep visits them in descending id and
 * uses neighbours nb > i.  On the GPU the sweep runs level by level (a cell's
 * level is one more than the highest level of the neighbours it depends on),
 * which gives exactly the sequential result.
 *
 * Layouts, pointer kinds and return values as in coloredlusgs.h.
 */
#ifndef UCFD_LUSGS_H
#define UCFD_LUSGS_H

#ifdef __cplusplus
extern "C" {
#endif

/* replaces: serial_pre_lusgs, called at examples/mpi3d.c:166 and
 * examples/mpi3d.py:71-72 (same arithmetic as parallel_pre_lusgs). */
int serial_pre_lusgs(int neles, int nface, double factor, double *fnorm_vol,
                     double *dt, double *diag, double *fspr);

/* C in-place form (9 args).  replaces: serial_ns_lower_sweep /
 * serial_ns_upper_sweep, called at examples/mpi3d.c:167-172. */
int serial_ns_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                          double *vec_fnorm, double *upts, double *dub,
                          double *diag, double *fspr);
int serial_ns_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                          double *vec_fnorm, double *upts, double *dub,
                          double *diag, double *fspr);

/* Python two-array form (10 args): lower rhsb -> dub, upper dub -> rhsb.
 * replaces: ns_serial_lower_sweep / ns_serial_upper_sweep, bound at
 * examples/mpi3d.py:23-34 and called at examples/mpi3d.py:73-80. */
int ns_serial_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                          double *vec_fnorm, double *upts, double *rhsb,
                          double *dub, double *diag, double *fspr);
int ns_serial_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                          double *vec_fnorm, double *upts, double *rhsb,
                          double *dub, double *diag, double *fspr);

/* RANS turbulence rows, serial (sequential-sweep) form: the
 * {rans}_{serial}_{lower,upper}_sweep family of doc/1_introduction.md:19-22,
 * 40-41 (no reference implementation or call site).  Signatures and
 * arithmetic as the coloured RANS sweeps of coloredlusgs.h, neighbour lower
 * iff nb < i (upper iff nb > i), as for the NS rows above. */
int serial_rans_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *dub,
                            double *tdiag, double *tfspr);
int serial_rans_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *dub,
                            double *tdiag, double *tfspr);
int rans_serial_lower_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *rhsb,
                            double *dub, double *tdiag, double *tfspr);
int rans_serial_upper_sweep(int neles, int nface, int *nei_ele, double *fnorm_vol,
                            double *vec_fnorm, double *upts, double *rhsb,
                            double *dub, double *tdiag, double *tfspr);

/* upts[k,i] += x[k,i] for k < NVARS.
 * replaces: serial_ns_lusgs_update (examples/mpi3d.c:173) and
 * lusgs_serial_ns_update (examples/mpi3d.py:21,26,81). */
int serial_ns_lusgs_update(int neles, double *upts, double *dub);
int lusgs_serial_ns_update(int neles, double *upts, double *rhsb);

#ifdef __cplusplus
}
#endif

#endif /* UCFD_LUSGS_H */
