/*
 * ORACLE — test infrastructure only (see lusgs_oracle.c for the rules).
 *
 * CPU statement of the LU-SGS system operator and of the FP64 defect
 * correction built around the coloured sweeps.  The reference has no
 * residual routine; the operator is the one LU-SGS factors approximately,
 * doc/3.1_lusgs.md:79-86,100:
 *   A(dq)_i = D_i dq_i + 1/2 sum_{f, nb != i} fnorm_vol_f [F(Q_nb + dq_nb; n_f)
 *             - F(Q_nb; n_f) - fspr_f dq_nb]
 * with D the pre diagonal (src/lusgs/lusgs.c:33,100-110) and the flux
 * difference of src/lusgs/lusgs.c:143-155 (F(u + du) - F(u), direct form).
 * The defect correction is e <- e + M^{-1}(b - A(e)) with M^{-1} one
 * coloured LU-SGS pass (orc_lower / orc_upper, lusgs.c:157-257) on base
 * state Q.
 *
 * Parity pins (tests/test_solve_oracle.py): after one LU-SGS pass the
 * residual vanishes on colour 1 (it is the lower sweep's own equation there,
 * which the reference fixtures pin); r = b at dq = 0; on an all-self mesh
 * r = b - D dq; the operator's neighbour term matches the sweep's on a
 * two-colour mesh; finite-difference consistency with the flux Jacobian.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define RGAMMA 1.4
#define RPMIN 1e-13
#define RMAXV 16

void orc_pre(int neles, int nface, double factor, const double *fnorm_vol,
             const double *dt, double *diag, const double *fspr);
void orc_lower(int mode, int n0, int ne, int neles, int nface, int ndims, int nfvars,
               const int32_t *nei, const int32_t *icolor, const int32_t *lcolor,
               const double *fnorm_vol, const double *vec_fnorm, const double *upts,
               const double *in, double *out, const double *diag, const double *fspr);
void orc_upper(int mode, int n0, int ne, int neles, int nface, int ndims, int nfvars,
               const int32_t *nei, const int32_t *icolor, const int32_t *lcolor,
               const double *fnorm_vol, const double *vec_fnorm, const double *upts,
               const double *in, double *out, const double *diag, const double *fspr);
void orc_flux(int nfvars, int ndims, const double *u_in, const double *nf, double *f);

/* r = b - A(dq); returns ||r||_2 */
ORC_EXPORT double orc_residual(int neles, int nface, int ndims, const int32_t *nei,
                               const double *fnorm_vol, const double *vec_fnorm,
                               const double *fspr, const double *diag, const double *q,
                               const double *b, const double *dq, double *r)
{
    const int64_t N = neles;
    const int nfv = ndims + 2;
    double ss = 0.0;
    #pragma omp parallel for schedule(static) reduction(+:ss)
    for (int64_t i = 0; i < N; ++i) {
        double acc[RMAXV] = {0};
        for (int f = 0; f < nface; ++f) {
            const int64_t nb = nei[f*N + i];
            if (nb == i) continue;
            double u[RMAXV], up[RMAXV], f0[RMAXV], f1[RMAXV], nf[3];
            for (int k = 0; k < nfv; ++k) {
                u[k] = q[k*N + nb];
                up[k] = dq[k*N + nb] + u[k];
            }
            for (int d = 0; d < ndims; ++d) nf[d] = vec_fnorm[(f*ndims + d)*N + i];
            orc_flux(nfv, ndims, u, nf, f0);
            orc_flux(nfv, ndims, up, nf, f1);
            const double fv = fnorm_vol[f*N + i], lam = fspr[f*N + i];
            for (int k = 0; k < nfv; ++k)
                acc[k] += ((f1[k] - f0[k]) - lam*dq[k*N + nb])*fv;
        }
        for (int k = 0; k < nfv; ++k) {
            const double v = b[k*N + i] - (diag[i]*dq[k*N + i] + 0.5*acc[k]);
            r[k*N + i] = v;
            ss += v*v;
        }
    }
    return sqrt(ss);
}

/*
 * FP64 defect correction on a coloured mesh: dq <- dq + LUSGS_Q(b - A(dq)),
 * starting from dq = 0, until ||b - A(dq)|| <= rtol ||b|| (checked before
 * each pass) or maxpass passes.  color_off[c]..color_off[c+1] index icolor.
 * Writes the relative residual of every check to hist (nullable, maxpass
 * entries) and returns the number of passes applied.
 */
ORC_EXPORT int orc_dc_solve(int neles, int nface, int ndims, int ncolors, const int32_t *color_off,
                            const int32_t *nei, const int32_t *icolor, const int32_t *lcolor,
                            const double *fnorm_vol, const double *vec_fnorm, const double *fspr,
                            const double *dt, const double *q, const double *b, double *dq,
                            double rtol, int maxpass, double *hist)
{
    const int64_t N = neles;
    const int nfv = ndims + 2;
    double *diag = malloc(sizeof(double)*N);
    double *s = malloc(sizeof(double)*nfv*N);
    double *d = malloc(sizeof(double)*nfv*N);
    orc_pre(neles, nface, 1.0, fnorm_vol, dt, diag, fspr);
    memset(dq, 0, sizeof(double)*nfv*N);
    double b0 = 0.0;
    int pass = 0;
    for (; pass < maxpass; ++pass) {
        const double rn = orc_residual(neles, nface, ndims, nei, fnorm_vol, vec_fnorm, fspr, diag,
                                       q, b, dq, s);
        if (pass == 0) b0 = rn;
        if (hist) hist[pass] = b0 > 0 ? rn/b0 : 0.0;
        if (b0 == 0.0 || rn <= rtol*b0) break;
        memcpy(d, s, sizeof(double)*nfv*N);
        for (int c = 0; c < ncolors; ++c)
            orc_lower(0, color_off[c], color_off[c + 1], neles, nface, ndims, nfv, nei, icolor,
                      lcolor, fnorm_vol, vec_fnorm, q, d, d, diag, fspr);
        for (int c = ncolors - 1; c >= 0; --c)
            orc_upper(0, color_off[c], color_off[c + 1], neles, nface, ndims, nfv, nei, icolor,
                      lcolor, fnorm_vol, vec_fnorm, q, d, d, diag, fspr);
        for (int64_t j = 0; j < (int64_t)nfv*N; ++j) dq[j] += d[j];
    }
    free(diag);
    free(s);
    free(d);
    return pass;
}
