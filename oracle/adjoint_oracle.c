/*
 * ORACLE — test infrastructure only (see lusgs_oracle.c for the rules).
 *
 * Adjoint (reverse-mode derivative) of one FP64 coloured LU-SGS pass
 * dQ = upper(lower(pre(dt, fspr), rhsb)) of lusgs_oracle.c (which follows
 * src/lusgs/lusgs.c:87-257 and is pinned bit for bit to the reference), on a
 * colour-contiguous numbering: colour c = cells off[c] .. off[c+1]-1.  The
 * reference has no adjoint; this is a build-defined restatement.
 *
 * Forward, per cell i with face data fv = fnorm_vol[f,i], lam = fspr[f,i]:
 *   diag_i  = 1/(dt_i factor) + sum_f 0.5 lam fv
 *   lower (colours ascending):  ds_i = (rhs_i - 0.5 sum_{nb lower} g) / diag_i
 *   upper (colours descending): dq_i = ds_i - 0.5 sum_{nb upper} h / diag_i
 *   g = [F(Q_nb + ds_nb; n) - F(Q_nb; n) - lam ds_nb] fv, h likewise with dq.
 * Reverse, given w = dJ/d dq:
 *   a_i = w_i + sum_{j lower nb of i} (-0.5 fv_j / diag_j) (A(Q_i+dq_i; n_j)^T
 *         a_j - lam_j a_j)                       (colours ascending)
 *   b_i = a_i + the same sum over upper nbs with b_j and A(Q_i+ds_i; n_j)
 *                                                 (colours descending)
 *   rhsbar_i = b_i / diag_i
 *   diagbar_i = a_i.(ds_i - dq_i)/diag_i - b_i.ds_i/diag_i
 *   dtbar_i = -diagbar_i / (dt_i^2 factor)
 *   fsprbar[f,i] = 0.5 fv diagbar_i + 0.5 fv (a_i.dq_nb) / diag_i (upper nb)
 *                                   + 0.5 fv (b_i.ds_nb) / diag_i (lower nb)
 * where j's face data (fv_j, lam_j, n_j) are those of the face of j that
 * points back to i (pair map).  Gathers only, in face order: deterministic.
 *
 * Pins: central differences of the forward oracle (tests/test_adjoint_oracle.py)
 * for every input; orc_flux_jt against central differences of orc_flux
 * including the PMIN-clamped branch.
 */
#include <stdint.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define AGAMMA 1.4
#define APMIN 1e-13

ORC_EXPORT void orc_flux_jt(int nd, const double *u, const double *n, const double *v, double *out)
{
    const int nv = nd + 2;
    const double g = AGAMMA, gm1 = AGAMMA - 1.0;
    const double rho = u[0], inv = 1.0/rho;
    double temp = 0.0, contrav = 0.0;
    for (int i = 0; i < nd; ++i) {
        contrav += u[i + 1]*n[i];
        temp += u[i + 1]*u[i + 1];
    }
    const double theta = contrav/rho;
    double et = u[nv - 1];
    double p = gm1*(et - 0.5*temp/rho);
    const int clamped = p < APMIN;
    if (clamped) {
        p = APMIN;
        et = p/gm1 + 0.5*temp/rho;
    }
    const double H = (et + p)*inv;
    double vel[3], um = 0.0, nm = 0.0, q2 = 0.0;
    for (int d = 0; d < nd; ++d) {
        vel[d] = u[d + 1]*inv;
        um += vel[d]*v[d + 1];
        nm += n[d]*v[d + 1];
        q2 += vel[d]*vel[d];
    }
    const double v0 = v[0], vE = v[nv - 1];
    if (!clamped) {
        const double phi = gm1*0.5*q2;
        out[0] = -theta*um + phi*nm + theta*(phi - H)*vE;
        for (int k = 0; k < nd; ++k)
            out[k + 1] = n[k]*v0 + theta*v[k + 1] + n[k]*um - gm1*vel[k]*nm
                         + (H*n[k] - gm1*theta*vel[k])*vE;
        out[nv - 1] = gm1*nm + g*theta*vE;
    } else {
        out[0] = -theta*um - theta*(0.5*q2 + H)*vE;
        for (int k = 0; k < nd; ++k)
            out[k + 1] = n[k]*v0 + theta*v[k + 1] + n[k]*um + (theta*vel[k] + H*n[k])*vE;
        out[nv - 1] = 0.0;
    }
}

/* one neighbour's contribution: acc += (-0.5 fv/dj) (A(U; n)^T x - lam x) */
static void accumulate(int nd, int64_t N, int64_t j, int f, const double *U, const double *fnv,
                       const double *vfn, const double *fspr, const double *diag,
                       const double *work, double *acc)
{
    const int nv = nd + 2;
    double n[3], x[5], jt[5];
    const double fv = fnv[f*N + j], lam = fspr[f*N + j];
    for (int d = 0; d < nd; ++d) n[d] = vfn[((int64_t)f*nd + d)*N + j];
    for (int k = 0; k < nv; ++k) x[k] = work[k*N + j];
    orc_flux_jt(nd, U, n, x, jt);
    const double coef = -0.5*fv/diag[j];
    for (int k = 0; k < nv; ++k) acc[k] += coef*(jt[k] - lam*x[k]);
}

ORC_EXPORT void orc_lusgs_adjoint(int64_t N, int F, int nd, int nc, const int32_t *off,
                                  const int32_t *nei, const int32_t *pair, const double *fnv,
                                  const double *vfn, const double *fspr, const double *dt,
                                  double factor, const double *diag, const double *upts,
                                  const double *dstar, const double *dq, const double *w,
                                  double *work, double *rhsbar, double *dtbar, double *fsprbar)
{
    const int nv = nd + 2;
    double U[5], acc[5], a[5];
    for (int c = 0; c < nc; ++c)                       /* adjoint of the upper sweep */
        for (int64_t i = off[c]; i < off[c + 1]; ++i) {
            for (int k = 0; k < nv; ++k) {
                acc[k] = w[k*N + i];
                U[k] = upts[k*N + i] + dq[k*N + i];
            }
            for (int f = 0; f < F; ++f) {
                const int64_t j = nei[f*N + i];
                if (j == i || j >= off[c]) continue;
                accumulate(nd, N, j, pair[f*N + i], U, fnv, vfn, fspr, diag, work, acc);
            }
            for (int k = 0; k < nv; ++k) work[k*N + i] = acc[k];
        }
    for (int c = nc - 1; c >= 0; --c)                  /* adjoint of the lower sweep */
        for (int64_t i = off[c]; i < off[c + 1]; ++i) {
            for (int k = 0; k < nv; ++k) {
                a[k] = work[k*N + i];
                acc[k] = a[k];
                U[k] = upts[k*N + i] + dstar[k*N + i];
            }
            for (int f = 0; f < F; ++f) {
                const int64_t j = nei[f*N + i];
                if (j == i || j < off[c + 1]) continue;
                accumulate(nd, N, j, pair[f*N + i], U, fnv, vfn, fspr, diag, work, acc);
            }
            const double di = diag[i];
            double dgb = 0.0;
            for (int k = 0; k < nv; ++k) {
                work[k*N + i] = acc[k];
                rhsbar[k*N + i] = acc[k]/di;
                dgb += a[k]*((dstar[k*N + i] - dq[k*N + i])/di) - acc[k]*(dstar[k*N + i]/di);
            }
            dtbar[i] = -dgb/(dt[i]*dt[i]*factor);
            for (int f = 0; f < F; ++f) {
                const double fv = fnv[f*N + i];
                const int64_t nb = nei[f*N + i];
                double val = dgb*0.5*fv;
                if (nb != i) {
                    double s = 0.0;
                    if (nb >= off[c + 1])
                        for (int k = 0; k < nv; ++k) s += a[k]*dq[k*N + nb];
                    else
                        for (int k = 0; k < nv; ++k) s += acc[k]*dstar[k*N + nb];
                    val += 0.5*fv*s/di;
                }
                fsprbar[f*N + i] = val;
            }
        }
}
