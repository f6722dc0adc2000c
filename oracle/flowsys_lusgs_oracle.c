/*
 * ORACLE — test infrastructure only (see lusgs_oracle.c for the rules).
 *
 * Single-threaded CPU restatement of the reference's FlowSys scalar LU-SGS
 * (src/lusgs/lusgs.c), in its operation order:
 *   orc_fsl_pack     lusgs_pack          lusgs.c:8-37 (factor = turb_factor
 *                    on the turbulence rows)
 *   orc_fsl_update   rank_lusgs_update   lusgs.c:54-70
 *   orc_fsl_prepare  pre_lusgs           lusgs.c:87-113 (rows nv0 .. nv1-1)
 *   orc_fsl_sweep    lower_sweep / upper_sweep lusgs.c:157-257 with diff_flux
 *                    lusgs.c:143-155 and ns_flux_container (NS rows) or
 *                    rans_flux_container (turbulence rows), flux.c:29-90
 * Layouts: element arrays [nvars, neles]; rank u / du / diag [nvars, nlocal];
 * face_normal [ndims, nfaces]; the spectral radii lambdaf per face slot.
 * Rows are visited ascending (lower) or descending (upper) and du is updated
 * in place, as in the reference.
 *
 * Pins: bit-exact against the reference compiled from src/lusgs/*.c
 * (oracle/_ref/libref_lusgs.so) driven through its own FlowSys API
 * (tests/golden/refflowsys.py run_lusgs, fixture
 * tests/golden/flowsys_lusgs_tet3p.npz); with one slot per (cell, face),
 * sides +1, kappa 1 and no turbulence rows it is the arithmetic of the
 * legacy oracle (lusgs_oracle.c), which tests/test_flowsys_oracle.py checks.
 */
#include <stdint.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define FSL_MAXV 16

void orc_ns_flux(int nfvars, int ndims, double *u, const double *nf, double *f);

/* flux.c:77-90 */
static void fsl_rans_flux(int nfvars, int nturbvars, int ndims, const double *u, const double *nf,
                          double *f)
{
    const double rho = u[0];
    double contrav = 0.0;
    for (int i = 0; i < ndims; ++i) contrav += u[i + 1]*nf[i];
    contrav /= rho;
    for (int i = 0; i < nturbvars; ++i) f[i] = u[nfvars + i]*contrav;
}

ORC_EXPORT void orc_fsl_pack(int nlocal, int neles, int nvars, int nfvars, double turb_factor,
                             double a0, const int32_t *cell_ids, const double *upts,
                             const double *rhs, const double *dt, const double *dsrc, double *u,
                             double *du, double *diag)
{
    for (int e = 0; e < neles; ++e) {
        const int64_t r = cell_ids[e];
        for (int k = 0; k < nvars; ++k) {
            const int64_t ri = r + (int64_t)k*nlocal, ei = e + (int64_t)k*neles;
            const double factor = k < nfvars ? 1.0 : turb_factor;
            u[ri] = upts[ei];
            du[ri] = rhs[ei];
            diag[ri] = 1.0/(dt[e]*factor) + a0 + dsrc[ei];
        }
    }
}

ORC_EXPORT void orc_fsl_update(int nlocal, int neles, int nvars, const int32_t *cell_ids,
                               const double *du, double *upts)
{
    for (int e = 0; e < neles; ++e) {
        const int64_t r = cell_ids[e];
        for (int k = 0; k < nvars; ++k) upts[e + (int64_t)k*neles] += du[r + (int64_t)k*nlocal];
    }
}

ORC_EXPORT void orc_fsl_prepare(int nlocal, int nv0, int nv1, double kappa, const int32_t *rowptr,
                                const int32_t *slots, const double *face_area,
                                const double *rcp_vol, const double *lambdaf, double *diag)
{
    for (int r = 0; r < nlocal; ++r) {
        double sd = 0.0;
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int64_t s = slots[p];
            const double lamf = lambdaf[s]*kappa;
            sd += 0.5*lamf*face_area[s]*rcp_vol[r];
        }
        for (int k = nv0; k < nv1; ++k) diag[r + (int64_t)k*nlocal] += sd;
    }
}

/* one row of lower_sweep (upper = 0) or upper_sweep (upper = 1) */
static void fsl_row(int upper, int turb, double kappa, const double *lambdaf, int nlocal,
                    int nvars, int nfvars, int ndims, int64_t nfaces, int r,
                    const int32_t *rowptr, const int32_t *colidx, const int8_t *sides,
                    const int32_t *slots, const double *face_area, const double *face_normal,
                    const double *rcp_vol, const double *diag, const double *u_r, double *du_r)
{
    const int nv0 = turb ? nfvars : 0, nv1 = turb ? nvars : nfvars, dnv = nv1 - nv0;
    const int ntv = nvars - nfvars;
    double df[FSL_MAXV], dfj[FSL_MAXV], f[FSL_MAXV], u[FSL_MAXV], du[FSL_MAXV], nf[3];
    for (int k = 0; k < dnv; ++k) df[k] = 0.0;
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
        const int64_t nb = colidx[p];
        if (upper ? !(nb > r) : !(nb < r)) continue;
        const int64_t s = slots[p];
        for (int k = 0; k < nvars; ++k) u[k] = u_r[nb + (int64_t)k*nlocal];
        const int8_t side = sides[p];
        for (int d = 0; d < ndims; ++d) nf[d] = side*face_normal[s + (int64_t)d*nfaces];
        const double fv = face_area[s]*rcp_vol[r];
        for (int k = 0; k < nvars; ++k) du[k] = 0.0;
        for (int k = nv0; k < nv1; ++k) du[k] = du_r[nb + (int64_t)k*nlocal];
        /* diff_flux, lusgs.c:143-155 */
        for (int k = 0; k < nvars; ++k) du[k] += u[k];
        if (turb) {
            fsl_rans_flux(nfvars, ntv, ndims, u, nf, f);
            fsl_rans_flux(nfvars, ntv, ndims, du, nf, dfj);
        } else {
            orc_ns_flux(nfvars, ndims, u, nf, f);
            orc_ns_flux(nfvars, ndims, du, nf, dfj);
        }
        for (int k = 0; k < dnv; ++k) dfj[k] -= f[k];
        for (int k = 0; k < dnv; ++k)
            df[k] += (dfj[k] - kappa*lambdaf[s]*du_r[nb + (int64_t)(k + nv0)*nlocal])*fv;
    }
    for (int k = 0; k < dnv; ++k) {
        const int64_t l = r + (int64_t)(k + nv0)*nlocal;
        if (upper) du_r[l] -= (0.5*df[k]/diag[l]);
        else du_r[l] = (du_r[l] - 0.5*df[k])/diag[l];
    }
}

ORC_EXPORT void orc_fsl_sweep(int upper, int turb, double kappa, const double *lambdaf, int nlocal,
                              int nvars, int nfvars, int ndims, int64_t nfaces,
                              const int32_t *rowptr, const int32_t *colidx, const int8_t *sides,
                              const int32_t *slots, const double *face_area,
                              const double *face_normal, const double *rcp_vol,
                              const double *diag, const double *u, double *du)
{
    if (upper) {
        for (int r = nlocal - 1; r > -1; --r)
            fsl_row(1, turb, kappa, lambdaf, nlocal, nvars, nfvars, ndims, nfaces, r, rowptr, colidx,
                    sides, slots, face_area, face_normal, rcp_vol, diag, u, du);
    } else {
        for (int r = 0; r < nlocal; ++r)
            fsl_row(0, turb, kappa, lambdaf, nlocal, nvars, nfvars, ndims, nfaces, r, rowptr, colidx,
                    sides, slots, face_area, face_normal, rcp_vol, diag, u, du);
    }
}
