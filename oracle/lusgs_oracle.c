/*
 * ORACLE — test infrastructure only.  A plain-C CPU restatement of the
 * reference's scalar LU-SGS arithmetic, used as the checker by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in
 * the product library links or calls this file.
 *
 * It restates, in the legacy array layout (nface-major, SoA rows of stride
 * neles), the arithmetic of
 *   src/lusgs/lusgs.c:8-37    lusgs_pack        (diag = 1/(dt*factor) + a0 + dsrc)
 *   src/lusgs/lusgs.c:87-113  pre_lusgs         (diag += sum_f 0.5*lambda*area*rcp_vol)
 *   src/lusgs/lusgs.c:143-155 diff_flux
 *   src/lusgs/lusgs.c:157-205 lower_sweep
 *   src/lusgs/lusgs.c:207-257 upper_sweep
 *   src/lusgs/lusgs.c:54-70   rank_lusgs_update
 *   src/lusgs/flux.c:29-67    ns_flux_container (PMIN clamp included)
 * with the operation order of those lines, so that with
 * -ffp-contract=off the results agree bit for bit with the reference's
 * sequential FlowSys sweeps fed through the renumbering bridge
 * (tests/golden/make_golden.py).
 *
 * "Lower" neighbour: lcolor[nb] < lcolor[i] (colored, cells icolor[n0:ne))
 * or nb < i (serial, all cells ascending); "upper" mirrors it.  Self
 * (boundary) faces are neither.
 *
 * Parity pins: golden fixtures from the compiled reference (lower / upper /
 * update / pre), closed forms for the flux and the first colour, the
 * all-self mesh, and brute-force checks in tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define ORC_GAMMA 1.4
#define ORC_PMIN 1e-13
#define ORC_MAXV 16

/* flux.c:29-67: Euler normal flux with the pressure floor.  u may be
 * modified (energy rewritten when clamped), as in the reference. */
void orc_ns_flux(int nfvars, int ndims, double *u, const double *nf, double *f)
{
    double rho = u[0];
    double et = u[nfvars - 1];
    double temp = 0.0;
    double contrav = 0.0;
    for (int i = 0; i < ndims; ++i) {
        contrav += u[i + 1]*nf[i];
        temp += u[i + 1]*u[i + 1];
    }
    contrav /= rho;
    double p = (ORC_GAMMA - 1.0)*(et - 0.5*temp/rho);
    if (p < ORC_PMIN) {
        p = ORC_PMIN;
        et = p/(ORC_GAMMA - 1.0) + 0.5*temp/rho;
        u[nfvars - 1] = et;
    }
    double ht = et + p;
    f[0] = rho*contrav;
    for (int i = 0; i < ndims; ++i) f[i + 1] = u[i + 1]*contrav + nf[i]*p;
    f[nfvars - 1] = ht*contrav;
}

ORC_EXPORT void orc_flux(int nfvars, int ndims, const double *u_in, const double *nf, double *f)
{
    double u[ORC_MAXV];
    for (int k = 0; k < nfvars; ++k) u[k] = u_in[k];
    orc_ns_flux(nfvars, ndims, u, nf, f);
}

/* lusgs.c:143-155: dfj = F(u + du) - F(u) */
static void orc_diff_flux(int nfvars, int ndims, const double *u0, const double *du0,
                          const double *nf, double *dfj)
{
    double u[ORC_MAXV], du[ORC_MAXV], f[ORC_MAXV];
    for (int k = 0; k < nfvars; ++k) { u[k] = u0[k]; du[k] = du0[k]; }
    for (int k = 0; k < nfvars; ++k) du[k] += u[k];
    orc_ns_flux(nfvars, ndims, u, nf, f);
    orc_ns_flux(nfvars, ndims, du, nf, dfj);
    for (int k = 0; k < nfvars; ++k) dfj[k] -= f[k];
}

/* lusgs.c:33 (pack, a0 = dsrc = 0) then lusgs.c:100-110 (kappa = 1, rcp_vol = 1) */
ORC_EXPORT void orc_pre(int neles, int nface, double factor, const double *fnorm_vol,
                        const double *dt, double *diag, const double *fspr)
{
    const int64_t N = neles;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double d = 1.0/(dt[i]*factor) + 0.0 + 0.0;
        double sd = 0.0;
        for (int f = 0; f < nface; ++f)
            sd += 0.5*fspr[f*N + i]*fnorm_vol[f*N + i]*1.0;
        diag[i] = d + sd;
    }
}

/* one cell of a sweep: df = sum over selected faces of (dF - lam*dQ_nb)*fv */
static inline void orc_face_sum(int64_t N, int nface, int nfvars, int ndims, int64_t i,
                                const int32_t *nei, const double *fnorm_vol,
                                const double *vec_fnorm, const double *upts,
                                const double *x, const double *fspr,
                                int mode, int lower, const int32_t *lcolor, double *df)
{
    double u[ORC_MAXV], du[ORC_MAXV], dfj[ORC_MAXV], nf[3];
    for (int k = 0; k < nfvars; ++k) df[k] = 0.0;
    for (int f = 0; f < nface; ++f) {
        const int64_t nb = nei[f*N + i];
        int take;
        if (mode == 0) take = lower ? (lcolor[nb] < lcolor[i]) : (lcolor[nb] > lcolor[i]);
        else           take = lower ? (nb < i) : (nb > i);
        if (!take) continue;
        for (int k = 0; k < nfvars; ++k) { u[k] = upts[k*N + nb]; du[k] = x[k*N + nb]; }
        for (int d = 0; d < ndims; ++d) nf[d] = 1.0*vec_fnorm[(f*ndims + d)*N + i];
        const double fv = fnorm_vol[f*N + i]*1.0;
        orc_diff_flux(nfvars, ndims, u, du, nf, dfj);
        for (int k = 0; k < nfvars; ++k)
            df[k] += (dfj[k] - 1.0*fspr[f*N + i]*x[k*N + nb])*fv;
    }
}

/*
 * Lower sweep (lusgs.c:175-203): out_i = (in_i - 0.5*df)/diag_i, df taken
 * over lower neighbours using out.  in == out gives the in-place C form.
 * mode 0: coloured over icolor[n0:ne) (OpenMP); mode 1: serial, all cells.
 */
ORC_EXPORT void orc_lower(int mode, int n0, int ne, int neles, int nface, int ndims, int nfvars,
                          const int32_t *nei, const int32_t *icolor, const int32_t *lcolor,
                          const double *fnorm_vol, const double *vec_fnorm, const double *upts,
                          const double *in, double *out, const double *diag, const double *fspr)
{
    const int64_t N = neles;
    if (mode == 0) {
        #pragma omp parallel for schedule(static)
        for (int64_t p = n0; p < ne; ++p) {
            const int64_t i = icolor[p];
            double df[ORC_MAXV];
            orc_face_sum(N, nface, nfvars, ndims, i, nei, fnorm_vol, vec_fnorm, upts, out,
                         fspr, 0, 1, lcolor, df);
            for (int k = 0; k < nfvars; ++k)
                out[k*N + i] = (in[k*N + i] - 0.5*df[k])/diag[i];
        }
    } else {
        for (int64_t i = 0; i < N; ++i) {
            double df[ORC_MAXV];
            orc_face_sum(N, nface, nfvars, ndims, i, nei, fnorm_vol, vec_fnorm, upts, out,
                         fspr, 1, 1, NULL, df);
            for (int k = 0; k < nfvars; ++k)
                out[k*N + i] = (in[k*N + i] - 0.5*df[k])/diag[i];
        }
    }
}

/*
 * Upper sweep (lusgs.c:227-255): out_i = in_i - (0.5*df)/diag_i, df over
 * upper neighbours using out.  in == out gives the in-place C form.
 */
ORC_EXPORT void orc_upper(int mode, int n0, int ne, int neles, int nface, int ndims, int nfvars,
                          const int32_t *nei, const int32_t *icolor, const int32_t *lcolor,
                          const double *fnorm_vol, const double *vec_fnorm, const double *upts,
                          const double *in, double *out, const double *diag, const double *fspr)
{
    const int64_t N = neles;
    if (mode == 0) {
        #pragma omp parallel for schedule(static)
        for (int64_t p = n0; p < ne; ++p) {
            const int64_t i = icolor[p];
            double df[ORC_MAXV];
            orc_face_sum(N, nface, nfvars, ndims, i, nei, fnorm_vol, vec_fnorm, upts, out,
                         fspr, 0, 0, lcolor, df);
            for (int k = 0; k < nfvars; ++k)
                out[k*N + i] = in[k*N + i] - (0.5*df[k]/diag[i]);
        }
    } else {
        for (int64_t i = N - 1; i >= 0; --i) {
            double df[ORC_MAXV];
            orc_face_sum(N, nface, nfvars, ndims, i, nei, fnorm_vol, vec_fnorm, upts, out,
                         fspr, 1, 0, NULL, df);
            for (int k = 0; k < nfvars; ++k)
                out[k*N + i] = in[k*N + i] - (0.5*df[k]/diag[i]);
        }
    }
}

/* lusgs.c:62-68: upts[k,i] += x[k,i] for k < nvars */
ORC_EXPORT void orc_update(int neles, int nvars, double *upts, const double *x)
{
    const int64_t n = (int64_t)neles*nvars;
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) upts[j] += x[j];
}

/*
 * One full coloured LU-SGS iteration in the C in-place form:
 * pre, lower over colours ascending, upper over colours descending, update.
 * color_off[c]..color_off[c+1] are the icolor ranges of colour c+1.
 * dub holds rhs on entry.  Used by the CPU baseline.
 */
ORC_EXPORT void orc_colored_iteration(int neles, int nface, int ndims, int nfvars, int nvars,
                                      int ncolors, const int32_t *color_off,
                                      const int32_t *nei, const int32_t *icolor,
                                      const int32_t *lcolor, const double *fnorm_vol,
                                      const double *vec_fnorm, double *upts, double *dub,
                                      double *diag, const double *fspr, const double *dt)
{
    orc_pre(neles, nface, 1.0, fnorm_vol, dt, diag, fspr);
    for (int c = 0; c < ncolors; ++c)
        orc_lower(0, color_off[c], color_off[c + 1], neles, nface, ndims, nfvars, nei, icolor,
                  lcolor, fnorm_vol, vec_fnorm, upts, dub, dub, diag, fspr);
    for (int c = ncolors - 1; c >= 0; --c)
        orc_upper(0, color_off[c], color_off[c + 1], neles, nface, ndims, nfvars, nei, icolor,
                  lcolor, fnorm_vol, vec_fnorm, upts, dub, dub, diag, fspr);
    orc_update(neles, nvars, upts, dub);
}
