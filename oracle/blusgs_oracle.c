/*
 * ORACLE — test infrastructure only (see lusgs_oracle.c for the rules).
 *
 * Single-threaded CPU restatement of the reference's FlowSys block LU-SGS
 * (src/lusgs/blusgs.c), in its operation order:
 *   orc_fs_pack      rank_blusgs_pack   blusgs.c:17-45
 *   orc_fs_tpack     rank_tblusgs_pack  blusgs.c:60-94 with the source
 *                    Jacobians of source_jacobian.c:6-18 (kind 1 = SA,
 *                    2 = k-omega SST)
 *   orc_fs_update    rank_blusgs_update blusgs.c:126-143
 *   orc_fs_subres    rank_sub_residual  blusgs.c:161-185
 *   orc_fs_prepare   pre_blusgs         blusgs.c:200-246 (LU of
 *                    inverse.h:13-48)
 *   orc_fs_sweep     lower_sweep / upper_sweep blusgs.c:273-373 (dir 0 =
 *                    rows ascending, 1 = descending; in place on du)
 * Layouts: element arrays [nvars, neles]; rank rhs / du / dup [nlocal, nvars];
 * diag [nlocal, nv, nv]; jmat [2][nv][nv][nfaces].
 *
 * Pins: bit-exact against the reference compiled from src/lusgs/*.c
 * (oracle/_ref/libref_lusgs.so) driven through its own FlowSys API
 * (tests/golden/refflowsys.py, fixture tests/golden/flowsys_tet3p.npz);
 * closed forms: one lower + upper pass from du = 0 solves
 * (D + L) D^-1 (D + U) du = rhs on the block matrix the positions define,
 * and the converged Gauss-Seidel iterate solves the full block system.
 */
#include <stdint.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define FMAXV 8

void orc_ludcmp(int block, double *A);
void orc_lusub(int block, const double *LU, double *b);

ORC_EXPORT void orc_fs_pack(int neles, int nvars, int nfvars, double a0, const int32_t *cell_ids,
                            const double *rhs, const double *dt, double *rank_rhs, double *diag)
{
    const int d2 = nfvars*nfvars;
    for (int e = 0; e < neles; ++e) {
        const int64_t r = cell_ids[e];
        for (int k = 0; k < nvars; ++k) rank_rhs[k + r*nvars] = rhs[e + (int64_t)k*neles];
        for (int row = 0; row < nfvars; ++row) {
            double *drow = diag + r*d2 + row*nfvars;
            for (int col = 0; col < nfvars; ++col) drow[col] = 0.0;
            drow[row] = 1/dt[e] + a0;
        }
    }
}

ORC_EXPORT void orc_fs_tpack(int neles, int nvars, int nfvars, double a0, double factor, int kind,
                             const int32_t *cell_ids, const double *uptsb, const double *dsrc,
                             const double *dt, double *tdiag)
{
    const int nt = nvars - nfvars, d2 = nt*nt;
    double u[16], d[16];
    for (int e = 0; e < neles; ++e) {
        const int64_t r = cell_ids[e];
        double *A = tdiag + r*d2;
        for (int q = 0; q < d2; ++q) A[q] = 0.0;
        for (int k = 0; k < nvars; ++k) {
            u[k] = uptsb[e + (int64_t)k*neles];
            d[k] = dsrc[e + (int64_t)k*neles];
        }
        if (kind == 1) {
            A[0] += d[nvars - 1];
        } else if (kind == 2) {
            const double kk = u[nvars - 2]/u[0];
            const double bk = 0.09*kk;
            A[0] += d[nvars - 2];
            A[1] += bk > 0.0 ? bk : 0.0;
            A[1 + nt] = d[nvars - 1];
        }
        for (int row = 0; row < nt; ++row) A[row + row*nt] += 1/(dt[e]*factor) + a0;
    }
}

ORC_EXPORT void orc_fs_update(int neles, int nvars, const int32_t *cell_ids, const double *du,
                              double *upts)
{
    for (int e = 0; e < neles; ++e)
        for (int k = 0; k < nvars; ++k)
            upts[e + (int64_t)k*neles] += du[k + (int64_t)cell_ids[e]*nvars];
}

ORC_EXPORT void orc_fs_subres(int neles, int nvars, const int32_t *cell_ids, const double *vol,
                              const double *du, double *dup, double *res)
{
    for (int e = 0; e < neles; ++e) {
        const int64_t r = cell_ids[e];
        for (int k = 0; k < nvars; ++k) {
            const int64_t o = k + r*nvars;
            const double diff = du[o] - dup[o];
            res[e + (int64_t)k*neles] = diff*diff*vol[e];
            dup[o] = du[o];
        }
    }
}

/* jmat entry (half, row, col) of face slot s */
static inline double jm(const double *jmat, int nv, int64_t nfaces, int half, int row, int col,
                        int64_t s)
{
    return jmat[s + col*nfaces + row*nv*nfaces + half*(int64_t)nv*nv*nfaces];
}

ORC_EXPORT void orc_fs_prepare(int nlocal, int nv, int64_t nfaces, const int32_t *rowptr,
                               const int32_t *slots, const int8_t *sides,
                               const double *face_area, const double *rcp_vol, double *diag,
                               const double *jmat)
{
    const int d2 = nv*nv;
    double dm[FMAXV*FMAXV];
    for (int64_t r = 0; r < nlocal; ++r) {
        memcpy(dm, diag + r*d2, sizeof(double)*d2);
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int64_t s = slots[p];
            const double fv = face_area[s]*rcp_vol[r];
            for (int row = 0; row < nv; ++row)
                for (int col = 0; col < nv; ++col) {
                    const double v = sides[p] == 1 ? jm(jmat, nv, nfaces, 0, row, col, s)
                                                   : -jm(jmat, nv, nfaces, 1, row, col, s);
                    dm[col + row*nv] += v*fv;
                }
        }
        orc_ludcmp(nv, dm);
        memcpy(diag + r*d2, dm, sizeof(double)*d2);
    }
}

static void fs_row(int64_t r, int var0, int nv, int nvars, int64_t nfaces, const int32_t *rowptr,
                   const int32_t *colidx, const int8_t *sides, const int32_t *slots,
                   const double *face_area, const double *rcp_vol, const double *rhsb,
                   const double *diag, const double *jmat, double *du)
{
    double x[FMAXV];
    for (int row = 0; row < nv; ++row) x[row] = rhsb[var0 + row + r*nvars];
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
        const int64_t nb = colidx[p], s = slots[p];
        const double fv = face_area[s]*rcp_vol[r];
        for (int row = 0; row < nv; ++row) {
            double val = 0.0;
            for (int col = 0; col < nv; ++col) {
                const double j = sides[p] == 1 ? jm(jmat, nv, nfaces, 1, row, col, s)
                                               : -jm(jmat, nv, nfaces, 0, row, col, s);
                val += j*du[var0 + col + nb*nvars];
            }
            x[row] -= val*fv;
        }
    }
    orc_lusub(nv, diag + r*nv*nv, x);
    for (int row = 0; row < nv; ++row) du[var0 + row + r*nvars] = x[row];
}

ORC_EXPORT void orc_fs_sweep(int dir, int var0, int nv, int nlocal, int nvars, int64_t nfaces,
                             const int32_t *rowptr, const int32_t *colidx, const int8_t *sides,
                             const int32_t *slots, const double *face_area,
                             const double *rcp_vol, const double *rhsb, const double *diag,
                             const double *jmat, double *du)
{
    if (dir == 0)
        for (int64_t r = 0; r < nlocal; ++r)
            fs_row(r, var0, nv, nvars, nfaces, rowptr, colidx, sides, slots, face_area, rcp_vol,
                   rhsb, diag, jmat, du);
    else
        for (int64_t r = nlocal - 1; r > -1; --r)
            fs_row(r, var0, nv, nvars, nfaces, rowptr, colidx, sides, slots, face_area, rcp_vol,
                   rhsb, diag, jmat, du);
}
