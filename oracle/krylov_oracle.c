/*
 * ORACLE — test infrastructure only (see lusgs_oracle.c for the rules).
 *
 * Single-threaded CPU restatement of the reference's Krylov path, in the
 * operation order of:
 *   src/krylov/matrix/sparsemat.c:57-101   SpMV_BSR  (y = alpha*A*x + beta*y)
 *   src/krylov/utils/inverse.c:7-71        ludcmp / lusub (Doolittle, no pivoting)
 *   src/krylov/precon/blusgs.c:7-100       BLUSGSPreconPrepare / Apply (sequential)
 *   src/krylov/precon/pblusgs.c:7-104      coloured lower / upper apply, colours
 *                                          ascending then descending
 *   src/krylov/utils/blas_basic.c:4-61     dot, nrm2, axpy, scal, gemv (col-major)
 *   src/krylov/solver/gmres.c:4-129        restarted GMRES, CGS2, Givens
 *   src/krylov/solver/bicgstab.c:5-131     BiCGSTAB with the fused 5-way reduction
 * The solvers also take the ILU(0) family as preconditioner (ilu_oracle.c:
 * factored values of a separate array, as the reference's preconditioner
 * factors its own copy).
 * Deviations, each recorded: the coloured lower apply uses the diagonal block
 * of its own row (pblusgs.c:41 indexes diagvalues[idx+blkdim], a defect);
 * GMRES has an additional right-preconditioned / relative-tolerance mode
 * (right = 1), which the reference lacks.
 *
 * Pins: bit-exact against oracle/_ref (the reference compiled from its own
 * sources, OpenMP at one thread) for SpMV, prepare, sequential apply, GMRES
 * and BiCGSTAB residual histories (with block LU-SGS, none, and block ILU(0)
 * preconditioners); scipy bsr_matrix / numpy.linalg.solve
 * closed forms; coloured apply = sequential apply on a colour-ordered matrix.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define KMAXB 8

/* sparsemat.c:57-101 */
ORC_EXPORT void orc_spmv_bsr(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                             const double *values, double alpha, const double *x, double beta,
                             double *y)
{
    const int blk2 = blk*blk;
    for (int idx = 0; idx < bn; ++idx) {
        double arr[KMAXB], submat[KMAXB*KMAXB], xprt[KMAXB];
        for (int k = 0; k < blk; ++k) arr[k] = 0.0;
        for (int j = rowptr[idx]; j < rowptr[idx + 1]; ++j) {
            const int ldx = colidx[j];
            for (int row = 0; row < blk; ++row) {
                for (int col = 0; col < blk; ++col)
                    submat[row*blk + col] = values[(int64_t)j*blk2 + row*blk + col];
                xprt[row] = x[(int64_t)ldx*blk + row];
            }
            for (int row = 0; row < blk; ++row) {
                double v = 0.0;
                for (int col = 0; col < blk; ++col) v += submat[row*blk + col]*xprt[col];
                arr[row] += v;
            }
        }
        for (int k = 0; k < blk; ++k)
            y[(int64_t)idx*blk + k] = alpha*arr[k] + beta*y[(int64_t)idx*blk + k];
    }
}

/* inverse.c:7-42 */
ORC_EXPORT void orc_ludcmp(int block, double *A)
{
    if (block == 1) {
        A[0] = 1.0/A[0];
        return;
    }
    for (int row = 1; row < block; ++row) {
        const int nrow = block*row;
        A[nrow] /= A[0];
        for (int col = 1; col < block; ++col) {
            if (row > col) {
                double val = 0.0;
                for (int k = 0; k < col; ++k) val += A[nrow + k]*A[col + block*k];
                A[nrow + col] = (A[nrow + col] - val)/A[(block + 1)*col];
            } else {
                double val = 0.0;
                for (int k = 0; k < row; ++k) val += A[nrow + k]*A[block*k + col];
                A[nrow + col] -= val;
            }
        }
    }
}

/* inverse.c:45-71 */
ORC_EXPORT void orc_lusub(int block, const double *LU, double *b)
{
    if (block == 1) {
        b[0] *= LU[0];
        return;
    }
    for (int row = 1; row < block; ++row) {
        const int nrow = row*block;
        double val = 0.0;
        for (int col = 0; col < row; ++col) val += LU[nrow + col]*b[col];
        b[row] -= val;
    }
    b[block - 1] /= LU[block*block - 1];
    for (int row = block - 2; row > -1; --row) {
        const int nrow = row*block;
        double val = 0.0;
        for (int col = row + 1; col < block; ++col) val += LU[nrow + col]*b[col];
        b[row] = (b[row] - val)/LU[nrow + row];
    }
}

/* precon/blusgs.c:7-26 */
ORC_EXPORT void orc_blusgs_prepare(int bn, int blk, const int32_t *diagslots, const double *values,
                                   double *diagvalues)
{
    const int blk2 = blk*blk;
    for (int idx = 0; idx < bn; ++idx) {
        memcpy(&diagvalues[(int64_t)idx*blk2], &values[(int64_t)diagslots[idx]*blk2],
               sizeof(double)*blk2);
        orc_ludcmp(blk, &diagvalues[(int64_t)idx*blk2]);
    }
}

static void lower_row(int idx, int blk, const int32_t *rowptr, const int32_t *colidx,
                      const int32_t *diagslots, const double *values, const double *diagvalues,
                      double *b)
{
    const int blk2 = blk*blk;
    double arr[KMAXB];
    for (int k = 0; k < blk; ++k) arr[k] = b[(int64_t)idx*blk + k];
    for (int j = rowptr[idx]; j < diagslots[idx]; ++j) {
        const int c = colidx[j];
        for (int row = 0; row < blk; ++row) {
            double v = 0.0;
            for (int col = 0; col < blk; ++col)
                v += values[(int64_t)j*blk2 + row*blk + col]*b[(int64_t)c*blk + col];
            arr[row] -= v;
        }
    }
    orc_lusub(blk, &diagvalues[(int64_t)idx*blk2], arr);
    for (int k = 0; k < blk; ++k) b[(int64_t)idx*blk + k] = arr[k];
}

static void upper_row(int idx, int blk, const int32_t *rowptr, const int32_t *colidx,
                      const int32_t *diagslots, const double *values, const double *diagvalues,
                      double *b)
{
    const int blk2 = blk*blk;
    double arr[KMAXB];
    for (int k = 0; k < blk; ++k) arr[k] = 0.0;
    for (int j = diagslots[idx] + 1; j < rowptr[idx + 1]; ++j) {
        const int c = colidx[j];
        for (int row = 0; row < blk; ++row) {
            double v = 0.0;
            for (int col = 0; col < blk; ++col)
                v += values[(int64_t)j*blk2 + row*blk + col]*b[(int64_t)c*blk + col];
            arr[row] += v;
        }
    }
    orc_lusub(blk, &diagvalues[(int64_t)idx*blk2], arr);
    for (int k = 0; k < blk; ++k) b[(int64_t)idx*blk + k] -= arr[k];
}

/* precon/blusgs.c:28-100 (sequential) when ncolors == 0; otherwise the
 * coloured form of pblusgs.c:90-104, colour c = rows icolors[c]..icolors[c+1]-1
 * (matrix numbered in colour order, columns sorted). */
ORC_EXPORT void orc_blusgs_apply(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                                 const int32_t *diagslots, const double *values,
                                 const double *diagvalues, int ncolors, const int32_t *icolors,
                                 double *b)
{
    if (ncolors == 0) {
        for (int idx = 0; idx < bn; ++idx)
            lower_row(idx, blk, rowptr, colidx, diagslots, values, diagvalues, b);
        for (int idx = bn - 1; idx > -1; --idx)
            upper_row(idx, blk, rowptr, colidx, diagslots, values, diagvalues, b);
        return;
    }
    for (int c = 0; c < ncolors; ++c)
        for (int idx = icolors[c]; idx < icolors[c + 1]; ++idx)
            lower_row(idx, blk, rowptr, colidx, diagslots, values, diagvalues, b);
    for (int c = ncolors - 1; c >= 0; --c)
        for (int idx = icolors[c]; idx < icolors[c + 1]; ++idx)
            upper_row(idx, blk, rowptr, colidx, diagslots, values, diagvalues, b);
}

/* ------------------------------------------------------------ BLAS-1/2 */
static double o_dot(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i]*y[i];
    return s;
}
static double o_nrm2(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i]*x[i];
    return sqrt(s);
}
static void o_axpy(int n, double a, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] += a*x[i];
}
static void o_scal(int n, double a, double *x)
{
    for (int i = 0; i < n; ++i) x[i] = a*x[i];
}
/* y = alpha*A*x + beta*y, A m x k column-major with lda (blas_basic.c:41-50) */
static void o_gemv(int m, int k, int lda, double alpha, const double *a, const double *x,
                   double beta, double *y)
{
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += a[i + (int64_t)j*lda]*x[j];
        y[i] = alpha*s + beta*y[i];
    }
}
/* y = alpha*A^T*x + beta*y (blas_basic.c:52-61) */
static void o_gemvt(int m, int k, int lda, double alpha, const double *a, const double *x,
                    double beta, double *y)
{
    for (int j = 0; j < k; ++j) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += a[i + (int64_t)j*lda]*x[i];
        y[j] = alpha*s + beta*y[j];
    }
}

typedef struct {
    int bn, blk;
    const int32_t *rowptr, *colidx, *diagslots;
    const double *values;
    const double *diagvalues;   /* NULL: no preconditioner */
    int ncolors;
    const int32_t *icolors;
    int pckind;                 /* 0 block LU-SGS (diagvalues), 1 block ILU(0), 2 ILU(0) */
    const double *pcvalues;     /* ILU kinds: the factored values */
} OrcSystem;

void orc_ilu_apply(int n, const int32_t *rowptr, const int32_t *colidx, const int32_t *diagslots,
                   const double *values, double *b);
void orc_bilu_apply(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                    const int32_t *diagslots, const double *values, double *b);
void orc_pbilu_apply(int bn, int blk, int ncolors, const int32_t *icolors, const int32_t *rowptr,
                     const int32_t *colidx, const int32_t *diagslots, const double *values,
                     double *b);

static void o_apply(const OrcSystem *S, double *v)
{
    if (S->pckind == 1) {
        if (S->ncolors > 0)
            orc_pbilu_apply(S->bn, S->blk, S->ncolors, S->icolors, S->rowptr, S->colidx,
                            S->diagslots, S->pcvalues, v);
        else
            orc_bilu_apply(S->bn, S->blk, S->rowptr, S->colidx, S->diagslots, S->pcvalues, v);
        return;
    }
    if (S->pckind == 2) {
        orc_ilu_apply(S->bn, S->rowptr, S->colidx, S->diagslots, S->pcvalues, v);
        return;
    }
    if (S->diagvalues)
        orc_blusgs_apply(S->bn, S->blk, S->rowptr, S->colidx, S->diagslots, S->values,
                         S->diagvalues, S->ncolors, S->icolors, v);
}

static void o_spmv(const OrcSystem *S, double alpha, const double *x, double beta, double *y)
{
    orc_spmv_bsr(S->bn, S->blk, S->rowptr, S->colidx, S->values, alpha, x, beta, y);
}

/*
 * gmres.c:42-129.  right == 0: the reference (left preconditioning, absolute
 * tolerance on the true residual at the start of each restart, maxiter counts
 * restarts).  right == 1: flexible right preconditioning (z_j = M^-1 v_j
 * kept, x += Z y), tolerance tol*||b||, early exit inside a cycle when the
 * Givens residual estimate drops below it.
 * Returns the restart count; stat: 0 converged, 1 happy breakdown,
 * 2 maxiter; hist[k] = residual at restart k (maxiter entries).
 */
ORC_EXPORT int orc_gmres(const OrcSystem *S, int m, int maxiter, double tol, int right,
                         double *x, const double *b, int *stat, double *res_out, double *hist)
{
    const int n = S->bn*S->blk;
    const double haptol = 1e-12;
    const int ld = m + 1;
    double *V = calloc((size_t)n*(m + 1), sizeof(double));
    double *H = calloc((size_t)(m + 1)*m, sizeof(double));
    double *y = calloc(m + 1, sizeof(double)), *cs = calloc(m + 1, sizeof(double));
    double *sn = calloc(m + 1, sizeof(double)), *h2 = calloc(m + 1, sizeof(double));
    double *r = calloc(n, sizeof(double)), *w = calloc(n, sizeof(double));
    double *Z = right ? calloc((size_t)n*m, sizeof(double)) : NULL;
    int iter = 0, st = -1;
    double abeta = 0.0;
    const double tolv = right ? tol*o_nrm2(n, b) : tol;
    while (iter < maxiter) {
        memcpy(r, b, sizeof(double)*n);
        o_spmv(S, -1.0, x, 1.0, r);
        abeta = o_nrm2(n, r);
        if (hist) hist[iter] = abeta;
        if (abeta <= tolv) { st = 0; break; }
        if (!right) o_apply(S, r);
        const double beta = o_nrm2(n, r);
        memset(y, 0, sizeof(double)*(m + 1));
        y[0] = beta;
        memcpy(V, r, sizeof(double)*n);
        o_scal(n, 1.0/beta, V);
        int k = 0, happy = 0;
        for (int j = 0; j < m; ++j) {
            const int kk = j + 1;
            double *vj = V + (size_t)j*n, *Hcol = H + (size_t)j*ld;
            if (right) {
                double *zj = Z + (size_t)j*n;
                memcpy(zj, vj, sizeof(double)*n);
                o_apply(S, zj);
                o_spmv(S, 1.0, zj, 0.0, w);
            } else {
                o_spmv(S, 1.0, vj, 0.0, w);
                o_apply(S, w);
            }
            o_gemvt(n, kk, n, 1.0, V, w, 0.0, Hcol);
            o_gemv(n, kk, n, -1.0, V, Hcol, 1.0, w);
            o_gemvt(n, kk, n, 1.0, V, w, 0.0, h2);
            o_gemv(n, kk, n, -1.0, V, h2, 1.0, w);
            o_axpy(kk, 1.0, h2, Hcol);
            const double hsub = o_nrm2(n, w);
            Hcol[j + 1] = hsub;
            double *vn = V + (size_t)(j + 1)*n;
            memcpy(vn, w, sizeof(double)*n);
            o_scal(n, 1.0/hsub, vn);
            for (int i = 0; i < j; ++i) {
                const double c = cs[i], s = sn[i];
                const double t = c*Hcol[i] + s*Hcol[i + 1];
                Hcol[i + 1] = -s*Hcol[i] + c*Hcol[i + 1];
                Hcol[i] = t;
            }
            const double h1 = Hcol[j], hh2 = Hcol[j + 1];
            const double rr = hypot(h1, hh2);
            cs[j] = h1/rr;
            sn[j] = hh2/rr;
            Hcol[j] = cs[j]*Hcol[j] + sn[j]*Hcol[j + 1];
            Hcol[j + 1] = 0.0;
            y[j + 1] = -sn[j]*y[j];
            y[j] = cs[j]*y[j];
            k = j + 1;
            if (hsub < haptol*beta) { happy = 1; break; }
            if (right && fabs(y[j + 1]) <= tolv) break;
        }
        for (int i = k - 1; i >= 0; --i) {
            double sum = y[i];
            for (int jj = i + 1; jj < k; ++jj) sum -= H[i + (size_t)jj*ld]*y[jj];
            y[i] = sum/H[i + (size_t)i*ld];
        }
        if (right) {
            o_gemv(n, k, n, 1.0, Z, y, 1.0, x);
        } else {
            o_gemv(n, k, n, 1.0, V, y, 1.0, x);
        }
        if (happy) st = 1;
        iter++;
    }
    if (iter == maxiter) st = 2;
    *stat = st;
    *res_out = abeta;
    free(V); free(H); free(y); free(cs); free(sn); free(h2); free(r); free(w); free(Z);
    return iter;
}

/* bicgstab.c:24-131.  relative != 0 uses tol*||b|| (an extension).  Returns
 * the iteration count; stat 0 converged, 2 maxiter, 3 rho breakdown,
 * 4 pi breakdown; hist[it] = residual estimate of iteration it. */
ORC_EXPORT int orc_bicgstab(const OrcSystem *S, int maxiter, double tol, int relative, double *x,
                            const double *b, int *stat, double *res_out, double *hist)
{
    const int n = S->bn*S->blk;
    double *r = calloc(n, 8), *rt = calloc(n, 8), *p = calloc(n, 8), *pt = calloc(n, 8);
    double *v = calloc(n, 8), *s = calloc(n, 8), *shat = calloc(n, 8), *t = calloc(n, 8);
    int iter = 0, st = -1;
    double rho, rhoprev, alpha, beta, omega, resnorm = 0.0;
    const double tolv = relative ? tol*o_nrm2(n, b) : tol;
    memcpy(r, b, 8*(size_t)n);
    o_spmv(S, -1.0, x, 1.0, r);
    memcpy(rt, r, 8*(size_t)n);
    rho = o_dot(n, rt, r);
    rhoprev = 1.0;
    alpha = 1.0;
    omega = 1.0;
    if (fabs(rho) < 1e-30) {
        st = 3;
        goto done;
    }
    while (iter < maxiter) {
        if (iter == 0) memcpy(p, r, 8*(size_t)n);
        else {
            beta = (rho/rhoprev)*(alpha/omega);
            o_axpy(n, -omega, v, p);
            o_scal(n, beta, p);
            o_axpy(n, 1.0, r, p);
        }
        memcpy(pt, p, 8*(size_t)n);
        o_apply(S, pt);
        o_spmv(S, 1.0, pt, 0.0, v);
        const double pi = o_dot(n, rt, v);
        if (fabs(pi) < 1e-30) { st = 4; break; }
        alpha = rho/pi;
        memcpy(s, r, 8*(size_t)n);
        o_axpy(n, -alpha, v, s);
        memcpy(shat, s, 8*(size_t)n);
        o_apply(S, shat);
        o_spmv(S, 1.0, shat, 0.0, t);
        double a = 0, bb = 0, c = 0, d = 0, e = 0;
        for (int k = 0; k < n; ++k) {
            const double sk = s[k], tk = t[k], rk = rt[k];
            a += rk*sk; bb += rk*tk; c += tk*sk; d += tk*tk; e += sk*sk;
        }
        omega = c/d;
        o_axpy(n, alpha, pt, x);
        o_axpy(n, omega, shat, x);
        memcpy(r, s, 8*(size_t)n);
        o_axpy(n, -omega, t, r);
        const double rho_new = a - omega*bb;
        resnorm = sqrt(fabs(e - 2.0*omega*c + omega*omega*d));
        if (hist) hist[iter] = resnorm;
        if (resnorm < tolv) { st = 0; goto done; }
        rhoprev = rho;
        rho = rho_new;
        iter++;
    }
done:
    if (iter == maxiter) st = 2;
    *stat = st;
    *res_out = resnorm;
    free(r); free(rt); free(p); free(pt); free(v); free(s); free(shat); free(t);
    return iter;
}

/* system handles for ctypes callers */
ORC_EXPORT int orc_gmres_solve(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                               const int32_t *diagslots, const double *values,
                               const double *diagvalues, int ncolors, const int32_t *icolors,
                               int m, int maxiter, double tol, int right, double *x,
                               const double *b, int *stat, double *res, double *hist)
{
    OrcSystem S = {bn, blk, rowptr, colidx, diagslots, values, diagvalues, ncolors, icolors};
    return orc_gmres(&S, m, maxiter, tol, right, x, b, stat, res, hist);
}

ORC_EXPORT int orc_bicgstab_solve(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                                  const int32_t *diagslots, const double *values,
                                  const double *diagvalues, int ncolors, const int32_t *icolors,
                                  int maxiter, double tol, int relative, double *x,
                                  const double *b, int *stat, double *res, double *hist)
{
    OrcSystem S = {bn, blk, rowptr, colidx, diagslots, values, diagvalues, ncolors, icolors};
    return orc_bicgstab(&S, maxiter, tol, relative, x, b, stat, res, hist);
}

/* the solvers with an ILU-family preconditioner: pckind 1 block ILU(0)
 * (coloured when ncolors > 0), 2 scalar ILU(0); pcvalues = factored values */
ORC_EXPORT int orc_gmres_solve_ilu(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                                   const int32_t *diagslots, const double *values, int pckind,
                                   const double *pcvalues, int ncolors, const int32_t *icolors,
                                   int m, int maxiter, double tol, int right, double *x,
                                   const double *b, int *stat, double *res, double *hist)
{
    OrcSystem S = {bn, blk, rowptr, colidx, diagslots, values, NULL, ncolors, icolors, pckind,
                   pcvalues};
    return orc_gmres(&S, m, maxiter, tol, right, x, b, stat, res, hist);
}

ORC_EXPORT int orc_bicgstab_solve_ilu(int bn, int blk, const int32_t *rowptr,
                                      const int32_t *colidx, const int32_t *diagslots,
                                      const double *values, int pckind, const double *pcvalues,
                                      int ncolors, const int32_t *icolors, int maxiter,
                                      double tol, int relative, double *x, const double *b,
                                      int *stat, double *res, double *hist)
{
    OrcSystem S = {bn, blk, rowptr, colidx, diagslots, values, NULL, ncolors, icolors, pckind,
                   pcvalues};
    return orc_bicgstab(&S, maxiter, tol, relative, x, b, stat, res, hist);
}
