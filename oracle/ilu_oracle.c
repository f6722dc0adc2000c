/*
 * ORACLE — test infrastructure only (see lusgs_oracle.c for the rules).
 *
 * Single-threaded CPU restatement of the reference's incomplete-LU
 * preconditioners of the Krylov API, in their operation order:
 *   orc_ilu_prepare    ILUPreconPrepare   precon/ilu.c:4-48 (ILU(0), in place)
 *   orc_ilu_apply      ILUPreconApply     precon/ilu.c:50-86, backward sweep
 *                      over the strictly upper entries (see below)
 *   orc_bilu_prepare   BILUPreconPrepare  precon/bilu.c:5-65 (block ILU(0),
 *                      lusubmattrans of utils/inverse.c:77-107, ludcmp of the
 *                      diagonal block at the end of each row)
 *   orc_bilu_apply     BILUPreconApply    precon/bilu.c:67-130
 *   orc_pbilu_apply    PBILUPreconApply   precon/pbilu.c:80-180 (lower colour
 *                      ranges ascending, upper colour ranges descending, rows
 *                      ascending inside a range)
 * PBILUPreconPrepare (pbilu.c:5-76) at one thread visits the rows of the
 * colour ranges in ascending order, i.e. it is orc_bilu_prepare.
 *
 * The reference's scalar backward sweep starts at the diagonal entry
 * (ilu.c:81: jdx = dd), subtracting values[dd]*b[i] before dividing by
 * values[dd]; the block form (bilu.c:108) and the definition of ILU(0)
 * start after it.  This restatement uses dd + 1 and the GPU library does the
 * same; tests/test_ilu_oracle.py shows that the reference's output is this
 * restatement's with the extra diagonal term and nothing else.
 *
 * Pins: factorisations and forward sweeps bit-exact against the reference
 * compiled from src/krylov (oracle/_ref/libref_krylov.so, fixture
 * tests/golden/ilu_tet3p.npz); applies equal to the dense triangular solves
 * with the factors (closed form).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))
#define ILU_MAXB 8

void orc_ludcmp(int block, double *A);
void orc_lusub(int block, const double *LU, double *b);

/* inverse.c:77-107: B := B inv(LU), LU the factored block */
ORC_EXPORT void orc_lusubmattrans(int block, const double *LU, double *B)
{
    if (block == 1) {
        B[0] *= LU[0];
        return;
    }
    for (int scol = 0; scol < block; ++scol) B[scol*block] /= LU[0];
    for (int row = 1; row < block; ++row)
        for (int scol = 0; scol < block; ++scol) {
            double val = 0.0;
            for (int col = 0; col < row; ++col) val += B[scol*block + col]*LU[col*block + row];
            B[scol*block + row] = (B[scol*block + row] - val)/LU[row*block + row];
        }
    for (int row = block - 2; row > -1; --row)
        for (int scol = 0; scol < block; ++scol) {
            double val = 0.0;
            for (int col = row + 1; col < block; ++col) val += B[scol*block + col]*LU[col*block + row];
            B[scol*block + row] -= val;
        }
}

/* iw[col] = the position of col in row ck (the last one when repeated), -1 otherwise */
ORC_EXPORT void orc_ilu_prepare(int n, const int32_t *rowptr, const int32_t *colidx,
                                const int32_t *diagslots, double *values)
{
    int32_t *iw = malloc(sizeof(int32_t)*(size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) iw[i] = -1;
    for (int i = 0; i < n; ++i) {
        const int st = rowptr[i], ed = diagslots[i], jed = rowptr[i + 1];
        for (int k = st; k < ed; ++k) {
            const int ck = colidx[k], kk = diagslots[ck], kst = rowptr[ck], ked = rowptr[ck + 1];
            values[k] /= values[kk];
            const double aik = values[k];
            for (int j = kst; j < ked; ++j) iw[colidx[j]] = j;
            for (int j = k + 1; j < jed; ++j) {
                const int iwj = iw[colidx[j]];
                if (iwj != -1) values[j] -= aik*values[iwj];
            }
            for (int j = kst; j < ked; ++j) iw[colidx[j]] = -1;
        }
    }
    free(iw);
}

ORC_EXPORT void orc_ilu_apply(int n, const int32_t *rowptr, const int32_t *colidx,
                              const int32_t *diagslots, const double *values, double *b)
{
    for (int i = 0; i < n; ++i) {
        double v = b[i];
        for (int j = rowptr[i]; j < diagslots[i]; ++j) v -= values[j]*b[colidx[j]];
        b[i] = v;
    }
    for (int i = n - 1; i > -1; --i) {
        const int dd = diagslots[i];
        double v = b[i];
        for (int j = dd + 1; j < rowptr[i + 1]; ++j) v -= values[j]*b[colidx[j]];
        b[i] = v/values[dd];
    }
}

ORC_EXPORT void orc_bilu_prepare(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                                 const int32_t *diagslots, double *values)
{
    const int b2 = blk*blk;
    double aik[ILU_MAXB*ILU_MAXB];
    int32_t *iw = malloc(sizeof(int32_t)*(size_t)(bn > 0 ? bn : 1));
    for (int i = 0; i < bn; ++i) iw[i] = -1;
    for (int i = 0; i < bn; ++i) {
        const int st = rowptr[i], ed = diagslots[i], jed = rowptr[i + 1];
        for (int k = st; k < ed; ++k) {
            const int ck = colidx[k], kk = diagslots[ck], kst = rowptr[ck], ked = rowptr[ck + 1];
            orc_lusubmattrans(blk, values + (int64_t)kk*b2, values + (int64_t)k*b2);
            memcpy(aik, values + (int64_t)k*b2, sizeof(double)*b2);
            for (int j = kst; j < ked; ++j) iw[colidx[j]] = j;
            for (int j = k + 1; j < jed; ++j) {
                const int iwj = iw[colidx[j]];
                if (iwj == -1) continue;
                for (int row = 0; row < blk; ++row)
                    for (int col = 0; col < blk; ++col) {
                        double v = 0.0;
                        for (int e = 0; e < blk; ++e)
                            v += aik[row*blk + e]*values[(int64_t)iwj*b2 + e*blk + col];
                        values[(int64_t)j*b2 + row*blk + col] -= v;
                    }
            }
            for (int j = kst; j < ked; ++j) iw[colidx[j]] = -1;
        }
        orc_ludcmp(blk, values + (int64_t)ed*b2);
    }
    free(iw);
}

static void bilu_lower_row(int blk, int i, const int32_t *rowptr, const int32_t *colidx,
                           const int32_t *diagslots, const double *values, double *b)
{
    const int b2 = blk*blk;
    double arr[ILU_MAXB];
    for (int k = 0; k < blk; ++k) arr[k] = b[k + (int64_t)i*blk];
    for (int j = rowptr[i]; j < diagslots[i]; ++j) {
        const int64_t c = colidx[j];
        for (int row = 0; row < blk; ++row) {
            double v = 0.0;
            for (int col = 0; col < blk; ++col)
                v += values[col + row*blk + (int64_t)j*b2]*b[col + c*blk];
            arr[row] -= v;
        }
    }
    for (int k = 0; k < blk; ++k) b[k + (int64_t)i*blk] = arr[k];
}

static void bilu_upper_row(int blk, int i, const int32_t *rowptr, const int32_t *colidx,
                           const int32_t *diagslots, const double *values, double *b)
{
    const int b2 = blk*blk;
    double arr[ILU_MAXB];
    for (int k = 0; k < blk; ++k) arr[k] = b[k + (int64_t)i*blk];
    const int dd = diagslots[i];
    for (int j = dd + 1; j < rowptr[i + 1]; ++j) {
        const int64_t c = colidx[j];
        for (int row = 0; row < blk; ++row) {
            double v = 0.0;
            for (int col = 0; col < blk; ++col)
                v += values[col + row*blk + (int64_t)j*b2]*b[col + c*blk];
            arr[row] -= v;
        }
    }
    orc_lusub(blk, values + (int64_t)dd*b2, arr);
    for (int k = 0; k < blk; ++k) b[(int64_t)i*blk + k] = arr[k];
}

ORC_EXPORT void orc_bilu_apply(int bn, int blk, const int32_t *rowptr, const int32_t *colidx,
                               const int32_t *diagslots, const double *values, double *b)
{
    for (int i = 0; i < bn; ++i) bilu_lower_row(blk, i, rowptr, colidx, diagslots, values, b);
    for (int i = bn - 1; i > -1; --i) bilu_upper_row(blk, i, rowptr, colidx, diagslots, values, b);
}

ORC_EXPORT void orc_pbilu_apply(int bn, int blk, int ncolors, const int32_t *icolors,
                                const int32_t *rowptr, const int32_t *colidx,
                                const int32_t *diagslots, const double *values, double *b)
{
    (void)bn;
    for (int c = 0; c < ncolors; ++c)
        for (int i = icolors[c]; i < icolors[c + 1]; ++i)
            bilu_lower_row(blk, i, rowptr, colidx, diagslots, values, b);
    for (int c = ncolors - 1; c >= 0; --c)
        for (int i = icolors[c]; i < icolors[c + 1]; ++i)
            bilu_upper_row(blk, i, rowptr, colidx, diagslots, values, b);
}
