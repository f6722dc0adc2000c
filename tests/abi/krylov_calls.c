/*
 * ABI check of the Krylov API: a plain C caller that includes ucfdsolver.h
 * and uses the reference's calls in their documented order
 * (UCFDMatCreateBSR, UCFDPreconCreatefromArrays + UCFDPreconSetBLUSGS /
 * UCFDPreconSetPBLUSGS + UCFDPreconPrepare, UCFDSolverCreateGMRES /
 * UCFDSolverCreateBICGSTAB, UCFDSolverTraceResidualHistory, UCFDSolve,
 * UCFDSolverGetResult) on host arrays.
 *
 * The matrix is a block chain of bn 5x5 blocks numbered colour by colour
 * (chain nodes 0, 2, 4, ... first, then 1, 3, 5, ...): row r couples to its
 * chain neighbours only, so the two colour ranges are internally uncoupled.
 *
 *   krylov_calls.x <out.bin> [bn]
 *
 * File layout (little endian): int32 bn, blk, nnz, ncolors; icolors[ncolors+1],
 * rowptr[bn+1], colidx[nnz], diagslots[bn] (int32); values[nnz*25], x[n],
 * b[n] (f64); then y = 1.5*A*x, the sequential and the coloured apply of b,
 * and for GMRES(4) and BiCGSTAB: x[n] (f64), int32 iterations, status,
 * f64 residual, f64 history[maxiter].
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ucfdsolver.h"

#define B 5

static double val(unsigned k)
{
    /* deterministic value in (-1, 1) */
    unsigned x = k*2654435761u + 12345u;
    x ^= x >> 13;
    x *= 0x5bd1e995u;
    x ^= x >> 15;
    return (double)(x % 200001u)/100000.0 - 1.0;
}

static int row_of(int node, int bn)
{
    const int h = (bn + 1)/2;
    return (node % 2 == 0) ? node/2 : h + node/2;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s out.bin [bn]\n", argv[0]);
        return 2;
    }
    const int bn = argc > 2 ? atoi(argv[2]) : 40;
    const int n = bn*B, h = (bn + 1)/2;
    int *rowptr = calloc(bn + 1, sizeof(int)), *colidx = calloc(3*bn, sizeof(int));
    int *diag = calloc(bn, sizeof(int));
    double *values = calloc((size_t)3*bn*B*B, sizeof(double));
    int nnz = 0;
    for (int r = 0; r < bn; ++r) {
        const int node = r < h ? 2*r : 2*(r - h) + 1;
        int cols[3], nc = 0;
        if (node > 0) cols[nc++] = row_of(node - 1, bn);
        cols[nc++] = r;
        if (node + 1 < bn) cols[nc++] = row_of(node + 1, bn);
        for (int a = 0; a < nc; ++a)       /* sort the columns */
            for (int c = a + 1; c < nc; ++c)
                if (cols[c] < cols[a]) { int t = cols[a]; cols[a] = cols[c]; cols[c] = t; }
        rowptr[r] = nnz;
        for (int a = 0; a < nc; ++a) {
            const int c = cols[a];
            if (c == r) diag[r] = nnz;
            for (int i = 0; i < B; ++i)
                for (int j = 0; j < B; ++j) {
                    double v = 0.3*val((unsigned)(nnz*B*B + i*B + j));
                    if (c == r && i == j) v += 4.0;
                    values[(size_t)nnz*B*B + i*B + j] = v;
                }
            colidx[nnz++] = c;
        }
    }
    rowptr[bn] = nnz;
    double *x = malloc(n*sizeof(double)), *b = malloc(n*sizeof(double));
    double *y = calloc(n, sizeof(double)), *as = malloc(n*sizeof(double));
    double *ac = malloc(n*sizeof(double));
    for (int i = 0; i < n; ++i) {
        x[i] = val(100000u + i);
        b[i] = val(200000u + i);
    }
    int icolors[3] = {0, h, bn};

    SpMat A;
    Precon seq, col;
    int rc = 0;
    rc |= UCFDMatCreateBSR(&A, bn, B, rowptr, colidx, values);
    rc |= UCFDMatMult(1.5, A, x, 0.0, y);
    rc |= UCFDPreconCreatefromArrays(&seq, rowptr, colidx, diag, values);
    rc |= UCFDPreconSetBLUSGS(&seq, bn, B);
    rc |= UCFDPreconPrepare(seq);
    rc |= UCFDPreconCreatefromArrays(&col, rowptr, colidx, diag, values);
    rc |= UCFDPreconSetPBLUSGS(&col, bn, B, 2, icolors);
    rc |= UCFDPreconPrepare(col);
    memcpy(as, b, n*sizeof(double));
    memcpy(ac, b, n*sizeof(double));
    rc |= UCFDPreconApply(seq, as);
    rc |= UCFDPreconApply(col, ac);

    FILE *f = fopen(argv[1], "wb");
    if (!f) return 3;
    int hdr[4] = {bn, B, nnz, 2};
    fwrite(hdr, sizeof(int), 4, f);
    fwrite(icolors, sizeof(int), 3, f);
    fwrite(rowptr, sizeof(int), bn + 1, f);
    fwrite(colidx, sizeof(int), nnz, f);
    fwrite(diag, sizeof(int), bn, f);
    fwrite(values, sizeof(double), (size_t)nnz*B*B, f);
    fwrite(x, sizeof(double), n, f);
    fwrite(b, sizeof(double), n, f);
    fwrite(y, sizeof(double), n, f);
    fwrite(as, sizeof(double), n, f);
    fwrite(ac, sizeof(double), n, f);

    const int maxiter = 30;
    for (int k = 0; k < 2; ++k) {
        Solver s;
        if (k == 0) rc |= UCFDSolverCreateGMRES(&s, n, 4, maxiter, 1e-10);
        else rc |= UCFDSolverCreateBICGSTAB(&s, n, maxiter, 1e-10);
        rc |= UCFDSolverTraceResidualHistory(s);
        double *xs = calloc(n, sizeof(double)), hist[30];
        rc |= UCFDSolve(s, seq, A, xs, b);
        UCFDInt st, it;
        UCFDReal res;
        rc |= UCFDSolverGetResult(s, &st, &it, &res);
        rc |= UCFDSolverGetResidualHistory(s, hist);
        int meta[2] = {it, st};
        fwrite(xs, sizeof(double), n, f);
        fwrite(meta, sizeof(int), 2, f);
        fwrite(&res, sizeof(double), 1, f);
        fwrite(hist, sizeof(double), maxiter, f);
        rc |= UCFDSolverDestroy(&s);
        free(xs);
    }
    fclose(f);
    rc |= UCFDPreconDestroy(&seq);
    rc |= UCFDPreconDestroy(&col);
    rc |= UCFDMatDestroy(&A);
    printf("krylov_calls: bn=%d nnz=%d status=%d\n", bn, nnz, rc);
    return rc ? 1 : 0;
}
