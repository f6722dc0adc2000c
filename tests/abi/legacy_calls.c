/*
 * ABI check of the legacy array API: a plain C caller that includes the
 * library headers and makes the calls of the reference examples with their
 * exact argument lists (examples/omp3d.c:180-189, examples/mpi3d.c:166-173),
 * on an m x n x l hex block with red-black colouring.  Inputs are filled
 * with non-degenerate deterministic values; every array is written to the
 * output file so the test can compare with the CPU oracle.
 *
 *   legacy_calls.x <out.bin> [m n l]
 *
 * File layout (all little endian): int32 m, n, l, N; then nei_ele[6N],
 * icolor[N], lcolor[N], fnorm_vol[6N], vec_fnorm[18N], fspr[6N], dt[N],
 * upts0[NV*N], rhs0[NV*N], then results of the coloured C form
 * (diag[N], dub_after_lower[NV*N], dub_after_upper[NV*N], upts[NV*N]),
 * the coloured two-array form (dub[NV*N], rhsb[NV*N], upts[NV*N]) and the
 * serial C form (dub_after_upper[NV*N], upts[NV*N]), NV = 5.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "coloredlusgs.h"
#include "lusgs.h"

#define NV 5
#define NF 6

static void fill_mesh(int m, int n, int l, int *nei, int *icolor, int *lcolor)
{
    const int N = m*n*l, mn = m*n;
    int ka = 0, kb = 0, half = 0;
    for (int i = 0; i < N; ++i) {
        const int x = i % m, y = (i / m) % n, z = i / mn;
        half += ((x + y + z) & 1) == 0;
    }
    for (int z = 0; z < l; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < m; ++x) {
                const int i = x + m*y + mn*z;
                nei[0*N + i] = x > 0 ? i - 1 : i;
                nei[1*N + i] = x + 1 < m ? i + 1 : i;
                nei[2*N + i] = y > 0 ? i - m : i;
                nei[3*N + i] = y + 1 < n ? i + m : i;
                nei[4*N + i] = z > 0 ? i - mn : i;
                nei[5*N + i] = z + 1 < l ? i + mn : i;
                if (((x + y + z) & 1) == 0) { icolor[ka++] = i; lcolor[i] = 1; }
                else { icolor[half + kb++] = i; lcolor[i] = 2; }
            }
}

static void fill_fields(int N, double *fnv, double *vfn, double *fspr, double *dt,
                        double *upts, double *rhs)
{
    const double nrm[NF][3] = {{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},{0,0,-1},{0,0,1}};
    for (int i = 0; i < N; ++i) {
        const double s = 0.37*i;
        const double rho = 1.0 + 0.2*sin(s), u = 0.3*cos(1.3*s), v = 0.2*sin(0.7*s);
        const double w = 0.1*cos(0.3*s), p = 1.0 + 0.25*cos(0.9*s);
        upts[0*N + i] = rho;
        upts[1*N + i] = rho*u;
        upts[2*N + i] = rho*v;
        upts[3*N + i] = rho*w;
        upts[4*N + i] = p/0.4 + 0.5*rho*(u*u + v*v + w*w);
        for (int k = 0; k < NV; ++k) rhs[k*N + i] = 1e-3*sin(1.7*s + k);
        dt[i] = 1e-3*(2.0 + sin(0.11*i));
        const double c = sqrt(1.4*p/rho);
        for (int f = 0; f < NF; ++f) {
            fnv[f*N + i] = 10.0*(1.0 + 0.05*cos(s + f));
            for (int d = 0; d < 3; ++d) vfn[(f*3 + d)*N + i] = nrm[f][d];
            const double un = u*nrm[f][0] + v*nrm[f][1] + w*nrm[f][2];
            fspr[f*N + i] = fabs(un) + c;
        }
    }
}

static void put(FILE *fp, const void *p, size_t bytes) { fwrite(p, 1, bytes, fp); }

int main(int argc, char **argv)
{
    if (argc < 2) { fprintf(stderr, "usage: %s out.bin [m n l]\n", argv[0]); return 2; }
    int m = 6, n = 5, l = 4;
    if (argc >= 5) { m = atoi(argv[2]); n = atoi(argv[3]); l = atoi(argv[4]); }
    const int N = m*n*l, half = (N + 1)/2;
    int *nei = malloc(sizeof(int)*NF*N), *icolor = malloc(sizeof(int)*N), *lcolor = malloc(sizeof(int)*N);
    double *fnv = malloc(8*NF*N), *vfn = malloc(8*NF*3*N), *fspr = malloc(8*NF*N), *dt = malloc(8*N);
    double *upts0 = malloc(8*NV*N), *rhs0 = malloc(8*NV*N), *diag = calloc(N, 8);
    double *upts = malloc(8*NV*N), *dub = malloc(8*NV*N), *rhsb = malloc(8*NV*N);
    double *dlow = malloc(8*NV*N);
    fill_mesh(m, n, l, nei, icolor, lcolor);
    fill_fields(N, fnv, vfn, fspr, dt, upts0, rhs0);
    int h = 0;
    for (int i = 0; i < N; ++i) h += lcolor[i] == 1;

    FILE *fp = fopen(argv[1], "wb");
    if (!fp) return 3;
    int hdr[4] = {m, n, l, N};
    put(fp, hdr, sizeof hdr);
    put(fp, nei, 4*NF*N); put(fp, icolor, 4*N); put(fp, lcolor, 4*N);
    put(fp, fnv, 8*NF*N); put(fp, vfn, 8*NF*3*N); put(fp, fspr, 8*NF*N); put(fp, dt, 8*N);
    put(fp, upts0, 8*NV*N); put(fp, rhs0, 8*NV*N);
    int rc = 0;

    /* coloured, C in-place form (examples/omp3d.c:180-189) */
    memcpy(upts, upts0, 8*NV*N);
    memcpy(dub, rhs0, 8*NV*N);
    rc |= parallel_pre_lusgs(N, NF, 1.0, fnv, dt, diag, fspr);
    rc |= parallel_ns_lower_sweep(0, h, N, NF, nei, icolor, lcolor, fnv, vfn, upts, dub, diag, fspr);
    rc |= parallel_ns_lower_sweep(h, N, N, NF, nei, icolor, lcolor, fnv, vfn, upts, dub, diag, fspr);
    memcpy(dlow, dub, 8*NV*N);
    rc |= parallel_ns_upper_sweep(h, N, N, NF, nei, icolor, lcolor, fnv, vfn, upts, dub, diag, fspr);
    rc |= parallel_ns_upper_sweep(0, h, N, NF, nei, icolor, lcolor, fnv, vfn, upts, dub, diag, fspr);
    rc |= lusgs_parallel_ns_update(N, upts, dub);
    put(fp, diag, 8*N); put(fp, dlow, 8*NV*N); put(fp, dub, 8*NV*N); put(fp, upts, 8*NV*N);

    /* coloured, Python two-array form (examples/omp3d.py:76-94) */
    memcpy(upts, upts0, 8*NV*N);
    memcpy(rhsb, rhs0, 8*NV*N);
    memset(dub, 0, 8*NV*N);
    rc |= parallel_pre_lusgs(N, NF, 1.0, fnv, dt, diag, fspr);
    rc |= ns_parallel_lower_sweep(0, h, N, NF, nei, icolor, lcolor, fnv, vfn, upts, rhsb, dub, diag, fspr);
    rc |= ns_parallel_lower_sweep(h, N, N, NF, nei, icolor, lcolor, fnv, vfn, upts, rhsb, dub, diag, fspr);
    rc |= ns_parallel_upper_sweep(h, N, N, NF, nei, icolor, lcolor, fnv, vfn, upts, rhsb, dub, diag, fspr);
    rc |= ns_parallel_upper_sweep(0, h, N, NF, nei, icolor, lcolor, fnv, vfn, upts, rhsb, dub, diag, fspr);
    rc |= lusgs_parallel_ns_update(N, upts, rhsb);
    put(fp, dub, 8*NV*N); put(fp, rhsb, 8*NV*N); put(fp, upts, 8*NV*N);

    /* serial, C in-place form (examples/mpi3d.c:166-173) */
    memcpy(upts, upts0, 8*NV*N);
    memcpy(dub, rhs0, 8*NV*N);
    rc |= serial_pre_lusgs(N, NF, 1.0, fnv, dt, diag, fspr);
    rc |= serial_ns_lower_sweep(N, NF, nei, fnv, vfn, upts, dub, diag, fspr);
    rc |= serial_ns_upper_sweep(N, NF, nei, fnv, vfn, upts, dub, diag, fspr);
    rc |= serial_ns_lusgs_update(N, upts, dub);
    put(fp, dub, 8*NV*N); put(fp, upts, 8*NV*N);
    fclose(fp);
    (void)half;
    printf("legacy_calls: N=%d status=%d\n", N, rc);
    return rc ? 1 : 0;
}
