/*
 * Seeded synthetic meshes and flow fields for the LU-SGS workloads.
 *
 * Every random draw is counter based: value = mix64(seed, stream, index), so a
 * draw depends only on (seed, stream, index) and never on evaluation order.
 * Flow-field draws are indexed by the ORIGINAL (pre-permutation) cell id, so a
 * physical cell keeps its state whatever numbering the mesh is stored in.
 *
 * Array layouts follow the legacy LU-SGS API (nface-major, C order):
 *   nei_ele[f*N + i], fnorm_vol[f*N + i], fspr[f*N + i],
 *   vec_fnorm[(f*ndims + d)*N + i], upts/rhsb[k*N + i], dt[i].
 * Boundary faces point to the cell itself (examples/utils/pbutils.c:43-84).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GEN_EXPORT __attribute__((visibility("default")))

enum {
    ST_PERM = 1, ST_VOLJIT = 2, ST_RHO = 3, ST_VEL = 4, ST_PRES = 5,
    ST_DT = 6, ST_RHSB = 7, ST_VERT = 8, ST_NU = 9, ST_PERT = 10,
    ST_AFACE = 11, ST_VEC = 12
};

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t ctr64(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return mix64(seed * 0x9E3779B97F4A7C15ULL
                 + mix64(stream * 0xD1B54A32D192ED03ULL + idx + 1));
}

/* uniform in [0, 1) with 53 random bits */
static inline double u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return (double)(ctr64(seed, stream, idx) >> 11) * 0x1.0p-53;
}

/* uniform in (0, 1) (never 0, for the logarithm of Box-Muller) */
static inline double u01_open(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return ((double)(ctr64(seed, stream, idx) >> 11) + 0.5) * 0x1.0p-53;
}

static inline double normal01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    const double a = u01_open(seed, stream, 2*idx);
    const double b = u01(seed, stream, 2*idx + 1);
    return sqrt(-2.0*log(a)) * cos(6.283185307179586476925*b);
}

GEN_EXPORT double ucfdgen_uniform(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return u01(seed, stream, idx);
}

GEN_EXPORT double ucfdgen_normal(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return normal01(seed, stream, idx);
}

/*
 * Random permutation of [0, n): perm[old] = new.  Fisher-Yates driven by the
 * counter stream ST_PERM; perm = identity when permute == 0.
 */
GEN_EXPORT void ucfdgen_permutation(int64_t n, uint64_t seed, int permute, int32_t *perm)
{
    int32_t *a = (int32_t *)malloc(sizeof(int32_t)*(size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) a[i] = (int32_t)i;
    if (permute) {
        for (int64_t i = n - 1; i > 0; --i) {
            int64_t j = (int64_t)(ctr64(seed, ST_PERM, (uint64_t)i) % (uint64_t)(i + 1));
            int32_t t = a[i]; a[i] = a[j]; a[j] = t;
        }
    }
    /* a[new] = old */
    for (int64_t q = 0; q < n; ++q) perm[a[q]] = (int32_t)q;
    free(a);
}

static uint64_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    uint64_t k = 0;
    for (int b = 0; b < 21; ++b) {
        k |= ((uint64_t)((x >> b) & 1u)) << (3*b);
        k |= ((uint64_t)((y >> b) & 1u)) << (3*b + 1);
        k |= ((uint64_t)((z >> b) & 1u)) << (3*b + 2);
    }
    return k;
}

/*
 * Structured m x n x l hex grid in the unit cube, nface = 6, faces ordered
 * -x,+x,-y,+y,-z,+z (pbutils.c:27-32), linear id x + m*y + m*n*z
 * (pbutils.py:6-8) before the optional seeded permutation.
 *
 * fnorm_vol = face_area / cell_volume with the volume jittered by a factor
 * U[1-voljit, 1+voljit] per cell.  vec_fnorm holds the matching outward
 * axis normals.  Outputs, all in the permuted numbering:
 *   nei_ele[6*N], fnorm_vol[6*N], vec_fnorm[6*3*N],
 *   orig_id[N]   (original id of each stored cell; nullable),
 *   key[N]       (Morton key of the cell position; nullable).
 */
GEN_EXPORT int ucfdgen_hex(int m, int n, int l, uint64_t seed, int permute, double voljit,
                           int32_t *nei_ele, double *fnorm_vol, double *vec_fnorm,
                           int32_t *orig_id, uint64_t *key)
{
    if (m <= 0 || n <= 0 || l <= 0) return 1;
    const int64_t N = (int64_t)m*n*l, mn = (int64_t)m*n;
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t)*N);
    if (!perm) return 2;
    ucfdgen_permutation(N, seed, permute, perm);

    const double hx = 1.0/m, hy = 1.0/n, hz = 1.0/l;
    const double area[6] = {hy*hz, hy*hz, hx*hz, hx*hz, hx*hy, hx*hy};
    const double nrm[6][3] = {{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},{0,0,-1},{0,0,1}};

    for (int z = 0; z < l; ++z)
    for (int y = 0; y < n; ++y)
    for (int x = 0; x < m; ++x) {
        const int64_t o = x + (int64_t)m*y + mn*z;
        const int64_t nb[6] = {
            x > 0     ? o - 1  : o,  x + 1 < m ? o + 1  : o,
            y > 0     ? o - m  : o,  y + 1 < n ? o + m  : o,
            z > 0     ? o - mn : o,  z + 1 < l ? o + mn : o};
        const int64_t i = perm[o];
        const double vol = hx*hy*hz*(1.0 + voljit*(2.0*u01(seed, ST_VOLJIT, (uint64_t)o) - 1.0));
        for (int f = 0; f < 6; ++f) {
            nei_ele[f*N + i] = perm[nb[f]];
            fnorm_vol[f*N + i] = area[f]/vol;
            for (int d = 0; d < 3; ++d) vec_fnorm[(f*3 + d)*N + i] = nrm[f][d];
        }
        if (orig_id) orig_id[i] = (int32_t)o;
        if (key) key[i] = morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
    }
    free(perm);
    return 0;
}

/*
 * 3-D Hilbert index of (x, y, z) with `bits` bits per axis, by Skilling's
 * transpose algorithm (J. Skilling, "Programming the Hilbert curve", AIP
 * Conf. Proc. 707, 2004): undo excess work, Gray-encode, then interleave
 * the transposed bits from the most significant bit down.
 */
static uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits)
{
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) X[0] ^= P;
            else { const uint32_t t = (X[0] ^ X[i]) & P; X[0] ^= t; X[i] ^= t; }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint64_t h = 0;
    for (int q = bits - 1; q >= 0; --q)
        for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> q) & 1u);
    return h;
}

static int bits_for(int m)
{
    int b = 1;
    while ((1 << b) < m) ++b;
    return b;
}

/*
 * Hilbert locality keys from the original cell ids: hex cells (kind 0,
 * id = x + m*y + m*n*z) or Kuhn tets (kind 1, id = 6*cube + s, cube on an
 * m^3 grid; key = 8*hilbert(cube) + s).
 */
GEN_EXPORT void ucfdgen_hilbert_keys(int kind, int m, int n, int l, int64_t N,
                                     const int32_t *orig_id, uint64_t *key)
{
    const int bits = bits_for(m > n ? (m > l ? m : l) : (n > l ? n : l));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        int64_t o = orig_id[i], s = 0;
        if (kind == 1) { s = o % 6; o /= 6; }
        const uint32_t x = (uint32_t)(o % m), y = (uint32_t)((o / m) % n);
        const uint32_t z = (uint32_t)(o / ((int64_t)m*n));
        const uint64_t h = hilbert3(x, y, z, bits);
        key[i] = kind == 1 ? 8*h + (uint64_t)s : h;
    }
}

/* Kuhn (Freudenthal) split of a cube into 6 tets: tet s walks the cube's
 * main diagonal adding unit steps in the axis order KUHN[s]. */
static const int KUHN[6][3] = {{0,1,2},{0,2,1},{1,0,2},{1,2,0},{2,0,1},{2,1,0}};

static int kuhn_index(int a, int b, int c)
{
    for (int s = 0; s < 6; ++s)
        if (KUHN[s][0] == a && KUHN[s][1] == b && KUHN[s][2] == c) return s;
    return -1;
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1]*b[2] - a[2]*b[1];
    c[1] = a[2]*b[0] - a[0]*b[2];
    c[2] = a[0]*b[1] - a[1]*b[0];
}

/*
 * m^3 hex grid of the unit cube, each cube split into 6 Kuhn tets
 * (N = 6 m^3, nface = 4, conforming).  Original tet id = 6*cube + s with
 * cube = x + m*y + m*m*z.  Face k of a tet is the face opposite its vertex
 * p_k.  Grid vertices are jittered by U[-jit*h, +jit*h] per axis, boundary
 * vertices only tangentially, so the domain stays the unit cube.  Normals,
 * areas and volumes come from the jittered coordinates.
 *
 * Outputs (permuted numbering): nei_ele[4N], fnorm_vol[4N], vec_fnorm[12N],
 * vol[N], centroid[3N], orig_id[N], key[N], face_area[4N]
 * (all but the first three nullable).
 */
GEN_EXPORT int ucfdgen_kuhn_tets(int m, uint64_t seed, int permute, double jit,
                                 int32_t *nei_ele, double *fnorm_vol, double *vec_fnorm,
                                 double *vol_out, double *centroid, int32_t *orig_id,
                                 uint64_t *key, double *face_area)
{
    if (m <= 0) return 1;
    const int64_t nc = (int64_t)m*m*m, N = 6*nc, mv = m + 1;
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t)*N);
    double *vx = (double *)malloc(sizeof(double)*3*(size_t)(mv*mv*mv));
    if (!perm || !vx) { free(perm); free(vx); return 2; }
    ucfdgen_permutation(N, seed, permute, perm);

    const double h = 1.0/m;
    for (int64_t k = 0; k < mv; ++k)
    for (int64_t j = 0; j < mv; ++j)
    for (int64_t i = 0; i < mv; ++i) {
        const int64_t v = i + mv*j + mv*mv*k;
        const int64_t g[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
            double dd = jit*h*(2.0*u01(seed, ST_VERT, (uint64_t)(3*v + d)) - 1.0);
            if (g[d] == 0 || g[d] == m) dd = 0.0;
            vx[3*v + d] = g[d]*h + dd;
        }
    }

    int rc = 0;
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        const int x = (int)(c % m), y = (int)((c / m) % m), z = (int)(c / ((int64_t)m*m));
        for (int s = 0; s < 6; ++s) {
            const int a = KUHN[s][0], b = KUHN[s][1], cc = KUHN[s][2];
            int pg[4][3];
            pg[0][0] = x; pg[0][1] = y; pg[0][2] = z;
            for (int d = 0; d < 3; ++d) pg[1][d] = pg[0][d] + (d == a);
            for (int d = 0; d < 3; ++d) pg[2][d] = pg[1][d] + (d == b);
            for (int d = 0; d < 3; ++d) pg[3][d] = pg[2][d] + (d == cc);
            double p[4][3];
            for (int q = 0; q < 4; ++q) {
                const int64_t v = pg[q][0] + mv*pg[q][1] + mv*mv*pg[q][2];
                for (int d = 0; d < 3; ++d) p[q][d] = vx[3*v + d];
            }
            const int64_t o = 6*c + s, i = perm[o];

            double e1[3], e2[3], e3[3], cr[3];
            for (int d = 0; d < 3; ++d) {
                e1[d] = p[1][d] - p[0][d]; e2[d] = p[2][d] - p[0][d]; e3[d] = p[3][d] - p[0][d];
            }
            cross3(e1, e2, cr);
            const double vol = fabs(cr[0]*e3[0] + cr[1]*e3[1] + cr[2]*e3[2])/6.0;
            if (!(vol > 0.0)) rc = 3;

            /* neighbours: (cube offset axis/sign, Kuhn order) per face */
            int ncube[4][3], nord[4];
            for (int d = 0; d < 3; ++d) {
                ncube[0][d] = pg[0][d] + (d == a);   /* opposite p0: cube + e_a, (b,c,a) */
                ncube[1][d] = pg[0][d];              /* opposite p1: same cube, (b,a,c) */
                ncube[2][d] = pg[0][d];              /* opposite p2: same cube, (a,c,b) */
                ncube[3][d] = pg[0][d] - (d == cc);  /* opposite p3: cube - e_c, (c,a,b) */
            }
            nord[0] = kuhn_index(b, cc, a);
            nord[1] = kuhn_index(b, a, cc);
            nord[2] = kuhn_index(a, cc, b);
            nord[3] = kuhn_index(cc, a, b);

            for (int f = 0; f < 4; ++f) {
                const int *q = ncube[f];
                int64_t nbo = o;
                if (q[0] >= 0 && q[0] < m && q[1] >= 0 && q[1] < m && q[2] >= 0 && q[2] < m)
                    nbo = 6*(q[0] + (int64_t)m*q[1] + (int64_t)m*m*q[2]) + nord[f];
                nei_ele[f*N + i] = perm[nbo];

                /* face = the three vertices other than p_f; outward from p_f */
                int t[3], nt = 0;
                for (int r = 0; r < 4; ++r) if (r != f) t[nt++] = r;
                double u1[3], u2[3], nn[3], w[3];
                for (int d = 0; d < 3; ++d) {
                    u1[d] = p[t[1]][d] - p[t[0]][d];
                    u2[d] = p[t[2]][d] - p[t[0]][d];
                    w[d]  = p[f][d] - p[t[0]][d];
                }
                cross3(u1, u2, nn);
                double sgn = (nn[0]*w[0] + nn[1]*w[1] + nn[2]*w[2]) > 0.0 ? -1.0 : 1.0;
                const double len = sqrt(nn[0]*nn[0] + nn[1]*nn[1] + nn[2]*nn[2]);
                const double area = 0.5*len;
                for (int d = 0; d < 3; ++d) vec_fnorm[(f*3 + d)*N + i] = sgn*nn[d]/len;
                fnorm_vol[f*N + i] = area/vol;
                if (face_area) face_area[f*N + i] = area;
            }
            if (vol_out) vol_out[i] = vol;
            if (centroid)
                for (int d = 0; d < 3; ++d)
                    centroid[d*N + i] = 0.25*(p[0][d] + p[1][d] + p[2][d] + p[3][d]);
            if (orig_id) orig_id[i] = (int32_t)o;
            if (key) key[i] = morton3((uint32_t)x, (uint32_t)y, (uint32_t)z)*8 + (uint64_t)s;
        }
    }
    free(perm);
    free(vx);
    return rc;
}

/*
 * Flow fields for the Euler LU-SGS workloads (gamma = 1.4):
 *   rho ~ U[0.8,1.2], u,v,w ~ U[-0.3,0.3], p ~ U[0.7,1.3]  -> conservative upts
 *   fspr[f,i] = |u.n_f| + sqrt(gamma p / rho)   (cell's own state)
 *   dt ~ U[1e-3, 1e-2],  rhsb ~ N(0, sigma = 1e-3)
 * Draws are indexed by orig_id[i] (identity if orig_id is NULL).
 * upts and rhsb have nvars rows of stride N; rows >= ndims+2 are set to 0.
 * Any output pointer may be NULL.
 */
GEN_EXPORT int ucfdgen_flow(int64_t N, int nface, int ndims, int nvars, uint64_t seed,
                            const int32_t *orig_id, const double *vec_fnorm,
                            double *upts, double *fspr, double *dt, double *rhsb)
{
    const double gamma = 1.4;
    const int nfv = ndims + 2;
    if (ndims < 2 || ndims > 3 || nvars < nfv) return 1;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const uint64_t o = orig_id ? (uint64_t)orig_id[i] : (uint64_t)i;
        const double rho = 0.8 + 0.4*u01(seed, ST_RHO, o);
        double vel[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < ndims; ++d) vel[d] = -0.3 + 0.6*u01(seed, ST_VEL, 3*o + d);
        const double p = 0.7 + 0.6*u01(seed, ST_PRES, o);
        const double c = sqrt(gamma*p/rho);
        double ke = 0.0;
        for (int d = 0; d < ndims; ++d) ke += vel[d]*vel[d];
        if (upts) {
            upts[i] = rho;
            for (int d = 0; d < ndims; ++d) upts[(1 + d)*N + i] = rho*vel[d];
            upts[(nfv - 1)*N + i] = p/(gamma - 1.0) + 0.5*rho*ke;
            for (int k = nfv; k < nvars; ++k) upts[k*N + i] = 0.0;
        }
        if (fspr) {
            for (int f = 0; f < nface; ++f) {
                double un = 0.0;
                for (int d = 0; d < ndims; ++d) un += vel[d]*vec_fnorm[(f*ndims + d)*N + i];
                fspr[f*N + i] = fabs(un) + c;
            }
        }
        if (dt) dt[i] = 1e-3 + 9e-3*u01(seed, ST_DT, o);
        if (rhsb) {
            for (int k = 0; k < nfv; ++k) rhsb[k*N + i] = 1e-3*normal01(seed, ST_RHSB, nfv*o + k);
            for (int k = nfv; k < nvars; ++k) rhsb[k*N + i] = 0.0;
        }
    }
    return 0;
}

/*
 * Face-pair map: pair[f*N + i] = the face index f' of cell nb = nei_ele[f*N+i]
 * with nei_ele[f'*N + nb] == i, or -1 for a boundary (self) face.  Returns
 * the number of interior faces with no matching back face (0 on a conforming
 * mesh).
 */
GEN_EXPORT int64_t ucfdgen_face_pairs(int64_t N, int nface, const int32_t *nei_ele, int32_t *pair)
{
    int64_t bad = 0;
    #pragma omp parallel for schedule(static) reduction(+:bad)
    for (int64_t i = 0; i < N; ++i) {
        for (int f = 0; f < nface; ++f) {
            const int64_t nb = nei_ele[f*N + i];
            int32_t pf = -1;
            if (nb != i) {
                for (int g = 0; g < nface; ++g)
                    if (nei_ele[g*N + nb] == i) { pf = g; break; }
                if (pf < 0) ++bad;
            }
            pair[f*N + i] = pf;
        }
    }
    return bad;
}

/* N(0,1) / U[0,1) fills (counter based, index = position) */
GEN_EXPORT void ucfdgen_normal_fill(uint64_t seed, uint64_t stream, int64_t n, double *out)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = normal01(seed, stream, (uint64_t)i);
}

/* number of BSR blocks of the face-neighbour system: one diagonal block per
 * cell plus one per interior face slot */
GEN_EXPORT int64_t ucfdgen_bsr_count(int64_t N, int nface, const int32_t *nei_ele)
{
    int64_t nnz = N;
    for (int f = 0; f < nface; ++f)
        for (int64_t i = 0; i < N; ++i) nnz += nei_ele[f*N + i] != i;
    return nnz;
}

/*
 * Block Jacobian-like system on a cell mesh (block size B = 5 for the Euler
 * block, B = 1 for a one-equation turbulence block), in the mesh's own cell
 * numbering, BSR with sorted columns:
 *   off-diagonal (i, nb):  0.5 fv (s A_f - lam I),  A_f entries N(0, 0.2)
 *       shared by both sides of the face, s = +1 on the side with the smaller
 *       original id and -1 on the other, fv = fnorm_vol[f,i], lam = fspr[f,i]
 *   diagonal i: (1/dt_i + sum_f fv (0.5 lam + nu_f / dist_f)) I + P_i,
 *       nu_f ~ U[1e-5, 1e-3] per face, dist_f = centroid distance (1/fv on
 *       boundary faces or without centroids), P_i entries U[-0.1, 0.1] times
 *       the diagonal value.
 * Draws are indexed by original ids (face pairs), so renumbering the mesh
 * permutes the matrix and nothing else.  Outputs: rowptr[N+1], colidx[nnz],
 * diagslots[N], values[nnz*B*B] (row-major blocks).
 */
GEN_EXPORT int ucfdgen_bsr_fill(int64_t N, int nface, int B, uint64_t seed, const int32_t *nei,
                                const double *fnorm_vol, const double *fspr, const double *dt,
                                const int32_t *orig_id, const double *centroid, int32_t *rowptr,
                                int32_t *colidx, int32_t *diagslots, double *values)
{
    if (B < 1 || B > 8 || nface > 16) return 1;
    const int B2 = B*B;
    rowptr[0] = 0;
    for (int64_t i = 0; i < N; ++i) {
        int cnt = 1;
        for (int f = 0; f < nface; ++f) cnt += nei[f*N + i] != i;
        rowptr[i + 1] = rowptr[i] + cnt;
    }
    const uint64_t STA = (uint64_t)ST_AFACE*16 + (uint64_t)B;
    const uint64_t STN = (uint64_t)ST_NU*16 + (uint64_t)B;
    const uint64_t STP = (uint64_t)ST_PERT*16 + (uint64_t)B;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        int32_t cols[17];
        int32_t fof[17];
        int n = 0;
        cols[n] = (int32_t)i; fof[n] = -1; ++n;
        for (int f = 0; f < nface; ++f) {
            const int32_t nb = nei[f*N + i];
            if (nb == i) continue;
            cols[n] = nb; fof[n] = f; ++n;
        }
        /* insertion sort by column */
        for (int a = 1; a < n; ++a) {
            int32_t c = cols[a], g = fof[a];
            int b = a - 1;
            while (b >= 0 && cols[b] > c) { cols[b + 1] = cols[b]; fof[b + 1] = fof[b]; --b; }
            cols[b + 1] = c; fof[b + 1] = g;
        }
        const int64_t base = rowptr[i];
        const uint64_t oi = (uint64_t)(orig_id ? orig_id[i] : i);
        double dval = 1.0/dt[i];
        for (int a = 0; a < n; ++a) {
            colidx[base + a] = cols[a];
            double *blk = values + (base + a)*B2;
            if (fof[a] < 0) { diagslots[i] = (int32_t)(base + a); continue; }
            const int f = fof[a];
            const int64_t nb = cols[a];
            const uint64_t on = (uint64_t)(orig_id ? orig_id[nb] : nb);
            const uint64_t lo = oi < on ? oi : on, hi = oi < on ? on : oi;
            const uint64_t pair = lo*(uint64_t)N + hi;
            const double s = oi < on ? 1.0 : -1.0;
            const double fv = fnorm_vol[f*N + i], lam = fspr[f*N + i];
            for (int r = 0; r < B; ++r)
                for (int c = 0; c < B; ++c) {
                    const double af = 0.2*normal01(seed, STA, pair*64 + (uint64_t)(r*B + c));
                    blk[r*B + c] = 0.5*fv*(s*af - (r == c ? lam : 0.0));
                }
            const double nu = 1e-5 + (1e-3 - 1e-5)*u01(seed, STN, pair);
            double dist = 1.0/fv;
            if (centroid) {
                double d2 = 0.0;
                for (int d = 0; d < 3; ++d) {
                    const double x = centroid[d*N + i] - centroid[d*N + nb];
                    d2 += x*x;
                }
                dist = sqrt(d2);
            }
            dval += fv*(0.5*lam + nu/dist);
        }
        for (int f = 0; f < nface; ++f)
            if (nei[f*N + i] == i) dval += fnorm_vol[f*N + i]*0.5*fspr[f*N + i];
        double *dg = values + (int64_t)diagslots[i]*B2;
        for (int r = 0; r < B; ++r)
            for (int c = 0; c < B; ++c) {
                const double pert = (2.0*u01(seed, STP, oi*64 + (uint64_t)(r*B + c)) - 1.0)*0.1;
                dg[r*B + c] = (r == c ? dval : 0.0) + pert*dval;
            }
    }
    return 0;
}

/*
 * Face-slot CSR system for the FlowSys block LU-SGS API
 * (src/lusgs/flowsys.h:34-59): one slot per interior face, numbered in the
 * order first met scanning cells, then faces, of the given numbering; each
 * cell lists its interior faces (face order) as positions with neighbour,
 * slot and side (+1 on the side with the smaller original id, -1 on the
 * other).  Boundary faces carry no position.  Per slot: face_area (the
 * owner side's), face_normal [ndims, nslots] (owner's outward normal) and
 *   jmat[0][r][c][slot] = A+ = 0.5 (A_f + lam I) + (nu/dist) I
 *   jmat[1][r][c][slot] = A- = 0.5 (A_f - lam I) - (nu/dist) I
 * with A_f, nu, dist drawn and defined as in ucfdgen_bsr_fill (lam = the
 * owner's fspr), so the block system matches that generator's off-diagonal
 * blocks.  Pass 1 (values == 0): counts only, *nslots and *npos.
 */
GEN_EXPORT int ucfdgen_flowsys_fill(int64_t N, int nface, int ndims, int nv, uint64_t seed,
                                    const int32_t *nei, const double *face_area_cf,
                                    const double *vec_fnorm, const double *fspr,
                                    const int32_t *orig_id, const double *centroid,
                                    const double *fnorm_vol, int64_t *nslots, int64_t *npos,
                                    int32_t *rowptr, int32_t *colidx, int32_t *slots,
                                    int8_t *sides, double *face_area, double *face_normal,
                                    double *jmat, int values)
{
    if (nv < 1 || nv > 8) return 1;
    int64_t ns = 0, np = 0;
    for (int64_t i = 0; i < N; ++i)
        for (int f = 0; f < nface; ++f) {
            const int64_t nb = nei[f*N + i];
            if (nb == i) continue;
            ++np;
            if (i < nb) ++ns;
        }
    *nslots = ns;
    *npos = np;
    if (!values) return 0;
    /* slot of the face (i, f) with i < nb: running count; the partner cell
     * finds it through its back face */
    int32_t *slot_of = (int32_t *)malloc(sizeof(int32_t)*(size_t)N*nface);
    if (!slot_of) return 2;
    int64_t s = 0;
    for (int64_t i = 0; i < N; ++i)
        for (int f = 0; f < nface; ++f) {
            const int64_t nb = nei[f*N + i];
            slot_of[f*N + i] = (nb != i && i < nb) ? (int32_t)s++ : -1;
        }
    int64_t p = 0;
    rowptr[0] = 0;
    for (int64_t i = 0; i < N; ++i) {
        for (int f = 0; f < nface; ++f) {
            const int64_t nb = nei[f*N + i];
            if (nb == i) continue;
            int32_t sl = slot_of[f*N + i];
            if (sl < 0)
                for (int g = 0; g < nface; ++g)
                    if (nei[g*N + nb] == i) { sl = slot_of[g*N + nb]; break; }
            if (sl < 0) { free(slot_of); return 3; }
            const uint64_t oi = (uint64_t)(orig_id ? orig_id[i] : i);
            const uint64_t on = (uint64_t)(orig_id ? orig_id[nb] : nb);
            colidx[p] = (int32_t)nb;
            slots[p] = sl;
            sides[p] = oi < on ? 1 : -1;
            ++p;
        }
        rowptr[i + 1] = (int32_t)p;
    }
    const int nv2 = nv*nv;
    const uint64_t STA = (uint64_t)ST_AFACE*16 + (uint64_t)nv;
    const uint64_t STN = (uint64_t)ST_NU*16 + (uint64_t)nv;
    for (int64_t i = 0; i < N; ++i)
        for (int f = 0; f < nface; ++f) {
            const int64_t nb = nei[f*N + i];
            if (nb == i) continue;
            const uint64_t oi = (uint64_t)(orig_id ? orig_id[i] : i);
            const uint64_t on = (uint64_t)(orig_id ? orig_id[nb] : nb);
            if (oi > on) continue;                 /* the owner writes the slot */
            int32_t sl = slot_of[f*N + i];
            if (sl < 0)
                for (int g = 0; g < nface; ++g)
                    if (nei[g*N + nb] == i) { sl = slot_of[g*N + nb]; break; }
            const uint64_t pair = oi*(uint64_t)N + on;
            face_area[sl] = face_area_cf ? face_area_cf[f*N + i] : 1.0;
            for (int d = 0; d < ndims; ++d)
                face_normal[(int64_t)d*ns + sl] = vec_fnorm[((int64_t)f*ndims + d)*N + i];
            const double lam = fspr[f*N + i];
            const double nu = 1e-5 + (1e-3 - 1e-5)*u01(seed, STN, pair);
            double dist = 1.0/fnorm_vol[f*N + i];
            if (centroid) {
                double d2 = 0.0;
                for (int d = 0; d < ndims; ++d) {
                    const double x = centroid[d*N + i] - centroid[d*N + nb];
                    d2 += x*x;
                }
                dist = sqrt(d2);
            }
            for (int r = 0; r < nv; ++r)
                for (int c = 0; c < nv; ++c) {
                    const double af = 0.2*normal01(seed, STA, pair*64 + (uint64_t)(r*nv + c));
                    const double dg = r == c ? 1.0 : 0.0;
                    jmat[((int64_t)(0*nv2) + r*nv + c)*ns + sl] = 0.5*(af + lam*dg) + nu/dist*dg;
                    jmat[((int64_t)(1*nv2) + r*nv + c)*ns + sl] = 0.5*(af - lam*dg) - nu/dist*dg;
                }
        }
    free(slot_of);
    return 0;
}
