// Kernels of the Krylov path: BSR / CSR SpMV, block LU-SGS (prepare, lower
// and upper sweeps in FP64 and in the mixed FP32/FP64 form), the ILU(0)
// family (level-scheduled factorisation, triangular solves), BLAS-1/2 with
// deterministic reductions, and the fused BiCGSTAB vector updates.
//
// FP64 kernels keep the operation order of the reference
// (src/krylov/matrix/sparsemat.c:57-101, src/krylov/utils/inverse.c:7-107,
// src/krylov/precon/blusgs.c:28-100, ilu.c:4-86, bilu.c:5-130, pbilu.c:5-180)
// and are compiled without FMA contraction, so SpMV, the factorisations and
// the FP64 preconditioner applies agree bit for bit with the CPU oracle.
// Reductions sum block partials in a fixed order, so every dot product is
// reproducible run to run.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "block_lu.h"
#include "krylov_internal.h"
#include "wave_rows.h"

namespace {

constexpr int KB = KRY_BLOCK;
constexpr int KMAXPART = 1024;

__device__ __forceinline__ int kxcd()
{
    const int b = blockIdx.x, nblk = gridDim.x;
    const int x = b % 8, k = b / 8, per = nblk / 8, rem = nblk % 8;
    return x*per + (x < rem ? x : rem) + k;
}

inline int nblocks(int64_t n)
{
    const int64_t b = (n + KB - 1)/KB;
    return (int)(b < 1 ? 1 : (b > KMAXPART ? KMAXPART : b));
}

// ------------------------------------------------------------------ SpMV
template <int B>
__global__ __launch_bounds__(KB) void k_spmv_bsr(int bn, const int *__restrict__ rowptr,
                                                 const int *__restrict__ colidx,
                                                 const double *__restrict__ values, double alpha,
                                                 const double *__restrict__ x, double beta,
                                                 double *__restrict__ y, int *flag)
{
    const int idx = kxcd()*KB + threadIdx.x;
    if (idx >= bn) return;
    bool bad = false;
    double arr[B];
#pragma unroll
    for (int k = 0; k < B; ++k) arr[k] = 0.0;
    const int st = rowptr[idx], ed = rowptr[idx + 1];
    for (int j = st; j < ed; ++j) {
        const int c = colidx[j];
        if ((unsigned)c >= (unsigned)bn) {
            bad = true;
            continue;
        }
        const double *blk = values + (int64_t)j*B*B;
        double xp[B];
#pragma unroll
        for (int r = 0; r < B; ++r) xp[r] = x[(int64_t)c*B + r];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < B; ++q) v += __builtin_nontemporal_load(blk + r*B + q)*xp[q];
            arr[r] += v;
        }
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t o = (int64_t)idx*B + k;
        y[o] = alpha*arr[k] + beta*y[o];
    }
    if (bad) atomicOr(flag, UCFD_KFLAG_INDEX);
}

// Block sizes 2..7: one wave per G = 64/B block rows (wave_rows.h).  The
// wave's blocks [rowptr[i0], rowptr[i0+G]) are contiguous; they are staged
// through LDS in chunks of CB = KSPMV/B^2 blocks with all of a chunk's loads
// in flight, the column indices and x rows of each lane's first KP blocks
// are fetched before the staging loads, and lane (g, r) sums scalar row r of
// block row g in the reference order (per block: v = sum_q A[r][q] x[q],
// then arr += v).
constexpr int KSTAGE = 1600;    // staged doubles per wave of the block sweeps (12.8 KB of LDS)
// SpMV chunks: 6.4 KB of LDS per wave leaves room for 4 waves per SIMD (the
// register limit), against 3 with 12.8 KB chunks; measured 5 % faster on
// 5x5 blocks
constexpr int KSPMV = 800;

template <int B>
__global__ __launch_bounds__(64) void k_spmv_bsr_w(int bn, const int *__restrict__ rowptr,
                                                   const int *__restrict__ colidx,
                                                   const double *__restrict__ values,
                                                   double alpha, const double *__restrict__ x,
                                                   double beta, double *__restrict__ y, int *flag)
{
    constexpr int B2 = B*B, G = 64/B, CB = KSPMV/B2, U = (CB*B2 + 63)/64, KP = 5;
    __shared__ double sv[CB*B2];
    const int lane = threadIdx.x;
    const int i0 = ucfd::wave_xcd_index()*G;
    if (i0 >= bn) return;
    const int iend = min(i0 + G, bn);
    const int g = lane/B, r = lane%B;
    const int i = i0 + g;
    const bool act = g < G && i < iend;
    const int s0 = rowptr[i0], s1 = rowptr[iend];
    const int st = act ? rowptr[i] : 0, ed = act ? rowptr[i + 1] : 0;
    bool bad = false;
    double xp[KP][B];
    int cp[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
        cp[k] = -1;
        if (st + k < ed) {
            const int c = colidx[st + k];
            if ((unsigned)c < (unsigned)bn) {
                cp[k] = c;
#pragma unroll
                for (int q = 0; q < B; ++q) xp[k][q] = x[(int64_t)c*B + q];
            } else {
                bad = true;
            }
        }
    }
    double arr = 0.0;
    for (int c0 = s0; c0 < s1; c0 += CB) {
        const int c1 = min(c0 + CB, s1);
        const double *src = values + (int64_t)c0*B2;
        const int ne = (c1 - c0)*B2;
        double t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = u*64 + lane;
            t[u] = e < ne ? __builtin_nontemporal_load(src + e) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = u*64 + lane;
            if (e < ne) sv[e] = t[u];
        }
        ucfd::wave_lds_sync();
        if (act) {
            const int ja = max(st, c0), jb = min(ed, c1);
#pragma unroll
            for (int k = 0; k < KP; ++k) {
                const int j = st + k;
                if (j >= ja && j < jb && cp[k] >= 0) {
                    const double *blk = sv + (j - c0)*B2 + r*B;
                    double v = 0.0;
#pragma unroll
                    for (int q = 0; q < B; ++q) v += blk[q]*xp[k][q];
                    arr += v;
                }
            }
            for (int j = max(ja, st + KP); j < jb; ++j) {
                const int c = colidx[j];
                if ((unsigned)c >= (unsigned)bn) {
                    bad = true;
                    continue;
                }
                const double *blk = sv + (j - c0)*B2 + r*B;
                double xq[B];
#pragma unroll
                for (int q = 0; q < B; ++q) xq[q] = x[(int64_t)c*B + q];
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < B; ++q) v += blk[q]*xq[q];
                arr += v;
            }
        }
        ucfd::wave_lds_sync();
    }
    if (act) {
        const int64_t o = (int64_t)i*B + r;
        y[o] = alpha*arr + beta*y[o];
    }
    if (bad) atomicOr(flag, UCFD_KFLAG_INDEX);
}

__global__ __launch_bounds__(KB) void k_spmv_csr(int n, const int *__restrict__ rowptr,
                                                 const int *__restrict__ colidx,
                                                 const double *__restrict__ values, double alpha,
                                                 const double *__restrict__ x, double beta,
                                                 double *__restrict__ y, int *flag)
{
    const int i = kxcd()*KB + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    bool bad = false;
    for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) {
        const int c = colidx[j];
        if ((unsigned)c < (unsigned)n) s += values[j]*x[c];
        else bad = true;
    }
    y[i] = alpha*s + beta*y[i];
    if (bad) atomicOr(flag, UCFD_KFLAG_INDEX);
}

template <int B>
__global__ __launch_bounds__(KB) void k_prepare(int bn, const int *__restrict__ diagslots,
                                                const double *__restrict__ values,
                                                double *__restrict__ dv, int *flag)
{
    const int i = blockIdx.x*KB + threadIdx.x;
    if (i >= bn) return;
    double a[B*B];
    const double *src = values + (int64_t)diagslots[i]*B*B;
#pragma unroll
    for (int q = 0; q < B*B; ++q) a[q] = src[q];
    if (!ucfd::ludcmp<B>(a)) atomicOr(flag, UCFD_KFLAG_PIVOT);
#pragma unroll
    for (int q = 0; q < B*B; ++q) dv[(int64_t)i*B*B + q] = a[q];
}

// ILU(0) factorisation, one level of rows per launch, one thread per row, in
// place on the values (ilu.c:4-48, bilu.c:5-65, pbilu.c:5-76 at one thread).
// For each lower slot k of row i, with ck = colidx[k]:
//   L_ik := A_ik U_kk^-1   (block: lusubmattrans with the factored diagonal
//                           block of row ck; scalar: division by its pivot)
//   A_ij -= L_ik A_{ck,j}  for each later slot j of row i whose column row ck
//                          also holds (the last such slot of row ck: the
//                          reference's marker array keeps the last write)
// then the block form factors its diagonal block in place (ludcmp).  dv gets
// the row's factored diagonal block (scalar: the pivot) for the solves.  A
// level holds rows that read only rows of earlier levels (complete) or of
// later ones (untouched), as the reference's row order has them.
template <int B, bool SCALAR>
__global__ __launch_bounds__(KB) void k_ilu_factor(int base, int cnt, const int *__restrict__ order,
                                                   const int *__restrict__ rowptr,
                                                   const int *__restrict__ colidx,
                                                   const int *__restrict__ diagslots,
                                                   double *values, double *__restrict__ dv,
                                                   int *flag)
{
    static_assert(!SCALAR || B == 1, "the scalar form has 1x1 blocks");
    constexpr int B2 = B*B;
    const int t = kxcd()*KB + threadIdx.x;
    if (t >= cnt) return;
    const int i = order ? order[base + t] : base + t;
    const int st = rowptr[i], ed = diagslots[i], jed = rowptr[i + 1];
    bool bad = false;
    for (int k = st; k < ed; ++k) {
        const int ck = colidx[k], kk = diagslots[ck], kst = rowptr[ck], ked = rowptr[ck + 1];
        double aik[B2];
        if (SCALAR) {
            const double p = values[kk];
            bad |= p == 0.0;
            aik[0] = values[k]/p;
            values[k] = aik[0];
        } else {
#pragma unroll
            for (int q = 0; q < B2; ++q) aik[q] = values[(int64_t)k*B2 + q];
            ucfd::lusubmattrans<B>(values + (int64_t)kk*B2, aik);
#pragma unroll
            for (int q = 0; q < B2; ++q) values[(int64_t)k*B2 + q] = aik[q];
        }
        for (int j = k + 1; j < jed; ++j) {
            const int col = colidx[j];
            int iwj = -1;
            for (int q = kst; q < ked; ++q)
                if (colidx[q] == col) iwj = q;
            if (iwj < 0) continue;
            if (SCALAR) {
                values[j] -= aik[0]*values[iwj];
                continue;
            }
            const double *u = values + (int64_t)iwj*B2;
            double *a = values + (int64_t)j*B2;
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c) {
                    double v = 0.0;
#pragma unroll
                    for (int e = 0; e < B; ++e) v += aik[r*B + e]*u[e*B + c];
                    a[r*B + c] -= v;
                }
        }
    }
    double d[B2];
#pragma unroll
    for (int q = 0; q < B2; ++q) d[q] = values[(int64_t)ed*B2 + q];
    if (SCALAR) {
        bad |= d[0] == 0.0;
    } else {
        bad |= !ucfd::ludcmp<B>(d);
#pragma unroll
        for (int q = 0; q < B2; ++q) values[(int64_t)ed*B2 + q] = d[q];
    }
#pragma unroll
    for (int q = 0; q < B2; ++q) dv[(int64_t)i*B2 + q] = d[q];
    if (bad) atomicOr(flag, UCFD_KFLAG_PIVOT);
}

// The block form on one wave per G = 64/B rows of a level: lane (g, r) owns
// scalar row r of block row g.  Both block operations of a lower slot act on
// rows independently: row r of L_ik = row r of A_ik times (L_kk U_kk)^-1
// (lusubmattrans restricted to one row, the same operations in the same
// order), and row r of A_ij -= (row r of L_ik) A_kj; so the lanes of a group
// run the row loop of the thread-per-row kernel side by side, reading the
// blocks of row ck whole (an earlier or a later level: not written here).
// The group then gathers its diagonal block with lane shuffles and every
// lane factors it; lane r writes row r.  Bit-identical to k_ilu_factor.
// Requires that no lower slot of a row refers to the row itself (the host
// keeps k_ilu_factor for such a matrix).
template <int B>
__global__ __launch_bounds__(64) void k_ilu_factor_w(int base, int cnt,
                                                     const int *__restrict__ order,
                                                     const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx,
                                                     const int *__restrict__ diagslots,
                                                     double *values, double *__restrict__ dv,
                                                     int *flag)
{
    constexpr int B2 = B*B, G = 64/B;
    const int lane = threadIdx.x;
    const int t0 = ucfd::wave_xcd_index()*G;
    if (t0 >= cnt) return;
    const int g = lane/B, r = lane%B;
    const bool act = g < G && t0 + g < cnt;
    int i = 0, st = 0, ed = 0, jed = 0;
    if (act) {
        i = order ? order[base + t0 + g] : base + t0 + g;
        st = rowptr[i];
        ed = diagslots[i];
        jed = rowptr[i + 1];
    }
    for (int k = st; k < ed; ++k) {
        const int ck = colidx[k], kk = diagslots[ck], kst = rowptr[ck], ked = rowptr[ck + 1];
        const double *LU = values + (int64_t)kk*B2;
        double m[B];
#pragma unroll
        for (int c = 0; c < B; ++c) m[c] = values[(int64_t)k*B2 + r*B + c];
        // row r of M (L U)^-1, in the order of ucfd::lusubmattrans
        m[0] /= LU[0];
#pragma unroll
        for (int c = 1; c < B; ++c) {
            double val = 0.0;
#pragma unroll
            for (int q = 0; q < c; ++q) val += m[q]*LU[q*B + c];
            m[c] = (m[c] - val)/LU[c*B + c];
        }
#pragma unroll
        for (int c = B - 2; c > -1; --c) {
            double val = 0.0;
#pragma unroll
            for (int q = c + 1; q < B; ++q) val += m[q]*LU[q*B + c];
            m[c] -= val;
        }
#pragma unroll
        for (int c = 0; c < B; ++c) values[(int64_t)k*B2 + r*B + c] = m[c];
        for (int j = k + 1; j < jed; ++j) {
            const int col = colidx[j];
            int iwj = -1;
            for (int q = kst; q < ked; ++q)
                if (colidx[q] == col) iwj = q;
            if (iwj < 0) continue;
            const double *u = values + (int64_t)iwj*B2;
            double *a = values + (int64_t)j*B2 + r*B;
#pragma unroll
            for (int c = 0; c < B; ++c) {
                double v = 0.0;
#pragma unroll
                for (int e = 0; e < B; ++e) v += m[e]*u[e*B + c];
                a[c] -= v;
            }
        }
    }
    // the group's diagonal block (each lane wrote its own row), then LU
    double dr[B];
#pragma unroll
    for (int c = 0; c < B; ++c) dr[c] = act ? values[(int64_t)ed*B2 + r*B + c] : 0.0;
    double d[B2];
#pragma unroll
    for (int rr = 0; rr < B; ++rr)
#pragma unroll
        for (int c = 0; c < B; ++c) d[rr*B + c] = __shfl(dr[c], g*B + rr, 64);
    if (!act) return;
    if (!ucfd::ludcmp<B>(d) && r == 0) atomicOr(flag, UCFD_KFLAG_PIVOT);
    double o[B];
#pragma unroll
    for (int c = 0; c < B; ++c) o[c] = d[c];
#pragma unroll
    for (int rr = 1; rr < B; ++rr)
        if (r == rr)
#pragma unroll
            for (int c = 0; c < B; ++c) o[c] = d[rr*B + c];
#pragma unroll
    for (int c = 0; c < B; ++c) {
        values[(int64_t)ed*B2 + r*B + c] = o[c];
        dv[(int64_t)i*B2 + r*B + c] = o[c];
    }
}

__global__ void k_to32(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t j = (int64_t)blockIdx.x*KB + threadIdx.x; j < n; j += st) b[j] = (float)a[j];
}

// ----------------------------------------------------------- block sweeps
//
// Sweep kinds (template parameter K) of the FP64 kernels:
//   SW_LUSGS  block LU-SGS (blusgs.c:28-100): lower x_i = D_i^-1 (b_i -
//             sum_{j<diag} A_ij x_j); upper x_i -= D_i^-1 sum_{j>diag} A_ij x_j
//   SW_BILU   block ILU(0) solve with the factored blocks (bilu.c:67-130):
//             lower x_i = b_i - sum_{j<diag} L_ij x_j; upper x_i = U_ii^-1
//             (x_i - sum_{j>diag} U_ij x_j), U_ii^-1 by LU substitution
//   SW_ILU    scalar ILU(0) solve (ilu.c:50-86, B = 1): the same with the
//             products subtracted one by one and a division by U_ii
constexpr int SW_LUSGS = KRY_SW_LUSGS, SW_BILU = KRY_SW_BILU, SW_ILU = KRY_SW_ILU;

template <int B, bool UPPER, int K>
__global__ __launch_bounds__(KB) void k_sweep64(int base, int cnt, const int *__restrict__ order,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ colidx,
                                                const int *__restrict__ diagslots,
                                                const double *__restrict__ values,
                                                const double *__restrict__ dv, double *b)
{
    static_assert(K != SW_ILU || B == 1, "the scalar ILU solve has 1x1 blocks");
    const int t = kxcd()*KB + threadIdx.x;
    if (t >= cnt) return;
    const int i = order ? order[base + t] : base + t;
    const int dd = diagslots[i];
    double arr[B];
    if (!UPPER || K != SW_LUSGS) {
#pragma unroll
        for (int k = 0; k < B; ++k) arr[k] = b[(int64_t)i*B + k];
    } else {
#pragma unroll
        for (int k = 0; k < B; ++k) arr[k] = 0.0;
    }
    const int j0 = UPPER ? dd + 1 : rowptr[i];
    const int j1 = UPPER ? rowptr[i + 1] : dd;
    for (int j = j0; j < j1; ++j) {
        const int c = colidx[j];
        if (K == SW_ILU) {
            arr[0] -= __builtin_nontemporal_load(values + j)*b[c];
            continue;
        }
        const double *blk = values + (int64_t)j*B*B;
        double xc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) xc[q] = b[(int64_t)c*B + q];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < B; ++q) v += __builtin_nontemporal_load(blk + r*B + q)*xc[q];
            if (UPPER && K == SW_LUSGS) arr[r] += v;
            else arr[r] -= v;
        }
    }
    if (K != SW_LUSGS && !UPPER) {
#pragma unroll
        for (int k = 0; k < B; ++k) b[(int64_t)i*B + k] = arr[k];
        return;
    }
    if (K == SW_ILU) {
        b[i] = arr[0]/dv[i];
        return;
    }
    double lu[B*B];
#pragma unroll
    for (int q = 0; q < B*B; ++q) lu[q] = __builtin_nontemporal_load(dv + (int64_t)i*B*B + q);
    ucfd::lusub<B>(lu, arr);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t o = (int64_t)i*B + k;
        if (UPPER && K == SW_LUSGS) b[o] -= arr[k];
        else b[o] = arr[k];
    }
}

template <int B, bool UPPER>
__global__ __launch_bounds__(KB) void k_sweep32(int base, int cnt, const int *__restrict__ order,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ colidx,
                                                const int *__restrict__ diagslots,
                                                const float *__restrict__ v32,
                                                const double *__restrict__ dv, float *x32,
                                                double *b)
{
    const int t = kxcd()*KB + threadIdx.x;
    if (t >= cnt) return;
    const int i = order ? order[base + t] : base + t;
    const int dd = diagslots[i];
    float acc[B];
#pragma unroll
    for (int k = 0; k < B; ++k) acc[k] = 0.0f;
    const int j0 = UPPER ? dd + 1 : rowptr[i];
    const int j1 = UPPER ? rowptr[i + 1] : dd;
    for (int j = j0; j < j1; ++j) {
        const int c = colidx[j];
        const float *blk = v32 + (int64_t)j*B*B;
        float xc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) xc[q] = x32[(int64_t)c*B + q];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            float v = 0.0f;
#pragma unroll
            for (int q = 0; q < B; ++q) v += __builtin_nontemporal_load(blk + r*B + q)*xc[q];
            acc[r] += v;
        }
    }
    double lu[B*B], arr[B];
#pragma unroll
    for (int q = 0; q < B*B; ++q) lu[q] = __builtin_nontemporal_load(dv + (int64_t)i*B*B + q);
    if (!UPPER) {
#pragma unroll
        for (int k = 0; k < B; ++k) arr[k] = b[(int64_t)i*B + k] - (double)acc[k];
        ucfd::lusub<B>(lu, arr);
#pragma unroll
        for (int k = 0; k < B; ++k) x32[(int64_t)i*B + k] = (float)arr[k];
    } else {
#pragma unroll
        for (int k = 0; k < B; ++k) arr[k] = (double)acc[k];
        ucfd::lusub<B>(lu, arr);
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int64_t o = (int64_t)i*B + k;
            const double v = (double)x32[o] - arr[k];
            x32[o] = (float)v;
            b[o] = v;
        }
    }
}

// Block sweeps on one wave per G = 64/B rows of a contiguous row range
// (block sizes 2..7).  Lane (g, r) accumulates scalar row r of row g over
// the row's lower (lower sweep) or upper (upper sweep) blocks in the order
// of k_sweep64 / k_sweep32 above; the group then exchanges its B sums with
// lane shuffles and every lane runs the row's diagonal solve (bit-identical
// to the thread-per-row kernels).  K selects the FP64 sweep kind (SW_LUSGS
// or SW_BILU; the mixed form is LU-SGS only).  When the wave's blocks fit
// the cap words of LDS the launch provides (sized from the level's largest
// wave) they are staged through LDS (per row, contiguous loads); otherwise
// the lanes read them from HBM directly.  The rows' LU blocks are read from
// HBM by the row's lanes (staging them costs LDS, which limits the waves per
// SIMD).
template <int B, bool UPPER, bool F32, bool STAGED, int K>
__device__ __forceinline__ void bsgs_rows(int r, int g, bool act, int i, int j0, int n, int off,
                                          const int *__restrict__ colidx,
                                          const typename std::conditional<F32, float, double>::type
                                              *blocks,
                                          const double *lu_src, float *x32, double *b)
{
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int B2 = B*B, KP = 5;
    static_assert(K == SW_LUSGS || (K == SW_BILU && !F32), "sweep kind");
    // accumulate with the sign of the reference: LU-SGS adds the upper
    // products and applies D^-1 to the sum, the others subtract from x_i
    constexpr bool ADD = UPPER && K == SW_LUSGS;
    // column x rows of the first KP blocks
    T xp[KP][B];
#pragma unroll
    for (int k = 0; k < KP; ++k)
        if (k < n) {
            const int c = colidx[j0 + k];
#pragma unroll
            for (int q = 0; q < B; ++q)
                xp[k][q] = F32 ? (T)x32[(int64_t)c*B + q] : (T)b[(int64_t)c*B + q];
        }
    T acc = T(0);
    double a = 0.0;
    if (!F32 && !ADD && act) a = b[(int64_t)i*B + r];
#pragma unroll
    for (int k = 0; k < KP; ++k)
        if (k < n) {
            const T *blk = blocks + (STAGED ? (int64_t)(off + k)*B2 : (int64_t)(j0 + k)*B2) + r*B;
            T v = T(0);
#pragma unroll
            for (int q = 0; q < B; ++q) v += blk[q]*xp[k][q];
            if (F32) acc += v;
            else if (ADD) a += (double)v;
            else a -= (double)v;
        }
    for (int k = KP; k < n; ++k) {
        const int c = colidx[j0 + k];
        const T *blk = blocks + (STAGED ? (int64_t)(off + k)*B2 : (int64_t)(j0 + k)*B2) + r*B;
        T xq[B];
#pragma unroll
        for (int q = 0; q < B; ++q) xq[q] = F32 ? (T)x32[(int64_t)c*B + q] : (T)b[(int64_t)c*B + q];
        T v = T(0);
#pragma unroll
        for (int q = 0; q < B; ++q) v += blk[q]*xq[q];
        if (F32) acc += v;
        else if (ADD) a += (double)v;
        else a -= (double)v;
    }
    if (K == SW_BILU && !UPPER) {
        if (act) b[(int64_t)i*B + r] = a;
        return;
    }
    if (F32) {
        if (UPPER) a = (double)acc;
        else if (act) a = b[(int64_t)i*B + r] - (double)acc;
    }
    // the group's B sums, then the diagonal solve in every lane of the group
    double arr[B], lu[B2];
#pragma unroll
    for (int k = 0; k < B; ++k) arr[k] = __shfl(a, g*B + k, 64);
    if (!act) return;
#pragma unroll
    for (int q = 0; q < B2; ++q) lu[q] = lu_src[q];
    ucfd::lusub<B>(lu, arr);
    const int64_t o = (int64_t)i*B + r;
    if (F32) {
        if (UPPER) {
            const double v = (double)x32[o] - arr[r];
            x32[o] = (float)v;
            b[o] = v;
        } else {
            x32[o] = (float)arr[r];
        }
    } else {
        if (ADD) b[o] -= arr[r];
        else b[o] = arr[r];
    }
}

template <int B, bool UPPER, bool F32, int K>
__global__ __launch_bounds__(64) void k_bsgs_w(int base, int cnt, int cap,
                                               const int *__restrict__ rowptr,
                                               const int *__restrict__ colidx,
                                               const int *__restrict__ diagslots,
                                               const typename std::conditional<F32, float,
                                                                               double>::type
                                                   *__restrict__ vals,
                                               const double *__restrict__ dv, float *x32,
                                               double *b)
{
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int B2 = B*B, G = 64/B, PER = (5*B2 + 63)/64;
    // dynamic LDS: cap staged block words (the rows' LU blocks are read from
    // HBM by the row's lanes)
    extern __shared__ __align__(16) unsigned char kdyn[];
    T *sv = (T *)kdyn;
    const int lane = threadIdx.x;
    const int t0 = ucfd::wave_xcd_index()*G;
    if (t0 >= cnt) return;
    const int g = lane/B, r = lane%B;
    const int i = base + t0 + g;
    const bool act = g < G && t0 + g < cnt;
    int j0 = 0, n = 0;
    if (act) {
        const int dd = diagslots[i];
        j0 = UPPER ? dd + 1 : rowptr[i];
        n = (UPPER ? rowptr[i + 1] : dd) - j0;
    }
    // block offset of each row in the staged list; the total decides the path
    const int lead = (act && r == 0) ? n : 0;
    const int ex = ucfd::wave_excl_scan(lead);
    const int total = __shfl(ex + lead, 63, 64);
    const int off = __shfl(ex, g*B < 64 ? g*B : 0, 64);
    if (total*B2 > cap) {
        bsgs_rows<B, UPPER, F32, false, K>(r, g, act, i, j0, n, 0, colidx, vals,
                                        dv + (int64_t)i*B2, x32, b);
        return;
    }
    // stage: every row's blocks [j0, j0+n) to sv[off*B2 ...]
    T t[G][PER];
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
        const int jg = __shfl(j0, gg*B, 64), ng = __shfl(n, gg*B, 64)*B2;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = k*64 + lane;
            t[gg][k] = e < ng ? __builtin_nontemporal_load(vals + (int64_t)jg*B2 + e) : T(0);
        }
    }
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
        const int jg = __shfl(j0, gg*B, 64), ng = __shfl(n, gg*B, 64)*B2;
        const int og = __shfl(ex, gg*B, 64)*B2;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = k*64 + lane;
            if (e < ng) sv[og + e] = t[gg][k];
        }
        for (int e = PER*64 + lane; e < ng; e += 64) sv[og + e] = vals[(int64_t)jg*B2 + e];
    }
    ucfd::wave_lds_sync();
    bsgs_rows<B, UPPER, F32, true, K>(r, g, act, i, j0, n, off, colidx, sv,
                                      dv + (int64_t)(act ? i : base + t0)*B2, x32, b);
}

// A level whose rows take no block in the sweep's direction (the first
// colour of a lower sweep, the last of an upper one): k_bsgs_w's row update
// with an empty product sum, its expressions kept literally (a = b - 0 for
// the FP32 lower sweep, a = 0 for the LU-SGS upper one), so the bits are
// the same.  The row's own entries and its LU block are requested together
// at the start, with no index loads in front of them: one memory round trip
// per wave instead of three.  (The lower BILU sweep of such a level leaves
// b unchanged and is not launched.)
template <int B, bool UPPER, bool F32, int K>
__global__ __launch_bounds__(64) void k_bsgs_diag(int base, int cnt, const double *__restrict__ dv,
                                                  float *x32, double *b)
{
    constexpr int B2 = B*B, G = 64/B;
    constexpr bool ADD = UPPER && K == SW_LUSGS;
    static_assert(K == SW_LUSGS || (K == SW_BILU && UPPER && !F32), "sweep kind");
    const int lane = threadIdx.x;
    const int t0 = ucfd::wave_xcd_index()*G;
    if (t0 >= cnt) return;
    const int g = lane/B, r = lane%B;
    const int i = base + t0 + g;
    const bool act = g < G && t0 + g < cnt;
    const int64_t o = (int64_t)i*B + r;
    double lu[B2];
    double bo = 0.0;
    float xo = 0.f;
    if (act) {
#pragma unroll
        for (int q = 0; q < B2; ++q) lu[q] = dv[(int64_t)i*B2 + q];
        if (!ADD) bo = b[o];
        if (F32 && UPPER) xo = x32[o];
    }
    double a = 0.0;
    if (!F32 && !ADD && act) a = bo;
    if (F32) {
        if (UPPER) a = (double)0.0f;
        else if (act) a = bo - (double)0.0f;
    }
    double arr[B];
#pragma unroll
    for (int k = 0; k < B; ++k) arr[k] = __shfl(a, g*B + k, 64);
    if (!act) return;
    ucfd::lusub<B>(lu, arr);
    if (F32) {
        if (UPPER) {
            const double v = (double)xo - arr[r];
            x32[o] = (float)v;
            b[o] = v;
        } else {
            x32[o] = (float)arr[r];
        }
    } else {
        if (ADD) b[o] = b[o] - arr[r];
        else b[o] = arr[r];
    }
}

// --------------------------------------------------------------- BLAS
__device__ __forceinline__ double block_reduce(double v)
{
    __shared__ double red[KB/64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x/64] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < KB/64; ++w) s += red[w];
    __syncthreads();
    return s;
}

// out[q] = sum of part[q*stride + 0..np-1], one block per q, fixed order
__global__ __launch_bounds__(1024) void k_sum_rows(const double *part, int np, int stride,
                                                   double *out)
{
    __shared__ double red[16];
    const double *p = part + (int64_t)blockIdx.x*stride;
    double s = 0.0;
    for (int j = threadIdx.x; j < np; j += 1024) s += p[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x/64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += red[w];
        out[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(KB) void k_dot(int64_t n, const double *__restrict__ x,
                                            const double *__restrict__ y, double *part)
{
    double s = 0.0;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) s += x[i]*y[i];
    s = block_reduce(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(KB) void k_dot5(int64_t n, const double *__restrict__ rt,
                                             const double *__restrict__ sv,
                                             const double *__restrict__ t, double *part)
{
    double a = 0, b = 0, c = 0, d = 0, e = 0;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        const double sk = sv[i], tk = t[i], rk = rt[i];
        a += rk*sk; b += rk*tk; c += tk*sk; d += tk*tk; e += sk*sk;
    }
    const double v[5] = {a, b, c, d, e};
    for (int q = 0; q < 5; ++q) {
        const double r = block_reduce(v[q]);
        if (threadIdx.x == 0) part[q*KMAXPART + blockIdx.x] = r;
    }
}

// partial dots of w with V columns j0..j0+7 (fewer at the end)
__global__ __launch_bounds__(KB) void k_gemvt8(int64_t n, int k, int j0, const double *__restrict__ V,
                                               int64_t ldv, const double *__restrict__ w,
                                               double *part)
{
    double acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
    const int nj = k - j0 < 8 ? k - j0 : 8;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        const double wi = w[i];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q < nj) acc[q] += V[(int64_t)(j0 + q)*ldv + i]*wi;
    }
    for (int q = 0; q < nj; ++q) {
        const double r = block_reduce(acc[q]);
        if (threadIdx.x == 0) part[(int64_t)(j0 + q)*KMAXPART + blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(KB) void k_gemv(int64_t n, int k, const double *__restrict__ V,
                                             int64_t ldv, double alpha, const double *__restrict__ h,
                                             double *__restrict__ w, double *part)
{
    __shared__ double hs[64];
    if (threadIdx.x < k) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double ss = 0.0;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        double sum = 0.0;
        for (int j = 0; j < k; ++j) sum += V[(int64_t)j*ldv + i]*hs[j];
        const double v = alpha*sum + 1.0*w[i];
        w[i] = v;
        ss += v*v;
    }
    if (part) {
        ss = block_reduce(ss);
        if (threadIdx.x == 0) part[blockIdx.x] = ss;
    }
}

// k_gemv for k <= KM columns with the row of V loaded into registers first
// (every load issued before the first product; the same sums in the same
// order)
template <int KM>
__global__ __launch_bounds__(KB) void k_gemv_r(int64_t n, int k, const double *__restrict__ V,
                                               int64_t ldv, double alpha,
                                               const double *__restrict__ h,
                                               double *__restrict__ w, double *part)
{
    __shared__ double hs[KM];
    if (threadIdx.x < k) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double ss = 0.0;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        double vr[KM];
#pragma unroll
        for (int q = 0; q < KM; ++q) vr[q] = q < k ? V[(int64_t)q*ldv + i] : 0.0;
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < KM; ++q)
            if (q < k) sum += vr[q]*hs[q];
        const double v = alpha*sum + 1.0*w[i];
        w[i] = v;
        ss += v*v;
    }
    if (part) {
        ss = block_reduce(ss);
        if (threadIdx.x == 0) part[blockIdx.x] = ss;
    }
}

// partial dots of w with all k <= KM columns of V in one pass over w (each
// column's sum runs over the same elements in the same order as k_gemvt8's)
template <int KM>
__global__ __launch_bounds__(KB) void k_gemvt_all(int64_t n, int k, const double *__restrict__ V,
                                                  int64_t ldv, const double *__restrict__ w,
                                                  double *part)
{
    double acc[KM];
#pragma unroll
    for (int q = 0; q < KM; ++q) acc[q] = 0.0;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        // the row of V first, every load issued before the first product
        double vr[KM];
#pragma unroll
        for (int q = 0; q < KM; ++q) vr[q] = q < k ? V[(int64_t)q*ldv + i] : 0.0;
        const double wi = w[i];
#pragma unroll
        for (int q = 0; q < KM; ++q)
            if (q < k) acc[q] += vr[q]*wi;
    }
#pragma unroll
    for (int q = 0; q < KM; ++q)
        if (q < k) {
            const double r = block_reduce(acc[q]);
            if (threadIdx.x == 0) part[(int64_t)q*KMAXPART + blockIdx.x] = r;
        }
}

// The middle of a CGS2 step in one pass over V: w_i += alpha*sum_j V_ji h_j
// (k_gemv's arithmetic), then partial dots of the new w with the same k <= KM
// columns (k_gemvt_all's), the row of V held in registers between the two.
template <int KM>
__global__ __launch_bounds__(KB) void k_gemv_gemvt(int64_t n, int k,
                                                   const double *__restrict__ V, int64_t ldv,
                                                   double alpha, const double *__restrict__ h,
                                                   double *__restrict__ w, double *part)
{
    __shared__ double hs[KM];
    if (threadIdx.x < k) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double acc[KM];
#pragma unroll
    for (int q = 0; q < KM; ++q) acc[q] = 0.0;
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        double vr[KM];
#pragma unroll
        for (int q = 0; q < KM; ++q) vr[q] = q < k ? V[(int64_t)q*ldv + i] : 0.0;
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < KM; ++q)
            if (q < k) sum += vr[q]*hs[q];
        const double v = alpha*sum + 1.0*w[i];
        w[i] = v;
#pragma unroll
        for (int q = 0; q < KM; ++q)
            if (q < k) acc[q] += vr[q]*v;
    }
#pragma unroll
    for (int q = 0; q < KM; ++q)
        if (q < k) {
            const double r = block_reduce(acc[q]);
            if (threadIdx.x == 0) part[(int64_t)q*KMAXPART + blockIdx.x] = r;
        }
}

__global__ __launch_bounds__(KB) void k_scale_isqrt(int64_t n, const double *__restrict__ src,
                                                    const double *__restrict__ nsq,
                                                    double *__restrict__ dst,
                                                    double *__restrict__ dst2)
{
    const double inv = 1.0/sqrt(*nsq);
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        const double v = inv*src[i];
        dst[i] = v;
        if (dst2) dst2[i] = v;
    }
}

__global__ __launch_bounds__(KB) void k_axpby(int64_t n, double a, const double *__restrict__ x,
                                              double b, double *__restrict__ y, int mode)
{
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        if (mode == 0) y[i] += a*x[i];          // axpy
        else if (mode == 1) y[i] = a*x[i] + b*y[i];
        else if (mode == 2) y[i] = x[i];        // copy
        else y[i] = a;                          // fill
    }
}

__global__ __launch_bounds__(KB) void k_bicg_p(int64_t n, double beta, double omega,
                                               const double *__restrict__ r,
                                               const double *__restrict__ v,
                                               double *__restrict__ p, double *__restrict__ pt)
{
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        double t = p[i] + (-omega)*v[i];
        t = beta*t;
        t = t + 1.0*r[i];
        p[i] = t;
        pt[i] = t;
    }
}

// s = r - alpha*v (as copy + axpy); shat = s
// alpha = rho/pi with pi read on the device (the host forms the same
// quotient once it has pi)
__global__ __launch_bounds__(KB) void k_bicg_s(int64_t n, double rho, const double *__restrict__ pi,
                                               const double *__restrict__ r,
                                               const double *__restrict__ v,
                                               double *__restrict__ sv, double *__restrict__ shat)
{
    const double alpha = rho/(*pi);
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        const double t = r[i] + (-alpha)*v[i];
        sv[i] = t;
        shat[i] = t;
    }
}

__global__ __launch_bounds__(KB) void k_cast32(int64_t n, const double *__restrict__ a,
                                               float *__restrict__ b)
{
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) b[i] = (float)a[i];
}

__global__ __launch_bounds__(KB) void k_bicg_xr(int64_t n, double alpha,
                                                const double *__restrict__ pt, double omega,
                                                const double *__restrict__ shat,
                                                const double *__restrict__ sv,
                                                const double *__restrict__ t,
                                                double *__restrict__ x, double *__restrict__ r)
{
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) {
        double xv = x[i] + alpha*pt[i];
        x[i] = xv + omega*shat[i];
        r[i] = sv[i] + (-omega)*t[i];
    }
}

int ok() { return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP; }

template <template <int> class F, typename... A>
int by_block(int blk, A... a)
{
    switch (blk) {
    case 1: return F<1>::run(a...);
    case 2: return F<2>::run(a...);
    case 3: return F<3>::run(a...);
    case 4: return F<4>::run(a...);
    case 5: return F<5>::run(a...);
    case 6: return F<6>::run(a...);
    case 7: return F<7>::run(a...);
    default: return UCFD_ERR_ARG;
    }
}

template <int B> struct SpmvL {
    static int run(hipStream_t s, int bn, const int *rp, const int *ci, const double *v, double al,
                   const double *x, double be, double *y, int *flag)
    {
        if (B == 1) {
            hipLaunchKernelGGL(k_spmv_bsr<B>, dim3((bn + KB - 1)/KB), dim3(KB), 0, s, bn, rp, ci,
                               v, al, x, be, y, flag);
        } else {
            constexpr int G = 64/B;
            hipLaunchKernelGGL(k_spmv_bsr_w<B>, dim3((bn + G - 1)/G), dim3(64), 0, s, bn, rp, ci,
                               v, al, x, be, y, flag);
        }
        return ok();
    }
};

template <int B> struct PrepL {
    static int run(hipStream_t s, int bn, const int *ds, const double *v, double *dv, int *flag)
    {
        hipLaunchKernelGGL(k_prepare<B>, dim3((bn + KB - 1)/KB), dim3(KB), 0, s, bn, ds, v, dv,
                           flag);
        return ok();
    }
};

template <int B> struct IluL {
    static int run(hipStream_t s, bool scalar, bool rows, int base, int cnt, const int *ord,
                   const int *rp, const int *ci, const int *ds, double *v, double *dv, int *flag)
    {
        const dim3 g((cnt + KB - 1)/KB);
        if (scalar) {
            if (B != 1) return UCFD_ERR_ARG;
            hipLaunchKernelGGL((k_ilu_factor<1, true>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci,
                               ds, v, dv, flag);
        } else if (B > 1 && rows) {
            if constexpr (B > 1)
                hipLaunchKernelGGL((k_ilu_factor_w<B>), dim3((cnt + 64/B - 1)/(64/B)), dim3(64), 0,
                                   s, base, cnt, ord, rp, ci, ds, v, dv, flag);
        } else {
            hipLaunchKernelGGL((k_ilu_factor<B, false>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci,
                               ds, v, dv, flag);
        }
        return ok();
    }
};

// LDS words for the staged blocks of a wave-staged sweep launch
template <int B, typename T> int stage_cap(int maxw)
{
    const int full = KSTAGE*8/(int)sizeof(T);
    return maxw < 0 ? full : (maxw*B*B < full ? maxw*B*B : full);
}

template <int B, typename T> size_t stage_bytes(int cap)
{
    return (size_t)(cap > 0 ? cap : 1)*sizeof(T);
}

template <int B, int K>
int sweep64_launch(hipStream_t s, int upper, int base, int cnt, const int *ord, const int *rp,
                   const int *ci, const int *ds, const double *v, const double *dv, double *b,
                   int maxw)
{
    if constexpr (B > 1) {
        if (!ord) {
            const dim3 g((cnt + 64/B - 1)/(64/B));
            if (maxw == 0 && K != SW_ILU) {
                // no row of the level takes a block in this direction
                if (K == SW_BILU && !upper) return UCFD_OK;
                if (upper)
                    hipLaunchKernelGGL((k_bsgs_diag<B, true, false, K>), g, dim3(64), 0, s, base,
                                       cnt, dv, nullptr, b);
                else if constexpr (K == SW_LUSGS)
                    hipLaunchKernelGGL((k_bsgs_diag<B, false, false, K>), g, dim3(64), 0, s, base,
                                       cnt, dv, nullptr, b);
                return ok();
            }
            const int cap = stage_cap<B, double>(maxw);
            const size_t lds = stage_bytes<B, double>(cap);
            if (upper)
                hipLaunchKernelGGL((k_bsgs_w<B, true, false, K>), g, dim3(64), lds, s, base, cnt,
                                   cap, rp, ci, ds, v, dv, nullptr, b);
            else
                hipLaunchKernelGGL((k_bsgs_w<B, false, false, K>), g, dim3(64), lds, s, base, cnt,
                                   cap, rp, ci, ds, v, dv, nullptr, b);
            return ok();
        }
    }
    const dim3 g((cnt + KB - 1)/KB);
    if (upper)
        hipLaunchKernelGGL((k_sweep64<B, true, K>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci, ds, v,
                           dv, b);
    else
        hipLaunchKernelGGL((k_sweep64<B, false, K>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci, ds,
                           v, dv, b);
    return ok();
}

template <int B> struct Sw64L {
    static int run(hipStream_t s, int kind, int upper, int base, int cnt, const int *ord,
                   const int *rp, const int *ci, const int *ds, const double *v, const double *dv,
                   double *b, int maxw)
    {
        switch (kind) {
        case SW_LUSGS:
            return sweep64_launch<B, SW_LUSGS>(s, upper, base, cnt, ord, rp, ci, ds, v, dv, b, maxw);
        case SW_BILU:
            return sweep64_launch<B, SW_BILU>(s, upper, base, cnt, ord, rp, ci, ds, v, dv, b, maxw);
        case SW_ILU:
            if (B != 1) return UCFD_ERR_ARG;
            return sweep64_launch<1, SW_ILU>(s, upper, base, cnt, ord, rp, ci, ds, v, dv, b, maxw);
        default:
            return UCFD_ERR_ARG;
        }
    }
};

template <int B> struct Sw32L {
    static int run(hipStream_t s, int upper, int base, int cnt, const int *ord, const int *rp,
                   const int *ci, const int *ds, const float *v, const double *dv, float *x32,
                   double *b, int maxw)
    {
        if (B > 1 && !ord) {
            const dim3 g((cnt + 64/B - 1)/(64/B));
            if (maxw == 0) {
                // no row of the level takes a block in this direction
                if (upper)
                    hipLaunchKernelGGL((k_bsgs_diag<B, true, true, SW_LUSGS>), g, dim3(64), 0, s,
                                       base, cnt, dv, x32, b);
                else
                    hipLaunchKernelGGL((k_bsgs_diag<B, false, true, SW_LUSGS>), g, dim3(64), 0, s,
                                       base, cnt, dv, x32, b);
                return ok();
            }
            const int cap = stage_cap<B, float>(maxw);
            const size_t lds = stage_bytes<B, float>(cap);
            if (upper)
                hipLaunchKernelGGL((k_bsgs_w<B, true, true, SW_LUSGS>), g, dim3(64), lds, s, base,
                                   cnt, cap, rp, ci, ds, v, dv, x32, b);
            else
                hipLaunchKernelGGL((k_bsgs_w<B, false, true, SW_LUSGS>), g, dim3(64), lds, s, base,
                                   cnt, cap, rp, ci, ds, v, dv, x32, b);
            return ok();
        }
        const dim3 g((cnt + KB - 1)/KB);
        if (upper)
            hipLaunchKernelGGL((k_sweep32<B, true>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci, ds, v,
                               dv, x32, b);
        else
            hipLaunchKernelGGL((k_sweep32<B, false>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci, ds,
                               v, dv, x32, b);
        return ok();
    }
};

}  // namespace

int kry_spmv_bsr(hipStream_t s, int bn, int blk, const int *rowptr, const int *colidx,
                 const double *values, double alpha, const double *x, double beta, double *y,
                 int *flag)
{
    if (bn <= 0) return UCFD_OK;
    return by_block<SpmvL>(blk, s, bn, rowptr, colidx, values, alpha, x, beta, y, flag);
}

int kry_spmv_csr(hipStream_t s, int n, const int *rowptr, const int *colidx, const double *values,
                 double alpha, const double *x, double beta, double *y, int *flag)
{
    if (n <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_spmv_csr, dim3((n + KB - 1)/KB), dim3(KB), 0, s, n, rowptr, colidx,
                       values, alpha, x, beta, y, flag);
    return ok();
}

int kry_bsgs_prepare(hipStream_t s, int bn, int blk, int64_t nnz, const int *diagslots,
                     const double *values, double *diagvalues, float *vals32, int *flag)
{
    if (bn <= 0) return UCFD_OK;
    int rc = by_block<PrepL>(blk, s, bn, diagslots, values, diagvalues, flag);
    if (rc || !vals32) return rc;
    const int64_t n = nnz*blk*blk;
    hipLaunchKernelGGL(k_to32, dim3(nblocks(n)), dim3(KB), 0, s, n, values, vals32);
    return ok();
}

int kry_ilu_factor(hipStream_t s, int scalar, int rows, int blk, int base, int cnt,
                   const int *order, const int *rowptr, const int *colidx, const int *diagslots,
                   double *values, double *diagvalues, int *flag)
{
    if (cnt <= 0) return UCFD_OK;
    return by_block<IluL>(blk, s, scalar != 0, rows != 0, base, cnt, order, rowptr, colidx,
                          diagslots, values, diagvalues, flag);
}

int kry_bsgs_sweep64(hipStream_t s, int kind, int upper, int blk, int base, int cnt,
                     const int *order, const int *rowptr, const int *colidx, const int *diagslots,
                     const double *values, const double *diagvalues, double *b, int maxw)
{
    if (cnt <= 0) return UCFD_OK;
    return by_block<Sw64L>(blk, s, kind, upper, base, cnt, order, rowptr, colidx, diagslots,
                           values, diagvalues, b, maxw);
}

int kry_bsgs_sweep32(hipStream_t s, int upper, int blk, int base, int cnt, const int *order,
                     const int *rowptr, const int *colidx, const int *diagslots,
                     const float *vals32, const double *diagvalues, float *x32, double *b,
                     int maxw)
{
    if (cnt <= 0) return UCFD_OK;
    return by_block<Sw32L>(blk, s, upper, base, cnt, order, rowptr, colidx, diagslots, vals32,
                           diagvalues, x32, b, maxw);
}

int kry_parts(int64_t n) { return KMAXPART; }

int kry_dot(hipStream_t s, int64_t n, const double *x, const double *y, double *part, double *out)
{
    const int nb = nblocks(n);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(KB), 0, s, n, x, y, part);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, s, part, nb, KMAXPART, out);
    return ok();
}

int kry_dot5(hipStream_t s, int64_t n, const double *rt, const double *sv, const double *t,
             double *part, double *out)
{
    const int nb = nblocks(n);
    hipLaunchKernelGGL(k_dot5, dim3(nb), dim3(KB), 0, s, n, rt, sv, t, part);
    hipLaunchKernelGGL(k_sum_rows, dim3(5), dim3(1024), 0, s, part, nb, KMAXPART, out);
    return ok();
}

int kry_gemvt(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv, const double *w,
              double *part, double *out)
{
    if (k <= 0 || k > 64) return UCFD_ERR_ARG;
    const int nb = nblocks(n);
    if (k <= 8)
        hipLaunchKernelGGL(k_gemvt_all<8>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, w, part);
    else if (k <= 16)
        hipLaunchKernelGGL(k_gemvt_all<16>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, w, part);
    else if (k <= 32)
        hipLaunchKernelGGL(k_gemvt_all<32>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, w, part);
    else
        for (int j0 = 0; j0 < k; j0 += 8)
            hipLaunchKernelGGL(k_gemvt8, dim3(nb), dim3(KB), 0, s, n, k, j0, V, ldv, w, part);
    hipLaunchKernelGGL(k_sum_rows, dim3(k), dim3(1024), 0, s, part, nb, KMAXPART, out);
    return ok();
}

int kry_gemv_gemvt(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv, double alpha,
                   const double *h, double *w, double *part, double *out)
{
    if (k <= 0 || k > 64) return UCFD_ERR_ARG;
    if (k > 32) {
        int rc = kry_gemv(s, n, k, V, ldv, alpha, h, w, part, nullptr);
        return rc ? rc : kry_gemvt(s, n, k, V, ldv, w, part, out);
    }
    const int nb = nblocks(n);
    if (k <= 8)
        hipLaunchKernelGGL(k_gemv_gemvt<8>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w,
                           part);
    else if (k <= 16)
        hipLaunchKernelGGL(k_gemv_gemvt<16>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w,
                           part);
    else
        hipLaunchKernelGGL(k_gemv_gemvt<32>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w,
                           part);
    hipLaunchKernelGGL(k_sum_rows, dim3(k), dim3(1024), 0, s, part, nb, KMAXPART, out);
    return ok();
}

int kry_gemv(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv, double alpha,
             const double *h, double *w, double *part, double *normsq)
{
    if (k <= 0 || k > 64) return UCFD_ERR_ARG;
    const int nb = nblocks(n);
    double *pp = normsq ? part : nullptr;
    if (k <= 8)
        hipLaunchKernelGGL(k_gemv_r<8>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w, pp);
    else if (k <= 16)
        hipLaunchKernelGGL(k_gemv_r<16>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w, pp);
    else if (k <= 32)
        hipLaunchKernelGGL(k_gemv_r<32>, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w, pp);
    else
        hipLaunchKernelGGL(k_gemv, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w, pp);
    if (normsq) hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, s, part, nb, KMAXPART, normsq);
    return ok();
}

int kry_scale_by_inv_sqrt(hipStream_t s, int64_t n, const double *src, const double *normsq,
                          double *dst, double *dst2)
{
    hipLaunchKernelGGL(k_scale_isqrt, dim3(nblocks(n)), dim3(KB), 0, s, n, src, normsq, dst, dst2);
    return ok();
}

int kry_axpy(hipStream_t s, int64_t n, double a, const double *x, double *y)
{
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(KB), 0, s, n, a, x, 0.0, y, 0);
    return ok();
}

int kry_axpby(hipStream_t s, int64_t n, double a, const double *x, double b, double *y)
{
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(KB), 0, s, n, a, x, b, y, 1);
    return ok();
}

int kry_copy(hipStream_t s, int64_t n, const double *x, double *y)
{
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(KB), 0, s, n, 0.0, x, 0.0, y, 2);
    return ok();
}

int kry_fill(hipStream_t s, int64_t n, double v, double *y)
{
    hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(KB), 0, s, n, v, (const double *)nullptr, 0.0,
                       y, 3);
    return ok();
}

int kry_bicg_p(hipStream_t s, int64_t n, double beta, double omega, const double *r,
               const double *v, double *p, double *pt)
{
    hipLaunchKernelGGL(k_bicg_p, dim3(nblocks(n)), dim3(KB), 0, s, n, beta, omega, r, v, p, pt);
    return ok();
}

int kry_bicg_s(hipStream_t s, int64_t n, double rho, const double *pi, const double *r,
               const double *v, double *sv, double *shat)
{
    hipLaunchKernelGGL(k_bicg_s, dim3(nblocks(n)), dim3(KB), 0, s, n, rho, pi, r, v, sv, shat);
    return ok();
}

int kry_cast32(hipStream_t s, int64_t n, const double *a, float *b)
{
    if (n <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_cast32, dim3(nblocks(n)), dim3(KB), 0, s, n, a, b);
    return ok();
}

int kry_bicg_xr(hipStream_t s, int64_t n, double alpha, const double *pt, double omega,
                const double *shat, const double *sv, const double *t, double *x, double *r)
{
    hipLaunchKernelGGL(k_bicg_xr, dim3(nblocks(n)), dim3(KB), 0, s, n, alpha, pt, omega, shat, sv,
                       t, x, r);
    return ok();
}
