// Kernels of the FlowSys block LU-SGS path (src/lusgs/blusgs.c): element
// pack / update / sub-residual, the block-diagonal assembly and LU with the
// per-position packing of the face Jacobians, and the block Gauss-Seidel
// sweep over all face neighbours (FP64, or FP32 block products with the
// FP64 diagonal solve).
//
// FP64 kernels keep the reference's operation order and are compiled
// without FMA contraction, so they agree bit for bit with the CPU oracle.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "block_lu.h"
#include "flowsys_internal.h"
#include "flux.h"
#include "wave_rows.h"

namespace {

constexpr int FB = 256;
constexpr int FSTAGE = 16000;   // LDS bytes of staged blocks per wave
constexpr int FKP = 4;          // positions per cell whose neighbour rows are fetched first

__device__ __forceinline__ int fxcd()
{
    const int b = blockIdx.x, nblk = gridDim.x;
    const int x = b % 8, k = b / 8, per = nblk / 8, rem = nblk % 8;
    return x*per + (x < rem ? x : rem) + k;
}

inline dim3 grid(int64_t n) { return dim3((unsigned)((n + FB - 1)/FB)); }

int ok() { return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP; }

// rank_blusgs_pack (blusgs.c:17-45): rank_rhs[r][k] = rhs[k][e] and the
// packed diagonal block of r = (1/dt + a0) I, r = cell_ids[e].  One wave per
// 64 elements: the wave reads rhs one variable at a time (consecutive lanes
// on consecutive elements) into LDS, then writes the rows of rank_rhs and
// the diagonal blocks as flat ranges, consecutive lanes on consecutive
// words of each row (one thread per element would scatter every store over
// up to 64 cache lines).
constexpr int PACK_MAXV = 16;
typedef double fs_d2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(64) void k_pack(int neles, int nlocal, int nvars, int nfvars, double a0,
                                             const int *__restrict__ cell_ids,
                                             const double *__restrict__ rhs,
                                             const double *__restrict__ dt, double *rank_rhs,
                                             double *diag, int *flag)
{
    __shared__ double srhs[64*PACK_MAXV];
    const int lane = threadIdx.x;
    const int e0 = ucfd::wave_xcd_index()*64;
    if (e0 >= neles) return;
    const int nact = min(64, neles - e0);
    const int e = e0 + lane;
    int r = -1;
    double dg = 0.0;
    if (lane < nact) {
        r = cell_ids[e];
        if ((unsigned)r >= (unsigned)nlocal) {
            atomicOr(flag, UCFD_KFLAG_INDEX);
            r = -1;
        }
        dg = 1/dt[e] + a0;
        for (int k = 0; k < nvars; ++k) srhs[lane*nvars + k] = rhs[e + (int64_t)k*neles];
    }
    ucfd::wave_lds_sync();
    // uniform trip counts: every lane takes part in every shuffle
    for (int f0 = 0; f0 < nact*nvars; f0 += 64) {
        const int f = f0 + lane;
        const int el = min(f/nvars, nact - 1), k = f - el*nvars;
        const int re = __shfl(r, el, 64);
        if (f < nact*nvars && re >= 0) rank_rhs[k + (int64_t)re*nvars] = srhs[f];
    }
    const int d2 = nfvars*nfvars;
    for (int f0 = 0; f0 < nact*d2; f0 += 64) {
        const int f = f0 + lane;
        const int el = min(f/d2, nact - 1), q = f - el*d2;
        const int re = __shfl(r, el, 64);
        const double de = __shfl(dg, el, 64);
        if (f < nact*d2 && re >= 0) {
            const int row = q/nfvars, col = q - row*nfvars;
            diag[(int64_t)re*d2 + q] = row == col ? de : 0.0;
        }
    }
}

// The same through one record per cell, for cell_ids that scatter elements
// over the rank cells: (1) element order, rhs read coalesced, the element's
// diagonal value 1/dt + a0 and rhs row written as one contiguous record
// [d, rhs_0 .. rhs_{nvars-1}] at its cell (and inv[r] = e); (2) rank order,
// a wave's 64 consecutive cells read their records and write the rank rows
// and diagonal blocks as flat ranges, consecutive lanes on consecutive words,
// skipping cells that no element of this type maps to (inv[r] < 0).  The
// scattered writes shrink from a rank row and a whole diagonal block per
// element to one record of (nvars + 2) & ~1 doubles (nvars 4-9; k_pack
// otherwise).
template <int NV>
__global__ __launch_bounds__(FB) void k_bpack_rec(int neles, int nlocal, double a0,
                                                  const int *__restrict__ cell_ids,
                                                  int *__restrict__ inv,
                                                  const double *__restrict__ rhs,
                                                  const double *__restrict__ dt,
                                                  double *__restrict__ rec, int *flag)
{
    constexpr int W = (NV + 2) & ~1;
    const int e = fxcd()*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    inv[r] = e;
    double v[W];
    v[0] = 1/dt[e] + a0;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[1 + k] = rhs[e + (int64_t)k*neles];
    if (W > NV + 1) v[W - 1] = 0.0;
    fs_d2 *o = (fs_d2 *)(rec + (int64_t)r*W);
#pragma unroll
    for (int q = 0; q < W/2; ++q) o[q] = fs_d2{v[2*q], v[2*q + 1]};
}

template <int NV>
__global__ __launch_bounds__(64) void k_bpack_from_rec(int nlocal, int nfvars,
                                                       const int *__restrict__ inv,
                                                       const double *__restrict__ rec,
                                                       double *rank_rhs, double *diag)
{
    constexpr int W = (NV + 2) & ~1;
    __shared__ double srhs[64*NV];
    const int lane = threadIdx.x;
    const int r0 = ucfd::wave_xcd_index()*64;
    if (r0 >= nlocal) return;
    const int nact = min(64, nlocal - r0);
    const int r = r0 + lane;
    int has = 0;
    double dg = 0.0;
    if (lane < nact && inv[r] >= 0) {
        has = 1;
        const fs_d2 *in = (const fs_d2 *)(rec + (int64_t)r*W);
        double v[W];
#pragma unroll
        for (int q = 0; q < W/2; ++q) {
            const fs_d2 t = in[q];
            v[2*q] = t.x;
            v[2*q + 1] = t.y;
        }
        dg = v[0];
#pragma unroll
        for (int k = 0; k < NV; ++k) srhs[lane*NV + k] = v[1 + k];
    }
    ucfd::wave_lds_sync();
    // uniform trip counts: every lane takes part in every shuffle
    for (int f0 = 0; f0 < nact*NV; f0 += 64) {
        const int f = f0 + lane;
        const int el = min(f/NV, nact - 1);
        const int h = __shfl(has, el, 64);
        if (f < nact*NV && h) rank_rhs[(int64_t)r0*NV + f] = srhs[f];
    }
    const int d2 = nfvars*nfvars;
    for (int f0 = 0; f0 < nact*d2; f0 += 64) {
        const int f = f0 + lane;
        const int el = min(f/d2, nact - 1), q = f - el*d2;
        const int h = __shfl(has, el, 64);
        const double de = __shfl(dg, el, 64);
        if (f < nact*d2 && h) {
            const int row = q/nfvars, col = q - row*nfvars;
            diag[(int64_t)r0*d2 + f] = row == col ? de : 0.0;
        }
    }
}

// rank_tblusgs_pack (blusgs.c:60-94) with sa_src_jacobian / kwsst_src_jacobian
// (source_jacobian.c:6-18); BETAST = 0.09 (flux.h:8)
// NT > 0: the block (NT = 1 Spalart-Allmaras, 2 k-omega SST) is formed in
// registers with the reference's operations in order and stored once; NT = 0
// is the general form on tdiag in place.
template <int NT>
__global__ __launch_bounds__(FB) void k_tpack(int neles, int nlocal, int nvars, int nfvars, double a0,
                                              double factor, int kind,
                                              const int *__restrict__ cell_ids,
                                              const double *__restrict__ uptsb,
                                              const double *__restrict__ dsrc,
                                              const double *__restrict__ dt, double *tdiag,
                                              int *flag)
{
    const int e = blockIdx.x*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    if (NT > 0) {
        constexpr int D2 = NT > 0 ? NT*NT : 1;
        double A[D2];
#pragma unroll
        for (int q = 0; q < D2; ++q) A[q] = 0.0;
        if (kind == 1) {
            A[0] += dsrc[e + (int64_t)(nvars - 1)*neles];
        } else if (kind == 2 && NT >= 2) {
            const double kk = uptsb[e + (int64_t)(nvars - 2)*neles]/uptsb[e];
            const double bk = 0.09*kk;
            A[0] += dsrc[e + (int64_t)(nvars - 2)*neles];
            A[1] += bk > 0.0 ? bk : 0.0;
            A[(1 + NT) % D2] = dsrc[e + (int64_t)(nvars - 1)*neles];
        }
        const double dd = 1/(dt[e]*factor) + a0;
#pragma unroll
        for (int row = 0; row < NT; ++row) A[row + row*NT] += dd;
#pragma unroll
        for (int q = 0; q < D2; ++q) tdiag[(int64_t)r*D2 + q] = A[q];
        return;
    }
    const int nt = nvars - nfvars, d2 = nt*nt;
    double *A = tdiag + (int64_t)r*d2;
    for (int q = 0; q < d2; ++q) A[q] = 0.0;
    if (kind == 1) {
        A[0] += dsrc[e + (int64_t)(nvars - 1)*neles];
    } else if (kind == 2) {
        const double kk = uptsb[e + (int64_t)(nvars - 2)*neles]/uptsb[e];
        const double bk = 0.09*kk;
        A[0] += dsrc[e + (int64_t)(nvars - 2)*neles];
        A[1] += bk > 0.0 ? bk : 0.0;
        A[1 + nt] = dsrc[e + (int64_t)(nvars - 1)*neles];
    }
    for (int row = 0; row < nt; ++row) A[row + row*nt] += 1/(dt[e]*factor) + a0;
}

// rank_blusgs_update (blusgs.c:126-143)
__global__ __launch_bounds__(FB) void k_update(int neles, int nlocal, int nvars,
                                               const int *__restrict__ cell_ids,
                                               const double *__restrict__ du,
                                               double *__restrict__ upts, int *flag)
{
    const int e = blockIdx.x*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    for (int k = 0; k < nvars; ++k) upts[e + (int64_t)k*neles] += du[k + (int64_t)r*nvars];
}

// rank_sub_residual (blusgs.c:161-185)
__global__ __launch_bounds__(FB) void k_subres(int neles, int nlocal, int nvars,
                                               const int *__restrict__ cell_ids,
                                               const double *__restrict__ vol,
                                               const double *__restrict__ du,
                                               double *__restrict__ dup, double *__restrict__ res,
                                               int *flag)
{
    const int e = blockIdx.x*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    for (int k = 0; k < nvars; ++k) {
        const int64_t o = k + (int64_t)r*nvars;
        const double diff = du[o] - dup[o];
        res[e + (int64_t)k*neles] = diff*diff*vol[e];
        dup[o] = du[o];
    }
}

// pre_blusgs (blusgs.c:200-246): per cell, D += (side ? A+ : -A-)*fv over
// its positions in order, then LU; plus the packing of every position's
// neighbour block (the sweep's jval, side ? A- : -A+, blusgs.c:297-300) and
// fv = face_area*rcp_vol.  Two passes:
// (1) Face slots, one lane per slot: consecutive lanes read consecutive
// slots of the face-major jmat and face_area (coalesced, no index load
// before them); the slot's (position, cell) pairs are sorted[slot_ptr[s] ..
// slot_ptr[s+1]) (built with the structure; two for an interior face).  Round k forms fv and the neighbour
// block of every slot's k-th position, the wave transposes the blocks through
// LDS and writes whole blocks with consecutive lanes on consecutive words.
template <int NV>
__global__ __launch_bounds__(64) void k_prep_pos(int64_t nfaces,
                                                 const int *__restrict__ slot_ptr,
                                                 const int2 *__restrict__ sorted,
                                                 const signed char *__restrict__ sides,
                                                 const double *__restrict__ face_area,
                                                 const double *__restrict__ rcp_vol,
                                                 const double *__restrict__ jmat, double *jp,
                                                 double *fvp, float *jp32, float *fv32)
{
    constexpr int D2 = NV*NV;
    __shared__ double sb[64*D2];
    __shared__ int sp[64];
    const int lane = threadIdx.x;
    const int64_t s0 = (int64_t)ucfd::wave_xcd_index()*64;
    if (s0 >= nfaces) return;
    const int64_t s = s0 + lane;
    const bool act = s < nfaces;
    int b0 = 0, nr = 0;
    double fa = 0.0;
    if (act) {
        b0 = slot_ptr[s];
        nr = slot_ptr[s + 1] - b0;
        fa = face_area[s];
    }
    const int64_t half = (int64_t)D2*nfaces;
    double a[D2], b[D2];
#pragma unroll
    for (int q = 0; q < D2; ++q) {
        a[q] = act ? jmat[s + (int64_t)q*nfaces] : 0.0;
        b[q] = act ? jmat[s + (int64_t)q*nfaces + half] : 0.0;
    }
    int mx = nr;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    for (int k = 0; k < mx; ++k) {
        int p = -1;
        bool plus = false;
        if (k < nr) {
            const int2 pc = sorted[b0 + k];
            p = pc.x;
            plus = sides[p] == 1;
            const double fv = fa*rcp_vol[pc.y];
            fvp[p] = fv;
            if (fv32) fv32[p] = (float)fv;
        }
        sp[lane] = p;
#pragma unroll
        for (int q = 0; q < D2; ++q) sb[lane*D2 + q] = plus ? b[q] : -a[q];
        ucfd::wave_lds_sync();
        for (int e = lane; e < 64*D2; e += 64) {
            const int blk = e/D2;
            const int pp = sp[blk];
            if (pp < 0) continue;
            const int64_t o = (int64_t)pp*D2 + (e - blk*D2);
            const double j = sb[e];
            jp[o] = j;
            if (jp32) jp32[o] = (float)j;
        }
        ucfd::wave_lds_sync();
    }
}

// (2) Cells, one wave per G = 64/NV cells: lane (g, row) adds row `row` of
// each position's diagonal contribution v*fv to the packed diagonal in
// position order, the group exchanges the rows with lane shuffles, every
// lane factors the block, and lane (g, row) writes row `row` of the LU.
// v = (side ? A+ : -A-) is the negated neighbour block that pass (1) packed
// for the partner position, the other side of the same face slot (its block
// is -A+ when this side is "+", A- otherwise): negation is exact, so v is
// the reference's value bit for bit.  A position without a partner of the
// other side (a boundary face) reads its half of jmat.  Each lane issues the
// loads of up to FPM positions together.
constexpr int FPM = 4;

template <int NV>
__device__ __forceinline__ void prep_jmat_row(const double *__restrict__ jmat, int64_t nfaces,
                                              int64_t s, bool plus, int row, double *v)
{
    const int64_t half = (int64_t)NV*NV*nfaces;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        const int64_t q = (int64_t)(row*NV + c)*nfaces + s;
        v[c] = plus ? jmat[q] : -jmat[q + half];
    }
}

template <int NV>
__global__ __launch_bounds__(64) void k_prep_diag(int nlocal, int64_t nfaces,
                                                  const int *__restrict__ rowptr,
                                                  const int *__restrict__ partner,
                                                  const int *__restrict__ slots,
                                                  const signed char *__restrict__ sides,
                                                  const double *__restrict__ jmat,
                                                  const double *__restrict__ jp,
                                                  const double *__restrict__ fvp, double *diag,
                                                  int *flag)
{
    constexpr int D2 = NV*NV, G = 64/NV;
    const int lane = threadIdx.x;
    const int t0 = ucfd::wave_xcd_index()*G;
    if (t0 >= nlocal) return;
    const int g = lane/NV, row = lane%NV;
    const int nrows = min(G, nlocal - t0);
    const int r = t0 + g;
    const bool act = g < nrows;
    const int st = act ? rowptr[r] : 0, ed = act ? rowptr[r + 1] : 0;
    double dm[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) dm[c] = act ? diag[(int64_t)r*D2 + row*NV + c] : 0.0;
    for (int p0 = st; p0 < ed; p0 += FPM) {
        const int n = min(FPM, ed - p0);
        int pp[FPM];
        bool pl[FPM];
        double fv[FPM];
#pragma unroll
        for (int k = 0; k < FPM; ++k)
            if (k < n) {
                pp[k] = partner[p0 + k];
                pl[k] = sides[p0 + k] == 1;
                fv[k] = fvp[p0 + k];
            }
        double v[FPM][NV];
        bool other[FPM];
#pragma unroll
        for (int k = 0; k < FPM; ++k)
            if (k < n) {
                other[k] = pp[k] >= 0;
                if (other[k]) {
                    other[k] = (sides[pp[k]] == 1) != pl[k];
#pragma unroll
                    for (int c = 0; c < NV; ++c) v[k][c] = -jp[(int64_t)pp[k]*D2 + row*NV + c];
                }
            }
#pragma unroll
        for (int k = 0; k < FPM; ++k)
            if (k < n) {
                if (!other[k])
                    prep_jmat_row<NV>(jmat, nfaces, slots[p0 + k], pl[k], row, v[k]);
#pragma unroll
                for (int c = 0; c < NV; ++c) dm[c] += v[k][c]*fv[k];
            }
    }
    double m[D2];
#pragma unroll
    for (int rr = 0; rr < NV; ++rr)
#pragma unroll
        for (int c = 0; c < NV; ++c) m[rr*NV + c] = __shfl(dm[c], g*NV + rr, 64);
    if (!act) return;
    if (!ucfd::ludcmp<NV>(m) && row == 0) atomicOr(flag, UCFD_KFLAG_PIVOT);
    // this lane's row, selected with constant indices (an index that depends
    // on the lane would place m in scratch memory)
    double o[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) o[c] = m[c];
#pragma unroll
    for (int rr = 1; rr < NV; ++rr)
        if (row == rr)
#pragma unroll
            for (int c = 0; c < NV; ++c) o[c] = m[rr*NV + c];
#pragma unroll
    for (int c = 0; c < NV; ++c) diag[(int64_t)r*D2 + row*NV + c] = o[c];
}

// block Gauss-Seidel row update over every position (blusgs.c:273-373)
template <int NV, bool F32>
__global__ __launch_bounds__(FB) void k_sweep(int base, int cnt, const int *__restrict__ order,
                                              int var0, int nvars,
                                              const int *__restrict__ rowptr,
                                              const int *__restrict__ colidx,
                                              const double *__restrict__ jp,
                                              const double *__restrict__ fvp,
                                              const float *__restrict__ jp32,
                                              const float *__restrict__ fv32,
                                              const double *__restrict__ rhs,
                                              const double *__restrict__ diag, double *du,
                                              float *du32)
{
    constexpr int D2 = NV*NV;
    const int t = fxcd()*FB + threadIdx.x;
    if (t >= cnt) return;
    const int r = order ? order[base + t] : base + t;
    double x[NV];
#pragma unroll
    for (int row = 0; row < NV; ++row)
        x[row] = __builtin_nontemporal_load(rhs + var0 + row + (int64_t)r*nvars);
    const int p0 = rowptr[r], p1 = rowptr[r + 1];
    for (int p = p0; p < p1; ++p) {
        const int64_t nb = colidx[p];
        if (F32) {
            const float fv = __builtin_nontemporal_load(fv32 + p);
            float xn[NV];
#pragma unroll
            for (int col = 0; col < NV; ++col) xn[col] = du32[var0 + col + nb*nvars];
#pragma unroll
            for (int row = 0; row < NV; ++row) {
                float val = 0.0f;
#pragma unroll
                for (int col = 0; col < NV; ++col)
                    val += __builtin_nontemporal_load(jp32 + (int64_t)p*D2 + row*NV + col)*xn[col];
                x[row] -= (double)(val*fv);
            }
        } else {
            const double fv = __builtin_nontemporal_load(fvp + p);
            double xn[NV];
#pragma unroll
            for (int col = 0; col < NV; ++col) xn[col] = du[var0 + col + nb*nvars];
#pragma unroll
            for (int row = 0; row < NV; ++row) {
                double val = 0.0;
#pragma unroll
                for (int col = 0; col < NV; ++col)
                    val += __builtin_nontemporal_load(jp + (int64_t)p*D2 + row*NV + col)*xn[col];
                x[row] -= val*fv;
            }
        }
    }
    double lu[D2];
#pragma unroll
    for (int q = 0; q < D2; ++q) lu[q] = __builtin_nontemporal_load(diag + (int64_t)r*D2 + q);
    ucfd::lusub<NV>(lu, x);
#pragma unroll
    for (int row = 0; row < NV; ++row) {
        const int64_t o = var0 + row + (int64_t)r*nvars;
        du[o] = x[row];
        if (F32) du32[o] = (float)x[row];
    }
}

// The same sweep on one wave per G = 64/NV rows of a contiguous row range:
// lane (g, row) = (lane / NV, lane % NV) accumulates component `row` of
// cell g over every position in the order of k_sweep above; the group
// exchanges its NV values with lane shuffles and every lane runs the cell's
// diagonal solve.  The wave's packed blocks and fv values are contiguous
// (positions of consecutive cells are consecutive) and are staged through
// LDS with coalesced loads when they fit the launch's cap positions (dynamic
// LDS, sized from the level's largest wave so that LDS does not limit the
// waves per SIMD more than needed; at most FSTAGE bytes of blocks);
// otherwise the lanes read them from HBM directly.  The LU blocks are read
// from HBM by the row's lanes (staging them would cost LDS, and one wave
// per SIMD).  Results equal k_sweep's bit for bit.

template <int NV, bool F32, bool STAGED>
__device__ __forceinline__ void fs_rows(int row, int g, bool act, int r, int st, int ed, int p0w,
                                        int var0, int nvars, const int *__restrict__ colidx,
                                        const typename std::conditional<F32, float, double>::type
                                            *blocks,
                                        const typename std::conditional<F32, float, double>::type
                                            *fvs,
                                        const double *lu_src, const double *__restrict__ rhs,
                                        double *du, float *du32)
{
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int D2 = NV*NV;
    T xn[FKP][NV];
#pragma unroll
    for (int k = 0; k < FKP; ++k)
        if (st + k < ed) {
            const int64_t nb = colidx[st + k];
#pragma unroll
            for (int c = 0; c < NV; ++c)
                xn[k][c] = F32 ? (T)du32[var0 + c + nb*nvars] : (T)du[var0 + c + nb*nvars];
        }
    double x = act ? rhs[var0 + row + (int64_t)r*nvars] : 0.0;
#pragma unroll
    for (int k = 0; k < FKP; ++k) {
        const int p = st + k;
        if (p < ed) {
            const int64_t q = STAGED ? p - p0w : p;
            const T *blk = blocks + q*D2 + row*NV;
            T val = T(0);
#pragma unroll
            for (int c = 0; c < NV; ++c) val += blk[c]*xn[k][c];
            x -= (double)(val*fvs[q]);
        }
    }
    for (int p = st + FKP; p < ed; ++p) {
        const int64_t nb = colidx[p];
        const int64_t q = STAGED ? p - p0w : p;
        const T *blk = blocks + q*D2 + row*NV;
        T val = T(0);
#pragma unroll
        for (int c = 0; c < NV; ++c)
            val += blk[c]*(F32 ? (T)du32[var0 + c + nb*nvars] : (T)du[var0 + c + nb*nvars]);
        x -= (double)(val*fvs[q]);
    }
    double xv[NV], lu[D2], myrow[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) xv[c] = __shfl(x, g*NV + c, 64);
    // each lane reads its own row of the LU block; the group's lanes then
    // exchange the rows, so every lane holds the whole block
#pragma unroll
    for (int c = 0; c < NV; ++c) myrow[c] = act ? lu_src[row*NV + c] : 0.0;
#pragma unroll
    for (int rr = 0; rr < NV; ++rr)
#pragma unroll
        for (int c = 0; c < NV; ++c) lu[rr*NV + c] = __shfl(myrow[c], g*NV + rr, 64);
    if (!act) return;
    ucfd::lusub<NV>(lu, xv);
    const int64_t o = var0 + row + (int64_t)r*nvars;
    du[o] = xv[row];
    if (F32) du32[o] = (float)xv[row];
}

// staged positions per wave: FSTAGE bytes of blocks, at most 8 per cell
template <int NV, typename T> constexpr int fs_stage_max()
{
    return FSTAGE/((int)sizeof(T)*NV*NV) < 8*(64/NV) ? FSTAGE/((int)sizeof(T)*NV*NV) : 8*(64/NV);
}

template <int NV, bool F32>
__global__ __launch_bounds__(64) void k_sweep_w(int base, int cnt, int cap, int var0, int nvars,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ colidx,
                                                const typename std::conditional<F32, float,
                                                                                double>::type
                                                    *__restrict__ jp,
                                                const typename std::conditional<F32, float,
                                                                                double>::type
                                                    *__restrict__ fvp,
                                                const double *__restrict__ rhs,
                                                const double *__restrict__ diag, double *du,
                                                float *du32)
{
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int D2 = NV*NV, G = 64/NV;
    constexpr int CP = fs_stage_max<NV, T>();
    constexpr int UJ = (CP*D2 + 63)/64, UF = (CP + 63)/64;
    // dynamic LDS: cap blocks, then cap fv values (the rows' LU blocks are
    // read from HBM by each row's lanes)
    extern __shared__ __align__(16) unsigned char fdyn[];
    T *sj = (T *)fdyn;
    T *sf = sj + (size_t)cap*D2;
    const int lane = threadIdx.x;
    const int t0 = ucfd::wave_xcd_index()*G;
    if (t0 >= cnt) return;
    const int g = lane/NV, row = lane%NV;
    const int nrows = min(G, cnt - t0);
    const int r0 = base + t0, r = r0 + g;
    const bool act = g < nrows;
    const int p0w = rowptr[r0], p1w = rowptr[r0 + nrows];
    const int st = act ? rowptr[r] : 0, ed = act ? rowptr[r + 1] : 0;
    if (p1w - p0w > cap) {
        fs_rows<NV, F32, false>(row, g, act, r, st, ed, 0, var0, nvars, colidx, jp, fvp,
                                diag + (int64_t)r*D2, rhs, du, du32);
        return;
    }
    const int nj = (p1w - p0w)*D2, nf = p1w - p0w;
    const T *js = jp + (int64_t)p0w*D2;
    const T *fs = fvp + p0w;
    T tj[UJ], tf[UF];
#pragma unroll
    for (int u = 0; u < UJ; ++u) {
        const int e = u*64 + lane;
        tj[u] = e < nj ? __builtin_nontemporal_load(js + e) : T(0);
    }
#pragma unroll
    for (int u = 0; u < UF; ++u) {
        const int e = u*64 + lane;
        tf[u] = e < nf ? __builtin_nontemporal_load(fs + e) : T(0);
    }
#pragma unroll
    for (int u = 0; u < UJ; ++u) {
        const int e = u*64 + lane;
        if (e < nj) sj[e] = tj[u];
    }
#pragma unroll
    for (int u = 0; u < UF; ++u) {
        const int e = u*64 + lane;
        if (e < nf) sf[e] = tf[u];
    }
    ucfd::wave_lds_sync();
    fs_rows<NV, F32, true>(row, g, act, r, st, ed, p0w, var0, nvars, colidx, sj, sf,
                           diag + (int64_t)(act ? r : r0)*D2, rhs, du, du32);
}

// ------------------------------------------------------ scalar LU-SGS
// Rank arrays u / du / diag are [nvars, nlocal] (src/lusgs/flowsys.h:20-24).

// lusgs_pack (lusgs.c:8-37) in rank order, for nvars without a record form
// below: inv[r] = the element whose cell_ids entry is r (k_lpack_inv, one
// 4-byte exchange per element; a second element on the same cell raises
// UCFD_KFLAG_DUP), then one thread per rank gathers the element's rows and
// writes the rank rows coalesced: u = uptsb, du = rhs, diag =
// 1/(dt*factor) + a0 + dsrc (factor = turb_factor on turbulence rows).
__global__ __launch_bounds__(FB) void k_lpack_inv(int neles, int nlocal,
                                                  const int *__restrict__ cell_ids,
                                                  int *__restrict__ inv, int *flag)
{
    const int e = fxcd()*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    if (atomicExch(inv + r, e) >= 0) atomicOr(flag, UCFD_KFLAG_DUP);
}

__global__ __launch_bounds__(FB) void k_lpack_rank(int neles, int nlocal, int nvars, int nfvars,
                                                   double turb_factor, double a0,
                                                   const int *__restrict__ inv,
                                                   const double *__restrict__ upts,
                                                   const double *__restrict__ rhs,
                                                   const double *__restrict__ dt,
                                                   const double *__restrict__ dsrc,
                                                   double *__restrict__ u,
                                                   double *__restrict__ du,
                                                   double *__restrict__ diag)
{
    const int r = fxcd()*FB + threadIdx.x;
    if (r >= nlocal) return;
    const int e = inv[r];
    if (e < 0) return;                  // a rank of another element type
    const double dte = dt[e];
    for (int k = 0; k < nvars; ++k) {
        const int64_t ri = r + (int64_t)k*nlocal, ei = e + (int64_t)k*neles;
        const double factor = k < nfvars ? 1.0 : turb_factor;
        u[ri] = upts[ei];
        du[ri] = rhs[ei];
        diag[ri] = 1.0/(dte*factor) + a0 + dsrc[ei];
    }
}

__global__ __launch_bounds__(FB) void k_fill_int(int64_t n, int *a, int v)
{
    const int64_t st = (int64_t)gridDim.x*FB;
    for (int64_t i = (int64_t)blockIdx.x*FB + threadIdx.x; i < n; i += st) a[i] = v;
}

// Record forms of lusgs_pack / rank_lusgs_update for a random element ->
// cell map.  Pack: (1) element order, every input read coalesced, the
// element's 3*NV output values written as one contiguous record at
// rec[r*W] (with the inverse-map exchange of k_lpack_inv); (2) rank order,
// each record read whole and written to the [nvars, nlocal] rows
// coalesced.  Update: (1) rank order, du's NV rows to records; (2) element
// order, one record per element gathered (one or two cache lines instead of
// NV) and added to uptsb.  Same values and expressions as k_lpack_rank /
// k_lupdate.  W = 3*NV (pack) or NV (update), rounded up to even, so the
// records are read and written as 16-byte pairs.
template <int NV> constexpr int rec_w(int per) { return (per*NV + 1) & ~1; }


template <int NV>
__global__ __launch_bounds__(FB) void k_lpack_rec(int neles, int nlocal, int nfvars,
                                                  double turb_factor, double a0,
                                                  const int *__restrict__ cell_ids,
                                                  int *__restrict__ inv,
                                                  const double *__restrict__ upts,
                                                  const double *__restrict__ rhs,
                                                  const double *__restrict__ dt,
                                                  const double *__restrict__ dsrc,
                                                  double *__restrict__ rec, int *flag)
{
    constexpr int W = rec_w<NV>(3);
    const int e = fxcd()*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    if (atomicExch(inv + r, e) >= 0) atomicOr(flag, UCFD_KFLAG_DUP);
    const double dte = dt[e];
    double v[W];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int64_t ei = e + (int64_t)k*neles;
        const double factor = k < nfvars ? 1.0 : turb_factor;
        v[k] = upts[ei];
        v[NV + k] = rhs[ei];
        v[2*NV + k] = 1.0/(dte*factor) + a0 + dsrc[ei];
    }
    if (W > 3*NV) v[W - 1] = 0.0;
    fs_d2 *o = (fs_d2 *)(rec + (int64_t)r*W);
#pragma unroll
    for (int q = 0; q < W/2; ++q) o[q] = fs_d2{v[2*q], v[2*q + 1]};
}

template <int NV>
__global__ __launch_bounds__(FB) void k_lpack_from_rec(int nlocal, const int *__restrict__ inv,
                                                       const double *__restrict__ rec,
                                                       double *__restrict__ u,
                                                       double *__restrict__ du,
                                                       double *__restrict__ diag)
{
    constexpr int W = rec_w<NV>(3);
    const int r = fxcd()*FB + threadIdx.x;
    if (r >= nlocal || inv[r] < 0) return;      // a cell of another element type
    const fs_d2 *in = (const fs_d2 *)(rec + (int64_t)r*W);
    double v[W];
#pragma unroll
    for (int q = 0; q < W/2; ++q) {
        const fs_d2 t = in[q];
        v[2*q] = t.x;
        v[2*q + 1] = t.y;
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int64_t ri = r + (int64_t)k*nlocal;
        u[ri] = v[k];
        du[ri] = v[NV + k];
        diag[ri] = v[2*NV + k];
    }
}

template <int NV>
__global__ __launch_bounds__(FB) void k_lupdate_rec(int nlocal, const double *__restrict__ du,
                                                    double *__restrict__ rec)
{
    constexpr int W = rec_w<NV>(1);
    const int r = fxcd()*FB + threadIdx.x;
    if (r >= nlocal) return;
    double v[W];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = du[r + (int64_t)k*nlocal];
    if (W > NV) v[W - 1] = 0.0;
    fs_d2 *o = (fs_d2 *)(rec + (int64_t)r*W);
#pragma unroll
    for (int q = 0; q < W/2; ++q) o[q] = fs_d2{v[2*q], v[2*q + 1]};
}

template <int NV>
__global__ __launch_bounds__(FB) void k_lupdate_from_rec(int neles, int nlocal,
                                                         const int *__restrict__ cell_ids,
                                                         const double *__restrict__ rec,
                                                         double *__restrict__ upts, int *flag)
{
    constexpr int W = rec_w<NV>(1);
    const int e = fxcd()*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    const fs_d2 *in = (const fs_d2 *)(rec + (int64_t)r*W);
    double v[W];
#pragma unroll
    for (int q = 0; q < W/2; ++q) {
        const fs_d2 t = in[q];
        v[2*q] = t.x;
        v[2*q + 1] = t.y;
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) upts[e + (int64_t)k*neles] += v[k];
}

// rank_lusgs_update (lusgs.c:54-70)
__global__ __launch_bounds__(FB) void k_lupdate(int neles, int nlocal, int nvars,
                                                const int *__restrict__ cell_ids,
                                                const double *__restrict__ du,
                                                double *__restrict__ upts, int *flag)
{
    const int e = fxcd()*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    for (int k = 0; k < nvars; ++k) upts[e + (int64_t)k*neles] += du[r + (int64_t)k*nlocal];
}

// One level of lower_sweep / upper_sweep (lusgs.c:157-257): rows
// base..base+cnt-1 (or order[...]), in place on du.  NS rows (NT = 0) with
// ns_flux_container, turbulence rows (NT turbulence variables) with
// rans_flux_container (flux.c:77-90); the flux difference of diff_flux
// (lusgs.c:143-155) in the reference's order.  The indices of a cell's first
// LKP positions are read together, then every gather those positions need,
// then the face terms are added in position order (one memory round trip for
// them instead of two per position); further positions follow one by one.
constexpr int LKP = 4;

template <int ND, int NT>
struct LFace {
    static constexpr int NF = ND + 2, DNV = NT > 0 ? NT : ND + 2;
    double nf[ND], fv, lam, uf[NF], dnb[DNV], ut[NT > 0 ? NT : 1];
};

template <int ND, int NT>
__device__ __forceinline__ void lface_load(LFace<ND, NT> &F, int64_t nb, int64_t sl, double side,
                                           double rv, double kappa, int nfvars, int nlocal,
                                           int64_t nfaces, const double *__restrict__ lambdaf,
                                           const double *__restrict__ face_area,
                                           const double *__restrict__ face_normal,
                                           const double *__restrict__ u, const double *du)
{
    constexpr int NF = LFace<ND, NT>::NF, DNV = LFace<ND, NT>::DNV;
    const int nv0 = NT > 0 ? nfvars : 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) F.nf[d] = side*face_normal[sl + (int64_t)d*nfaces];
    F.fv = face_area[sl]*rv;
    F.lam = kappa*lambdaf[sl];
#pragma unroll
    for (int k = 0; k < NF; ++k) F.uf[k] = u[nb + (int64_t)k*nlocal];
#pragma unroll
    for (int k = 0; k < DNV; ++k) F.dnb[k] = du[nb + (int64_t)(k + nv0)*nlocal];
    if (NT > 0) {
#pragma unroll
        for (int k = 0; k < DNV; ++k) F.ut[k] = u[nb + (int64_t)(nfvars + k)*nlocal];
    }
}

template <int ND, int NT>
__device__ __forceinline__ void lface_add(const LFace<ND, NT> &F, double *df)
{
    constexpr int NF = LFace<ND, NT>::NF, DNV = LFace<ND, NT>::DNV;
    double dfj[DNV];
    if (NT == 0) {
        double uf[NF], dnb[DNV], nf[ND];
#pragma unroll
        for (int k = 0; k < NF; ++k) uf[k] = F.uf[k];
#pragma unroll
        for (int k = 0; k < DNV; ++k) dnb[k] = F.dnb[k];
#pragma unroll
        for (int d = 0; d < ND; ++d) nf[d] = F.nf[d];
        ucfd::diff_flux<double, ND>(uf, dnb, nf, dfj);
    } else {
        // flow rows of du are zero: du + u = u for them
        double c0 = 0.0, c1 = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) c0 += F.uf[d + 1]*F.nf[d];
        c0 /= F.uf[0];
        double up[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) up[k] = 0.0 + F.uf[k];
#pragma unroll
        for (int d = 0; d < ND; ++d) c1 += up[d + 1]*F.nf[d];
        c1 /= up[0];
#pragma unroll
        for (int k = 0; k < DNV; ++k) dfj[k] = (F.dnb[k] + F.ut[k])*c1 - F.ut[k]*c0;
    }
#pragma unroll
    for (int k = 0; k < DNV; ++k) df[k] += (dfj[k] - F.lam*F.dnb[k])*F.fv;
}

template <int ND, int NT, bool UPPER>
__global__ __launch_bounds__(FB) void k_lsweep(int base, int cnt, const int *__restrict__ order,
                                               int nvars, int nfvars, double kappa,
                                               const double *__restrict__ lambdaf, int nlocal,
                                               int64_t nfaces, const int *__restrict__ rowptr,
                                               const int *__restrict__ colidx,
                                               const signed char *__restrict__ sides,
                                               const int *__restrict__ slots,
                                               const double *__restrict__ face_area,
                                               const double *__restrict__ face_normal,
                                               const double *__restrict__ rcp_vol,
                                               const double *__restrict__ diag,
                                               const double *__restrict__ u, double *du)
{
    constexpr int DNV = LFace<ND, NT>::DNV;    // rows this sweep updates
    const int t = fxcd()*FB + threadIdx.x;
    if (t >= cnt) return;
    const int r = order ? order[base + t] : base + t;
    const int nv0 = NT > 0 ? nfvars : 0;
    const double rv = rcp_vol[r];
    const int p0 = rowptr[r], p1 = rowptr[r + 1];
    double df[DNV];
#pragma unroll
    for (int k = 0; k < DNV; ++k) df[k] = 0.0;
    int64_t nbk[LKP], slk[LKP];
    double sdk[LKP];
    bool tk[LKP];
#pragma unroll
    for (int k = 0; k < LKP; ++k) {
        tk[k] = false;
        if (p0 + k < p1) {
            nbk[k] = colidx[p0 + k];
            slk[k] = slots[p0 + k];
            sdk[k] = (double)sides[p0 + k];
            tk[k] = UPPER ? nbk[k] > r : nbk[k] < r;
        }
    }
    LFace<ND, NT> F[LKP];
#pragma unroll
    for (int k = 0; k < LKP; ++k)
        if (tk[k])
            lface_load<ND, NT>(F[k], nbk[k], slk[k], sdk[k], rv, kappa, nfvars, nlocal, nfaces,
                               lambdaf, face_area, face_normal, u, du);
#pragma unroll
    for (int k = 0; k < LKP; ++k)
        if (tk[k]) lface_add<ND, NT>(F[k], df);
    for (int p = p0 + LKP; p < p1; ++p) {
        const int64_t nb = colidx[p];
        if (UPPER ? !(nb > r) : !(nb < r)) continue;
        LFace<ND, NT> G;
        lface_load<ND, NT>(G, nb, slots[p], (double)sides[p], rv, kappa, nfvars, nlocal, nfaces,
                           lambdaf, face_area, face_normal, u, du);
        lface_add<ND, NT>(G, df);
    }
#pragma unroll
    for (int k = 0; k < DNV; ++k) {
        const int64_t l = r + (int64_t)(k + nv0)*nlocal;
        if (UPPER) du[l] -= (0.5*df[k]/diag[l]);
        else du[l] = (du[l] - 0.5*df[k])/diag[l];
    }
}

// pre_lusgs (lusgs.c:87-113) with the per-position face records of the
// scalar method: the face sum sd = sum_p 0.5*kappa*lambdaf*face_area*rcp_vol
// is added to rows nv0..nv1-1 of diag, and the same pass writes geo[p] =
// {side*n (ND values), face_area[slot]*rcp_vol[cell]} padded to 4 doubles
// (when geo is given) and lam[p] = lambdaf[slot].  The sweeps then read a
// cell's records in position order (coalesced) instead of gathering five
// values through the slot index; every stored value is the product the
// sweep itself forms (lusgs.c:185-195), so the results do not change by a
// bit.  A wave's 64 cells own one contiguous range of positions: the
// records of its first LPCAP positions are staged in LDS and written out
// with consecutive lanes on consecutive records (one lane per cell would
// write 32-byte pieces at the position stride, partial cache lines that
// measured 2.15x the record bytes in HBM writes); positions beyond LPCAP
// (cells with many faces) are written directly.
constexpr int LPCAP = 256;   // staged positions per wave (64 tets)

template <int ND>
__global__ __launch_bounds__(FB) void k_lprep_pos(int nlocal, int64_t nfaces, int nv0, int nv1,
                                                  double kappa, const int *__restrict__ rowptr,
                                                  const int *__restrict__ slots,
                                                  const signed char *__restrict__ sides,
                                                  const double *__restrict__ face_area,
                                                  const double *__restrict__ face_normal,
                                                  const double *__restrict__ rcp_vol,
                                                  const double *__restrict__ lambdaf,
                                                  double *diag, double4 *__restrict__ geo,
                                                  double *__restrict__ lam)
{
    __shared__ double4 sgeo[FB/64][LPCAP];
    __shared__ double slam[FB/64][LPCAP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w0 = fxcd()*FB + wv*64;   // the wave's first cell
    if (w0 >= nlocal) return;
    const int r = w0 + lane;
    const bool act = r < nlocal;
    const int P0 = rowptr[w0], P1 = rowptr[min(w0 + 64, nlocal)];
    const int p0 = act ? rowptr[r] : 0, p1 = act ? rowptr[r + 1] : 0;
    const double rv = act ? rcp_vol[r] : 0.0;
    // the slot indices of the first LKP positions together, then all their
    // gathers, then the face terms in position order
    int64_t sl[LKP];
    double lf[LKP], fa[LKP], fn[LKP][ND], side[LKP];
#pragma unroll
    for (int k = 0; k < LKP; ++k) {
        sl[k] = 0;
        lf[k] = fa[k] = side[k] = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) fn[k][d] = 0.0;
        if (p0 + k < p1) sl[k] = slots[p0 + k];
    }
#pragma unroll
    for (int k = 0; k < LKP; ++k)
        if (p0 + k < p1) {
            lf[k] = lambdaf[sl[k]];
            fa[k] = face_area[sl[k]];
            if (geo) {
                side[k] = (double)sides[p0 + k];
#pragma unroll
                for (int d = 0; d < ND; ++d) fn[k][d] = face_normal[sl[k] + (int64_t)d*nfaces];
            }
        }
    double sd = 0.0;
    auto face = [&](int p, double lfk, double fak, const double *fnk, double sidek) {
        sd += 0.5*(lfk*kappa)*fak*rv;
        const int j = p - P0;
        if (geo) {
            double g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int d = 0; d < ND; ++d) g[d] = sidek*fnk[d];
            g[ND] = fak*rv;
            const double4 gv = double4{g[0], g[1], g[2], g[3]};
            if (j < LPCAP) sgeo[wv][j] = gv;
            else geo[p] = gv;
        }
        if (j < LPCAP) slam[wv][j] = lfk;
        else lam[p] = lfk;
    };
#pragma unroll
    for (int k = 0; k < LKP; ++k)
        if (p0 + k < p1) face(p0 + k, lf[k], fa[k], fn[k], side[k]);
    for (int p = p0 + LKP; p < p1; ++p) {
        const int64_t s2 = slots[p];
        double f2[ND] = {};
        double sd2 = 0.0;
        if (geo) {
            sd2 = (double)sides[p];
#pragma unroll
            for (int d = 0; d < ND; ++d) f2[d] = face_normal[s2 + (int64_t)d*nfaces];
        }
        face(p, lambdaf[s2], face_area[s2], f2, sd2);
    }
    if (act)
        for (int k = nv0; k < nv1; ++k) diag[r + (int64_t)k*nlocal] += sd;
    ucfd::wave_lds_sync();
    const int nst = min(P1 - P0, LPCAP);
    for (int j = lane; j < nst; j += 64) {
        if (geo) geo[P0 + j] = sgeo[wv][j];
        lam[P0 + j] = slam[wv][j];
    }
}

// k_lsweep reading the per-position records of k_lprep_pos
template <int ND, int NT, bool UPPER>
__global__ __launch_bounds__(FB) void k_lsweep_p(int base, int cnt, const int *__restrict__ order,
                                                 int nfvars, double kappa, int nlocal,
                                                 const int *__restrict__ rowptr,
                                                 const int *__restrict__ colidx,
                                                 const double4 *__restrict__ geo,
                                                 const double *__restrict__ lam,
                                                 const double *__restrict__ diag,
                                                 const double *__restrict__ u, double *du)
{
    constexpr int NF = ND + 2, DNV = LFace<ND, NT>::DNV;
    const int t = fxcd()*FB + threadIdx.x;
    if (t >= cnt) return;
    const int r = order ? order[base + t] : base + t;
    const int nv0 = NT > 0 ? nfvars : 0;
    const int p0 = rowptr[r], p1 = rowptr[r + 1];
    // the cell's own rows first, so their latency overlaps the gathers
    double odu[DNV], odg[DNV];
#pragma unroll
    for (int k = 0; k < DNV; ++k) {
        const int64_t l = r + (int64_t)(k + nv0)*nlocal;
        odu[k] = du[l];
        odg[k] = diag[l];
    }
    double df[DNV];
#pragma unroll
    for (int k = 0; k < DNV; ++k) df[k] = 0.0;
    int64_t nbk[LKP];
    bool tk[LKP];
#pragma unroll
    for (int k = 0; k < LKP; ++k) {
        tk[k] = false;
        if (p0 + k < p1) {
            nbk[k] = colidx[p0 + k];
            tk[k] = UPPER ? nbk[k] > r : nbk[k] < r;
        }
    }
    LFace<ND, NT> F[LKP];
    auto load = [&](LFace<ND, NT> &G, int64_t nb, int p) {
        const double4 g = geo[p];
        const double gv[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
        for (int d = 0; d < ND; ++d) G.nf[d] = gv[d];
        G.fv = gv[ND];
        G.lam = kappa*lam[p];
#pragma unroll
        for (int k = 0; k < NF; ++k) G.uf[k] = u[nb + (int64_t)k*nlocal];
#pragma unroll
        for (int k = 0; k < DNV; ++k) G.dnb[k] = du[nb + (int64_t)(k + nv0)*nlocal];
        if (NT > 0) {
#pragma unroll
            for (int k = 0; k < DNV; ++k) G.ut[k] = u[nb + (int64_t)(nfvars + k)*nlocal];
        }
    };
#pragma unroll
    for (int k = 0; k < LKP; ++k)
        if (tk[k]) load(F[k], nbk[k], p0 + k);
#pragma unroll
    for (int k = 0; k < LKP; ++k)
        if (tk[k]) lface_add<ND, NT>(F[k], df);
    for (int p = p0 + LKP; p < p1; ++p) {
        const int64_t nb = colidx[p];
        if (UPPER ? !(nb > r) : !(nb < r)) continue;
        LFace<ND, NT> G;
        load(G, nb, p);
        lface_add<ND, NT>(G, df);
    }
#pragma unroll
    for (int k = 0; k < DNV; ++k) {
        const int64_t l = r + (int64_t)(k + nv0)*nlocal;
        if (UPPER) du[l] = odu[k] - (0.5*df[k]/odg[k]);
        else du[l] = (odu[k] - 0.5*df[k])/odg[k];
    }
}

// The first level of a scalar sweep schedule: cells with no neighbour
// visited before them (no lower neighbour in the lower sweep, no upper one
// in the upper sweep), so no face is taken.  The update is k_lsweep_p's with
// its face sum still +0.0, written out literally so the bits agree (0.5*0.0
// divided by diag keeps diag's sign); no index or face data is read.
template <bool UPPER>
__global__ __launch_bounds__(FB) void k_lsweep_free(int base, int cnt,
                                                    const int *__restrict__ order, int nv0,
                                                    int nv1, int nlocal,
                                                    const double *__restrict__ diag, double *du)
{
    const int t = fxcd()*FB + threadIdx.x;
    if (t >= cnt) return;
    const int r = order ? order[base + t] : base + t;
    const double df = 0.0;
    for (int k = nv0; k < nv1; ++k) {
        const int64_t l = r + (int64_t)k*nlocal;
        if (UPPER) du[l] -= (0.5*df/diag[l]);
        else du[l] = (du[l] - 0.5*df)/diag[l];
    }
}

__global__ __launch_bounds__(FB) void k_zero(int64_t n, double *a, double *b, float *c)
{
    const int64_t st = (int64_t)gridDim.x*FB;
    for (int64_t i = (int64_t)blockIdx.x*FB + threadIdx.x; i < n; i += st) {
        a[i] = 0.0;
        b[i] = 0.0;
        if (c) c[i] = 0.0f;
    }
}

template <int NV> struct PrepL {
    static int run(hipStream_t s, int nlocal, int64_t nfaces, int64_t npos, const int *rowptr,
                   const int *slot_ptr, const int2 *sorted, const int *partner,
                   const int *slots,
                   const signed char *sides, const double *face_area, const double *rcp_vol,
                   double *diag, const double *jmat, double *jp, double *fvp, float *jp32,
                   float *fv32, int *flag)
    {
        if (npos > 0 && nfaces > 0)
            hipLaunchKernelGGL(k_prep_pos<NV>, dim3((unsigned)((nfaces + 63)/64)), dim3(64), 0, s,
                               nfaces, slot_ptr, sorted, sides, face_area, rcp_vol,
                               jmat, jp, fvp, jp32, fv32);
        constexpr int G = 64/NV;
        hipLaunchKernelGGL(k_prep_diag<NV>, dim3((unsigned)((nlocal + G - 1)/G)), dim3(64), 0, s,
                           nlocal, nfaces, rowptr, partner, slots, sides, jmat, jp, fvp, diag,
                           flag);
        return ok();
    }
};

template <int NV, typename T> size_t sweep_lds(int cap)
{
    return (size_t)cap*(NV*NV + 1)*sizeof(T);
}

template <int NV> struct SweepL {
    static int run(hipStream_t s, bool f32, int base, int cnt, const int *order, int var0,
                   int nvars, const int *rowptr, const int *colidx, const double *jp,
                   const double *fvp, const float *jp32, const float *fv32, const double *rhs,
                   const double *diag, double *du, float *du32, int maxw)
    {
        if (!order) {
            const dim3 g((unsigned)((cnt + 64/NV - 1)/(64/NV)));
            if (f32) {
                constexpr int CP = fs_stage_max<NV, float>();
                const int cap = maxw < 0 || maxw > CP ? CP : maxw;
                const size_t lds = sweep_lds<NV, float>(cap);
                hipLaunchKernelGGL((k_sweep_w<NV, true>), g, dim3(64), lds, s, base, cnt, cap, var0,
                                   nvars, rowptr, colidx, jp32, fv32, rhs, diag, du, du32);
            } else {
                constexpr int CP = fs_stage_max<NV, double>();
                const int cap = maxw < 0 || maxw > CP ? CP : maxw;
                const size_t lds = sweep_lds<NV, double>(cap);
                hipLaunchKernelGGL((k_sweep_w<NV, false>), g, dim3(64), lds, s, base, cnt, cap,
                                   var0, nvars, rowptr, colidx, jp, fvp, rhs, diag, du, du32);
            }
            return ok();
        }
        if (f32)
            hipLaunchKernelGGL((k_sweep<NV, true>), grid(cnt), dim3(FB), 0, s, base, cnt, order,
                               var0, nvars, rowptr, colidx, jp, fvp, jp32, fv32, rhs, diag, du,
                               du32);
        else
            hipLaunchKernelGGL((k_sweep<NV, false>), grid(cnt), dim3(FB), 0, s, base, cnt, order,
                               var0, nvars, rowptr, colidx, jp, fvp, jp32, fv32, rhs, diag, du,
                               du32);
        return ok();
    }
};

template <template <int> class F, typename... A>
int by_nv(int nv, A... a)
{
    switch (nv) {
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

}  // namespace

int fs_launch_pack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars, double a0,
                   const int *cell_ids, const double *rhs, const double *dt, double *rank_rhs,
                   double *diag, int *flag, int *inv, double *rec)
{
    if (neles <= 0) return UCFD_OK;
    if (nvars > PACK_MAXV) return UCFD_ERR_ARG;
    if (inv && rec && nvars >= 4 && nvars <= 9) {
        const int64_t nb = ((int64_t)nlocal + FB - 1)/FB;
        hipLaunchKernelGGL(k_fill_int, dim3((unsigned)(nb > 4096 ? 4096 : nb)), dim3(FB), 0, s,
                           (int64_t)nlocal, inv, -1);
#define FS_BPACK(NVC) \
        case NVC: \
            hipLaunchKernelGGL(k_bpack_rec<NVC>, grid(neles), dim3(FB), 0, s, neles, nlocal, a0, \
                               cell_ids, inv, rhs, dt, rec, flag); \
            hipLaunchKernelGGL(k_bpack_from_rec<NVC>, dim3((unsigned)((nlocal + 63)/64)), dim3(64), \
                               0, s, nlocal, nfvars, inv, rec, rank_rhs, diag); \
            return ok();
        switch (nvars) {
        FS_BPACK(4) FS_BPACK(5) FS_BPACK(6) FS_BPACK(7) FS_BPACK(8) FS_BPACK(9)
        default: break;
        }
#undef FS_BPACK
    }
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((neles + 63)/64)), dim3(64), 0, s, neles, nlocal,
                       nvars, nfvars, a0, cell_ids, rhs, dt, rank_rhs, diag, flag);
    return ok();
}

int fs_launch_tpack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars, double a0,
                    double factor, int kind, const int *cell_ids, const double *uptsb,
                    const double *dsrc, const double *dt, double *tdiag, int *flag)
{
    if (neles <= 0) return UCFD_OK;
    const int nt = nvars - nfvars;
    if (nt == 1)
        hipLaunchKernelGGL(k_tpack<1>, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, nfvars,
                           a0, factor, kind, cell_ids, uptsb, dsrc, dt, tdiag, flag);
    else if (nt == 2)
        hipLaunchKernelGGL(k_tpack<2>, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, nfvars,
                           a0, factor, kind, cell_ids, uptsb, dsrc, dt, tdiag, flag);
    else
        hipLaunchKernelGGL(k_tpack<0>, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, nfvars,
                           a0, factor, kind, cell_ids, uptsb, dsrc, dt, tdiag, flag);
    return ok();
}

int fs_launch_update(hipStream_t s, int neles, int nlocal, int nvars, const int *cell_ids,
                     const double *du, double *upts, int *flag)
{
    if (neles <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_update, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, cell_ids, du,
                       upts, flag);
    return ok();
}

int fs_launch_subres(hipStream_t s, int neles, int nlocal, int nvars, const int *cell_ids,
                     const double *vol, const double *du, double *dup, double *res, int *flag)
{
    if (neles <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_subres, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, cell_ids, vol,
                       du, dup, res, flag);
    return ok();
}

int fs_launch_prepare(hipStream_t s, int nv, int nlocal, int64_t nfaces, int64_t npos,
                      const int *rowptr, const int *slot_ptr, const int2 *sorted,
                      const int *partner, const int *slots, const signed char *sides,
                      const double *face_area, const double *rcp_vol, double *diag,
                      const double *jmat, double *jp, double *fvp, float *jp32, float *fv32,
                      int *flag)
{
    if (nlocal <= 0) return UCFD_OK;
    return by_nv<PrepL>(nv, s, nlocal, nfaces, npos, rowptr, slot_ptr, sorted, partner, slots,
                        sides, face_area, rcp_vol, diag, jmat, jp, fvp, jp32, fv32, flag);
}

int fs_launch_sweep(hipStream_t s, int nv, bool f32, int base, int cnt, const int *order,
                    int var0, int nvars, const int *rowptr, const int *colidx, const double *jp,
                    const double *fvp, const float *jp32, const float *fv32, const double *rhs,
                    const double *diag, double *du, float *du32, int maxw)
{
    if (cnt <= 0) return UCFD_OK;
    return by_nv<SweepL>(nv, s, f32, base, cnt, order, var0, nvars, rowptr, colidx, jp, fvp, jp32,
                         fv32, rhs, diag, du, du32, maxw);
}

int fs_launch_zero(hipStream_t s, int64_t n, double *a, double *b, float *c)
{
    if (n <= 0) return UCFD_OK;
    const int64_t nb = (n + FB - 1)/FB;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)(nb > 2048 ? 2048 : nb)), dim3(FB), 0, s, n, a, b, c);
    return ok();
}

int fs_launch_lpack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars, double turb_factor,
                    double a0, const int *cell_ids, const double *upts, const double *rhs,
                    const double *dt, const double *dsrc, double *u, double *du, double *diag,
                    int *flag, int *inv, double *rec)
{
    if (neles <= 0) return UCFD_OK;
    if (!inv) return UCFD_ERR_ARG;
    const int64_t nb = ((int64_t)nlocal + FB - 1)/FB;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)(nb > 4096 ? 4096 : nb)), dim3(FB), 0, s,
                       (int64_t)nlocal, inv, -1);
#define FS_PACK_REC(NVC) \
    case NVC: \
        hipLaunchKernelGGL(k_lpack_rec<NVC>, grid(neles), dim3(FB), 0, s, neles, nlocal, nfvars, \
                           turb_factor, a0, cell_ids, inv, upts, rhs, dt, dsrc, rec, flag); \
        hipLaunchKernelGGL(k_lpack_from_rec<NVC>, grid(nlocal), dim3(FB), 0, s, nlocal, inv, rec, \
                           u, du, diag); \
        return ok();
    if (rec) {
        switch (nvars) {
        FS_PACK_REC(4) FS_PACK_REC(5) FS_PACK_REC(6) FS_PACK_REC(7) FS_PACK_REC(8) FS_PACK_REC(9)
        default: break;
        }
    }
#undef FS_PACK_REC
    hipLaunchKernelGGL(k_lpack_inv, grid(neles), dim3(FB), 0, s, neles, nlocal, cell_ids, inv,
                       flag);
    hipLaunchKernelGGL(k_lpack_rank, grid(nlocal), dim3(FB), 0, s, neles, nlocal, nvars, nfvars,
                       turb_factor, a0, inv, upts, rhs, dt, dsrc, u, du, diag);
    return ok();
}

int fs_launch_lupdate(hipStream_t s, int neles, int nlocal, int nvars, const int *cell_ids,
                      const double *du, double *upts, int *flag, double *rec)
{
    if (neles <= 0) return UCFD_OK;
#define FS_UPD_REC(NVC) \
    case NVC: \
        hipLaunchKernelGGL(k_lupdate_rec<NVC>, grid(nlocal), dim3(FB), 0, s, nlocal, du, rec); \
        hipLaunchKernelGGL(k_lupdate_from_rec<NVC>, grid(neles), dim3(FB), 0, s, neles, nlocal, \
                           cell_ids, rec, upts, flag); \
        return ok();
    if (rec) {
        switch (nvars) {
        FS_UPD_REC(4) FS_UPD_REC(5) FS_UPD_REC(6) FS_UPD_REC(7) FS_UPD_REC(8) FS_UPD_REC(9)
        default: break;
        }
    }
#undef FS_UPD_REC
    hipLaunchKernelGGL(k_lupdate, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, cell_ids, du,
                       upts, flag);
    return ok();
}

namespace {

template <int ND, int NT>
int lsweep_launch(hipStream_t s, bool upper, int base, int cnt, const int *order, int nvars,
                  int nfvars, double kappa, const double *lambdaf, int nlocal, int64_t nfaces,
                  const int *rowptr, const int *colidx, const signed char *sides,
                  const int *slots, const double *face_area, const double *face_normal,
                  const double *rcp_vol, const double *diag, const double *u, double *du)
{
    if (upper)
        hipLaunchKernelGGL((k_lsweep<ND, NT, true>), grid(cnt), dim3(FB), 0, s, base, cnt, order,
                           nvars, nfvars, kappa, lambdaf, nlocal, nfaces, rowptr, colidx, sides,
                           slots, face_area, face_normal, rcp_vol, diag, u, du);
    else
        hipLaunchKernelGGL((k_lsweep<ND, NT, false>), grid(cnt), dim3(FB), 0, s, base, cnt, order,
                           nvars, nfvars, kappa, lambdaf, nlocal, nfaces, rowptr, colidx, sides,
                           slots, face_area, face_normal, rcp_vol, diag, u, du);
    return ok();
}

template <int ND>
int lsweep_nt(hipStream_t s, bool upper, int nt, int base, int cnt, const int *order, int nvars,
              int nfvars, double kappa, const double *lambdaf, int nlocal, int64_t nfaces,
              const int *rowptr, const int *colidx, const signed char *sides, const int *slots,
              const double *face_area, const double *face_normal, const double *rcp_vol,
              const double *diag, const double *u, double *du)
{
#define LS(NT) return lsweep_launch<ND, NT>(s, upper, base, cnt, order, nvars, nfvars, kappa, \
                                            lambdaf, nlocal, nfaces, rowptr, colidx, sides, slots, \
                                            face_area, face_normal, rcp_vol, diag, u, du)
    switch (nt) {
    case 0: LS(0);
    case 1: LS(1);
    case 2: LS(2);
    case 3: LS(3);
    case 4: LS(4);
    default: return UCFD_ERR_ARG;
    }
#undef LS
}

}  // namespace

int fs_launch_lsweep(hipStream_t s, bool upper, int ndims, int nt, int base, int cnt,
                     const int *order, int nvars, int nfvars, double kappa, const double *lambdaf,
                     int nlocal, int64_t nfaces, const int *rowptr, const int *colidx,
                     const signed char *sides, const int *slots, const double *face_area,
                     const double *face_normal, const double *rcp_vol, const double *diag,
                     const double *u, double *du)
{
    if (cnt <= 0) return UCFD_OK;
    if (ndims == 3)
        return lsweep_nt<3>(s, upper, nt, base, cnt, order, nvars, nfvars, kappa, lambdaf, nlocal,
                            nfaces, rowptr, colidx, sides, slots, face_area, face_normal, rcp_vol,
                            diag, u, du);
    if (ndims == 2)
        return lsweep_nt<2>(s, upper, nt, base, cnt, order, nvars, nfvars, kappa, lambdaf, nlocal,
                            nfaces, rowptr, colidx, sides, slots, face_area, face_normal, rcp_vol,
                            diag, u, du);
    return UCFD_ERR_ARG;
}

int fs_launch_lprep_pos(hipStream_t s, int ndims, int nlocal, int64_t nfaces, int nv0, int nv1,
                        double kappa, const int *rowptr, const int *slots, const signed char *sides,
                        const double *face_area, const double *face_normal, const double *rcp_vol,
                        const double *lambdaf, double *diag, void *geo, double *lam)
{
    if (nlocal <= 0) return UCFD_OK;
    if (ndims == 3)
        hipLaunchKernelGGL(k_lprep_pos<3>, grid(nlocal), dim3(FB), 0, s, nlocal, nfaces, nv0, nv1,
                           kappa, rowptr, slots, sides, face_area, face_normal, rcp_vol, lambdaf,
                           diag, (double4 *)geo, lam);
    else if (ndims == 2)
        hipLaunchKernelGGL(k_lprep_pos<2>, grid(nlocal), dim3(FB), 0, s, nlocal, nfaces, nv0, nv1,
                           kappa, rowptr, slots, sides, face_area, face_normal, rcp_vol, lambdaf,
                           diag, (double4 *)geo, lam);
    else
        return UCFD_ERR_ARG;
    return ok();
}

namespace {

template <int ND>
int lsweep_p_nt(hipStream_t s, bool upper, int nt, int base, int cnt, const int *order,
                int nfvars, double kappa, int nlocal, const int *rowptr, const int *colidx,
                const double4 *geo, const double *lam, const double *diag, const double *u,
                double *du)
{
#define LP(NT) do { \
    if (upper) hipLaunchKernelGGL((k_lsweep_p<ND, NT, true>), grid(cnt), dim3(FB), 0, s, base, \
                                  cnt, order, nfvars, kappa, nlocal, rowptr, colidx, geo, lam, \
                                  diag, u, du); \
    else hipLaunchKernelGGL((k_lsweep_p<ND, NT, false>), grid(cnt), dim3(FB), 0, s, base, cnt, \
                            order, nfvars, kappa, nlocal, rowptr, colidx, geo, lam, diag, u, du); \
    return ok(); } while (0)
    switch (nt) {
    case 0: LP(0);
    case 1: LP(1);
    case 2: LP(2);
    case 3: LP(3);
    case 4: LP(4);
    default: return UCFD_ERR_ARG;
    }
#undef LP
}

}  // namespace

int fs_launch_lsweep_packed(hipStream_t s, bool upper, int ndims, int nt, int base, int cnt,
                            const int *order, int nfvars, double kappa, int nlocal,
                            const int *rowptr, const int *colidx, const void *geo,
                            const double *lam, const double *diag, const double *u, double *du)
{
    if (cnt <= 0) return UCFD_OK;
    if (ndims == 3)
        return lsweep_p_nt<3>(s, upper, nt, base, cnt, order, nfvars, kappa, nlocal, rowptr,
                              colidx, (const double4 *)geo, lam, diag, u, du);
    if (ndims == 2)
        return lsweep_p_nt<2>(s, upper, nt, base, cnt, order, nfvars, kappa, nlocal, rowptr,
                              colidx, (const double4 *)geo, lam, diag, u, du);
    return UCFD_ERR_ARG;
}

int fs_launch_lsweep_free(hipStream_t s, bool upper, int base, int cnt, const int *order, int nv0,
                          int nv1, int nlocal, const double *diag, double *du)
{
    if (cnt <= 0) return UCFD_OK;
    if (upper)
        hipLaunchKernelGGL(k_lsweep_free<true>, grid(cnt), dim3(FB), 0, s, base, cnt, order, nv0,
                           nv1, nlocal, diag, du);
    else
        hipLaunchKernelGGL(k_lsweep_free<false>, grid(cnt), dim3(FB), 0, s, base, cnt, order, nv0,
                           nv1, nlocal, diag, du);
    return ok();
}
