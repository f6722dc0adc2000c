// Kernels of the FlowSys block LU-SGS path (src/lusgs/blusgs.c): element
// pack / update / sub-residual, the block-diagonal assembly and LU with the
// per-position packing of the face Jacobians, and the block Gauss-Seidel
// sweep over all face neighbours (FP64, or FP32 block products with the
// FP64 diagonal solve).
//
// FP64 kernels keep the reference's operation order and are compiled
// without FMA contraction, so they agree bit for bit with the CPU oracle.
#include <hip/hip_runtime.h>

#include "block_lu.h"
#include "flowsys_internal.h"

namespace {

constexpr int FB = 256;

__device__ __forceinline__ int fxcd()
{
    const int b = blockIdx.x, nblk = gridDim.x;
    const int x = b % 8, k = b / 8, per = nblk / 8, rem = nblk % 8;
    return x*per + (x < rem ? x : rem) + k;
}

inline dim3 grid(int64_t n) { return dim3((unsigned)((n + FB - 1)/FB)); }

int ok() { return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP; }

// rank_blusgs_pack (blusgs.c:17-45)
__global__ __launch_bounds__(FB) void k_pack(int neles, int nlocal, int nvars, int nfvars, double a0,
                                             const int *__restrict__ cell_ids,
                                             const double *__restrict__ rhs,
                                             const double *__restrict__ dt, double *rank_rhs,
                                             double *diag, int *flag)
{
    const int e = blockIdx.x*FB + threadIdx.x;
    if (e >= neles) return;
    const int r = cell_ids[e];
    if ((unsigned)r >= (unsigned)nlocal) {
        atomicOr(flag, UCFD_KFLAG_INDEX);
        return;
    }
    for (int k = 0; k < nvars; ++k) rank_rhs[k + (int64_t)r*nvars] = rhs[e + (int64_t)k*neles];
    const int d2 = nfvars*nfvars;
    for (int row = 0; row < nfvars; ++row) {
        double *drow = diag + (int64_t)r*d2 + row*nfvars;
        for (int col = 0; col < nfvars; ++col) drow[col] = 0.0;
        drow[row] = 1/dt[e] + a0;
    }
}

// rank_tblusgs_pack (blusgs.c:60-94) with sa_src_jacobian / kwsst_src_jacobian
// (source_jacobian.c:6-18); BETAST = 0.09 (flux.h:8)
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
                                               const double *__restrict__ du, double *upts,
                                               int *flag)
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
                                               const double *__restrict__ du, double *dup,
                                               double *res, int *flag)
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

// pre_blusgs (blusgs.c:200-246) plus the packing of every position's
// off-diagonal block (the sweep's jval, blusgs.c:297-300) and fv
template <int NV>
__global__ __launch_bounds__(FB) void k_prepare(int nlocal, int64_t nfaces,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ slots,
                                                const signed char *__restrict__ sides,
                                                const double *__restrict__ face_area,
                                                const double *__restrict__ rcp_vol,
                                                double *diag, const double *__restrict__ jmat,
                                                double *jp, double *fvp, float *jp32, float *fv32,
                                                int *flag)
{
    constexpr int D2 = NV*NV;
    const int r = blockIdx.x*FB + threadIdx.x;
    if (r >= nlocal) return;
    double dm[D2];
#pragma unroll
    for (int q = 0; q < D2; ++q) dm[q] = diag[(int64_t)r*D2 + q];
    const int64_t half = (int64_t)D2*nfaces;
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
        const int64_t s = slots[p];
        const bool plus = sides[p] == 1;
        const double fv = face_area[s]*rcp_vol[r];
        fvp[p] = fv;
        if (fv32) fv32[p] = (float)fv;
#pragma unroll
        for (int row = 0; row < NV; ++row)
#pragma unroll
            for (int col = 0; col < NV; ++col) {
                const int64_t o = s + (int64_t)col*nfaces + (int64_t)row*NV*nfaces;
                const double a = jmat[o], b = jmat[o + half];
                const double v = plus ? a : -b;       // diagonal contribution
                const double j = plus ? b : -a;       // neighbour coupling
                dm[col + row*NV] += v*fv;
                jp[(int64_t)p*D2 + row*NV + col] = j;
                if (jp32) jp32[(int64_t)p*D2 + row*NV + col] = (float)j;
            }
    }
    if (!ucfd::ludcmp<NV>(dm)) atomicOr(flag, UCFD_KFLAG_PIVOT);
#pragma unroll
    for (int q = 0; q < D2; ++q) diag[(int64_t)r*D2 + q] = dm[q];
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
    static int run(hipStream_t s, int nlocal, int64_t nfaces, const int *rowptr, const int *slots,
                   const signed char *sides, const double *face_area, const double *rcp_vol,
                   double *diag, const double *jmat, double *jp, double *fvp, float *jp32,
                   float *fv32, int *flag)
    {
        hipLaunchKernelGGL(k_prepare<NV>, grid(nlocal), dim3(FB), 0, s, nlocal, nfaces, rowptr,
                           slots, sides, face_area, rcp_vol, diag, jmat, jp, fvp, jp32, fv32,
                           flag);
        return ok();
    }
};

template <int NV> struct SweepL {
    static int run(hipStream_t s, bool f32, int base, int cnt, const int *order, int var0,
                   int nvars, const int *rowptr, const int *colidx, const double *jp,
                   const double *fvp, const float *jp32, const float *fv32, const double *rhs,
                   const double *diag, double *du, float *du32)
    {
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
                   double *diag, int *flag)
{
    if (neles <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_pack, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, nfvars, a0,
                       cell_ids, rhs, dt, rank_rhs, diag, flag);
    return ok();
}

int fs_launch_tpack(hipStream_t s, int neles, int nlocal, int nvars, int nfvars, double a0,
                    double factor, int kind, const int *cell_ids, const double *uptsb,
                    const double *dsrc, const double *dt, double *tdiag, int *flag)
{
    if (neles <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_tpack, grid(neles), dim3(FB), 0, s, neles, nlocal, nvars, nfvars, a0,
                       factor, kind, cell_ids, uptsb, dsrc, dt, tdiag, flag);
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

int fs_launch_prepare(hipStream_t s, int nv, int nlocal, int64_t nfaces, const int *rowptr,
                      const int *slots, const signed char *sides, const double *face_area,
                      const double *rcp_vol, double *diag, const double *jmat, double *jp,
                      double *fvp, float *jp32, float *fv32, int *flag)
{
    if (nlocal <= 0) return UCFD_OK;
    return by_nv<PrepL>(nv, s, nlocal, nfaces, rowptr, slots, sides, face_area, rcp_vol, diag,
                        jmat, jp, fvp, jp32, fv32, flag);
}

int fs_launch_sweep(hipStream_t s, int nv, bool f32, int base, int cnt, const int *order,
                    int var0, int nvars, const int *rowptr, const int *colidx, const double *jp,
                    const double *fvp, const float *jp32, const float *fv32, const double *rhs,
                    const double *diag, double *du, float *du32)
{
    if (cnt <= 0) return UCFD_OK;
    return by_nv<SweepL>(nv, s, f32, base, cnt, order, var0, nvars, rowptr, colidx, jp, fvp, jp32,
                         fv32, rhs, diag, du, du32);
}

int fs_launch_zero(hipStream_t s, int64_t n, double *a, double *b, float *c)
{
    if (n <= 0) return UCFD_OK;
    const int64_t nb = (n + FB - 1)/FB;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)(nb > 2048 ? 2048 : nb)), dim3(FB), 0, s, n, a, b, c);
    return ok();
}
