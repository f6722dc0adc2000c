// Adjoint of one FP64 coloured LU-SGS pass (pre + lower + upper) on a
// colour-contiguous numbering: the transposed sweeps.  The upper sweep's
// adjoint runs colours ascending, the lower sweep's colours descending; a
// cell gathers the adjoint of its already-processed neighbours through the
// face-pair map (the neighbour's face pointing back to it), in face order,
// so there are no atomics and results are deterministic.  The lower-sweep
// launch also finishes the cell's gradients with respect to rhsb, dt (via
// the pre diagonal) and fspr.  Operation order = oracle/adjoint_oracle.c.
//
// rhsbar may alias dstar: the lower-sweep adjoint reads a cell's own dstar
// before writing its rhsbar, and reads dstar only of cells in lower colours,
// which it processes later (colours descending).
#include <hip/hip_runtime.h>

#include "adjoint_internal.h"
#include "flux.h"

namespace {

constexpr int AB = 256;

__device__ __forceinline__ int axcd()
{
    const int b = blockIdx.x, nblk = gridDim.x;
    const int x = b % 8, k = b / 8, per = nblk / 8, rem = nblk % 8;
    return x*per + (x < rem ? x : rem) + k;
}

int ok() { return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP; }

__global__ __launch_bounds__(AB) void k_pairs(int64_t N, int F, const int *__restrict__ nei,
                                              int *__restrict__ pair, int *flag)
{
    const int64_t i = (int64_t)blockIdx.x*AB + threadIdx.x;
    if (i >= N) return;
    for (int f = 0; f < F; ++f) {
        const int64_t nb = nei[f*N + i];
        int pf = -1;
        if ((uint64_t)nb >= (uint64_t)N) {
            atomicOr(flag, UCFD_KFLAG_INDEX);
        } else if (nb != i) {
            for (int g = 0; g < F; ++g)
                if (nei[g*N + nb] == i) { pf = g; break; }
            if (pf < 0) atomicOr(flag, UCFD_KFLAG_INDEX);
        }
        pair[f*N + i] = pf;
    }
}

// acc += (-0.5 fv_j / diag_j) (A(U; n_j)^T x_j - lam_j x_j)
template <int ND>
__device__ __forceinline__ void gather(int64_t N, int64_t j, int f, const double *U,
                                       const double *__restrict__ fnv,
                                       const double *__restrict__ vfn,
                                       const double *__restrict__ fspr,
                                       const double *__restrict__ diag, const double *work,
                                       double *acc)
{
    constexpr int NV = ND + 2;
    double n[ND], x[NV], jt[NV];
    const double fv = fnv[f*N + j], lam = fspr[f*N + j];
#pragma unroll
    for (int d = 0; d < ND; ++d) n[d] = vfn[((int64_t)f*ND + d)*N + j];
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] = work[k*N + j];
    ucfd::ns_flux_jt<double, ND>(U, n, x, jt);
    const double coef = -0.5*fv/diag[j];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += coef*(jt[k] - lam*x[k]);
}

template <int ND, int FC>
__global__ __launch_bounds__(AB) void k_adj_upper(int n0, int ne, int64_t N,
                                                  const int *__restrict__ nei,
                                                  const int *__restrict__ pair,
                                                  const double *__restrict__ fnv,
                                                  const double *__restrict__ vfn,
                                                  const double *__restrict__ fspr,
                                                  const double *__restrict__ diag,
                                                  const double *__restrict__ upts,
                                                  const double *__restrict__ dq,
                                                  const double *__restrict__ w, double *work)
{
    constexpr int NV = ND + 2;
    const int64_t i = n0 + (int64_t)axcd()*AB + threadIdx.x;
    if (i >= ne) return;
    int nb[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) nb[f] = nei[f*N + i];
    double acc[NV], U[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        acc[k] = w[k*N + i];
        U[k] = upts[k*N + i] + dq[k*N + i];
    }
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const int64_t j = nb[f];
        if (j == i || j >= n0) continue;
        gather<ND>(N, j, pair[f*N + i], U, fnv, vfn, fspr, diag, work, acc);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) work[k*N + i] = acc[k];
}

template <int ND, int FC>
__global__ __launch_bounds__(AB) void k_adj_lower(int n0, int ne, int64_t N,
                                                  const int *__restrict__ nei,
                                                  const int *__restrict__ pair,
                                                  const double *__restrict__ fnv,
                                                  const double *__restrict__ vfn,
                                                  const double *__restrict__ fspr,
                                                  const double *__restrict__ dt, double factor,
                                                  const double *__restrict__ diag,
                                                  const double *__restrict__ upts,
                                                  const double *dstar,
                                                  const double *__restrict__ dq, double *work,
                                                  double *rhsbar,
                                                  double *__restrict__ dtbar,
                                                  double *__restrict__ fsprbar)
{
    constexpr int NV = ND + 2;
    const int64_t i = n0 + (int64_t)axcd()*AB + threadIdx.x;
    if (i >= ne) return;
    int nb[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) nb[f] = nei[f*N + i];
    double a[NV], acc[NV], U[NV], ds[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        a[k] = work[k*N + i];
        acc[k] = a[k];
        ds[k] = dstar[k*N + i];
        U[k] = upts[k*N + i] + ds[k];
    }
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const int64_t j = nb[f];
        if (j == i || j < ne) continue;
        gather<ND>(N, j, pair[f*N + i], U, fnv, vfn, fspr, diag, work, acc);
    }
    const double di = diag[i];
    double dgb = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        work[k*N + i] = acc[k];
        rhsbar[k*N + i] = acc[k]/di;
        dgb += a[k]*((ds[k] - dq[k*N + i])/di) - acc[k]*(ds[k]/di);
    }
    dtbar[i] = -dgb/(dt[i]*dt[i]*factor);
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const double fv = fnv[f*N + i];
        const int64_t j = nb[f];
        double val = dgb*0.5*fv;
        if (j != i) {
            double s = 0.0;
            if (j >= ne) {
#pragma unroll
                for (int k = 0; k < NV; ++k) s += a[k]*dq[k*N + j];
            } else {
#pragma unroll
                for (int k = 0; k < NV; ++k) s += acc[k]*dstar[k*N + j];
            }
            val += 0.5*fv*s/di;
        }
        fsprbar[f*N + i] = val;
    }
}

template <int ND, int FC>
int launch(hipStream_t s, int lower, int n0, int ne, int64_t N, const int *nei, const int *pair,
           const double *fnv, const double *vfn, const double *fspr, const double *dt,
           double factor, const double *diag, const double *upts, const double *dstar,
           const double *dq, const double *w, double *work, double *rhsbar, double *dtbar,
           double *fsprbar)
{
    const dim3 g((unsigned)((ne - n0 + AB - 1)/AB));
    if (lower)
        hipLaunchKernelGGL((k_adj_lower<ND, FC>), g, dim3(AB), 0, s, n0, ne, N, nei, pair, fnv, vfn,
                           fspr, dt, factor, diag, upts, dstar, dq, work, rhsbar, dtbar, fsprbar);
    else
        hipLaunchKernelGGL((k_adj_upper<ND, FC>), g, dim3(AB), 0, s, n0, ne, N, nei, pair, fnv, vfn,
                           fspr, diag, upts, dq, w, work);
    return ok();
}

}  // namespace

int adj_launch_pairs(hipStream_t s, int64_t N, int F, const int *nei, int *pair, int *flag)
{
    if (N <= 0) return UCFD_OK;
    hipLaunchKernelGGL(k_pairs, dim3((unsigned)((N + AB - 1)/AB)), dim3(AB), 0, s, N, F, nei, pair,
                       flag);
    return ok();
}

int adj_launch_colour(hipStream_t s, int lower, int n0, int ne, int64_t N, int nface, int ndims,
                      const int *nei, const int *pair, const double *fnv, const double *vfn,
                      const double *fspr, const double *dt, double factor, const double *diag,
                      const double *upts, const double *dstar, const double *dq, const double *w,
                      double *work, double *rhsbar, double *dtbar, double *fsprbar)
{
    if (ne <= n0) return UCFD_OK;
#define ADJ_CASE(D, FC)                                                                          \
    if (ndims == D && nface == FC)                                                               \
        return launch<D, FC>(s, lower, n0, ne, N, nei, pair, fnv, vfn, fspr, dt, factor, diag,  \
                             upts, dstar, dq, w, work, rhsbar, dtbar, fsprbar);
    ADJ_CASE(3, 4)
    ADJ_CASE(3, 6)
    ADJ_CASE(3, 5)
    ADJ_CASE(2, 3)
    ADJ_CASE(2, 4)
#undef ADJ_CASE
    return UCFD_ERR_ARG;
}
