// Kernels of the Krylov path: BSR / CSR SpMV, block LU-SGS (prepare, lower
// and upper sweeps in FP64 and in the mixed FP32/FP64 form), BLAS-1/2 with
// deterministic reductions, and the fused BiCGSTAB vector updates.
//
// FP64 kernels keep the operation order of the reference
// (src/krylov/matrix/sparsemat.c:57-101, src/krylov/utils/inverse.c:7-71,
// src/krylov/precon/blusgs.c:28-100) and are compiled without FMA
// contraction, so SpMV and the FP64 preconditioner apply agree bit for bit
// with the CPU oracle.  Reductions sum block partials in a fixed order, so
// every dot product is reproducible run to run.
#include <hip/hip_runtime.h>

#include "block_lu.h"
#include "krylov_internal.h"

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

__global__ void k_to32(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t j = (int64_t)blockIdx.x*KB + threadIdx.x; j < n; j += st) b[j] = (float)a[j];
}

// ----------------------------------------------------------- block sweeps
template <int B, bool UPPER>
__global__ __launch_bounds__(KB) void k_sweep64(int base, int cnt, const int *__restrict__ order,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ colidx,
                                                const int *__restrict__ diagslots,
                                                const double *__restrict__ values,
                                                const double *__restrict__ dv, double *b)
{
    const int t = kxcd()*KB + threadIdx.x;
    if (t >= cnt) return;
    const int i = order ? order[base + t] : base + t;
    const int dd = diagslots[i];
    double arr[B];
    if (!UPPER) {
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
        const double *blk = values + (int64_t)j*B*B;
        double xc[B];
#pragma unroll
        for (int q = 0; q < B; ++q) xc[q] = b[(int64_t)c*B + q];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < B; ++q) v += __builtin_nontemporal_load(blk + r*B + q)*xc[q];
            if (UPPER) arr[r] += v;
            else arr[r] -= v;
        }
    }
    double lu[B*B];
#pragma unroll
    for (int q = 0; q < B*B; ++q) lu[q] = __builtin_nontemporal_load(dv + (int64_t)i*B*B + q);
    ucfd::lusub<B>(lu, arr);
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int64_t o = (int64_t)i*B + k;
        if (UPPER) b[o] -= arr[k];
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

__global__ __launch_bounds__(KB) void k_scale_isqrt(int64_t n, const double *__restrict__ src,
                                                    const double *__restrict__ nsq,
                                                    double *__restrict__ dst)
{
    const double inv = 1.0/sqrt(*nsq);
    const int64_t st = (int64_t)gridDim.x*KB;
    for (int64_t i = (int64_t)blockIdx.x*KB + threadIdx.x; i < n; i += st) dst[i] = inv*src[i];
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
__global__ __launch_bounds__(KB) void k_bicg_s(int64_t n, double alpha, const double *__restrict__ r,
                                               const double *__restrict__ v,
                                               double *__restrict__ sv, double *__restrict__ shat)
{
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
        hipLaunchKernelGGL(k_spmv_bsr<B>, dim3((bn + KB - 1)/KB), dim3(KB), 0, s, bn, rp, ci, v, al,
                           x, be, y, flag);
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

template <int B> struct Sw64L {
    static int run(hipStream_t s, int upper, int base, int cnt, const int *ord, const int *rp,
                   const int *ci, const int *ds, const double *v, const double *dv, double *b)
    {
        const dim3 g((cnt + KB - 1)/KB);
        if (upper)
            hipLaunchKernelGGL((k_sweep64<B, true>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci, ds, v,
                               dv, b);
        else
            hipLaunchKernelGGL((k_sweep64<B, false>), g, dim3(KB), 0, s, base, cnt, ord, rp, ci, ds,
                               v, dv, b);
        return ok();
    }
};

template <int B> struct Sw32L {
    static int run(hipStream_t s, int upper, int base, int cnt, const int *ord, const int *rp,
                   const int *ci, const int *ds, const float *v, const double *dv, float *x32,
                   double *b)
    {
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

int kry_bsgs_sweep64(hipStream_t s, int upper, int blk, int base, int cnt, const int *order,
                     const int *rowptr, const int *colidx, const int *diagslots,
                     const double *values, const double *diagvalues, double *b)
{
    if (cnt <= 0) return UCFD_OK;
    return by_block<Sw64L>(blk, s, upper, base, cnt, order, rowptr, colidx, diagslots, values,
                           diagvalues, b);
}

int kry_bsgs_sweep32(hipStream_t s, int upper, int blk, int base, int cnt, const int *order,
                     const int *rowptr, const int *colidx, const int *diagslots,
                     const float *vals32, const double *diagvalues, float *x32, double *b)
{
    if (cnt <= 0) return UCFD_OK;
    return by_block<Sw32L>(blk, s, upper, base, cnt, order, rowptr, colidx, diagslots, vals32,
                           diagvalues, x32, b);
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
    for (int j0 = 0; j0 < k; j0 += 8)
        hipLaunchKernelGGL(k_gemvt8, dim3(nb), dim3(KB), 0, s, n, k, j0, V, ldv, w, part);
    hipLaunchKernelGGL(k_sum_rows, dim3(k), dim3(1024), 0, s, part, nb, KMAXPART, out);
    return ok();
}

int kry_gemv(hipStream_t s, int64_t n, int k, const double *V, int64_t ldv, double alpha,
             const double *h, double *w, double *part, double *normsq)
{
    if (k <= 0 || k > 64) return UCFD_ERR_ARG;
    const int nb = nblocks(n);
    hipLaunchKernelGGL(k_gemv, dim3(nb), dim3(KB), 0, s, n, k, V, ldv, alpha, h, w,
                       normsq ? part : nullptr);
    if (normsq) hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(1024), 0, s, part, nb, KMAXPART, normsq);
    return ok();
}

int kry_scale_by_inv_sqrt(hipStream_t s, int64_t n, const double *src, const double *normsq,
                          double *dst)
{
    hipLaunchKernelGGL(k_scale_isqrt, dim3(nblocks(n)), dim3(KB), 0, s, n, src, normsq, dst);
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

int kry_bicg_s(hipStream_t s, int64_t n, double alpha, const double *r, const double *v,
               double *sv, double *shat)
{
    hipLaunchKernelGGL(k_bicg_s, dim3(nblocks(n)), dim3(KB), 0, s, n, alpha, r, v, sv, shat);
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
