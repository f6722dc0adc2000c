// Defect-correction LU-SGS solver on a colour-contiguous mesh: the FP64 form
// and the mixed FP32/FP64 form (extension; the reference has no solver
// around its sweeps - the operator is the one LU-SGS factors approximately,
// doc/3.1_lusgs.md:79-100).
//
// Operator.  For an increment e on the base state Q:
//   A(e)_i = D_i e_i + 1/2 sum_{f, nb != i} [ dF_S(Q_nb; e_nb) - L_f e_nb ]
// with D_i = 1/dt_i + 1/2 sum_f fspr_f fnorm_vol_f (the pre diagonal),
// S_f = fnorm_vol_f * vec_fnorm_f and L_f = fnorm_vol_f * fspr_f.  The Euler
// flux is linear in the normal, so fnorm_vol*[F(U+d; n) - F(U; n)] =
// F(U+d; S) - F(U; S) =: dF_S(U; d).  dF_S is evaluated in the exact
// difference form below (no subtraction of two O(1) fluxes), which keeps
// ~7 significant digits of dF in FP32 when |d| << |U|; where the pressure
// floor PMIN is active on either state the two fluxes are evaluated and
// subtracted as in src/lusgs/flux.c:29-67.
//
// Both forms solve A(dq) = b by defect correction with one LU-SGS pass as
// the preconditioner:
//   e <- e + M^{-1}(b - A(e)),  M^{-1} = upper sweep o lower sweep at base Q.
// The residual of iteration j is computed inside the lower-sweep launch of
// each colour and e += delta inside the upper-sweep launch.
// Mixed form (iterative refinement): the outer loop evaluates r = b - A(dq)
// in FP64, rounds r and the base state Q + dq to FP32, runs the same inner
// defect correction in FP32 on the correction equation A_dq(e) = r until its
// residual has dropped by the inner tolerance, and adds e to dq in FP64.
#include <algorithm>
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "flux.h"
#include "ucfd_ext.h"
#include "ucfd_internal.h"

namespace {

constexpr int kB = 256;
// sweep and residual kernels: 256-thread workgroups (FP64 pass 0.595 -> 0.574 ms,
// FP32 0.297 -> 0.281 ms against 128 threads at 5.3M cells)
constexpr int kSB = 256;
constexpr int NV = 5;

template <typename T> struct alignas(4*sizeof(T)) Face4 {
    T s0, s1, s2, l;    // S = fnorm_vol * n, L = fnorm_vol * fspr
};

template <typename T> __device__ __forceinline__ T ldg_nt(const T *p)
{
    return __builtin_nontemporal_load(p);
}

// face records are read as 16-byte vectors
__device__ __forceinline__ Face4<float> ldg_nt(const Face4<float> *p)
{
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load((const v4 *)p);
    return Face4<float>{v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ Face4<double> ldg_nt(const Face4<double> *p)
{
    typedef double v2 __attribute__((ext_vector_type(2)));
    const v2 a = __builtin_nontemporal_load((const v2 *)p);
    const v2 b = __builtin_nontemporal_load((const v2 *)p + 1);
    return Face4<double>{a.x, a.y, b.x, b.y};
}

// Solver-owned 5-vectors (the base state, the increment and the sweep
// intermediates) use a split layout, one buffer of 5N values: components 0-3
// of cell i at p[4i .. 4i+3], component 4 at p[4N + i].  A gathered
// neighbour state is then one 16-byte load (two in FP64) and one scalar load
// instead of five scalar loads from five rows; the bytes are the same.
// Arrays the caller passes (right-hand side, dq of the public residual) and
// cell-own arrays keep the [var][cell] layout.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <bool NT> __device__ __forceinline__ void ld_split(const float *p, int N, int i, float *v)
{
    const f4v *a = (const f4v *)p + i;
    const f4v x = NT ? __builtin_nontemporal_load(a) : *a;
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    v[4] = NT ? __builtin_nontemporal_load(p + 4*(size_t)N + i) : p[4*(size_t)N + i];
}

template <bool NT> __device__ __forceinline__ void ld_split(const double *p, int N, int i, double *v)
{
    const d2v *a = (const d2v *)p + 2*(size_t)i;
    const d2v x = NT ? __builtin_nontemporal_load(a) : a[0];
    const d2v y = NT ? __builtin_nontemporal_load(a + 1) : a[1];
    v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
    v[4] = NT ? __builtin_nontemporal_load(p + 4*(size_t)N + i) : p[4*(size_t)N + i];
}

template <bool NT> __device__ __forceinline__ void st_split(float *p, int N, int i, const float *v)
{
    const f4v x = {v[0], v[1], v[2], v[3]};
    if (NT) {
        __builtin_nontemporal_store(x, (f4v *)p + i);
        __builtin_nontemporal_store(v[4], p + 4*(size_t)N + i);
    } else {
        ((f4v *)p)[i] = x;
        p[4*(size_t)N + i] = v[4];
    }
}

template <bool NT> __device__ __forceinline__ void st_split(double *p, int N, int i, const double *v)
{
    const d2v x = {v[0], v[1]}, y = {v[2], v[3]};
    d2v *a = (d2v *)p + 2*(size_t)i;
    if (NT) {
        __builtin_nontemporal_store(x, a);
        __builtin_nontemporal_store(y, a + 1);
        __builtin_nontemporal_store(v[4], p + 4*(size_t)N + i);
    } else {
        a[0] = x;
        a[1] = y;
        p[4*(size_t)N + i] = v[4];
    }
}

// a 5-vector of cell i from a split (SPLIT) or [var][cell] array
template <bool SPLIT, typename T>
__device__ __forceinline__ void ld_vec(const T *p, int N, int i, T *v)
{
    if (SPLIT) {
        ld_split<false>(p, N, i, v);
    } else {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = p[(size_t)q*N + i];
    }
}

// Workgroup ids are dealt round-robin over the 8 XCDs (each with its own
// L2); each XCD takes runs of 32 consecutive workgroups (4096 cells), the
// XCDs neighbouring runs, so adjacent cells share one L2 and the XCDs work
// on nearby rows (measured 3 % faster than one contiguous eighth per XCD).
// The last incomplete group of runs keeps the identity: a permutation.
__device__ __forceinline__ int xcd_remap()
{
    constexpr int RUN = 32;
    const int b = blockIdx.x, nblk = gridDim.x;
    const int full = nblk/(8*RUN)*(8*RUN);
    if (b >= full) return b;
    const int x = b % 8, k = b / 8;
    return (k/RUN)*(8*RUN) + x*RUN + k%RUN;
}

// out = dF_S(U; d) - L d   (5 components, 3-D Euler)
template <typename T>
__device__ __forceinline__ void face_term(const T *U, const T *d, const Face4<T> &g, T *out)
{
    using C = ucfd::Const<T>;
    const T gm1 = C::gamma - T(1);
    const T rho = U[0], rhop = U[0] + d[0];
    const T mS = U[1]*g.s0 + U[2]*g.s1 + U[3]*g.s2;
    const T dmS = d[1]*g.s0 + d[2]*g.s1 + d[3]*g.s2;
    const T th = mS/rho;
    const T dth = (dmS - th*d[0])/rhop;
    const T thp = th + dth;
    const T mm = U[1]*U[1] + U[2]*U[2] + U[3]*U[3];
    const T mdm = U[1]*d[1] + U[2]*d[2] + U[3]*d[3];
    const T dd = d[1]*d[1] + d[2]*d[2] + d[3]*d[3];
    const T k = mm/rho;
    const T p = gm1*(U[4] - T(0.5)*k);
    const T dp = gm1*(d[4] - T(0.5)*(T(2)*mdm + dd - k*d[0])/rhop);
    if (p < C::pmin || p + dp < C::pmin) {
        // floor active: subtract the two clamped fluxes (flux.c:50-56)
        const T nf[3] = {g.s0, g.s1, g.s2};
        T up[NV], f0[NV], f1[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) up[q] = U[q] + d[q];
        ucfd::ns_flux<T, 3>(U, nf, f0);
        ucfd::ns_flux<T, 3>(up, nf, f1);
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = (f1[q] - f0[q]) - g.l*d[q];
        return;
    }
    out[0] = dmS - g.l*d[0];
    out[1] = d[1]*thp + U[1]*dth + g.s0*dp - g.l*d[1];
    out[2] = d[2]*thp + U[2]*dth + g.s1*dp - g.l*d[2];
    out[3] = d[3]*thp + U[3]*dth + g.s2*dp - g.l*d[3];
    out[4] = (d[4] + dp)*thp + (U[4] + p)*dth - g.l*d[4];
}

// block sum of v (FP64) written to part[slot]; deterministic order
template <int BS>
__device__ __forceinline__ void block_sum(double v, double *part, int slot)
{
    __shared__ double red[BS/64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x/64, l = threadIdx.x%64;
    if (l == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < BS/64; ++q) s += red[q];
        part[slot] = s;
    }
}

// Every kernel below issues all loads a thread needs (its own rows, the face
// records and the neighbour states of every face it uses) before the first
// arithmetic, so the gathers of all faces are in flight together.

// ---- lower sweep of colour [n0, n0+cnt), with the defect-correction residual
//   s_i = rhs_i - A(e)_i        (WITH_RES; otherwise s_i = rhs_i, e = 0)
//   dstar_i = (s_i - 1/2 sum_{nb < n0} term(Qb_nb; dstar_nb)) / D_i
// Writes sum_i |s_i|^2 of this launch's blocks to part[pofs + block].
// LAST: [n0, n0+cnt) is the last colour, which the upper sweep visits first
// and where it takes no face (every neighbour lies in an earlier colour):
// the launch also does that colour's upper step, delta = dstar - 1/2*0/D
// and e += delta (the general upper step's expressions with no face term),
// and stores delta and e instead of dstar, which nothing else reads.
template <typename T, int FC, bool WITH_RES, bool LAST = false>
__global__ __launch_bounds__(kSB) void k_lower_res(
    int n0, int cnt, int N, const int *__restrict__ nei, const Face4<T> *__restrict__ geo,
    const T *__restrict__ diag, const T *__restrict__ qb, const T *__restrict__ rhs,
    T *e, T *dstar, double *part, int pofs, T *delta = nullptr)
{
    const int b = xcd_remap();
    const int t = b*kSB + threadIdx.x;
    double ss = 0.0;
    if (t < cnt) {
        const int i = n0 + t;
        int nbv[FC];
#pragma unroll
        for (int f = 0; f < FC; ++f) nbv[f] = ldg_nt(nei + (size_t)f*N + i);
        const T dg = ldg_nt(diag + i);
        T s[NV], ei[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) s[q] = ldg_nt(rhs + (size_t)q*N + i);
        if (WITH_RES) ld_split<true>(e, N, i, ei);
        bool res_f[FC], low_f[FC];
        Face4<T> g[FC];
        T U[FC][NV], dr[FC][NV], dl[FC][NV];
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const int nb = nbv[f];
            res_f[f] = WITH_RES && nb != i;
            low_f[f] = nb < n0;
            if (!(res_f[f] || low_f[f])) continue;
            g[f] = ldg_nt(geo + (size_t)f*N + i);
            ld_split<false>(qb, N, nb, U[f]);
            if (res_f[f]) ld_split<false>(e, N, nb, dr[f]);
            if (low_f[f]) ld_split<false>(dstar, N, nb, dl[f]);
        }
        if (WITH_RES) {
            T res[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) res[q] = T(0);
#pragma unroll
            for (int f = 0; f < FC; ++f) {
                if (!res_f[f]) continue;
                T o[NV];
                face_term<T>(U[f], dr[f], g[f], o);
#pragma unroll
                for (int q = 0; q < NV; ++q) res[q] += o[q];
            }
#pragma unroll
            for (int q = 0; q < NV; ++q) s[q] = s[q] - (dg*ei[q] + T(0.5)*res[q]);
        }
#pragma unroll
        for (int q = 0; q < NV; ++q) ss += (double)s[q]*(double)s[q];
        T acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] = T(0);
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            if (!low_f[f]) continue;
            T o[NV];
            face_term<T>(U[f], dl[f], g[f], o);
#pragma unroll
            for (int q = 0; q < NV; ++q) acc[q] += o[q];
        }
        T o[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) o[q] = (s[q] - T(0.5)*acc[q])/dg;
        if (LAST) {
            T ev[NV], dlt[NV], en[NV];
            if (WITH_RES) {
#pragma unroll
                for (int q = 0; q < NV; ++q) ev[q] = ei[q];
            } else {
                ld_split<true>(e, N, i, ev);
            }
            const T acc0 = T(0);
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                dlt[q] = o[q] - T(0.5)*acc0/dg;
                en[q] = ev[q] + dlt[q];
            }
            // temporal stores: the next launch gathers them
            st_split<false>(delta, N, i, dlt);
            st_split<false>(e, N, i, en);
        } else {
            st_split<true>(dstar, N, i, o);
        }
    }
    block_sum<kSB>(ss, part, pofs + b);
}

// ---- upper sweep of colour [n0, n0+cnt), then e_i += delta_i
//   delta_i = dstar_i - 1/2 sum_{nb >= ne} term(Qb_nb; delta_nb) / D_i
template <typename T, int FC>
__global__ __launch_bounds__(kSB) void k_upper_upd(
    int n0, int cnt, int N, const int *__restrict__ nei, const Face4<T> *__restrict__ geo,
    const T *__restrict__ diag, const T *__restrict__ qb, const T *__restrict__ dstar,
    T *delta, T *e)
{
    const int t = xcd_remap()*kSB + threadIdx.x;
    if (t >= cnt) return;
    const int i = n0 + t, ne = n0 + cnt;
    int nbv[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) nbv[f] = ldg_nt(nei + (size_t)f*N + i);
    const T dg = ldg_nt(diag + i);
    T ds[NV], ei[NV];
    ld_split<true>(dstar, N, i, ds);
    ld_split<true>(e, N, i, ei);
    bool up_f[FC];
    Face4<T> g[FC];
    T U[FC][NV], d[FC][NV];
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const int nb = nbv[f];
        up_f[f] = nb >= ne;
        if (!up_f[f]) continue;
        g[f] = ldg_nt(geo + (size_t)f*N + i);
        ld_split<false>(qb, N, nb, U[f]);
        ld_split<false>(delta, N, nb, d[f]);
    }
    T acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = T(0);
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        if (!up_f[f]) continue;
        T o[NV];
        face_term<T>(U[f], d[f], g[f], o);
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] += o[q];
    }
    T dl[NV], en[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        dl[q] = ds[q] - T(0.5)*acc[q]/dg;
        en[q] = ei[q] + dl[q];
    }
    st_split<true>(delta, N, i, dl);
    st_split<true>(e, N, i, en);
}

// ---- the first colour's upper sweep and update (k_upper_upd on [0, cnt)),
// followed by the defect-correction residual of the next pass for the same
// cells, s_i = rhs_i - A(e)_i with the updated e.  The first colour is the
// last one the upper sweep visits, so every neighbour's e is final here,
// and all its neighbours lie in later colours: the faces the upper sweep
// takes are exactly the faces the residual takes (nb != i), and the
// neighbour states qb_nb it gathers are the residual's.  Same expressions
// and order as k_lower_res's residual, so s and the block partials
// (part[pofs + block], same blocks) are bit for bit those the next pass's
// lower launch would compute.  No face is lower than the first colour, so
// that launch's step for these cells is dstar = (s - 1/2*0) / D
// (k_lower_res's expression with an empty face sum): it is done here, over
// this thread's own dstar (read above as the current pass's), and the next
// lower pass starts at the second colour.
template <typename T, int FC>
__global__ __launch_bounds__(kSB) void k_upper_upd_res(
    int cnt, int N, const int *__restrict__ nei, const Face4<T> *__restrict__ geo,
    const T *__restrict__ diag, const T *__restrict__ qb, T *dstar, T *delta, T *e,
    const T *__restrict__ rhs, double *part, int pofs)
{
    const int b = xcd_remap();
    const int t = b*kSB + threadIdx.x;
    double ss = 0.0;
    if (t < cnt) {
        const int i = t, ne = cnt;
        int nbv[FC];
#pragma unroll
        for (int f = 0; f < FC; ++f) nbv[f] = ldg_nt(nei + (size_t)f*N + i);
        const T dg = ldg_nt(diag + i);
        T ds[NV], ei[NV], rh[NV];
        ld_split<true>(dstar, N, i, ds);
        ld_split<true>(e, N, i, ei);
#pragma unroll
        for (int q = 0; q < NV; ++q) rh[q] = ldg_nt(rhs + (size_t)q*N + i);
        bool up_f[FC];
        Face4<T> g[FC];
        T U[FC][NV], d[FC][NV], en[FC][NV];
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const int nb = nbv[f];
            up_f[f] = nb >= ne;
            if (!up_f[f]) continue;
            g[f] = ldg_nt(geo + (size_t)f*N + i);
            ld_split<false>(qb, N, nb, U[f]);
            ld_split<false>(delta, N, nb, d[f]);
            ld_split<false>(e, N, nb, en[f]);
        }
        T acc[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] = T(0);
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            if (!up_f[f]) continue;
            T o[NV];
            face_term<T>(U[f], d[f], g[f], o);
#pragma unroll
            for (int q = 0; q < NV; ++q) acc[q] += o[q];
        }
        T enew[NV], dl[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            dl[q] = ds[q] - T(0.5)*acc[q]/dg;
            enew[q] = ei[q] + dl[q];
        }
        // delta of the first colour is not stored: only the upper launches
        // of earlier colours gather a colour's delta, and there are none
        st_split<true>(e, N, i, enew);
        T res[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) res[q] = T(0);
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            if (!up_f[f]) continue;
            T o[NV];
            face_term<T>(U[f], en[f], g[f], o);
#pragma unroll
            for (int q = 0; q < NV; ++q) res[q] += o[q];
        }
        const T acc0 = T(0);
        T o[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const T sq = rh[q] - (dg*enew[q] + T(0.5)*res[q]);
            ss += (double)sq*(double)sq;
            o[q] = (sq - T(0.5)*acc0)/dg;
        }
        st_split<false>(dstar, N, i, o);
    }
    block_sum<kSB>(ss, part, pofs + b);
}

// ---- FP64 residual r = b - A(dq) on base Q (split layout) over all cells;
// dq in the split layout (DQS: the solver's own increment) or [var][cell]
// (a caller's array).  Optionally writes fl32(r) ([var][cell]) and the FP32
// base state fl32(Q + dq) (split) for the inner loop.
// Launched once per colour range [n0, n0+cnt) like the sweeps: all XCDs
// then work on slices of one colour and its neighbours at a time.
// ZERO: dq is known to be zero (the first outer step).  Every face term of a
// zero increment is +0 for finite states (a flux difference of identical
// states, or products with zero summed onto +0), so r = b - (D*0 + 0.5*0)
// bit for bit without the neighbour gathers; the partial sums keep the
// same blocks and order, so the norm is identical too.
template <int FC, bool OUT32, bool ZERO, bool DQS>
__global__ __launch_bounds__(kSB) void k_residual64(
    int n0, int cnt, int N, const int *__restrict__ nei, const Face4<double> *__restrict__ geo,
    const double *__restrict__ diag, const double *__restrict__ q, const double *__restrict__ b,
    const double *__restrict__ dq, double *__restrict__ r, float *__restrict__ r32,
    float *__restrict__ qb32, double *part, int pofs)
{
    const int blk = xcd_remap();
    const int t = blk*kSB + threadIdx.x;
    const int i = n0 + t;
    double ss = 0.0;
    if (t < cnt) {
        int nbv[FC];
#pragma unroll
        for (int f = 0; f < FC; ++f) nbv[f] = ldg_nt(nei + (size_t)f*N + i);
        const double dg = ldg_nt(diag + i);
        double bi[NV], dqi[NV], qi[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) bi[k] = ldg_nt(b + (size_t)k*N + i);
        ld_vec<DQS>(dq, N, i, dqi);
        if (OUT32) ld_split<false>(q, N, i, qi);
        bool in_f[FC];
        Face4<double> g[FC];
        double U[FC][NV], d[FC][NV];
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const int nb = nbv[f];
            in_f[f] = !ZERO && nb != i;
            if (!in_f[f]) continue;
            g[f] = ldg_nt(geo + (size_t)f*N + i);
            ld_split<false>(q, N, nb, U[f]);
            ld_vec<DQS>(dq, N, nb, d[f]);
        }
        double res[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) res[k] = 0.0;
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            if (!in_f[f]) continue;
            double o[NV];
            face_term<double>(U[f], d[f], g[f], o);
#pragma unroll
            for (int k = 0; k < NV; ++k) res[k] += o[k];
        }
        float qo[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const size_t o = (size_t)k*N + i;
            const double rv = bi[k] - (dg*dqi[k] + 0.5*res[k]);
            ss += rv*rv;
            if (r) r[o] = rv;
            if (OUT32) {
                r32[o] = (float)rv;
                qo[k] = (float)(qi[k] + dqi[k]);
            }
        }
        if (OUT32) st_split<false>(qb32, N, i, qo);
    }
    block_sum<kSB>(ss, part, pofs + blk);
}

// Sum of the block partials in a fixed order (thread t sums t, t+1024, ...
// with eight loads in flight, then a fixed tree), so the norm and every
// convergence decision are reproducible run to run.
// The result goes to out (device) and to hout (pinned host memory, read by
// the host after a stream synchronisation: no copy kernel).
__global__ __launch_bounds__(1024) void k_sum_parts(const double *part, int n, double *out,
                                                    double *hout)
{
    constexpr int U = 8;
    __shared__ double red[16];
    double s = 0.0;
    for (int j0 = threadIdx.x; j0 < n; j0 += 1024*U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u*1024;
            v[u] = j < n ? part[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) s += v[u];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x % 64 == 0) red[threadIdx.x/64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += red[w];
        *out = t;
        if (hout) *hout = t;
    }
}

// ---- r = b - A(dq) read straight from the legacy arrays (no packed face
// records): the face record and the diagonal are formed in registers with
// the expressions of k_pack, so r and the block partials equal k_residual64's
// bit for bit.  Cells of colour range [n0, n0+cnt); a neighbour index out of
// range or inside the cell's own colour raises a bit in flag (that face is
// skipped).
template <int FC>
__global__ __launch_bounds__(kSB) void k_residual_raw(
    int n0, int cnt, int N, const int *__restrict__ nei, const double *__restrict__ fnv,
    const double *__restrict__ vfn, const double *__restrict__ fspr,
    const double *__restrict__ dt, const double *__restrict__ q, const double *__restrict__ b,
    const double *__restrict__ dq, double *__restrict__ r, double *part, int pofs, int *flag)
{
    const int blk = xcd_remap();
    const int t = blk*kSB + threadIdx.x;
    const int i = n0 + t;
    double ss = 0.0;
    if (t < cnt) {
        int nbv[FC];
        double fv[FC], lam[FC], nrm[FC][3];
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const size_t o = (size_t)f*N + i;
            nbv[f] = ldg_nt(nei + o);
            fv[f] = ldg_nt(fnv + o);
            lam[f] = ldg_nt(fspr + o);
#pragma unroll
            for (int d = 0; d < 3; ++d) nrm[f][d] = ldg_nt(vfn + (size_t)(3*f + d)*N + i);
        }
        const double dti = ldg_nt(dt + i);
        double bi[NV], dqi[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const size_t o = (size_t)k*N + i;
            bi[k] = ldg_nt(b + o);
            dqi[k] = dq[o];
        }
        bool in_f[FC];
        double U[FC][NV], d[FC][NV];
        int bad = 0;
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            const int nb = nbv[f];
            if ((unsigned)nb >= (unsigned)N) {
                bad |= UCFD_KFLAG_INDEX;
                in_f[f] = false;
            } else if (nb != i && nb >= n0 && nb < n0 + cnt) {
                bad |= UCFD_KFLAG_LAYOUT;
                in_f[f] = false;
            } else {
                in_f[f] = nb != i;
            }
            if (!in_f[f]) continue;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                U[f][k] = q[(size_t)k*N + nb];
                d[f][k] = dq[(size_t)k*N + nb];
            }
        }
        if (bad) atomicOr(flag, bad);
        double sd = 0.0;
#pragma unroll
        for (int f = 0; f < FC; ++f) sd += 0.5*lam[f]*fv[f];
        const double dg = 1.0/dti + sd;
        double res[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) res[k] = 0.0;
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            if (!in_f[f]) continue;
            const Face4<double> g{fv[f]*nrm[f][0], fv[f]*nrm[f][1], fv[f]*nrm[f][2],
                                  fv[f]*lam[f]};
            double o[NV];
            face_term<double>(U[f], d[f], g, o);
#pragma unroll
            for (int k = 0; k < NV; ++k) res[k] += o[k];
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const size_t o = (size_t)k*N + i;
            const double rv = bi[k] - (dg*dqi[k] + 0.5*res[k]);
            ss += rv*rv;
            if (r) r[o] = rv;
        }
    }
    block_sum<kSB>(ss, part, pofs + blk);
}

__global__ void k_add32(int64_t n, double *__restrict__ y, const float *__restrict__ x)
{
    const int64_t st = (int64_t)gridDim.x*kB;
    for (int64_t j = (int64_t)blockIdx.x*kB + threadIdx.x; j < n; j += st) y[j] += (double)x[j];
}

template <typename T> __global__ void k_fill(int64_t n, T *y, T v)
{
    const int64_t st = (int64_t)gridDim.x*kB;
    for (int64_t j = (int64_t)blockIdx.x*kB + threadIdx.x; j < n; j += st) y[j] = v;
}

// [var][cell] <-> split layout of the solver's 5-vectors
__global__ void k_rows_to_split(int N, const double *__restrict__ rows, double *__restrict__ sp)
{
    const int i = blockIdx.x*kB + threadIdx.x;
    if (i >= N) return;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = rows[(size_t)q*N + i];
    st_split<false>(sp, N, i, v);
}

__global__ void k_split_to_rows(int N, const double *__restrict__ sp, double *__restrict__ rows)
{
    const int i = blockIdx.x*kB + threadIdx.x;
    if (i >= N) return;
    double v[NV];
    ld_split<false>(sp, N, i, v);
#pragma unroll
    for (int q = 0; q < NV; ++q) rows[(size_t)q*N + i] = v[q];
}

// packs S = fnv*n, L = fnv*lam and D = 1/dt + 1/2 sum fspr*fnv (FP64 and FP32)
__global__ void k_pack(int N, int F, const double *__restrict__ fnv, const double *__restrict__ vfn,
                       const double *__restrict__ fspr, const double *__restrict__ dt,
                       Face4<double> *g64, Face4<float> *g32, double *d64, float *d32)
{
    const int i = blockIdx.x*kB + threadIdx.x;
    if (i >= N) return;
    double sd = 0.0;
    for (int f = 0; f < F; ++f) {
        const size_t o = (size_t)f*N + i;
        const double fv = fnv[o], lam = fspr[o];
        Face4<double> g;
        g.s0 = fv*vfn[(size_t)(3*f + 0)*N + i];
        g.s1 = fv*vfn[(size_t)(3*f + 1)*N + i];
        g.s2 = fv*vfn[(size_t)(3*f + 2)*N + i];
        g.l = fv*lam;
        g64[o] = g;
        g32[o] = Face4<float>{(float)g.s0, (float)g.s1, (float)g.s2, (float)g.l};
        sd += 0.5*lam*fv;
    }
    const double d = 1.0/dt[i] + sd;
    d64[i] = d;
    d32[i] = (float)d;
}

inline dim3 grid1(int64_t n)
{
    int64_t b = (n + kB - 1)/kB;
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20)));
}

}  // namespace

// ============================================================== host driver

struct ucfd_lusgs_solver {
    int N = 0, F = 0, nc = 0;
    std::vector<int> off;
    int *nei = nullptr;
    Face4<double> *g64 = nullptr;
    Face4<float> *g32 = nullptr;
    // q64 (base state), dq64 (the solve's increment), dstar*/delta*, e32 and
    // qb32: split layout; r32: [var][cell]
    double *d64 = nullptr, *q64 = nullptr, *dq64 = nullptr, *dstar64 = nullptr,
           *delta64 = nullptr;
    float *d32 = nullptr, *qb32 = nullptr, *r32 = nullptr, *e32 = nullptr, *dstar32 = nullptr,
          *delta32 = nullptr;
    double *part = nullptr, *norm = nullptr, *hnorm = nullptr;
    double *hnorm_dev = nullptr;   // device address of the pinned hnorm
    int nparts = 0;
    std::vector<void *> allocs;
};

namespace {

template <typename T> int dalloc(ucfd_lusgs_solver *s, T **p, size_t n, const char *fn)
{
    int rc = ucfd_check_hip(hipMalloc((void **)p, sizeof(T)*std::max<size_t>(n, 1)), fn, "hipMalloc");
    if (!rc) s->allocs.push_back((void *)*p);
    return rc;
}

void free_all(ucfd_lusgs_solver *s)
{
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->hnorm) (void)hipHostFree(s->hnorm);
    s->allocs.clear();
}

int copy_in(void *dst, const void *src, size_t bytes, hipStream_t st, const char *fn)
{
    return ucfd_check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, st), fn,
                          "hipMemcpyAsync(in)");
}

// lower (with residual) over all colours, norm of s to *norm.  pre0: the
// previous pass's upper launch formed the first colour's residual, block
// partials and dstar (k_upper_upd_res), so the pass starts at the second
// colour.  With more than one colour the last colour's launch also does that
// colour's upper step, writing delta and e (upper_pass then skips it).
template <typename T>
int lower_pass(ucfd_lusgs_solver *s, hipStream_t st, bool with_res, const T *qb, const T *rhs,
               T *e, T *dstar, T *delta, const Face4<T> *geo, const T *diag, bool pre0 = false)
{
    int pofs = 0;
    for (int c = 0; c < s->nc; ++c) {
        const int n0 = s->off[c], cnt = s->off[c + 1] - n0;
        if (cnt <= 0) continue;
        const dim3 g((cnt + kSB - 1)/kSB);
        if (c == 0 && pre0) {
            pofs += (int)g.x;
            continue;
        }
        const bool last = s->nc > 1 && c == s->nc - 1;
#define LR(FC, RES, LAST) hipLaunchKernelGGL((k_lower_res<T, FC, RES, LAST>), g, dim3(kSB), 0, st, \
                                             n0, cnt, s->N, s->nei, geo, diag, qb, rhs, e, dstar, \
                                             s->part, pofs, delta)
        if (s->F == 4) {
            if (last) { if (with_res) LR(4, true, true); else LR(4, false, true); }
            else      { if (with_res) LR(4, true, false); else LR(4, false, false); }
        } else {
            if (last) { if (with_res) LR(6, true, true); else LR(6, false, true); }
            else      { if (with_res) LR(6, true, false); else LR(6, false, false); }
        }
#undef LR
        pofs += (int)g.x;
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, st, s->part, pofs, s->norm,
                       s->hnorm_dev);
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

// upper sweep and update over all colours (last colour first).  rhs given:
// the first colour's launch also forms the next pass's residual of its
// cells, their block partials (at the start of part) and their dstar for
// the next pass.  The last colour's upper step was done by the lower pass
// (more than one colour).
template <typename T>
int upper_pass(ucfd_lusgs_solver *s, hipStream_t st, const T *qb, T *dstar, T *delta, T *e,
               const Face4<T> *geo, const T *diag, const T *rhs = nullptr)
{
    for (int c = s->nc - 1; c >= 0; --c) {
        const int n0 = s->off[c], cnt = s->off[c + 1] - n0;
        if (cnt <= 0) continue;
        if (s->nc > 1 && c == s->nc - 1) continue;
        const dim3 g((cnt + kSB - 1)/kSB);
        if (c == 0 && rhs) {
            if (s->F == 4)
                hipLaunchKernelGGL((k_upper_upd_res<T, 4>), g, dim3(kSB), 0, st, cnt, s->N, s->nei,
                                   geo, diag, qb, dstar, delta, e, rhs, s->part, 0);
            else
                hipLaunchKernelGGL((k_upper_upd_res<T, 6>), g, dim3(kSB), 0, st, cnt, s->N, s->nei,
                                   geo, diag, qb, dstar, delta, e, rhs, s->part, 0);
            continue;
        }
        if (s->F == 4)
            hipLaunchKernelGGL((k_upper_upd<T, 4>), g, dim3(kSB), 0, st, n0, cnt, s->N, s->nei, geo,
                               diag, qb, dstar, delta, e);
        else
            hipLaunchKernelGGL((k_upper_upd<T, 6>), g, dim3(kSB), 0, st, n0, cnt, s->N, s->nei, geo,
                               diag, qb, dstar, delta, e);
    }
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

// zero_dq: dq is known to be zero (only with out32, the first outer step);
// dq_split: dq is in the split layout (the solver's dq64), else [var][cell]
int residual64(ucfd_lusgs_solver *s, hipStream_t st, const double *b, const double *dq, double *r,
               bool out32, bool zero_dq, bool dq_split)
{
    int pofs = 0;
    for (int c = 0; c < s->nc; ++c) {
        const int n0 = s->off[c], cnt = s->off[c + 1] - n0;
        if (cnt <= 0) continue;
        const dim3 g((cnt + kSB - 1)/kSB);
#define R(FC, O, Z, S) hipLaunchKernelGGL((k_residual64<FC, O, Z, S>), g, dim3(kSB), 0, st, n0, \
                                          cnt, s->N, s->nei, s->g64, s->d64, s->q64, b, dq, r, \
                                          s->r32, s->qb32, s->part, pofs)
        if (s->F == 4) {
            if (out32 && zero_dq) R(4, true, true, true);
            else if (out32) R(4, true, false, true);
            else if (dq_split) R(4, false, false, true);
            else R(4, false, false, false);
        } else {
            if (out32 && zero_dq) R(6, true, true, true);
            else if (out32) R(6, true, false, true);
            else if (dq_split) R(6, false, false, true);
            else R(6, false, false, false);
        }
#undef R
        pofs += (int)g.x;
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, st, s->part, pofs, s->norm,
                       s->hnorm_dev);
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

int read_norm(ucfd_lusgs_solver *s, hipStream_t st, double *out, const char *fn)
{
    // the last k_sum_parts wrote the norm into the pinned hnorm
    int rc = ucfd_check_hip(hipStreamSynchronize(st), fn, "hipStreamSynchronize");
    *out = sqrt(*s->hnorm);
    return rc;
}

}  // namespace

UCFD_API int ucfd_lusgs_solver_create(ucfd_lusgs_solver **out, int neles, int nface,
                                      int ncolors, const int *color_off, const int *nei_ele,
                                      const double *fnorm_vol, const double *vec_fnorm,
                                      const double *fspr, const double *dt, const double *upts)
{
    const char *fn = "ucfd_lusgs_solver_create";
    if (!out || neles <= 0 || (nface != 4 && nface != 6) || ncolors <= 0 || !color_off ||
        !nei_ele || !fnorm_vol || !vec_fnorm || !fspr || !dt || !upts)
        return ucfd_report(UCFD_ERR_ARG, fn, "bad arguments (nface must be 4 or 6, 3-D Euler)");
    if (ucfd_config().ndims != 3)
        return ucfd_report(UCFD_ERR_ARG, fn, "the solver supports ndims = 3");
    if (color_off[0] != 0 || color_off[ncolors] != neles)
        return ucfd_report(UCFD_ERR_ARG, fn, "color_off must run from 0 to neles");
    for (int c = 0; c < ncolors; ++c)
        if (color_off[c + 1] < color_off[c])
            return ucfd_report(UCFD_ERR_ARG, fn, "color_off must be non-decreasing");
    int rc = ucfd_ensure_device(fn);
    if (rc) return rc;
    const size_t N = (size_t)neles, F = (size_t)nface;

    // the neighbour table is validated on the host: in range, and no
    // neighbour inside the cell's own colour range (a valid colouring)
    std::vector<int> hn(F*N);
    if ((rc = ucfd_check_hip(hipMemcpy(hn.data(), nei_ele, F*N*4, hipMemcpyDefault), fn,
                             "hipMemcpy(nei_ele)")))
        return rc;
    for (int c = 0; c < ncolors; ++c)
        for (int i = color_off[c]; i < color_off[c + 1]; ++i)
            for (size_t f = 0; f < F; ++f) {
                const int nb = hn[f*N + i];
                if (nb < 0 || nb >= neles)
                    return ucfd_report(UCFD_ERR_INDEX, fn, "nei_ele entry out of range");
                if (nb != i && nb >= color_off[c] && nb < color_off[c + 1])
                    return ucfd_report(UCFD_ERR_ARG, fn, "a face joins two cells of one colour");
            }

    ucfd_lusgs_solver *s = new ucfd_lusgs_solver();
    s->N = neles;
    s->F = nface;
    s->nc = ncolors;
    s->off.assign(color_off, color_off + ncolors + 1);
    int maxparts = 0;
    for (int c = 0; c < ncolors; ++c) maxparts += (s->off[c + 1] - s->off[c] + kSB - 1)/kSB;
    s->nparts = std::max(maxparts, (neles + kSB - 1)/kSB) + 8;
    const hipStream_t st = ucfd_stream();
    double *tfnv = nullptr, *tvfn = nullptr, *tsp = nullptr, *tdt = nullptr, *tq = nullptr;
    if ((rc = dalloc(s, &s->nei, F*N, fn)) || (rc = dalloc(s, &s->g64, F*N, fn)) ||
        (rc = dalloc(s, &s->g32, F*N, fn)) || (rc = dalloc(s, &s->d64, N, fn)) ||
        (rc = dalloc(s, &s->d32, N, fn)) || (rc = dalloc(s, &s->q64, NV*N, fn)) ||
        (rc = dalloc(s, &s->dq64, NV*N, fn)) || (rc = dalloc(s, &s->dstar64, NV*N, fn)) ||
        (rc = dalloc(s, &s->delta64, NV*N, fn)) || (rc = dalloc(s, &s->qb32, NV*N, fn)) ||
        (rc = dalloc(s, &s->r32, NV*N, fn)) || (rc = dalloc(s, &s->e32, NV*N, fn)) ||
        (rc = dalloc(s, &s->dstar32, NV*N, fn)) || (rc = dalloc(s, &s->delta32, NV*N, fn)) ||
        (rc = dalloc(s, &s->part, (size_t)s->nparts, fn)) || (rc = dalloc(s, &s->norm, 8, fn)) ||
        (rc = dalloc(s, &tfnv, F*N, fn)) || (rc = dalloc(s, &tvfn, 3*F*N, fn)) ||
        (rc = dalloc(s, &tsp, F*N, fn)) || (rc = dalloc(s, &tdt, N, fn)) ||
        (rc = dalloc(s, &tq, NV*N, fn))) {
        free_all(s);
        delete s;
        return rc;
    }
    if ((rc = ucfd_check_hip(hipHostMalloc((void **)&s->hnorm, 64, hipHostMallocDefault), fn,
                             "hipHostMalloc")) ||
        (rc = ucfd_check_hip(hipHostGetDevicePointer((void **)&s->hnorm_dev, s->hnorm, 0), fn,
                             "hipHostGetDevicePointer"))) {
        free_all(s);
        delete s;
        return rc;
    }
    rc = copy_in(s->nei, hn.data(), F*N*4, st, fn);
    if (!rc) rc = copy_in(tfnv, fnorm_vol, F*N*8, st, fn);
    if (!rc) rc = copy_in(tvfn, vec_fnorm, 3*F*N*8, st, fn);
    if (!rc) rc = copy_in(tsp, fspr, F*N*8, st, fn);
    if (!rc) rc = copy_in(tdt, dt, N*8, st, fn);
    if (!rc) rc = copy_in(tq, upts, NV*N*8, st, fn);
    if (!rc) {
        hipLaunchKernelGGL(k_rows_to_split, dim3((neles + kB - 1)/kB), dim3(kB), 0, st, neles,
                           tq, s->q64);
        hipLaunchKernelGGL(k_pack, dim3((neles + kB - 1)/kB), dim3(kB), 0, st, neles, nface, tfnv,
                           tvfn, tsp, tdt, s->g64, s->g32, s->d64, s->d32);
        rc = ucfd_check_hip(hipStreamSynchronize(st), fn, "pack");
    }
    // the FP64 staging copies of the raw geometry and state are no longer needed
    for (void *p : {(void *)tfnv, (void *)tvfn, (void *)tsp, (void *)tdt, (void *)tq}) {
        (void)hipFree(p);
        s->allocs.erase(std::remove(s->allocs.begin(), s->allocs.end(), p), s->allocs.end());
    }
    if (rc) {
        free_all(s);
        delete s;
        return rc;
    }
    *out = s;
    return UCFD_OK;
}

UCFD_API int ucfd_lusgs_solver_destroy(ucfd_lusgs_solver *s)
{
    if (!s) return UCFD_OK;
    (void)hipStreamSynchronize(ucfd_stream());
    free_all(s);
    delete s;
    return UCFD_OK;
}

// r = b - A(dq) in FP64 (b, dq, r device or host [5][N]); *rnorm = ||r||_2.
UCFD_API int ucfd_lusgs_residual(ucfd_lusgs_solver *s, const double *b, const double *dq,
                                 double *r, double *rnorm)
{
    const char *fn = "ucfd_lusgs_residual";
    if (!s || !b || !dq) return ucfd_report(UCFD_ERR_ARG, fn, "bad arguments");
    const size_t n = (size_t)NV*s->N;
    UcfdStage sg(fn);
    const int ab = sg.add(b, n*8, UCFD_IN), aq = sg.add(dq, n*8, UCFD_IN);
    const int ar = r ? sg.add(r, n*8, UCFD_OUT) : -1;
    int rc = sg.enter();
    if (rc) return rc;
    const hipStream_t st = ucfd_stream();
    rc = residual64(s, st, (const double *)sg.dev(ab), (const double *)sg.dev(aq),
                    ar >= 0 ? (double *)sg.dev(ar) : nullptr, false, false, false);
    double nrm = 0.0;
    if (!rc) rc = read_norm(s, st, &nrm, fn);
    if (rnorm) *rnorm = nrm;
    return sg.leave(rc);
}

// Solve A(dq) = b to ||b - A(dq)|| <= rtol*||b|| by defect correction.
// mode UCFD_SOLVE_FP64: FP64 sweeps; UCFD_SOLVE_MIXED: FP32 inner sweeps with
// FP64 outer residual and FP64 accumulation of dq.  dq is output (the start
// value is zero).  maxiter bounds the total number of LU-SGS passes.
UCFD_API int ucfd_lusgs_solve(ucfd_lusgs_solver *s, int mode, const double *b, double *dq,
                              double rtol, int maxiter, ucfd_solve_stats *stats)
{
    const char *fn = "ucfd_lusgs_solve";
    if (!s || !b || !dq || maxiter <= 0 || (mode != UCFD_SOLVE_FP64 && mode != UCFD_SOLVE_MIXED))
        return ucfd_report(UCFD_ERR_ARG, fn, "bad arguments");
    const size_t n = (size_t)NV*s->N;
    UcfdStage sg(fn);
    const int ab = sg.add(b, n*8, UCFD_IN), aq = sg.add(dq, n*8, UCFD_OUT);
    int rc = sg.enter();
    if (rc) return rc;
    const hipStream_t st = ucfd_stream();
    const double *B = (const double *)sg.dev(ab);
    double *DQ = (double *)sg.dev(aq);
    ucfd_solve_stats info;
    memset(&info, 0, sizeof info);
    // the increment is accumulated in the split dq64 and written to DQ at the end
    double *E = s->dq64;
    hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(kB), 0, st, (int64_t)n, E, 0.0);
    double bnorm = 0.0, rn = 0.0;

    auto record = [&](double v) {
        if (info.nhist < UCFD_SOLVE_HIST) info.history[info.nhist++] = v;
    };

    if (mode == UCFD_SOLVE_FP64) {
        // e = dq; iteration j: s = b - A(e) (in the lower launches), e += M^-1 s
        for (int it = 0; it < maxiter && !rc; ++it) {
            rc = lower_pass<double>(s, st, it > 0, s->q64, B, E, s->dstar64, s->delta64, s->g64,
                                    s->d64, it > 0);
            if (!rc) rc = upper_pass<double>(s, st, s->q64, s->dstar64, s->delta64, E, s->g64,
                                             s->d64, B);
            if (!rc) rc = read_norm(s, st, &rn, fn);
            if (rc) break;
            if (it == 0) bnorm = rn;
            record(bnorm > 0 ? rn/bnorm : 0.0);
            info.passes++;
            if (bnorm == 0.0 || rn <= rtol*bnorm) break;
        }
        // the last check saw the iterate before its update; report the
        // residual of the dq actually returned
        if (!rc) rc = residual64(s, st, B, E, nullptr, false, false, true);
        if (!rc) rc = read_norm(s, st, &rn, fn);
        info.outer = info.passes + 1;
    } else {
        // outer FP64 residual; inner FP32 defect correction on the correction
        while (!rc && info.passes < maxiter) {
            rc = residual64(s, st, B, E, nullptr, true, info.outer == 0, true);
            if (!rc) rc = read_norm(s, st, &rn, fn);
            if (rc) break;
            if (info.outer == 0) bnorm = rn;
            info.outer++;
            record(bnorm > 0 ? rn/bnorm : 0.0);
            if (bnorm == 0.0 || rn <= rtol*bnorm) break;
            // the inner loop only needs to reach the outer target, within the
            // accuracy FP32 can deliver
            const double inner_tol = std::max(1e-5, 0.5*rtol*bnorm/rn);
            hipLaunchKernelGGL(k_fill<float>, grid1(n), dim3(kB), 0, st, (int64_t)n, s->e32, 0.0f);
            double s0 = 0.0;
            for (int j = 0; info.passes < maxiter && !rc; ++j) {
                rc = lower_pass<float>(s, st, j > 0, s->qb32, s->r32, s->e32, s->dstar32,
                                       s->delta32, s->g32, s->d32, j > 0);
                if (!rc) rc = upper_pass<float>(s, st, s->qb32, s->dstar32, s->delta32, s->e32,
                                                s->g32, s->d32, s->r32);
                double sn = 0.0;
                if (!rc) rc = read_norm(s, st, &sn, fn);
                info.passes++;
                if (j == 0) s0 = sn;
                if (s0 == 0.0 || sn <= inner_tol*s0) break;
            }
            // dq64 and e32 share the split layout: elementwise over 5N
            hipLaunchKernelGGL(k_add32, grid1(n), dim3(kB), 0, st, (int64_t)n, E, s->e32);
        }
        if (!rc && info.passes >= maxiter) {
            // report the residual of the returned dq
            rc = residual64(s, st, B, E, nullptr, false, false, true);
            if (!rc) rc = read_norm(s, st, &rn, fn);
            record(bnorm > 0 ? rn/bnorm : 0.0);
        }
    }
    if (!rc) {
        hipLaunchKernelGGL(k_split_to_rows, dim3((s->N + kB - 1)/kB), dim3(kB), 0, st, s->N, E,
                           DQ);
        rc = ucfd_check_hip(hipGetLastError(), fn, "kernel launch");
    }
    info.rnorm0 = bnorm;
    info.rel_residual = bnorm > 0 ? rn/bnorm : 0.0;
    info.converged = bnorm == 0.0 || rn <= rtol*bnorm;
    if (stats) *stats = info;
    return sg.leave(rc);
}

// Measurement extension: HIP-event times of the solver's kernels, one full
// pass each (every colour), averaged over reps, on the library stream:
// ms[0] lower sweep with residual (FP64; the first colour's residual comes
// from the previous pass, as in the solve), ms[1] upper sweep with update
// and the first colour's next residual (FP64), ms[2] / ms[3] the same in
// FP32 (the inner kernels of the mixed form), ms[4] the FP64 residual (all
// colours) and its norm.  The state is a
// first outer step (dq = 0 with right-hand side b); the solver's work arrays
// are overwritten.
UCFD_API int ucfd_lusgs_solver_time_kernels(ucfd_lusgs_solver *s, const double *b, int reps,
                                            double *ms)
{
    const char *fn = "ucfd_lusgs_solver_time_kernels";
    if (!s || !b || !ms || reps <= 0) return ucfd_report(UCFD_ERR_ARG, fn, "bad arguments");
    if (!ucfd_is_device_ptr(b)) return ucfd_report(UCFD_ERR_ARG, fn, "b must be device memory");
    const hipStream_t st = ucfd_stream();
    const size_t n = (size_t)NV*s->N;
    hipEvent_t ev[6];
    for (int k = 0; k < 6; ++k)
        if (int rc = ucfd_check_hip(hipEventCreate(&ev[k]), fn, "hipEventCreate")) {
            for (int j = 0; j < k; ++j) (void)hipEventDestroy(ev[j]);
            return rc;
        }
    // e = dq = 0 (dq64 serves as both): FP64 and FP32 state of a first step
    hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(kB), 0, st, (int64_t)n, s->dq64, 0.0);
    hipLaunchKernelGGL(k_fill<float>, grid1(n), dim3(kB), 0, st, (int64_t)n, s->e32, 0.0f);
    int rc = residual64(s, st, b, s->dq64, nullptr, true, true, true);
    // one untimed pass of each form, so the first colour's residual exists
    if (!rc) rc = lower_pass<double>(s, st, true, s->q64, b, s->dq64, s->dstar64, s->delta64,
                                     s->g64, s->d64);
    if (!rc) rc = upper_pass<double>(s, st, s->q64, s->dstar64, s->delta64, s->dq64, s->g64,
                                     s->d64, b);
    if (!rc) rc = lower_pass<float>(s, st, true, s->qb32, s->r32, s->e32, s->dstar32, s->delta32,
                                    s->g32, s->d32);
    if (!rc) rc = upper_pass<float>(s, st, s->qb32, s->dstar32, s->delta32, s->e32, s->g32, s->d32,
                                    s->r32);
    double acc[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r < reps && !rc; ++r) {
        (void)hipEventRecord(ev[0], st);
        rc = lower_pass<double>(s, st, true, s->q64, b, s->dq64, s->dstar64, s->delta64, s->g64,
                                s->d64, true);
        (void)hipEventRecord(ev[1], st);
        if (!rc) rc = upper_pass<double>(s, st, s->q64, s->dstar64, s->delta64, s->dq64, s->g64,
                                         s->d64, b);
        (void)hipEventRecord(ev[2], st);
        if (!rc) rc = lower_pass<float>(s, st, true, s->qb32, s->r32, s->e32, s->dstar32,
                                        s->delta32, s->g32, s->d32, true);
        (void)hipEventRecord(ev[3], st);
        if (!rc) rc = upper_pass<float>(s, st, s->qb32, s->dstar32, s->delta32, s->e32, s->g32,
                                        s->d32, s->r32);
        (void)hipEventRecord(ev[4], st);
        if (!rc) rc = residual64(s, st, b, s->dq64, nullptr, false, false, true);
        (void)hipEventRecord(ev[5], st);
        if (!rc) rc = ucfd_check_hip(hipEventSynchronize(ev[5]), fn, "hipEventSynchronize");
        for (int k = 0; k < 5 && !rc; ++k) {
            float t = 0.f;
            rc = ucfd_check_hip(hipEventElapsedTime(&t, ev[k], ev[k + 1]), fn,
                                "hipEventElapsedTime");
            acc[k] += t;
        }
    }
    for (int k = 0; k < 6; ++k) (void)hipEventDestroy(ev[k]);
    if (rc) return rc;
    for (int k = 0; k < 5; ++k) ms[k] = acc[k]/reps;
    return UCFD_OK;
}

// ===================================================== no-allocation residual
//
// ucfd_lusgs_residual_noalloc: the residual of ucfd_lusgs_residual without a
// solver handle and without any allocation, for callers that cannot allocate
// device memory.  Its only scratch (the block partial sums, the norm and an
// error word) is the file-scope array below: one copy per GPU, alive until
// the process ends, shared by every call on that GPU (overlapping calls on
// one GPU race).  Host code never passes the array's name to a kernel: its
// device address is looked up once per device with hipGetSymbolAddress.

namespace {

constexpr int kNoAllocParts = 131072;   // block partials: 1 MiB
struct NoAllocScratch {
    double part[kNoAllocParts];
    double norm[8];
    int flag[8];
};
static __device__ __align__(256) NoAllocScratch scratch;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
NoAllocScratch *g_scratch[kMaxDevices] = {};

__global__ void k_clear_flag(int *flag)
{
    if (threadIdx.x == 0) flag[0] = 0;
}

// device address of this device's scratch (looked up on the first call)
int scratch_ptr(NoAllocScratch **out, const char *fn)
{
    int dev = -1;
    int rc = ucfd_check_hip(hipGetDevice(&dev), fn, "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices)
        return ucfd_report(UCFD_ERR_HIP, fn, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        rc = ucfd_check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), fn,
                            "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        g_scratch[dev] = (NoAllocScratch *)p;
    }
    *out = g_scratch[dev];
    return UCFD_OK;
}

}  // namespace

UCFD_API int ucfd_lusgs_residual_noalloc(int neles, int nface, int ncolors, const int *color_off,
                                         const int *nei_ele, const double *fnorm_vol,
                                         const double *vec_fnorm, const double *fspr,
                                         const double *dt, const double *upts, const double *b,
                                         const double *dq, double *r, double *rnorm)
{
    const char *fn = "ucfd_lusgs_residual_noalloc";
    if (neles <= 0 || (nface != 4 && nface != 6) || ncolors <= 0 || !color_off || !nei_ele ||
        !fnorm_vol || !vec_fnorm || !fspr || !dt || !upts || !b || !dq)
        return ucfd_report(UCFD_ERR_ARG, fn, "bad arguments (nface must be 4 or 6, 3-D Euler)");
    if (color_off[0] != 0 || color_off[ncolors] != neles)
        return ucfd_report(UCFD_ERR_ARG, fn, "color_off must run from 0 to neles");
    long long parts = 0;
    for (int c = 0; c < ncolors; ++c) {
        if (color_off[c + 1] < color_off[c])
            return ucfd_report(UCFD_ERR_ARG, fn, "color_off must be non-decreasing");
        parts += (color_off[c + 1] - color_off[c] + kSB - 1)/kSB;
    }
    if (parts > kNoAllocParts)
        return ucfd_report(UCFD_ERR_ARG, fn, "mesh too large for the fixed scratch: the colour "
                                             "ranges need more than 131072 workgroups of 256 "
                                             "cells (about 33.5M cells)");
    if (ucfd_config().ndims != 3)
        return ucfd_report(UCFD_ERR_ARG, fn, "3-D Euler only (ndims = 3)");
    const void *arrays[] = {nei_ele, fnorm_vol, vec_fnorm, fspr, dt, upts, b, dq, r};
    for (const void *a : arrays)
        if (a && !ucfd_is_device_ptr(a))
            return ucfd_report(UCFD_ERR_ARG, fn, "arrays must be device (or managed) memory: "
                                                 "host arrays would need a staging allocation");
    NoAllocScratch *S = nullptr;
    int rc = scratch_ptr(&S, fn);
    if (rc) return rc;
    const hipStream_t st = ucfd_stream();
    int *flag = S->flag;
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, flag);
    int pofs = 0;
    for (int c = 0; c < ncolors; ++c) {
        const int n0 = color_off[c], cnt = color_off[c + 1] - n0;
        if (cnt <= 0) continue;
        const dim3 g((cnt + kSB - 1)/kSB);
        if (nface == 4)
            hipLaunchKernelGGL((k_residual_raw<4>), g, dim3(kSB), 0, st, n0, cnt, neles, nei_ele,
                               fnorm_vol, vec_fnorm, fspr, dt, upts, b, dq, r, S->part, pofs,
                               flag);
        else
            hipLaunchKernelGGL((k_residual_raw<6>), g, dim3(kSB), 0, st, n0, cnt, neles, nei_ele,
                               fnorm_vol, vec_fnorm, fspr, dt, upts, b, dq, r, S->part, pofs,
                               flag);
        pofs += (int)g.x;
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, st, S->part, pofs, S->norm,
                       (double *)nullptr);
    if ((rc = ucfd_check_hip(hipGetLastError(), fn, "kernel launch"))) return rc;
    double nrm2 = 0.0;
    int f = 0;
    if ((rc = ucfd_check_hip(hipMemcpyAsync(&nrm2, S->norm, sizeof nrm2, hipMemcpyDeviceToHost,
                                            st), fn, "hipMemcpyAsync(norm)")) ||
        (rc = ucfd_check_hip(hipMemcpyAsync(&f, flag, sizeof f, hipMemcpyDeviceToHost, st), fn,
                             "hipMemcpyAsync(flag)")) ||
        (rc = ucfd_check_hip(hipStreamSynchronize(st), fn, "hipStreamSynchronize")))
        return rc;
    if (f & UCFD_KFLAG_INDEX)
        return ucfd_report(UCFD_ERR_INDEX, fn, "nei_ele holds an entry outside [0, neles)");
    if (f & UCFD_KFLAG_LAYOUT)
        return ucfd_report(UCFD_ERR_ARG, fn, "a face joins two cells of one colour");
    if (rnorm) *rnorm = sqrt(nrm2);
    return UCFD_OK;
}
