// FP64 LU-SGS kernels behind the legacy array API: pre (diagonal), the
// lower / upper sweep over one colour (or one dependency level) and the
// solution update.
//
// Sweep: one thread per cell of the launched range; the cell reads its own
// row of every nface-major array (coalesced when the cells of a colour are
// stored contiguously) and gathers neighbour states through nei_ele.  Cells
// of one colour have no face neighbour of the same colour, so a launch has no
// intra-launch dependency and no atomics.  Arithmetic follows
// src/lusgs/lusgs.c:157-257 in operation order; this file is compiled with
// -ffp-contract=off so no FMA is introduced.
//
// Every gathered index is bounds checked; an out-of-range entry raises
// UCFD_KFLAG_INDEX in the device error word and the face (or cell) is
// skipped, so a malformed index array cannot make the kernel fault.
#include "flux.h"
#include "ucfd_internal.h"

namespace {

constexpr int kBlock = 256;
// sweeps: 4 waves per workgroup; at the sweep's 3 waves per SIMD a CU holds
// three such workgroups (12 waves).  Measured on the 2.65M-cell heavy
// launch: 201 us against 210 us with 128 threads and 218 us with 64.
constexpr int kSweepBlock = 256;

__device__ __forceinline__ void raise_flag(int *flag, int bit)
{
    atomicOr(flag, bit);
}

__global__ __launch_bounds__(kBlock) void k_pre(int N, int nface, double factor,
                                                const double *__restrict__ fnv,
                                                const double *__restrict__ dt,
                                                double *__restrict__ diag,
                                                const double *__restrict__ fspr)
{
    const int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= N) return;
    // lusgs.c:33 with a0 = dsrc = 0, then lusgs.c:102-110 with kappa = rcp_vol = 1
    const double d = 1.0/(dt[i]*factor) + 0.0 + 0.0;
    double sd = 0.0;
    for (int f = 0; f < nface; ++f)
        sd += 0.5*fspr[(size_t)f*N + i]*fnv[(size_t)f*N + i];
    diag[i] = d + sd;
}

// The same for a compile-time face count, CPT cells per thread (cells
// t, t + S, ..., S = threads in the grid): every load of a thread is issued
// before the first sum, so more bytes are in flight per wave.
template <int FC, int CPT>
__global__ __launch_bounds__(kBlock) void k_pre_f(int N, double factor,
                                                  const double *__restrict__ fnv,
                                                  const double *__restrict__ dt,
                                                  double *__restrict__ diag,
                                                  const double *__restrict__ fspr)
{
    const int S = gridDim.x*kBlock;
    const int i0 = blockIdx.x*kBlock + threadIdx.x;
    double t[CPT], a[CPT][FC], b[CPT][FC];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int i = i0 + c*S;
        if (i >= N) continue;
        t[c] = __builtin_nontemporal_load(dt + i);
#pragma unroll
        for (int f = 0; f < FC; ++f) {
            a[c][f] = __builtin_nontemporal_load(fspr + (size_t)f*N + i);
            b[c][f] = __builtin_nontemporal_load(fnv + (size_t)f*N + i);
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int i = i0 + c*S;
        if (i >= N) continue;
        const double d = 1.0/(t[c]*factor) + 0.0 + 0.0;
        double sd = 0.0;
#pragma unroll
        for (int f = 0; f < FC; ++f) sd += 0.5*a[c][f]*b[c][f];
        __builtin_nontemporal_store(d + sd, diag + i);
    }
}

// Workgroups are dealt round-robin over the 8 XCDs of the MI355X, each with
// its own L2.  Remapping the hardware workgroup id so that each XCD owns one
// contiguous eighth of the launched cell range keeps the neighbour rows that
// adjacent cells share inside one L2 (spatially sorted cells share most of
// their neighbours' cache lines).  Runs of 16 or 64 workgroups per XCD,
// interleaved, measured the same pass time within 0.3 % but fetched 25 % /
// 13 % more from beyond L2 in the heavy sweep launches; the identity mapping
// was 4 % slower.
__device__ __forceinline__ int xcd_block()
{
    const int b = blockIdx.x, nblk = gridDim.x;
    const int x = b % 8, k = b / 8, per = nblk / 8, rem = nblk % 8;
    return x*per + (x < rem ? x : rem) + k;
}

template <typename T> __device__ __forceinline__ T ld_stream(const T *p)
{
    return __builtin_nontemporal_load(p);
}

// DIR 0: lower, out_i = (in_i - 0.5 df)/diag_i
// DIR 1: upper, out_i = in_i - (0.5 df)/diag_i
// MODE 0 (colored):  cell i = order[n0 + t]; neighbour taken iff
//                    lcolor[nb] </> lcolor[i]
// MODE 1 (ordered):  cell i = order[n0 + t]; neighbour taken iff nb </> i
// MODE 2 (range):    colour-contiguous layout, cell i = n0 + t; neighbour
//                    taken iff nb < n0 (lower) / nb >= ne (upper), which is
//                    the colour test when [n0, ne) is one whole colour of a
//                    layout with non-decreasing lcolor
//
// Memory schedule: all neighbour indices of the cell are loaded first; then,
// G faces at a time, every load a taken face needs (the 2*NV gathered
// neighbour values and the face's own geometry) is issued under that face's
// predicate with no wait in between, and only then is the flux arithmetic
// done, face by face in index order (the reference's accumulation order).
// Rows streamed once per launch (indices, geometry, diag, the cell's own
// input and output) use non-temporal accesses so they do not evict the
// gathered neighbour lines from L2.
template <int DIR, int MODE, int ND, int FC>
__global__ __launch_bounds__(kSweepBlock) void k_sweep(
    int n0, int cnt, int N, int nface, const int *__restrict__ nei,
    const int *__restrict__ order, const int *__restrict__ lcolor,
    const double *__restrict__ fnv, const double *__restrict__ vfn,
    const double *__restrict__ upts, const double *in, double *out,
    const double *__restrict__ diag, const double *__restrict__ fspr, int *flag, int lce)
{
    constexpr int NV = ND + 2;
    // all faces' loads in flight at once for tets; pairs of faces for hexes
    constexpr int G = FC == 4 ? 4 : 2;
    static_assert(FC % G == 0, "face count must be a multiple of the group size");
    // MODE 3: a whole colour whose every neighbour lies in the sweep direction
    // (the last colour of a lower sweep, the first of an upper sweep): the
    // face geometry of all faces is requested before the neighbour indices
    // arrive; the gathers stay predicated.  Same arithmetic as MODE 2.
    constexpr bool RANGE = MODE >= 2, EAGER = MODE == 3;
    const int t = xcd_block()*kSweepBlock + threadIdx.x;
    if (t >= cnt) return;
    // range modes: the workgroup's first cell samples the declared layout
    // (uniform address: scalar loads, consumed after the stores)
    int sic = 0, slc = 0;
    const int ib = n0 + xcd_block()*kSweepBlock;
    if (RANGE) { sic = order[ib]; slc = lcolor[ib]; }
    const int i = RANGE ? n0 + t : ld_stream(order + n0 + t);
    if ((unsigned)i >= (unsigned)N) {
        raise_flag(flag, UCFD_KFLAG_INDEX);
        return;
    }
    int ci = 0;
    if (MODE == 0) ci = lcolor[i];

    int nbv[FC];
    bool tk[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) nbv[f] = ld_stream(nei + (size_t)f*N + i);
    double enf[EAGER ? FC : 1][ND], efv[EAGER ? FC : 1], elam[EAGER ? FC : 1];
    if (EAGER) {
#pragma unroll
        for (int f = 0; f < (EAGER ? FC : 0); ++f) {
#pragma unroll
            for (int d = 0; d < ND; ++d) enf[f][d] = ld_stream(vfn + (size_t)(f*ND + d)*N + i);
            efv[f] = ld_stream(fnv + (size_t)f*N + i);
            elam[f] = ld_stream(fspr + (size_t)f*N + i);
        }
    }
    // the cell's own rows are loaded before the neighbour gathers so their
    // latency overlaps (in may alias out: only other cells' out is gathered)
    const double dg = ld_stream(diag + i);
    double xin[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) xin[k] = ld_stream(in + (size_t)k*N + i);
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const int nb = nbv[f];
        if ((unsigned)nb >= (unsigned)N) {
            raise_flag(flag, UCFD_KFLAG_INDEX);
            tk[f] = false;
        } else if (MODE == 0) {
            const int cn = lcolor[nb];
            tk[f] = DIR == 0 ? (cn < ci) : (cn > ci);
        } else if (MODE == 1) {
            tk[f] = DIR == 0 ? (nb < i) : (nb > i);
        } else {
            tk[f] = DIR == 0 ? (nb < n0) : (nb >= n0 + cnt);
        }
    }

    double df[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) df[k] = 0.0;
#pragma unroll
    for (int f0 = 0; f0 < FC; f0 += G) {
        double u[G][NV], du[G][NV], nf[G][ND], fv[G], lam[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int f = f0 + g;
            if (!tk[f]) continue;
            const int nb = nbv[f];
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                u[g][k] = upts[(size_t)k*N + nb];
                du[g][k] = out[(size_t)k*N + nb];
            }
            if (!EAGER) {
#pragma unroll
                for (int d = 0; d < ND; ++d) nf[g][d] = ld_stream(vfn + (size_t)(f*ND + d)*N + i);
                fv[g] = ld_stream(fnv + (size_t)f*N + i);
                lam[g] = ld_stream(fspr + (size_t)f*N + i);
            }
        }
        // the eagerly loaded geometry is taken only here, after every gather
        // of the group is issued (taking it in the loop above would wait for
        // those loads between the faces' gathers)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int f = f0 + g;
            if (!tk[f]) continue;
            if (EAGER) {
#pragma unroll
                for (int d = 0; d < ND; ++d) nf[g][d] = enf[EAGER ? f : 0][d];
                fv[g] = efv[EAGER ? f : 0];
                lam[g] = elam[EAGER ? f : 0];
            }
            double dfj[NV];
            ucfd::diff_flux<double, ND>(u[g], du[g], nf[g], dfj);
#pragma unroll
            for (int k = 0; k < NV; ++k) df[k] += (dfj[k] - lam[g]*du[g][k])*fv[g];
        }
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const size_t o = (size_t)k*N + i;
        if (DIR == 0) __builtin_nontemporal_store((xin[k] - 0.5*df[k])/dg, out + o);
        else          __builtin_nontemporal_store(xin[k] - (0.5*df[k]/dg), out + o);
    }
    if (RANGE && threadIdx.x == 0 && (sic != ib || slc != lce)) raise_flag(flag, UCFD_KFLAG_LAYOUT);
}

// A colour range where no face can be taken in range mode: the first colour
// of a lower sweep (n0 == 0: no neighbour id lies below it) and the last of
// an upper sweep (the range ends at N).  The face sum is empty, so the
// neighbour indices are not needed; the update keeps k_sweep's expressions
// with df = 0 (so -0.0 inputs and a zero diagonal give the same bits).
// CPT cells per thread (t, t + S, ...), every load issued before the first
// store.
// The workgroup's first cell samples the declared layout as in k_sweep; the
// neighbour indices of this range are neither read nor bounds checked (the
// sweep in the other direction reads and checks them).
template <int DIR, int NV, int CPT>
__global__ __launch_bounds__(kBlock) void k_sweep_none(int n0, int cnt, int N,
                                                       const double *in, double *out,
                                                       const double *__restrict__ diag,
                                                       const int *__restrict__ icolor,
                                                       const int *__restrict__ lcolor, int lce,
                                                       int *flag)
{
    const int S = gridDim.x*kBlock;
    const int t0 = blockIdx.x*kBlock + threadIdx.x;
    const int ib = n0 + blockIdx.x*kBlock;
    const int sic = icolor[ib], slc = lcolor[ib];
    const double df = 0.0;
    double dg[CPT], xin[CPT][NV];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int t = t0 + c*S;
        if (t >= cnt) continue;
        const int i = n0 + t;
        dg[c] = ld_stream(diag + i);
#pragma unroll
        for (int k = 0; k < NV; ++k) xin[c][k] = ld_stream(in + (size_t)k*N + i);
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int t = t0 + c*S;
        if (t >= cnt) continue;
        const int i = n0 + t;
        // plain (temporal) stores: the next launch of the pass gathers these
        // rows (the lower sweep's heavy colour reads dQ* of this one; the
        // upper sweep's reads dQ), and they are still cached then
        // (measured: 223 -> 218 us for that launch, 45 -> 42 us for this one)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const size_t o = (size_t)k*N + i;
            if (DIR == 0) out[o] = (xin[c][k] - 0.5*df)/dg[c];
            else          out[o] = xin[c][k] - (0.5*df/dg[c]);
        }
    }
    if (threadIdx.x == 0 && (sic != ib || slc != lce)) raise_flag(flag, UCFD_KFLAG_LAYOUT);
}

// Any other face count: one face at a time.
template <int DIR, int MODE, int ND>
__global__ __launch_bounds__(kBlock) void k_sweep_any(
    int n0, int cnt, int N, int nface, const int *__restrict__ nei,
    const int *__restrict__ order, const int *__restrict__ lcolor,
    const double *__restrict__ fnv, const double *__restrict__ vfn,
    const double *__restrict__ upts, const double *in, double *out,
    const double *__restrict__ diag, const double *__restrict__ fspr, int *flag, int lce)
{
    constexpr int NV = ND + 2;
    const int t = xcd_block()*kBlock + threadIdx.x;
    if (t >= cnt) return;
    if (MODE >= 2 && threadIdx.x == 0) {
        const int ib = n0 + xcd_block()*kBlock;
        if (order[ib] != ib || lcolor[ib] != lce) raise_flag(flag, UCFD_KFLAG_LAYOUT);
    }
    const int i = MODE >= 2 ? n0 + t : order[n0 + t];
    if ((unsigned)i >= (unsigned)N) {
        raise_flag(flag, UCFD_KFLAG_INDEX);
        return;
    }
    int ci = 0;
    if (MODE == 0) ci = lcolor[i];
    double df[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) df[k] = 0.0;
    for (int f = 0; f < nface; ++f) {
        const int nb = nei[(size_t)f*N + i];
        if ((unsigned)nb >= (unsigned)N) {
            raise_flag(flag, UCFD_KFLAG_INDEX);
            continue;
        }
        bool take;
        if (MODE == 0) {
            const int cn = lcolor[nb];
            take = DIR == 0 ? (cn < ci) : (cn > ci);
        } else if (MODE == 1) {
            take = DIR == 0 ? (nb < i) : (nb > i);
        } else {
            take = DIR == 0 ? (nb < n0) : (nb >= n0 + cnt);
        }
        if (!take) continue;
        double u[NV], du[NV], nf[ND], dfj[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            u[k] = upts[(size_t)k*N + nb];
            du[k] = out[(size_t)k*N + nb];
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) nf[d] = vfn[(size_t)(f*ND + d)*N + i];
        const double fv = fnv[(size_t)f*N + i];
        const double lam = fspr[(size_t)f*N + i];
        ucfd::diff_flux<double, ND>(u, du, nf, dfj);
#pragma unroll
        for (int k = 0; k < NV; ++k) df[k] += (dfj[k] - lam*du[k])*fv;
    }
    const double dg = diag[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const size_t o = (size_t)k*N + i;
        if (DIR == 0) out[o] = (in[o] - 0.5*df[k])/dg;
        else          out[o] = in[o] - (0.5*df[k]/dg);
    }
}

// Turbulence rows of the legacy RANS sweeps (rans_* / *_rans_* entry
// points): rows NFV .. NFV+NT-1 of upts / in / out (NFV = ND + 2), the
// turbulence diagonal tdiag[N] and spectral radius tfspr[F][N].  The face
// term follows src/lusgs/lusgs.c:274-321 with rans_flux_container
// (flux.c:77-90): theta = (m . n)/rho of the neighbour state (the momentum
// rows of u + du equal those of u, so one theta serves both fluxes),
// dF_t = (du_t + u_t) theta - u_t theta, df_t += (dF_t - tfspr du_t) fnorm_vol.
// Same neighbour tests, layout sampling and bounds checks as k_sweep; all
// faces' gathers are issued before the arithmetic.
template <int DIR, int MODE, int ND, int NT, int FC>
__global__ __launch_bounds__(kSweepBlock) void k_rans_sweep(
    int n0, int cnt, int N, const int *__restrict__ nei, const int *__restrict__ order,
    const int *__restrict__ lcolor, const double *__restrict__ fnv,
    const double *__restrict__ vfn, const double *__restrict__ upts, const double *in,
    double *out, const double *__restrict__ tdiag, const double *__restrict__ tfspr, int *flag,
    int lce)
{
    constexpr int NFV = ND + 2;
    const int t = xcd_block()*kSweepBlock + threadIdx.x;
    if (t >= cnt) return;
    int sic = 0, slc = 0;
    const int ib = n0 + xcd_block()*kSweepBlock;
    if (MODE == 2) { sic = order[ib]; slc = lcolor[ib]; }
    const int i = MODE == 2 ? n0 + t : ld_stream(order + n0 + t);
    if ((unsigned)i >= (unsigned)N) {
        raise_flag(flag, UCFD_KFLAG_INDEX);
        return;
    }
    const int ci = MODE == 0 ? lcolor[i] : 0;
    int nbv[FC];
    bool tk[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) nbv[f] = ld_stream(nei + (size_t)f*N + i);
    const double dg = ld_stream(tdiag + i);
    double xin[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) xin[k] = ld_stream(in + (size_t)(NFV + k)*N + i);
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        const int nb = nbv[f];
        if ((unsigned)nb >= (unsigned)N) {
            raise_flag(flag, UCFD_KFLAG_INDEX);
            tk[f] = false;
        } else if (MODE == 0) {
            const int cn = lcolor[nb];
            tk[f] = DIR == 0 ? (cn < ci) : (cn > ci);
        } else if (MODE == 1) {
            tk[f] = DIR == 0 ? (nb < i) : (nb > i);
        } else {
            tk[f] = DIR == 0 ? (nb < n0) : (nb >= n0 + cnt);
        }
    }
    double rho[FC], m[FC][ND], ut[FC][NT], dut[FC][NT], nf[FC][ND], fv[FC], lam[FC];
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        if (!tk[f]) continue;
        const int nb = nbv[f];
        rho[f] = upts[nb];
#pragma unroll
        for (int d = 0; d < ND; ++d) m[f][d] = upts[(size_t)(d + 1)*N + nb];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            ut[f][k] = upts[(size_t)(NFV + k)*N + nb];
            dut[f][k] = out[(size_t)(NFV + k)*N + nb];
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) nf[f][d] = ld_stream(vfn + (size_t)(f*ND + d)*N + i);
        fv[f] = ld_stream(fnv + (size_t)f*N + i);
        lam[f] = ld_stream(tfspr + (size_t)f*N + i);
    }
    double df[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) df[k] = 0.0;
#pragma unroll
    for (int f = 0; f < FC; ++f) {
        if (!tk[f]) continue;
        double th = 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) th += m[f][d]*nf[f][d];
        th /= rho[f];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const double dfj = (dut[f][k] + ut[f][k])*th - ut[f][k]*th;
            df[k] += (dfj - lam[f]*dut[f][k])*fv[f];
        }
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const size_t o = (size_t)(NFV + k)*N + i;
        if (DIR == 0) __builtin_nontemporal_store((xin[k] - 0.5*df[k])/dg, out + o);
        else          __builtin_nontemporal_store(xin[k] - (0.5*df[k]/dg), out + o);
    }
    if (MODE == 2 && threadIdx.x == 0 && (sic != ib || slc != lce))
        raise_flag(flag, UCFD_KFLAG_LAYOUT);
}

template <int DIR, int MODE, int ND, int NT>
int launch_rans_f(hipStream_t s, int n0, int cnt, int N, int nface, const int *nei,
                  const int *order, const int *lcolor, const double *fnv, const double *vfn,
                  const double *upts, const double *in, double *out, const double *tdiag,
                  const double *tfspr, int *flag, int lce)
{
    const dim3 g((cnt + kSweepBlock - 1)/kSweepBlock), b(kSweepBlock);
#define RL(FC) hipLaunchKernelGGL((k_rans_sweep<DIR, MODE, ND, NT, FC>), g, b, 0, s, n0, cnt, N, \
                                  nei, order, lcolor, fnv, vfn, upts, in, out, tdiag, tfspr, flag, lce)
    switch (nface) {
    case 3: RL(3); break;
    case 4: RL(4); break;
    case 6: RL(6); break;
    default: return UCFD_ERR_ARG;
    }
#undef RL
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

template <int DIR, int MODE>
int launch_rans_m(hipStream_t s, int ndims, int nt, int n0, int cnt, int N, int nface,
                  const int *nei, const int *order, const int *lcolor, const double *fnv,
                  const double *vfn, const double *upts, const double *in, double *out,
                  const double *tdiag, const double *tfspr, int *flag, int lce)
{
#define RD(ND, NT) return launch_rans_f<DIR, MODE, ND, NT>(s, n0, cnt, N, nface, nei, order, \
                          lcolor, fnv, vfn, upts, in, out, tdiag, tfspr, flag, lce)
    if (ndims == 3 && nt == 1) RD(3, 1);
    if (ndims == 3 && nt == 2) RD(3, 2);
    if (ndims == 2 && nt == 1) RD(2, 1);
    if (ndims == 2 && nt == 2) RD(2, 2);
#undef RD
    return UCFD_ERR_ARG;
}

// Update y += x over nrows rows of n cells ([nrows][n]).  One thread per
// cell, cells in descending order, every row of a cell issued together: the
// last lines written are the first cells' rows.  On a colour-contiguous
// numbering the next pass's lower sweep of the later colours gathers the
// first colour's solution rows, which are then the newest lines in the
// device's last-level cache (measured at 5.3M tets, 2 colours: that launch
// 219 -> 211 us, the pass 0.700 -> 0.685 ms against a flat ascending y += x).
__global__ __launch_bounds__(kBlock) void k_update_desc(int n, int nrows, double *__restrict__ y,
                                                        const double *__restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x*kBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t j = n - 1 - t;
    for (int r0 = 0; r0 < nrows; r0 += 8) {
        double xv[8], yv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (r0 + q >= nrows) break;
            xv[q] = x[(int64_t)(r0 + q)*n + j];
            yv[q] = y[(int64_t)(r0 + q)*n + j];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (r0 + q >= nrows) break;
            y[(int64_t)(r0 + q)*n + j] = yv[q] + xv[q];
        }
    }
}

template <int DIR, int MODE, int ND>
int launch_sweep_f(hipStream_t s, int n0, int cnt, int N, int nface, const int *nei,
                   const int *order, const int *lcolor, const double *fnv, const double *vfn,
                   const double *upts, const double *in, double *out, const double *diag,
                   const double *fspr, int *flag, int lce)
{
    const dim3 grid((cnt + kBlock - 1)/kBlock), block(kBlock);
    const dim3 sgrid((cnt + kSweepBlock - 1)/kSweepBlock), sblock(kSweepBlock);
    switch (nface) {
    case 4:
        hipLaunchKernelGGL((k_sweep<DIR, MODE, ND, 4>), sgrid, sblock, 0, s, n0, cnt, N, nface, nei,
                           order, lcolor, fnv, vfn, upts, in, out, diag, fspr, flag, lce);
        break;
    case 6:
        hipLaunchKernelGGL((k_sweep<DIR, MODE, ND, 6>), sgrid, sblock, 0, s, n0, cnt, N, nface, nei,
                           order, lcolor, fnv, vfn, upts, in, out, diag, fspr, flag, lce);
        break;
    default:
        hipLaunchKernelGGL((k_sweep_any<DIR, MODE, ND>), grid, block, 0, s, n0, cnt, N, nface,
                           nei, order, lcolor, fnv, vfn, upts, in, out, diag, fspr, flag, lce);
        break;
    }
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

template <int DIR, int MODE>
int launch_sweep_d(hipStream_t s, int ndims, int n0, int cnt, int N, int nface, const int *nei,
                   const int *order, const int *lcolor, const double *fnv, const double *vfn,
                   const double *upts, const double *in, double *out, const double *diag,
                   const double *fspr, int *flag, int lce)
{
    if (ndims == 3)
        return launch_sweep_f<DIR, MODE, 3>(s, n0, cnt, N, nface, nei, order, lcolor, fnv, vfn,
                                            upts, in, out, diag, fspr, flag, lce);
    return launch_sweep_f<DIR, MODE, 2>(s, n0, cnt, N, nface, nei, order, lcolor, fnv, vfn, upts,
                                        in, out, diag, fspr, flag, lce);
}

}  // namespace

int ucfd_launch_pre(hipStream_t s, int neles, int nface, double factor, const double *fnorm_vol,
                    const double *dt, double *diag, const double *fspr)
{
    if (neles <= 0) return UCFD_OK;
    // one cell per thread: 75 -> 70 us against two (cells t and t + S) at
    // 5.3M cells, the pass 0.687 -> 0.682 ms
    constexpr int CPT = 1;
    const dim3 g2((unsigned)((neles + CPT*kBlock - 1)/(CPT*kBlock)));
    if (nface == 4)
        hipLaunchKernelGGL((k_pre_f<4, CPT>), g2, dim3(kBlock), 0, s, neles, factor, fnorm_vol, dt,
                           diag, fspr);
    else if (nface == 6)
        hipLaunchKernelGGL((k_pre_f<6, CPT>), g2, dim3(kBlock), 0, s, neles, factor, fnorm_vol, dt,
                           diag, fspr);
    else
        hipLaunchKernelGGL(k_pre, dim3((neles + kBlock - 1)/kBlock), dim3(kBlock), 0, s, neles,
                           nface, factor, fnorm_vol, dt, diag, fspr);
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

int ucfd_launch_sweep(hipStream_t s, int dir, int mode, int n0, int ne, int neles, int nface,
                      int ndims, int nfvars, const int *nei, const int *order, const int *lcolor,
                      const double *fnorm_vol, const double *vec_fnorm, const double *upts,
                      const double *in, double *out, const double *diag, const double *fspr,
                      int *flag, int lc_expect)
{
    const int cnt = ne - n0;
    if (cnt <= 0) return UCFD_OK;
    (void)nfvars;
    if (dir == 0 && mode == UCFD_SWEEP_COLORED)
        return launch_sweep_d<0, 0>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    if (dir == 1 && mode == UCFD_SWEEP_COLORED)
        return launch_sweep_d<1, 0>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    // a range with no neighbour in the sweep direction: nothing to gather
    if (mode == UCFD_SWEEP_RANGE && (dir == 0 ? n0 == 0 : ne == neles)) {
        constexpr int CPT = 1;   // one cell per thread (within 1 us of two)
        const dim3 g((cnt + CPT*kBlock - 1)/(CPT*kBlock)), b(kBlock);
        if (dir == 0 && ndims == 3)
            hipLaunchKernelGGL((k_sweep_none<0, 5, CPT>), g, b, 0, s, n0, cnt, neles, in, out, diag,
                               order, lcolor, lc_expect, flag);
        else if (dir == 0)
            hipLaunchKernelGGL((k_sweep_none<0, 4, CPT>), g, b, 0, s, n0, cnt, neles, in, out, diag,
                               order, lcolor, lc_expect, flag);
        else if (ndims == 3)
            hipLaunchKernelGGL((k_sweep_none<1, 5, CPT>), g, b, 0, s, n0, cnt, neles, in, out, diag,
                               order, lcolor, lc_expect, flag);
        else
            hipLaunchKernelGGL((k_sweep_none<1, 4, CPT>), g, b, 0, s, n0, cnt, neles, in, out, diag,
                               order, lcolor, lc_expect, flag);
        return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
    }
    // a whole colour with every neighbour in the sweep direction: the last
    // colour of a lower sweep, the first of an upper sweep (tets and hexes)
    const bool all_taken = mode == UCFD_SWEEP_RANGE && (nface == 4 || nface == 6) &&
                           (dir == 0 ? ne == neles : n0 == 0);
    if (dir == 0 && all_taken)
        return launch_sweep_d<0, 3>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    if (dir == 1 && all_taken)
        return launch_sweep_d<1, 3>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    if (dir == 0 && mode == UCFD_SWEEP_RANGE)
        return launch_sweep_d<0, 2>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    if (dir == 1 && mode == UCFD_SWEEP_RANGE)
        return launch_sweep_d<1, 2>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    if (dir == 0)
        return launch_sweep_d<0, 1>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                    vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
    return launch_sweep_d<1, 1>(s, ndims, n0, cnt, neles, nface, nei, order, lcolor, fnorm_vol,
                                vec_fnorm, upts, in, out, diag, fspr, flag, lc_expect);
}

// Content hash of an int array: sum over k of splitmix64(k, a[k]) mod 2^64.
// The sum is order independent, so the block partials can be added with
// atomics and the result is still deterministic.
__device__ __forceinline__ unsigned long long hash_mix(unsigned long long k, int v)
{
    unsigned long long z = k*0x9E3779B97F4A7C15ull + (unsigned long long)(unsigned)v;
    z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27))*0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_hash_zero(unsigned long long *out)
{
    if (threadIdx.x == 0) out[0] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_hash(const int *__restrict__ a, int64_t n,
                                                 unsigned long long *out)
{
    unsigned long long s = 0ull;
    const int64_t st = (int64_t)gridDim.x*kBlock;
    for (int64_t k = (int64_t)blockIdx.x*kBlock + threadIdx.x; k < n; k += st)
        s += hash_mix((unsigned long long)k, a[k]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

int ucfd_launch_hash(hipStream_t s, const int *a, int64_t n, unsigned long long *out)
{
    hipLaunchKernelGGL(k_hash_zero, dim3(1), dim3(64), 0, s, out);
    if (n > 0) {
        int64_t blocks = (n + kBlock - 1)/kBlock;
        if (blocks > 2048) blocks = 2048;
        hipLaunchKernelGGL(k_hash, dim3((unsigned)blocks), dim3(kBlock), 0, s, a, n, out);
    }
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

int ucfd_launch_update(hipStream_t s, int neles, int nrows, double *y, const double *x)
{
    if (neles <= 0 || nrows <= 0) return UCFD_OK;
    const unsigned blocks = (unsigned)(((int64_t)neles + kBlock - 1)/kBlock);
    hipLaunchKernelGGL(k_update_desc, dim3(blocks), dim3(kBlock), 0, s, neles, nrows, y, x);
    return hipGetLastError() == hipSuccess ? UCFD_OK : UCFD_ERR_HIP;
}

int ucfd_launch_rans_sweep(hipStream_t s, int dir, int mode, int n0, int ne, int neles, int nface,
                           int ndims, int nt, const int *nei, const int *order, const int *lcolor,
                           const double *fnorm_vol, const double *vec_fnorm, const double *upts,
                           const double *in, double *out, const double *tdiag,
                           const double *tfspr, int *flag, int lc_expect)
{
    const int cnt = ne - n0;
    if (cnt <= 0) return UCFD_OK;
#define RM(D, M) return launch_rans_m<D, M>(s, ndims, nt, n0, cnt, neles, nface, nei, order, \
                                            lcolor, fnorm_vol, vec_fnorm, upts, in, out, tdiag, \
                                            tfspr, flag, lc_expect)
    if (mode == UCFD_SWEEP_COLORED) { if (dir == 0) RM(0, 0); RM(1, 0); }
    if (mode == UCFD_SWEEP_ORDERED) { if (dir == 0) RM(0, 1); RM(1, 1); }
    if (dir == 0) RM(0, 2);
    RM(1, 2);
#undef RM
}
