// Euler normal flux and its variants, shared by every sweep kernel.
//
// ns_flux follows src/lusgs/flux.c:29-67 operation for operation (contravariant
// velocity, pressure with the PMIN floor that rewrites the energy, total
// enthalpy), so that the FP64 drop-in kernels reproduce the reference's
// rounding when compiled without FMA contraction.  NV = ND + 2 conservative
// variables (rho, rho*u_d, E).
#pragma once

#include <hip/hip_runtime.h>

#ifndef UCFD_GAMMA
#define UCFD_GAMMA 1.4
#endif
#ifndef UCFD_PMIN
#define UCFD_PMIN 1e-13
#endif

namespace ucfd {

template <typename T> struct Const;
template <> struct Const<double> {
    static constexpr double gamma = UCFD_GAMMA;
    static constexpr double pmin = UCFD_PMIN;
};
template <> struct Const<float> {
    static constexpr float gamma = (float)UCFD_GAMMA;
    static constexpr float pmin = (float)UCFD_PMIN;
};

// F(u; n).  u is taken by value: the reference rewrites u[NV-1] when the
// pressure is clamped, which only affects the flux itself.
template <typename T, int ND>
__host__ __device__ __forceinline__ void ns_flux(const T *uin, const T *nf, T *f)
{
    constexpr int NV = ND + 2;
    const T rho = uin[0];
    T et = uin[NV - 1];
    T temp = T(0);
    T contrav = T(0);
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        contrav += uin[i + 1]*nf[i];
        temp += uin[i + 1]*uin[i + 1];
    }
    contrav /= rho;
    T p = (Const<T>::gamma - T(1))*(et - T(0.5)*temp/rho);
    if (p < Const<T>::pmin) {
        p = Const<T>::pmin;
        et = p/(Const<T>::gamma - T(1)) + T(0.5)*temp/rho;
    }
    const T ht = et + p;
    f[0] = rho*contrav;
#pragma unroll
    for (int i = 0; i < ND; ++i) f[i + 1] = uin[i + 1]*contrav + nf[i]*p;
    f[NV - 1] = ht*contrav;
}

// out = A(u; n)^T v with A = dF/du the Jacobian of ns_flux (conservative
// variables).  Unclamped (p >= PMIN), with theta = (m.n)/rho, u = m/rho,
// phi = (gamma-1)|u|^2/2, H = (E+p)/rho:
//   out_rho = -theta (u.v_m) + phi (n.v_m) + theta (phi - H) v_E
//   out_m   = n v_rho + theta v_m + n (u.v_m) - (gamma-1) u (n.v_m)
//             + (H n - (gamma-1) theta u) v_E
//   out_E   = (gamma-1) (n.v_m) + gamma theta v_E
// Clamped (p = PMIN, E replaced by PMIN/(gamma-1) + |m|^2/(2 rho)): the
// pressure no longer depends on u and the flux on E:
//   out_rho = -theta (u.v_m) - theta (|u|^2/2 + H) v_E
//   out_m   = n v_rho + theta v_m + n (u.v_m) + (theta u + H n) v_E
//   out_E   = 0
template <typename T, int ND>
__host__ __device__ __forceinline__ void ns_flux_jt(const T *uin, const T *nf, const T *v, T *out)
{
    constexpr int NV = ND + 2;
    constexpr T g = Const<T>::gamma, gm1 = Const<T>::gamma - T(1);
    const T rho = uin[0], inv = T(1)/rho;
    T temp = T(0), contrav = T(0);
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        contrav += uin[i + 1]*nf[i];
        temp += uin[i + 1]*uin[i + 1];
    }
    const T theta = contrav/rho;
    T et = uin[NV - 1];
    T p = gm1*(et - T(0.5)*temp/rho);
    const bool clamped = p < Const<T>::pmin;
    if (clamped) {
        p = Const<T>::pmin;
        et = p/gm1 + T(0.5)*temp/rho;
    }
    const T H = (et + p)*inv;
    T vel[ND], um = T(0), nm = T(0), q2 = T(0);
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        vel[d] = uin[d + 1]*inv;
        um += vel[d]*v[d + 1];
        nm += nf[d]*v[d + 1];
        q2 += vel[d]*vel[d];
    }
    const T v0 = v[0], vE = v[NV - 1];
    if (!clamped) {
        const T phi = gm1*T(0.5)*q2;
        out[0] = -theta*um + phi*nm + theta*(phi - H)*vE;
#pragma unroll
        for (int k = 0; k < ND; ++k)
            out[k + 1] = nf[k]*v0 + theta*v[k + 1] + nf[k]*um - gm1*vel[k]*nm
                         + (H*nf[k] - gm1*theta*vel[k])*vE;
        out[NV - 1] = gm1*nm + g*theta*vE;
    } else {
        out[0] = -theta*um - theta*(T(0.5)*q2 + H)*vE;
#pragma unroll
        for (int k = 0; k < ND; ++k)
            out[k + 1] = nf[k]*v0 + theta*v[k + 1] + nf[k]*um + (theta*vel[k] + H*nf[k])*vE;
        out[NV - 1] = T(0);
    }
}

// dfj = F(u + du) - F(u)  (src/lusgs/lusgs.c:143-155)
template <typename T, int ND>
__host__ __device__ __forceinline__ void diff_flux(const T *u, const T *du, const T *nf, T *dfj)
{
    constexpr int NV = ND + 2;
    T up[NV], f0[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) up[k] = du[k] + u[k];
    ns_flux<T, ND>(u, nf, f0);
    ns_flux<T, ND>(up, nf, dfj);
#pragma unroll
    for (int k = 0; k < NV; ++k) dfj[k] -= f0[k];
}

}  // namespace ucfd
