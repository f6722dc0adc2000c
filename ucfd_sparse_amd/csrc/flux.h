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
