// Per-cell arithmetic shared by every stencil kernel.  The association order
// is the reference's (diffusion_2D/diffusion.cpp:48-53): with FMA contraction
// off, any kernel built from these functions is bitwise identical to a serial
// CPU evaluation, whatever its tiling or launch shape.
#pragma once

#include <hip/hip_runtime.h>

namespace ceda {

// t = -((Dw+De)+(Ds+Dn))*uc + Dw*uw + De*ue + Ds*us + Dn*un ; f = 0 + t
__device__ inline double cell(double Dw, double De, double Ds, double Dn, double uc,
                              double uw, double ue, double us, double un)
{
  double t = -((Dw + De) + (Ds + Dn)) * uc;
  t = t + Dw * uw;
  t = t + De * ue;
  t = t + Ds * us;
  t = t + Dn * un;
  return 0.0 + t;
}

// manufactured forcing b = T1*sx*sy + T2*(cx*sy + cy*sx) (README.md:28)
__device__ inline double forcing(double T1, double T2, double sxi, double cxi, double syj,
                                 double cyj)
{
  return (T1 * sxi) * syj + T2 * (cxi * syj + cyj * sxi);
}

// one STS stage combination, N_VLinearCombination(5) order:
//   c0*F + c1*Y_{j-2} + c2*Y_n + c3*Y_{j-1} + c4*F_n
__device__ inline double lc5(const double* c, double F, double a, double b, double u, double g)
{
  double acc = c[0] * F;
  acc = acc + c[1] * a;
  acc = acc + c[2] * b;
  acc = acc + c[3] * u;
  acc = acc + c[4] * g;
  return acc;
}

typedef double d2v __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ inline double2 ldv2(const double* p)
{
  if (NT) {
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    return make_double2(v.x, v.y);
  }
  return *reinterpret_cast<const double2*>(p);
}
template <bool NT>
__device__ inline void stv2(double* p, double a, double b)
{
  if (NT) {
    d2v w;
    w.x = a;
    w.y = b;
    __builtin_nontemporal_store(w, reinterpret_cast<d2v*>(p));
  } else *reinterpret_cast<double2*>(p) = make_double2(a, b);
}
template <bool NT>
__device__ inline double lds1(const double* p)
{
  return NT ? __builtin_nontemporal_load(p) : *p;
}
template <bool NT>
__device__ inline void sts1(double* p, double v)
{
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

}  // namespace ceda
