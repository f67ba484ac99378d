// Device gas model and face fluxes of the BES solver core (FP64).
//
// Written for one thread evaluating one face: all state lives in registers,
// the dimension ND (2 or 3) is a compile-time constant so every loop unrolls.
// Arithmetic follows the reference formulas in their evaluation order
// (the library is compiled with -ffp-contract=off), so the FP64 results equal
// the reference's bit for bit:
//   gas relations        GasPerfect.tpp:21-291, Gas.tpp:29-87
//   Roe + Harten fix     FluxRoe.tpp:61-333
//   far field            FluxFreestream.tpp:34-226
//   slip wall            FluxSlipWall.tpp:32-53
// The scalar type T is double, or Dual (value + one directional derivative)
// whose rules are those of the reference's Jacobian machinery (Eigen 3.3.4
// AutoDiffScalar: product, quotient, sqrt, abs with sign(0)=+1, comparisons
// on values), so one thread per Jacobian column reproduces the reference
// Jacobian exactly.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_common.hpp"

namespace bes {

struct GasConst {
  double gamma, R, harten, cv, cp;
};

struct Dual {
  double v, d;
  __device__ __forceinline__ Dual() {}
  __device__ __forceinline__ Dual(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + b.d * a.v); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  return Dual(a.v / b.v, (a.d * b.v - b.d * a.v) * (1.0 / (b.v * b.v)));
}
__device__ __forceinline__ Dual operator+(Dual a, double s) { return Dual(a.v + s, a.d); }
__device__ __forceinline__ Dual operator+(double s, Dual a) { return Dual(s + a.v, a.d); }
__device__ __forceinline__ Dual operator-(Dual a, double s) { return Dual(a.v - s, a.d); }
__device__ __forceinline__ Dual operator-(double s, Dual a) { return Dual(s - a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, double s) { return Dual(a.v * s, a.d * s); }
__device__ __forceinline__ Dual operator*(double s, Dual a) { return Dual(a.v * s, a.d * s); }
__device__ __forceinline__ Dual operator/(Dual a, double s) { return Dual(a.v / s, a.d * (1.0 / s)); }
__device__ __forceinline__ Dual operator/(double s, Dual a) { return Dual(s / a.v, a.d * ((-s) / (a.v * a.v))); }
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
__device__ __forceinline__ Dual &operator/=(Dual &a, double s) { a = a / s; return a; }
__device__ __forceinline__ Dual dsqrt(Dual a) { double s = ::sqrt(a.v); return Dual(s, a.d * (0.5 / s)); }
__device__ __forceinline__ Dual dabs(Dual a) { return Dual(::fabs(a.v), a.d * (a.v < 0 ? -1.0 : 1.0)); }
__device__ __forceinline__ double val(Dual a) { return a.v; }
__device__ __forceinline__ double dsqrt(double a) { return ::sqrt(a); }
__device__ __forceinline__ double dabs(double a) { return ::fabs(a); }
__device__ __forceinline__ double val(double a) { return a; }

template <int ND, typename T>
struct Gas {
  static constexpr int NV = ND + 2, IE = ND + 1, IP = ND + 1, IT = ND + 2;
  const GasConst &c;
  __device__ __forceinline__ explicit Gas(const GasConst &g) : c(g) {}

  __device__ __forceinline__ T pressure(const T *U) const {
    T t = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) t += U[1 + d] * U[1 + d];
    return (c.gamma - 1.0) * (U[IE] - t / (2.0 * U[0]));
  }
  __device__ __forceinline__ T temperature(const T *U) const {
    T t = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) t += U[1 + d] * U[1 + d];
    return (1.0 / (c.cv * U[0])) * (U[IE] - t / (2.0 * U[0]));
  }
  __device__ __forceinline__ T velocity(const T *U, int d) const { return U[1 + d] / U[0]; }
  __device__ __forceinline__ void primitive(const T *U, T *P) const {
    P[0] = U[0];
#pragma unroll
    for (int d = 0; d < ND; d++) P[1 + d] = velocity(U, d);
    P[IP] = pressure(U);
    P[IT] = temperature(U);
  }
  __device__ __forceinline__ T velocity_squared(const T *U) const {
    T t = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) t += velocity(U, d) * velocity(U, d);
    return t;
  }
  __device__ __forceinline__ T total_enthalpy(const T *U) const { return U[IE] / U[0] + pressure(U) / U[0]; }
  __device__ __forceinline__ T sound_speed(const T *U) const {
    return dsqrt(c.gamma * (c.gamma - 1) * (total_enthalpy(U) - velocity_squared(U) / 2.0));
  }
  __device__ __forceinline__ T sound_speed_T(const T &temp) const { return dsqrt(c.gamma * c.R * temp); }
  __device__ __forceinline__ T contravariant_cons(const T *U, const double *n) const {
    T t = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) t += velocity(U, d) * n[d];
    return t;
  }
  __device__ __forceinline__ T contravariant_prim(const T *P, const double *n) const {
    T t = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) t += P[1 + d] * n[d];
    return t;
  }
  // mass, momentum and energy flux of a primitive state through unit normal n
  __device__ __forceinline__ void flux(const T *P, const double *n, T *F) const {
    T V = contravariant_prim(P, n);
    F[0] = P[0] * V;
#pragma unroll
    for (int d = 0; d < ND; d++) F[1 + d] = P[0] * P[1 + d] * V + P[IP] * n[d];
    T q2 = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) q2 += P[1 + d] * P[1 + d];
    F[IE] = P[0] * (c.cp * P[IT] + q2 / 2.0) * V;
  }
};

// Quotients by one divisor: for double the shared-reciprocal form (RcpDiv,
// bitwise the division), for Dual the division itself.
template <typename T>
struct Divider {
  T b;
  __device__ __forceinline__ explicit Divider(const T &b_) : b(b_) {}
  __device__ __forceinline__ T operator()(const T &a) const { return a / b; }
};
template <>
struct Divider<double> {
  RcpDiv r;
  __device__ __forceinline__ explicit Divider(double b_) : r(b_) {}
  __device__ __forceinline__ double operator()(double a) const { return r(a); }
};

// One side of the Roe flux: the primitive state, total enthalpy and normal
// velocity, with the five divisions by rho of Gas::primitive, total_enthalpy
// and contravariant_cons done by one Divider (the same values)
template <int ND, typename T>
__device__ __forceinline__ void roe_side(const Gas<ND, T> &g, const T *U, const double *n, T *P, T &H, T &V) {
  constexpr int IE = ND + 1, IP = ND + 1, IT = ND + 2;
  const Divider<T> by_rho(U[0]);
  P[0] = U[0];
#pragma unroll
  for (int d = 0; d < ND; d++) P[1 + d] = by_rho(U[1 + d]);
  P[IP] = g.pressure(U);
  P[IT] = g.temperature(U);
  H = by_rho(U[IE]) + by_rho(P[IP]);
  V = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) V += P[1 + d] * n[d];
}

// Harten entropy fix on an eigenvalue magnitude
template <typename T>
__device__ __forceinline__ T harten_fix(const T &lam, const T &delta) {
  if (val(lam) > val(delta)) return lam;
  return (lam * lam + delta * delta) / (2.0 * delta);
}

// Roe flux per unit area, left/right conserved states, unit normal left->right.
template <int ND, typename T>
__device__ __forceinline__ void roe_flux(const GasConst &gc, const T *UL, const T *UR,
                                         const double *n, T *F) {
  constexpr int NV = ND + 2, IP = ND + 1;
  Gas<ND, T> g(gc);
  // extended primitive states: [rho, u_d, p, T] + H, V (c and q^2 of the
  // sides enter only through the deltas, which the dissipation never uses)
  T PL[ND + 3], PR[ND + 3], HL, HR, VL, VR;
  roe_side<ND, T>(g, UL, n, PL, HL, VL);
  roe_side<ND, T>(g, UR, n, PR, HR, VR);
  T dRho = PR[0] - PL[0], dP = PR[IP] - PL[IP], dV = VR - VL;
  T du[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) du[d] = PR[1 + d] - PL[1 + d];
  // Roe averages
  T sl = dsqrt(PL[0]), sr = dsqrt(PR[0]);
  T den = sl + sr;
  const Divider<T> by_den(den);
  T aRho = dsqrt(PL[0] * PR[0]);
  T au[ND];
  T aQ = 0.0, aV = 0.0;
#pragma unroll
  for (int d = 0; d < ND; d++) {
    au[d] = by_den(PL[1 + d] * dsqrt(PL[0]) + PR[1 + d] * dsqrt(PR[0]));
    aQ += au[d] * au[d];
    aV += au[d] * n[d];
  }
  T aH = by_den(HL * dsqrt(PL[0]) + HR * dsqrt(PR[0]));
  T aC = dsqrt((T(gc.gamma) - 1.0) * (aH - aQ / 2.0));
  // central part
  T tmp[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) F[i] = 0.0;
  g.flux(PL, n, tmp);
#pragma unroll
  for (int i = 0; i < NV; i++) F[i] += tmp[i];
  g.flux(PR, n, tmp);
#pragma unroll
  for (int i = 0; i < NV; i++) F[i] += tmp[i];
  T delta = dabs(aC * gc.harten);
  const Divider<T> by_2c2(2.0 * aC * aC);
  {  // acoustic wave V - c
    T cc = by_2c2(harten_fix(dabs(aV - aC), delta) * (dP - aRho * aC * dV));
    F[0] -= cc;
#pragma unroll
    for (int d = 0; d < ND; d++) F[1 + d] -= cc * (au[d] - aC * n[d]);
    F[NV - 1] -= cc * (aH - aC * aV);
  }
  {  // entropy and shear waves V
    T lam = harten_fix(dabs(aV), delta);
    T c1 = lam * (dRho - dP / (aC * aC));
    T c2 = lam * aRho;
    T t = 0.0;
    F[0] -= c1;
#pragma unroll
    for (int d = 0; d < ND; d++) {
      F[1 + d] -= (c1 * au[d] + c2 * (du[d] - dV * n[d]));
      t += du[d] * au[d];
    }
    F[NV - 1] -= (c1 * aQ / 2.0 + c2 * (t - dV * aV));
  }
  {  // acoustic wave V + c
    T cc = by_2c2(harten_fix(dabs(aV + aC), delta) * (dP + aRho * aC * dV));
    F[0] -= cc;
#pragma unroll
    for (int d = 0; d < ND; d++) F[1 + d] -= cc * (au[d] + aC * n[d]);
    F[NV - 1] -= cc * (aH + aC * aV);
  }
#pragma unroll
  for (int i = 0; i < NV; i++) F[i] /= 2.0;
}

template <int ND, typename T>
__device__ __forceinline__ void slip_flux(const GasConst &gc, const T *UL, const double *n, T *F) {
  Gas<ND, T> g(gc);
#pragma unroll
  for (int i = 0; i < ND + 2; i++) F[i] = 0.0;
  T p = g.pressure(UL);
#pragma unroll
  for (int d = 0; d < ND; d++) F[1 + d] = p * n[d];
}

// Characteristic far field; Pinf = freestream primitive [rho, u_d, p, T].
template <int ND, typename T>
__device__ __forceinline__ void freestream_flux(const GasConst &gc, const T *UL, const T *Pinf,
                                                const double *n, T *F) {
  constexpr int IP = ND + 1, IT = ND + 2;
  Gas<ND, T> g(gc);
  T PL[ND + 3], PB[ND + 3];
  g.primitive(UL, PL);
  T VL = g.contravariant_prim(PL, n);
  T ML = dabs(VL) / g.sound_speed_T(PL[IT]);
  const double m = val(ML), v = val(VL);
  if (m >= 1 && v <= 0) {
    g.flux(Pinf, n, F);
  } else if (m >= 1 && v > 0) {
    g.flux(PL, n, F);
  } else if (m < 1 && v <= 0) {
    T cL = g.sound_speed_T(PL[IT]);
    T t = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) t += n[d] * (Pinf[1 + d] - PL[1 + d]);
    T c1 = PL[0] * cL;
    PB[IP] = (Pinf[IP] + PL[IP] - c1 * t) / 2.0;
    T c2 = (Pinf[IP] - PB[IP]) / c1;
    PB[0] = Pinf[0] + (PB[IP] - Pinf[IP]) / (cL * cL);
    PB[IT] = PB[IP] / (gc.R * PB[0]);
#pragma unroll
    for (int d = 0; d < ND; d++) PB[1 + d] = Pinf[1 + d] - c2 * n[d];
    g.flux(PB, n, F);
  } else if (m < 1 && v > 0) {
    T cL = g.sound_speed_T(PL[IT]);
    T c1 = PL[0] * cL;
    PB[IP] = Pinf[IP];
    T c2 = (PL[IP] - PB[IP]) / c1;
    PB[0] = PL[0] + (PB[IP] - PL[IP]) / (cL * cL);
    PB[IT] = PB[IP] / (gc.R * PB[0]);
#pragma unroll
    for (int d = 0; d < ND; d++) PB[1 + d] = PL[1 + d] + c2 * n[d];
    g.flux(PB, n, F);
  } else {
#pragma unroll
    for (int i = 0; i < ND + 2; i++) F[i] = __builtin_nan("");
  }
}

}  // namespace bes
