// sx_qfunctions.hpp -- device quadrature-point functions (the compiled
// integrand registry) and the single-direction dual number they run on.
//
// Each functor evaluates physical-space (source, flux) from physical-space
// (value, grad) for a scalar type T in {double, sxk::d1}; forward-mode
// tangents come from running the same code on d1 with one seeded input per
// pass (fn/dual.hpp:28-147 propagates every gradient slot by this same
// arithmetic, so each pass yields one column of the reference's tangent).
#pragma once

#include <hip/hip_runtime.h>

namespace sxk {

// ---------------------------------------------------------------------------
// single-direction dual number: value v, directional derivative d
// ---------------------------------------------------------------------------
struct d1 {
  double v, d;
  __device__ __forceinline__ d1() : v(0.0), d(0.0) {}
  __device__ __forceinline__ d1(double x) : v(x), d(0.0) {}
  __device__ __forceinline__ d1(double x, double y) : v(x), d(y) {}
};
__device__ __forceinline__ d1 operator+(d1 a, d1 b) { return d1(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ d1 operator-(d1 a, d1 b) { return d1(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ d1 operator-(d1 a) { return d1(-a.v, -a.d); }
__device__ __forceinline__ d1 operator*(d1 a, d1 b) { return d1(a.v * b.v, a.v * b.d + a.d * b.v); }
__device__ __forceinline__ d1 operator*(double a, d1 b) { return d1(a * b.v, a * b.d); }
__device__ __forceinline__ d1 operator*(d1 b, double a) { return d1(a * b.v, a * b.d); }
__device__ __forceinline__ d1 operator+(d1 a, double b) { return d1(a.v + b, a.d); }
__device__ __forceinline__ d1 operator+(double b, d1 a) { return d1(a.v + b, a.d); }
__device__ __forceinline__ d1 operator-(d1 a, double b) { return d1(a.v - b, a.d); }
__device__ __forceinline__ d1 operator/(d1 a, d1 b)
{
  return d1(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
__device__ __forceinline__ d1 operator/(d1 a, double b) { return d1(a.v / b, a.d / b); }
__device__ __forceinline__ d1 operator/(double a, d1 b) { return d1(a / b.v, -(a / (b.v * b.v)) * b.d); }
__device__ __forceinline__ d1 dlog1p(d1 a) { return d1(::log1p(a.v), a.d / (1.0 + a.v)); }
__device__ __forceinline__ double dlog1p(double a) { return ::log1p(a); }
__device__ __forceinline__ double vpart(double a) { return a; }
__device__ __forceinline__ double vpart(d1 a) { return a.v; }
__device__ __forceinline__ double dpart(double) { return 0.0; }
__device__ __forceinline__ double dpart(d1 a) { return a.d; }

// ---------------------------------------------------------------------------
// 3x3 helpers on row-major arrays, restating fn/tensor.hpp:1242-1290 (det,
// detApIm1) and :1610-1630 (inv)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T det3(const T* A)
{
  return A[0] * A[4] * A[8] + A[1] * A[5] * A[6] + A[2] * A[3] * A[7] - A[0] * A[5] * A[7] - A[1] * A[3] * A[8] -
         A[2] * A[4] * A[6];
}

// det(A + I) - 1 = tr A + I2(A) + det A, evaluated without cancellation
template <typename T>
__device__ __forceinline__ T det_ap_i_m1(const T* A)
{
  return A[0] + A[4] + A[8] - A[1] * A[3] * (1.0 + A[8]) + A[0] * A[4] * (1.0 + A[8]) -
         A[2] * A[6] * (1.0 + A[4]) - A[5] * A[7] * (1.0 + A[0]) + A[0] * A[8] + A[4] * A[8] +
         A[1] * A[5] * A[6] + A[2] * A[3] * A[7];
}

template <typename T>
__device__ __forceinline__ void inv3(const T* A, T* R)
{
  T id = 1.0 / det3(A);
  R[0] = (A[4] * A[8] - A[5] * A[7]) * id;
  R[1] = (A[2] * A[7] - A[1] * A[8]) * id;
  R[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  R[3] = (A[5] * A[6] - A[3] * A[8]) * id;
  R[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  R[5] = (A[2] * A[3] - A[0] * A[5]) * id;
  R[6] = (A[3] * A[7] - A[4] * A[6]) * id;
  R[7] = (A[1] * A[6] - A[0] * A[7]) * id;
  R[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// ---------------------------------------------------------------------------
// qfunctions.  HAS_SOURCE: the source output is nonzero; USES_VALUE: the
// outputs depend on the field value (not only its gradient); SYMMETRIC: the
// tangent is major-symmetric, so half of it is stored.
// ---------------------------------------------------------------------------

// steady heat conduction: ThermalMaterialIntegrand<LinearIsotropicConductor>
// plus setSource (phy/heat_transfer.hpp:381-406, 493-503;
// phy/materials/thermal_material.hpp:50-55).  p = {k, s}
struct QFHeat {
  static constexpr int  NC         = 1;
  static constexpr bool HAS_SOURCE = true;
  static constexpr bool USES_VALUE = false;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const T*, const T* g, T* src, T* flux)
  {
    src[0]  = T(-p[1]);
    flux[0] = p[0] * g[0];
    flux[1] = p[0] * g[1];
    flux[2] = p[0] * g[2];
  }
};

// mass + diffusion {a u, b grad u} (fn/tests/functional_comparisons.cpp:32-33)
template <int C>
struct QFMassDiffusion {
  static constexpr int  NC         = C;
  static constexpr bool HAS_SOURCE = true;
  static constexpr bool USES_VALUE = true;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const T* v, const T* g, T* src, T* flux)
  {
#pragma unroll
    for (int c = 0; c < C; c++) {
      src[c] = p[0] * v[c];
#pragma unroll
      for (int d = 0; d < 3; d++) flux[c * 3 + d] = p[1] * g[c * 3 + d];
    }
  }
};

// MaterialStressFunctor<NeoHookean> with geometric nonlinearity
// (phy/solid_mechanics.hpp:834-852; phy/materials/solid_material.hpp:112-122):
//   sigma = (lambda log1p(J-1) I + G (H H^T + H^T + H)) / J,  lambda = K - 2G/3
//   flux  = sigma . inv(I + H)^T . det(I + H);   source = density * 0 (quasi-static)
// p = {density, K, G}
struct QFNeoHookean {
  static constexpr int  NC         = 3;
  static constexpr bool HAS_SOURCE = false;
  static constexpr bool USES_VALUE = false;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const T*, const T* H, T*, T* flux)
  {
    const double K = p[1], G = p[2];
    const double lambda = K - (2.0 / 3.0) * G;
    T            Jm1    = det_ap_i_m1(H);
    T            J      = Jm1 + 1.0;
    T            L      = lambda * dlog1p(Jm1);
    T            sig[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        T hht = H[i * 3 + 0] * H[j * 3 + 0] + H[i * 3 + 1] * H[j * 3 + 1] + H[i * 3 + 2] * H[j * 3 + 2];
        T v   = G * (hht + H[j * 3 + i] + H[i * 3 + j]);
        if (i == j) v = v + L;
        sig[i * 3 + j] = v / J;
      }
    T F[9];
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] = H[k];
    F[0] = F[0] + 1.0;
    F[4] = F[4] + 1.0;
    F[8] = F[8] + 1.0;
    T Fi[9];
    inv3(F, Fi);
    T dF = det3(F);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        flux[i * 3 + j] = (sig[i * 3 + 0] * Fi[j * 3 + 0] + sig[i * 3 + 1] * Fi[j * 3 + 1] +
                           sig[i * 3 + 2] * Fi[j * 3 + 2]) *
                          dF;
  }
};

// small-strain linear elasticity: sigma = lambda tr(eps) I + 2 mu eps.  p = {lambda, mu}
struct QFLinearElastic {
  static constexpr int  NC         = 3;
  static constexpr bool HAS_SOURCE = false;
  static constexpr bool USES_VALUE = false;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const T*, const T* g, T*, T* flux)
  {
    const double lam = p[0], mu = p[1];
    T            tr  = g[0] + g[4] + g[8];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        T v = mu * (g[i * 3 + j] + g[j * 3 + i]);
        if (i == j) v = v + lam * tr;
        flux[i * 3 + j] = v;
      }
  }
};

// placeholder for kernels that do not call a qfunction (J.v, diagonal)
template <int C>
struct QFNone {
  static constexpr int  NC         = C;
  static constexpr bool HAS_SOURCE = false;
  static constexpr bool USES_VALUE = false;
  template <typename T>
  __device__ __forceinline__ static void eval(const double*, const T*, const T*, T*, T*)
  {
  }
};

}  // namespace sxk
