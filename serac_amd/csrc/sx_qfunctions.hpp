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
__device__ __forceinline__ d1 dsqrt(d1 a) { return d1(::sqrt(a.v), a.d / (2.0 * ::sqrt(a.v))); }
__device__ __forceinline__ double dsqrt(double a) { return ::sqrt(a); }
// the derivative of a root x(q) of f(x, q) = 0 with df/dq = 1: x' = -q' / df_dx
__device__ __forceinline__ double root_dual(double x, double, double) { return x; }
__device__ __forceinline__ d1 root_dual(double x, d1 q, double df_dx) { return d1(x, -q.d / df_dx); }
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
  static constexpr int  STATE      = 0;
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
  static constexpr int  STATE      = 0;
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
  static constexpr int  STATE      = 0;
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

// gradient of the quantity of interest  int u_c dV  (Functional<double(trial)>,
// fn/functional_qoi.inl): source e_c, no flux.  p = {c}
template <int C>
struct QFQoiValue {
  static constexpr int  NC         = C;
  static constexpr bool HAS_SOURCE = true;
  static constexpr bool USES_VALUE = false;
  static constexpr int  STATE      = 0;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const T*, const T*, T* src, T* flux)
  {
    const int c0 = int(p[0]);
#pragma unroll
    for (int c = 0; c < C; c++) {
      src[c] = T(c == c0 ? 1.0 : 0.0);
#pragma unroll
      for (int d = 0; d < 3; d++) flux[c * 3 + d] = T(0.0);
    }
  }
};

// small-strain linear elasticity: sigma = lambda tr(eps) I + 2 mu eps.  p = {lambda, mu}
struct QFLinearElastic {
  static constexpr int  NC         = 3;
  static constexpr bool HAS_SOURCE = false;
  static constexpr bool USES_VALUE = false;
  static constexpr int  STATE      = 0;
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

// flux = sigma . inv(F)^T . det(F), F = I + H (MaterialStressFunctor with
// geometric nonlinearity, phy/solid_mechanics.hpp:843-849)
template <typename T>
__device__ __forceinline__ void stress_to_flux_large(const T* sig, const T* H, T* flux)
{
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
      flux[i * 3 + j] =
          (sig[i * 3 + 0] * Fi[j * 3 + 0] + sig[i * 3 + 1] * Fi[j * 3 + 1] + sig[i * 3 + 2] * Fi[j * 3 + 2]) * dF;
}

// MaterialStressFunctor<J2SmallStrain<LinearHardening>>, quasi-static
// (phy/materials/solid_material.hpp:132-148, 199-258; phy/solid_mechanics.hpp:834-852).
// p = {E, nu, sigma_y, Hi, Hk, density, geom_nonlin}.  st = {plastic strain
// [3][3] row-major, accumulated plastic strain}: the caller's copy of the
// point's state, updated in place when the return map runs.
//
// The return map solves the consistency residual
//   f(x) = q - (3G + Hk) x - (sigma_y + Hi (eqps_old + x))
// with the bracketed Newton/bisection of solve_scalar_equation
// (fn/tuple_tensor_dual_functions.hpp:573-672) on values; the dual part of
// the root is -q' / df_dx.  A root that is not bracketed (the reference
// aborts there) returns NaN stresses.
struct QFJ2 {
  static constexpr int    NC         = 3;
  static constexpr bool   HAS_SOURCE = false;
  static constexpr bool   USES_VALUE = false;
  static constexpr int    STATE      = 10;
  static constexpr double TOL        = 1e-10;

  __device__ __forceinline__ static double solve(double q, double a, double sy, double Hi, double e0, double hi,
                                                 double* df_out)
  {
    const double dfdx = -a - Hi;
    const double lo   = 0.0;
    const double rtol = TOL * sy;
    auto         f    = [&](double x) { return q - a * x - (sy + Hi * (e0 + x)); };
    const double fl = f(lo), fh = f(hi);
    *df_out         = dfdx;
    if (fl * fh > 0) return __builtin_nan("");
    if (fl == 0) return lo;
    if (fh == 0) return hi;
    double xl = lo, xh = hi;
    if (fl > 0) {
      xl = hi;
      xh = lo;
    }
    double x = 0.0;  // initial guess, inside [lo, hi]
    double dx_old = fabs(hi - lo), dx = dx_old, fval = f(x);
    bool   conv = false;
    for (int it = 0; it < 25 && !conv; it++) {
      if ((x - xh) * dfdx - fval > 0 || (x - xl) * dfdx - fval < 0 || fabs(2. * fval) > fabs(dx_old * dfdx)) {
        dx_old = dx;
        dx     = 0.5 * (xh - xl);
        x      = xl + dx;
        conv   = (x == xl);
      } else {
        dx_old         = dx;
        dx             = fval / dfdx;
        const double t = x;
        x -= dx;
        conv = (x == t);
      }
      fval = f(x);
      conv = conv || (fabs(fval) < rtol);
      if (fval < 0)
        xl = x;
      else
        xh = x;
    }
    return x;
  }

  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const T*, const T* H, T*, T* flux, double* st)
  {
    const double E = p[0], nu = p[1], sy = p[2], Hi = p[3], Hk = p[4];
    const double K = E / (3.0 * (1.0 - 2.0 * nu));
    const double G = 0.5 * E / (1.0 + nu);
    T            s[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) s[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]) - st[i * 3 + j];
    const T tr = s[0] + s[4] + s[8];
    const T pr = K * tr;
#pragma unroll
    for (int i = 0; i < 3; i++) s[i * 4] = s[i * 4] - tr / 3.0;
#pragma unroll
    for (int k = 0; k < 9; k++) s[k] = (2.0 * G) * s[k];
    T eta[9];
    T sq = T(0.0);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      eta[k] = s[k] - ((2.0 / 3.0) * Hk) * st[k];
      sq     = sq + eta[k] * eta[k];
    }
    const T      q  = 1.2247448713915890491 * dsqrt(sq);  // sqrt(1.5)
    const double qv = vpart(q), e0 = st[9], a = 3.0 * G + Hk;
    if (qv - a * 0.0 - (sy + Hi * (e0 + 0.0)) > TOL * sy) {
      double       df_dx;
      const double x     = solve(qv, a, sy, Hi, e0, (qv - (sy + Hi * e0)) / a, &df_dx);
      const T      delta = root_dual(x, q, df_dx);
      const T      c     = (2.0 * G) * delta;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const T Np = (1.5 * eta[k]) / q;
        s[k]       = s[k] - c * Np;
        st[k] += x * vpart(Np);
      }
      st[9] += x;
    }
    s[0] = s[0] + pr;
    s[4] = s[4] + pr;
    s[8] = s[8] + pr;
    if (p[6] != 0.0) {
      stress_to_flux_large(s, H, flux);
    } else {
#pragma unroll
      for (int k = 0; k < 9; k++) flux[k] = s[k];
    }
  }
};

// placeholder for kernels that do not call a qfunction (J.v, diagonal)
template <int C>
struct QFNone {
  static constexpr int  NC         = C;
  static constexpr bool HAS_SOURCE = false;
  static constexpr bool USES_VALUE = false;
  static constexpr int  STATE      = 0;
  template <typename T>
  __device__ __forceinline__ static void eval(const double*, const T*, const T*, T*, T*)
  {
  }
};

}  // namespace sxk
