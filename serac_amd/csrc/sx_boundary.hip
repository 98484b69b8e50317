// sx_boundary.hip -- boundary (face) integrals on hexahedral meshes, the
// device path of AddBoundaryIntegral (fn/functional.hpp:328-354,
// fn/boundary_integral_kernels.hpp:106-205).
//
// One workgroup per boundary face.  A face is the trace of its element's H1
// space: Q x Q nodes with Q = p + 1, parametrized by (xi1, xi2) so that
// dX/dxi1 x dX/dxi2 points out of the element.  Per quadrature point the
// qfunction receives the reference position X, dX/dxi (3 x 2), the field value
// u and its face-parent derivative du/dxi (nc x 2), and returns a source that
// is multiplied by |dX/dxi1 x dX/dxi2| and integrated against the test
// function values (the reference's batch_apply_qf, :106-130).  Faces are
// launched one colour at a time (no two faces of a colour share a node), so
// the read-modify-write adds are race-free and reproducible.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sx_boundary.hpp"
#include "sx_qfunctions.hpp"

namespace sxk {

namespace {

// ---------------------------------------------------------------------------
// boundary qfunctions: physical (X, dX/dxi, u, du/dxi) -> source
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// setTraction (phy/solid_mechanics.hpp:1023-1037) with traction(X, n) = A n + b:
//   source = -(A n_hat + b),  n_hat = normalize(dX/dxi1 x dX/dxi2).
// p = {A (3 x 3 row-major), b (3)}
struct BQFTraction {
  static constexpr int NC = 3;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const double*, const double* dX, const T*, const T*,
                                              T* src)
  {
    const double t1[3] = {dX[0], dX[2], dX[4]}, t2[3] = {dX[1], dX[3], dX[5]};
    double       n[3];
    cross3(t1, t2, n);
    const double nn = ::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double nh[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
#pragma unroll
    for (int i = 0; i < 3; i++) src[i] = T(-1.0 * (p[i * 3 + 0] * nh[0] + p[i * 3 + 1] * nh[1] + p[i * 3 + 2] * nh[2] + p[9 + i]));
  }
};

// setPressure (phy/solid_mechanics.hpp:1071-1101): with geometric
// nonlinearity the normal follows the deformed surface x = X + u:
//   source = pressure * n / |dX/dxi1 x dX/dxi2|,  n = dx/dxi1 x dx/dxi2.
// p = {pressure, geom_nonlin (0/1)}
struct BQFPressure {
  static constexpr int NC = 3;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const double*, const double* dX, const T*, const T* du,
                                              T* src)
  {
    const bool g = p[1] != 0.0;
    T          t1[3], t2[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      t1[i] = g ? T(dX[i * 2 + 0]) + du[i * 2 + 0] : T(dX[i * 2 + 0]);
      t2[i] = g ? T(dX[i * 2 + 1]) + du[i * 2 + 1] : T(dX[i * 2 + 1]);
    }
    T n[3];
    cross3(t1, t2, n);
    const double r1[3] = {dX[0], dX[2], dX[4]}, r2[3] = {dX[1], dX[3], dX[5]};
    double       n0[3];
    cross3(r1, r2, n0);
    const double nref = ::sqrt(n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) src[i] = p[0] * (n[i] / nref);
  }
};

// setFluxBCs (phy/heat_transfer.hpp:536-552) with flux(X, n, t, T) =
// q0 + h (T - T_inf): a constant flux (h = 0) or a Robin condition
// (phy/tests/thermal_robin_condition.cpp).  p = {q0, h, T_inf}
struct BQFFlux {
  static constexpr int NC = 1;
  template <typename T>
  __device__ __forceinline__ static void eval(const double* p, const double*, const double*, const T* u, const T*,
                                              T* src)
  {
    src[0] = p[0] + p[1] * (u[0] - p[2]);
  }
};

__device__ __forceinline__ int64_t lidx_b(int64_t node, int c, int nc, int64_t nnodes, int ordering)
{
  return ordering ? node * nc + c : int64_t(c) * nnodes + node;
}

// bilinear face geometry at (x1, x2) from corners C[a + 2b][3]
__device__ __forceinline__ void face_geometry(const double* C, double x1, double x2, double* X, double* dX)
{
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double c00 = C[0 * 3 + i], c10 = C[1 * 3 + i], c01 = C[2 * 3 + i], c11 = C[3 * 3 + i];
    X[i]          = (1 - x1) * (1 - x2) * c00 + x1 * (1 - x2) * c10 + (1 - x1) * x2 * c01 + x1 * x2 * c11;
    dX[i * 2 + 0] = (1 - x2) * (c10 - c00) + x2 * (c11 - c01);
    dX[i * 2 + 1] = (1 - x1) * (c01 - c00) + x1 * (c11 - c10);
  }
}

// ---------------------------------------------------------------------------
// one workgroup (64 lanes) per face; lane k < N*N is quadrature point k in
// the qpoint stage and face node k in the integrate stage
// ---------------------------------------------------------------------------
template <int N, int NC, class BQF, int MODE, typename T>
__global__ void __launch_bounds__(64) k_face(const FaceLaunch P)
{
  constexpr int NF = N * N;
  constexpr int M  = 3 * NC;  // inputs: value (NC), then du/dxi (NC x 2)
  constexpr int DS = NC * M;  // stored tangent entries per point, [r][s]
  __shared__ double stab[4 * N * N + N];
  __shared__ T      su[NC * NF];
  __shared__ T      ssrc[NC * NF];
  __shared__ double sD[(MODE == FMODE_DIAG || MODE == FMODE_ASSEMBLE) ? DS * NF : 1];
  __shared__ double sp[3 * NF];  // face trace of the shape displacement

  const int     k    = threadIdx.x;
  const int64_t slot = P.slot_begin + blockIdx.x;
  for (int t = k; t < 4 * N * N + N; t += 64) stab[t] = P.tables[t];
  const int* fmap = P.fmap + slot * NF;
  if (k < NF && MODE != FMODE_DIAG && MODE != FMODE_ASSEMBLE) {
    const T* u = static_cast<const T*>(P.u);
#pragma unroll
    for (int c = 0; c < NC; c++) su[c * NF + k] = u[lidx_b(fmap[k], c, NC, P.nnodes, P.ordering)];
  }
  if constexpr (MODE == FMODE_DIAG || MODE == FMODE_ASSEMBLE) {
    for (int t = k; t < DS * NF; t += 64) sD[t] = P.D[(slot * DS + t / NF) * NF + t % NF];
  }
  if (P.shape != nullptr && k < NF) {
#pragma unroll
    for (int c = 0; c < 3; c++) sp[c * NF + k] = P.shape[lidx_b(fmap[k], c, 3, P.nnodes, P.ordering)];
  }
  __syncthreads();
  const double* tB  = stab;
  const double* tG  = stab + N * N;
  const double* tBw = stab + 2 * N * N;
  const double* txq = stab + 4 * N * N;

  if (MODE != FMODE_DIAG && MODE != FMODE_ASSEMBLE && k < NF) {
    const int q1 = k % N, q2 = k / N;
    T         in[M];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      T v = T(0), g1 = T(0), g2 = T(0);
#pragma unroll
      for (int b = 0; b < N; b++)
#pragma unroll
        for (int a = 0; a < N; a++) {
          const T uu = su[c * NF + b * N + a];
          v += T(tB[q1 * N + a] * tB[q2 * N + b]) * uu;
          g1 += T(tG[q1 * N + a] * tB[q2 * N + b]) * uu;
          g2 += T(tB[q1 * N + a] * tG[q2 * N + b]) * uu;
        }
      in[c]              = v;
      in[NC + c * 2 + 0] = g1;
      in[NC + c * 2 + 1] = g2;
    }
    T out[NC];
    if constexpr (MODE == FMODE_JVP) {
#pragma unroll
      for (int r = 0; r < NC; r++) {
        T acc = T(0);
#pragma unroll
        for (int s = 0; s < M; s++) acc += T(P.D[(slot * DS + r * M + s) * NF + k]) * in[s];
        out[r] = acc;
      }
    } else {
      double X[3], dX[6];
      face_geometry(P.corners + slot * 12, txq[q1], txq[q2], X, dX);
      if (P.shape != nullptr) {
        // the configuration X + p (fn/shape_aware_functional.hpp:150-170)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          double v = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
          for (int b = 0; b < N; b++)
#pragma unroll
            for (int a = 0; a < N; a++) {
              const double pp = sp[c * NF + b * N + a];
              v += tB[q1 * N + a] * tB[q2 * N + b] * pp;
              g1 += tG[q1 * N + a] * tB[q2 * N + b] * pp;
              g2 += tB[q1 * N + a] * tG[q2 * N + b] * pp;
            }
          X[c] += v;
          dX[c * 2 + 0] += g1;
          dX[c * 2 + 1] += g2;
        }
      }
      const double t1[3] = {dX[0], dX[2], dX[4]}, t2[3] = {dX[1], dX[3], dX[5]};
      double       n0[3];
      cross3(t1, t2, n0);
      const double scale = ::sqrt(n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]);
      double       vin[M];
#pragma unroll
      for (int s = 0; s < M; s++) vin[s] = double(in[s]);
      double src[NC];
      BQF::eval(P.params, X, dX, vin, vin + NC, src);
#pragma unroll
      for (int r = 0; r < NC; r++) out[r] = T(src[r] * scale);
      if constexpr (MODE == FMODE_LINEARIZE) {
        // forward-mode columns: d(source * scale) / d in_s
#pragma unroll
        for (int s = 0; s < M; s++) {
          d1 din[M];
#pragma unroll
          for (int t = 0; t < M; t++) din[t] = d1(vin[t], t == s ? 1.0 : 0.0);
          d1 ds[NC];
          BQF::eval(P.params, X, dX, din, din + NC, ds);
#pragma unroll
          for (int r = 0; r < NC; r++) P.D[(slot * DS + r * M + s) * NF + k] = (ds[r] * scale).d;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NC; r++) ssrc[r * NF + k] = out[r];
  }
  __syncthreads();
  if (k >= NF) return;
  const int a = k % N, b = k / N;  // this lane's face node
  if constexpr (MODE == FMODE_DIAG) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    for (int q = 0; q < NF; q++) {
      const int    q1 = q % N, q2 = q / N;
      const double tv = tBw[q1 * N + a] * tBw[q2 * N + b];
      const double uv = tB[q1 * N + a] * tB[q2 * N + b];
      const double u1 = tG[q1 * N + a] * tB[q2 * N + b], u2 = tB[q1 * N + a] * tG[q2 * N + b];
#pragma unroll
      for (int c = 0; c < NC; c++)
        acc[c] += tv * (sD[(c * M + c) * NF + q] * uv + sD[(c * M + NC + 2 * c) * NF + q] * u1 +
                        sD[(c * M + NC + 2 * c + 1) * NF + q] * u2);
    }
    double* r = static_cast<double*>(P.r);
#pragma unroll
    for (int c = 0; c < NC; c++) r[lidx_b(fmap[k], c, NC, P.nnodes, P.ordering)] += acc[c];
  } else if constexpr (MODE == FMODE_ASSEMBLE) {
#pragma unroll 1
    for (int j = 0; j < NF; j++) {
      const int ja = j % N, jb = j / N;
      double    acc[NC][NC];
#pragma unroll
      for (int c = 0; c < NC; c++)
#pragma unroll
        for (int c2 = 0; c2 < NC; c2++) acc[c][c2] = 0.0;
#pragma unroll 1
      for (int q = 0; q < NF; q++) {
        const int    q1 = q % N, q2 = q / N;
        const double tv = tBw[q1 * N + a] * tBw[q2 * N + b];
        const double uv = tB[q1 * N + ja] * tB[q2 * N + jb];
        const double u1 = tG[q1 * N + ja] * tB[q2 * N + jb], u2 = tB[q1 * N + ja] * tG[q2 * N + jb];
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
          for (int c2 = 0; c2 < NC; c2++)
            acc[c][c2] += tv * (sD[(c * M + c2) * NF + q] * uv + sD[(c * M + NC + 2 * c2) * NF + q] * u1 +
                                sD[(c * M + NC + 2 * c2 + 1) * NF + q] * u2);
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int64_t row = lidx_b(fmap[k], c, NC, P.nnodes, P.ordering);
#pragma unroll
        for (int c2 = 0; c2 < NC; c2++) {
          const int32_t colv = int32_t(lidx_b(fmap[j], c2, NC, P.nnodes, P.ordering));
          int64_t       lo = P.csr_row[row], hi = P.csr_row[row + 1];
          while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (P.csr_col[mid] <= colv)
              lo = mid;
            else
              hi = mid;
          }
          P.csr_val[lo] += acc[c][c2];
        }
      }
    }
  } else {
    T* r = static_cast<T*>(P.r);
#pragma unroll
    for (int c = 0; c < NC; c++) {
      T acc = T(0);
      for (int q = 0; q < NF; q++) {
        const int q1 = q % N, q2 = q / N;
        acc += T(tBw[q1 * N + a] * tBw[q2 * N + b]) * ssrc[c * NF + q];
      }
      r[lidx_b(fmap[k], c, NC, P.nnodes, P.ordering)] += acc;
    }
  }
}

template <int N, int NC, class BQF>
int launch_bqf(const FaceLaunch& P, hipError_t* err)
{
  const int64_t n = P.slot_end - P.slot_begin;
  if (n <= 0) return 1;
  const dim3 grid{static_cast<unsigned>(n)}, block{64u};
  switch (P.mode) {
    case FMODE_RESIDUAL: hipLaunchKernelGGL((k_face<N, NC, BQF, FMODE_RESIDUAL, double>), grid, block, 0, P.stream, P); break;
    case FMODE_LINEARIZE: hipLaunchKernelGGL((k_face<N, NC, BQF, FMODE_LINEARIZE, double>), grid, block, 0, P.stream, P); break;
    case FMODE_JVP:
      if (P.vprec == 1)
        hipLaunchKernelGGL((k_face<N, NC, BQF, FMODE_JVP, float>), grid, block, 0, P.stream, P);
      else
        hipLaunchKernelGGL((k_face<N, NC, BQF, FMODE_JVP, double>), grid, block, 0, P.stream, P);
      break;
    case FMODE_DIAG: hipLaunchKernelGGL((k_face<N, NC, BQF, FMODE_DIAG, double>), grid, block, 0, P.stream, P); break;
    case FMODE_ASSEMBLE: hipLaunchKernelGGL((k_face<N, NC, BQF, FMODE_ASSEMBLE, double>), grid, block, 0, P.stream, P); break;
    default: return 0;
  }
  *err = hipGetLastError();
  return 1;
}

template <int N>
int launch_order(const FaceLaunch& P, hipError_t* err)
{
  switch (P.qf) {
    case BQF_ID_TRACTION: return P.nc == 3 ? launch_bqf<N, 3, BQFTraction>(P, err) : 0;
    case BQF_ID_PRESSURE: return P.nc == 3 ? launch_bqf<N, 3, BQFPressure>(P, err) : 0;
    case BQF_ID_FLUX: return P.nc == 1 ? launch_bqf<N, 1, BQFFlux>(P, err) : 0;
  }
  return 0;
}

}  // namespace

int launch_face(const FaceLaunch& P, hipError_t* err)
{
  switch (P.order) {
    case 1: return launch_order<2>(P, err);
    case 2: return launch_order<3>(P, err);
    case 3: return launch_order<4>(P, err);
    case 4: return launch_order<5>(P, err);
  }
  return 0;
}

}  // namespace sxk
