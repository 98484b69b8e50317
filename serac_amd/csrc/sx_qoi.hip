// sx_qoi.hip -- quantities of interest, Functional<double(trial)>
// (fn/functional_qoi.inl:59-477): the value  sum_e sum_q q(X, u, grad u) detJ w
// per element, reduced in a fixed order.  The gradient dq/dU is a residual of a
// derived qfunction and runs on the element kernels of sx_element.hpp.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sx_qfunctions.hpp"
#include "sx_qoi.hpp"

namespace sxk {

namespace {

__device__ __forceinline__ int64_t node_dof(int64_t node, int c, int nc, int64_t nnodes, int ordering)
{
  return ordering ? node * nc + c : int64_t(c) * nnodes + node;
}

// q at one quadrature point: X position, v value [NC], H physical gradient
// [NC][3], v2 the second argument's value (QOI_ID_INNER)
template <int QOI, int NC>
__device__ __forceinline__ double qoi_point(const double* p, const double* X, const double* v, const double* H,
                                           const double* v2)
{
  if constexpr (QOI == QOI_ID_INNER) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < NC; c++) s += v[c] * v2[c];
    return s;
  } else if constexpr (QOI == QOI_ID_MEASURE) {
    return 1.0;
  } else if constexpr (QOI == QOI_ID_POSITION) {
    const int d = int(p[0]);
    return X[d < 0 ? 0 : (d > 2 ? 2 : d)];
  } else if constexpr (QOI == QOI_ID_VALUE) {
    const int c = int(p[0]);
    return v[c < 0 ? 0 : (c >= NC ? NC - 1 : c)];
  } else {
    // strain energy of MaterialStressFunctor<NeoHookean> (its Cauchy stress
    // sigma = (lambda log J I + G (B - I)) / J is d psi / dF F^T / J for)
    //   psi = G/2 (tr B - 3) - G log J + lambda/2 (log J)^2,  lambda = K - 2G/3
    // with tr B - 3 = 2 tr H + |H|^2 (F = I + H).  p = {density, K, G}
    const double K = p[1], G = p[2], lambda = K - (2.0 / 3.0) * G;
    double       h2 = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) h2 += H[k] * H[k];
    const double trB3 = 2.0 * (H[0] + H[4] + H[8]) + h2;
    const double lnJ  = ::log1p(det_ap_i_m1(H));
    return 0.5 * G * trB3 - G * lnJ + 0.5 * lambda * lnJ * lnJ;
  }
}

// One element per NQ lanes (lane li <-> node / quadrature point (i2, i1, i0)),
// the sum-factorized interpolation of sx_element.hpp, q detJ w per point, and
// a fixed-order sum over the element's points into partial[slot].
template <int N, int NC, int QOI>
__global__ void __launch_bounds__(256) k_qoi(const QoiLaunch P)
{
  constexpr int NQ  = N * N * N;
  constexpr int EB  = 256 / NQ;
  constexpr int R   = NC * NQ;
  constexpr int PER = 6 * R;
  constexpr int TAB = 4 * N * N + N;

  __shared__ double smem[EB * PER];
  __shared__ double stab[TAB];
  __shared__ double sxv[EB * 24];
  __shared__ double sq[EB * NQ];

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;
  const int     li     = tid - el * NQ;
  const int     i0     = li % N;
  const int     i1     = (li / N) % N;
  const int     i2     = li / (N * N);
  const int64_t slot   = int64_t(blockIdx.x) * EB + el;
  const bool    active = (el < EB) && (slot < P.nslots);
  double*       S_     = smem + (el < EB ? el : 0) * PER;

  for (int k = tid; k < TAB; k += blockDim.x) stab[k] = P.tables[k];
  const double* tB  = stab;
  const double* tG  = stab + N * N;
  const double* tBw = stab + 2 * N * N;
  const double* txq = stab + 4 * N * N;

  if (active)
    for (int k = li; k < 24; k += NQ) sxv[el * 24 + k] = P.coords[size_t(slot) * 24 + k];
  // value and parent gradient at this lane's point of a field, by the
  // sum factorization (every lane calls it: it holds the barriers)
  auto interpolate = [&](const double* uin, double* val, double* gp) {
    if (active) {
      const int64_t node = P.map[size_t(slot) * NQ + li];
#pragma unroll
      for (int c = 0; c < NC; c++) S_[c * NQ + li] = uin[node_dof(node, c, NC, P.nnodes, P.ordering)];
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const double* uu = S_ + c * NQ + (i2 * N + i1) * N;
        double        s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int x = 0; x < N; x++) {
          s0 += tB[i0 * N + x] * uu[x];
          s1 += tG[i0 * N + x] * uu[x];
        }
        S_[1 * R + c * NQ + li] = s0;
        S_[2 * R + c * NQ + li] = s1;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const double* a10 = S_ + 1 * R + c * NQ;
        const double* a11 = S_ + 2 * R + c * NQ;
        double        s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int y = 0; y < N; y++) {
          const int    idx = (i2 * N + y) * N + i0;
          const double b = tB[i1 * N + y], g = tG[i1 * N + y];
          s0 += b * a10[idx];
          s1 += b * a11[idx];
          s2 += g * a10[idx];
        }
        S_[3 * R + c * NQ + li] = s0;
        S_[4 * R + c * NQ + li] = s1;
        S_[5 * R + c * NQ + li] = s2;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const double* a20 = S_ + 3 * R + c * NQ;
        const double* a21 = S_ + 4 * R + c * NQ;
        const double* a22 = S_ + 5 * R + c * NQ;
        double        v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
        for (int z = 0; z < N; z++) {
          const int    idx = (z * N + i1) * N + i0;
          const double b = tB[i2 * N + z], g = tG[i2 * N + z];
          v += b * a20[idx];
          g0 += b * a21[idx];
          g1 += b * a22[idx];
          g2 += g * a20[idx];
        }
        val[c]        = v;
        gp[c * 3 + 0] = g0;
        gp[c * 3 + 1] = g1;
        gp[c * 3 + 2] = g2;
      }
    }
    __syncthreads();  // the next field reuses the regions
  };
  double val[NC], gp[3 * NC], val2[NC] = {}, gp2[3 * NC];
  interpolate(P.u, val, gp);
  if constexpr (QOI == QOI_ID_INNER) interpolate(P.u2, val2, gp2);
  double contrib = 0.0;
  if (active) {
    // geometry: x(xi) = sum_c N_c(xi) X_c, J = dx/dxi, grad_x u = grad_xi u J^-1
    const double* Xc = sxv + el * 24;
    const double  x0 = txq[i0], x1 = txq[i1], x2 = txq[i2];
    double        X[3] = {0.0, 0.0, 0.0}, J[9] = {0.0}, Ji[9];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int    a = c & 1, b = (c >> 1) & 1, d = (c >> 2) & 1;
      const double N0 = a ? x0 : 1.0 - x0, N1 = b ? x1 : 1.0 - x1, N2 = d ? x2 : 1.0 - x2;
      const double D0 = a ? 1.0 : -1.0, D1 = b ? 1.0 : -1.0, D2 = d ? 1.0 : -1.0;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double xi = Xc[c * 3 + i];
        X[i] += N0 * N1 * N2 * xi;
        J[i * 3 + 0] += xi * D0 * N1 * N2;
        J[i * 3 + 1] += xi * N0 * D1 * N2;
        J[i * 3 + 2] += xi * N0 * N1 * D2;
      }
    }
    inv3(J, Ji);
    const double dJ = det3(J);
    double       H[3 * NC];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int j = 0; j < 3; j++)
        H[c * 3 + j] = gp[c * 3 + 0] * Ji[0 * 3 + j] + gp[c * 3 + 1] * Ji[1 * 3 + j] + gp[c * 3 + 2] * Ji[2 * 3 + j];
    // the quadrature weight: the rows of the weighted basis table sum to w_q
    double wx = 0.0, wy = 0.0, wz = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      wx += tBw[i0 * N + k];
      wy += tBw[i1 * N + k];
      wz += tBw[i2 * N + k];
    }
    contrib = qoi_point<QOI, NC>(P.params, X, val, H, val2) * dJ * (wx * wy * wz);
  }
  if (el < EB) sq[el * NQ + li] = contrib;
  __syncthreads();
  if (active && li == 0) {
    double s = 0.0;
    for (int k = 0; k < NQ; k++) s += sq[el * NQ + k];
    P.partial[slot] = s;
  }
}

// Boundary faces: one thread per face sums q dA w over its Q x Q points.  The
// face's nodes are fmap[slot][b n + a] (a along xi1), its bilinear geometry the
// corners [slot][a + 2 b][3]; dA = |dX/dxi1 x dX/dxi2|.
template <int N, int NC, int QOI>
__global__ void __launch_bounds__(256) k_qoi_face(const QoiLaunch P)
{
  const int64_t slot = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (slot >= P.nslots) return;
  const double* tB  = P.tables;
  const double* tBw = P.tables + 2 * N * N;
  const double* txq = P.tables + 4 * N * N;
  const double* C   = P.coords + size_t(slot) * 12;
  const int*    fm  = P.map + size_t(slot) * N * N;
  double        ue[NC][N * N];
#pragma unroll
  for (int k = 0; k < N * N; k++)
#pragma unroll
    for (int c = 0; c < NC; c++) ue[c][k] = P.u[node_dof(fm[k], c, NC, P.nnodes, P.ordering)];
  double sum = 0.0;
  for (int q2 = 0; q2 < N; q2++)
    for (int q1 = 0; q1 < N; q1++) {
      const double x1 = txq[q1], x2 = txq[q2];
      double       X[3], t1[3], t2[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const double c00 = C[0 * 3 + i], c10 = C[1 * 3 + i], c01 = C[2 * 3 + i], c11 = C[3 * 3 + i];
        X[i]  = (1 - x1) * (1 - x2) * c00 + x1 * (1 - x2) * c10 + (1 - x1) * x2 * c01 + x1 * x2 * c11;
        t1[i] = (1 - x2) * (c10 - c00) + x2 * (c11 - c01);
        t2[i] = (1 - x1) * (c01 - c00) + x1 * (c11 - c10);
      }
      const double n0 = t1[1] * t2[2] - t1[2] * t2[1], n1 = t1[2] * t2[0] - t1[0] * t2[2];
      const double n2 = t1[0] * t2[1] - t1[1] * t2[0];
      const double dA = ::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      double       w1 = 0.0, w2 = 0.0;
#pragma unroll
      for (int k = 0; k < N; k++) {
        w1 += tBw[q1 * N + k];
        w2 += tBw[q2 * N + k];
      }
      double v[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < N; b++)
#pragma unroll
          for (int a = 0; a < N; a++) acc += tB[q1 * N + a] * tB[q2 * N + b] * ue[c][b * N + a];
        v[c] = acc;
      }
      double q;
      if constexpr (QOI == QOI_ID_MEASURE) {
        q = 1.0;
      } else if constexpr (QOI == QOI_ID_POSITION) {
        const int d = int(P.params[0]);
        q           = X[d < 0 ? 0 : (d > 2 ? 2 : d)];
      } else {
        const int c = int(P.params[0]);
        q           = v[c < 0 ? 0 : (c >= NC ? NC - 1 : c)];
      }
      sum += q * dA * (w1 * w2);
    }
  P.partial[slot] = sum;
}

template <int N, int NC, int QOI>
hipError_t launch_face3(const QoiLaunch& P)
{
  if (P.nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_qoi_face<N, NC, QOI>), dim3(unsigned((P.nslots + 255) / 256)), dim3(256), 0, P.stream, P);
  return hipGetLastError();
}

template <int N, int NC>
hipError_t launch_face2(const QoiLaunch& P)
{
  switch (P.qoi) {
    case QOI_ID_MEASURE: return launch_face3<N, NC, QOI_ID_MEASURE>(P);
    case QOI_ID_POSITION: return launch_face3<N, NC, QOI_ID_POSITION>(P);
    case QOI_ID_VALUE: return launch_face3<N, NC, QOI_ID_VALUE>(P);
  }
  return hipErrorInvalidValue;
}

template <int N>
hipError_t launch_face1(const QoiLaunch& P)
{
  if (P.nc == 1) return launch_face2<N, 1>(P);
  if (P.nc == 3) return launch_face2<N, 3>(P);
  return hipErrorInvalidValue;
}

template <int N, int NC, int QOI>
hipError_t launch3(const QoiLaunch& P)
{
  constexpr int EB = 256 / (N * N * N);
  const int64_t blocks = (P.nslots + EB - 1) / EB;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_qoi<N, NC, QOI>), dim3(unsigned(blocks)), dim3(256), 0, P.stream, P);
  return hipGetLastError();
}

template <int N, int NC>
hipError_t launch2(const QoiLaunch& P)
{
  switch (P.qoi) {
    case QOI_ID_MEASURE: return launch3<N, NC, QOI_ID_MEASURE>(P);
    case QOI_ID_POSITION: return launch3<N, NC, QOI_ID_POSITION>(P);
    case QOI_ID_VALUE: return launch3<N, NC, QOI_ID_VALUE>(P);
    case QOI_ID_NEOHOOKEAN_ENERGY:
      if constexpr (NC == 3) return launch3<N, NC, QOI_ID_NEOHOOKEAN_ENERGY>(P);
      break;
    case QOI_ID_INNER: return launch3<N, NC, QOI_ID_INNER>(P);
  }
  return hipErrorInvalidValue;
}

template <int N>
hipError_t launch1(const QoiLaunch& P)
{
  if (P.nc == 1) return launch2<N, 1>(P);
  if (P.nc == 3) return launch2<N, 3>(P);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_qoi_face(const QoiLaunch& P)
{
  switch (P.order) {
    case 1: return launch_face1<2>(P);
    case 2: return launch_face1<3>(P);
    case 3: return launch_face1<4>(P);
    case 4: return launch_face1<5>(P);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_qoi(const QoiLaunch& P)
{
  switch (P.order) {
    case 1: return launch1<2>(P);
    case 2: return launch1<3>(P);
    case 3: return launch1<4>(P);
    case 4: return launch1<5>(P);
  }
  return hipErrorInvalidValue;
}

}  // namespace sxk
