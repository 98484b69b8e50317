// sx_element.hpp -- fused element kernels for H1 hexahedra on gfx950.
//
// One kernel launch processes one colour of elements (no two elements of a
// colour share a node), so the scatter-add into the L-vector is a plain
// read-modify-write: deterministic, no floating-point atomics.
//
// Per element the pipeline of fn/domain_integral_kernels.hpp:155-304 runs in
// one workgroup-resident pass:
//   gather L -> LDS                           (fn/element_restriction.cpp:443-454)
//   sum-factorized interpolate (3 stages)     (fn/detail/hexahedron_H1.inl:169-228)
//   quadrature point: qfunction + parent/physical maps (+ dual tangents)
//                     or stored-tangent chain rule for J.v
//   sum-factorized integrate (3 stages)       (fn/detail/hexahedron_H1.inl:230-280)
//   scatter-add LDS -> L                      (fn/element_restriction.cpp:456-467)
//
// Thread mapping: the workgroup has 256 lanes (4 wave64s) and holds EB =
// floor(256 / Q^3) elements; lane li of element el owns the index triple
// (i2, i1, i0) = li in every stage (node (z,y,x), partial sums, or quadrature
// point (qz,qy,qx)), which is possible because every form here uses
// Q = p + 1 = nodes per direction (fn/functional.hpp:193).  LDS per element is
// (6 + has_source) * nc * Q^3 doubles, reused across stages.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sx_launch.hpp"
#include "sx_qfunctions.hpp"

namespace sxk {

template <int M>
__host__ __device__ constexpr int packed_index(int r, int s)
{
  // upper triangle (r <= s), row-major
  return r * M - (r * (r - 1)) / 2 + (s - r);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// With remap on, XCD x instead works through a contiguous run of logical
// blocks, so neighbouring elements (which share nodes) hit the same L2.
__device__ __forceinline__ int64_t logical_block(int remap)
{
  const int64_t b = blockIdx.x, G = gridDim.x;
  if (!remap) return b;
  const int64_t per = G / 8;
  if (b >= per * 8) return b;
  return (b % 8) * per + b / 8;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also fences global
// memory, which makes every wave drain its outstanding global loads (vmcnt 0)
// at each barrier; the stages here exchange data through LDS alone, so the
// stored-tangent stream may stay in flight across the barriers.
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ void scatter_add(double* r, int64_t o, double v, int atomic)
{
  if (atomic)
    unsafeAtomicAdd(r + o, v);
  else
    r[o] += v;
}

__device__ __forceinline__ int64_t lidx(int64_t node, int c, int nc, int64_t nnodes, int ordering)
{
  return ordering ? node * nc + c : int64_t(c) * nnodes + node;
}

// Jacobian of the trilinear map at parent point xi from the 8 corner
// coordinates X[corner][3] (corner = a + 2b + 4c lexicographic):
// J[i][j] = d x_i / d xi_j  (fn/geometric_factors.cpp:17-58 interpolates the
// same order-1 nodes).
__device__ __forceinline__ void trilinear_jacobian(const double* X, double x0, double x1, double x2, double* J)
{
#pragma unroll
  for (int k = 0; k < 9; k++) J[k] = 0.0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int    a = c & 1, b = (c >> 1) & 1, d = (c >> 2) & 1;
    const double N0 = a ? x0 : 1.0 - x0, N1 = b ? x1 : 1.0 - x1, N2 = d ? x2 : 1.0 - x2;
    const double D0 = a ? 1.0 : -1.0, D1 = b ? 1.0 : -1.0, D2 = d ? 1.0 : -1.0;
    const double g0 = D0 * N1 * N2, g1 = N0 * D1 * N2, g2 = N0 * N1 * D2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double xi = X[c * 3 + i];
      J[i * 3 + 0] += xi * g0;
      J[i * 3 + 1] += xi * g1;
      J[i * 3 + 2] += xi * g2;
    }
  }
}

// Quadrature-point stage of the residual / linearization: parent-space
// (value, grad) in, parent-space (source, flux) out, following
// fn/finite_element.hpp:246-322 (grad_x = grad_xi inv(J);
// source *= detJ; flux = flux inv(J)^T... * detJ).
template <class QF, int NC, bool FULL>
__device__ __forceinline__ void qpoint_eval(const double* prm, const double* J, const double* val,
                                            const double* gp, double* src, double* fp)
{
  double Ji[9];
  inv3(J, Ji);
  const double dJ = det3(J);
  double       g[3 * NC], s[NC], f[3 * NC];
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
      g[c * 3 + k] = gp[c * 3 + 0] * Ji[0 * 3 + k] + gp[c * 3 + 1] * Ji[1 * 3 + k] + gp[c * 3 + 2] * Ji[2 * 3 + k];
#pragma unroll
  for (int c = 0; c < NC; c++) s[c] = 0.0;
  QF::eval(prm, val, g, s, f);
#pragma unroll
  for (int c = 0; c < NC; c++) {
    src[c] = s[c] * dJ;
#pragma unroll
    for (int j = 0; j < 3; j++)
      fp[c * 3 + j] = (f[c * 3 + 0] * Ji[j * 3 + 0] + f[c * 3 + 1] * Ji[j * 3 + 1] + f[c * 3 + 2] * Ji[j * 3 + 2]) * dJ;
  }
}

// One forward-mode pass seeded at input s (the inputs are [value(NC)] then
// [parent grad (NC x 3)] when FULL, else only the parent grad); returns the
// derivative of [source(NC)] [parent flux (NC x 3)] (or only the flux).
template <class QF, int NC, bool FULL>
__device__ __forceinline__ void qpoint_tangent_column(const double* prm, const double* Ji, double dJ,
                                                      const double* val, const double* gp, int s, double* col)
{
  constexpr int VOFF = FULL ? NC : 0;
  d1            v[NC], g[3 * NC], sr[NC], f[3 * NC];
#pragma unroll
  for (int c = 0; c < NC; c++) v[c] = d1(val[c], (FULL && s == c) ? 1.0 : 0.0);
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double gv = gp[c * 3 + 0] * Ji[0 * 3 + k] + gp[c * 3 + 1] * Ji[1 * 3 + k] + gp[c * 3 + 2] * Ji[2 * 3 + k];
      // seeded parent-gradient entry (c, j) contributes Ji[j][k] to d grad_x(c, k)
      double gd = 0.0;
#pragma unroll
      for (int j = 0; j < 3; j++)
        if (s == VOFF + c * 3 + j) gd = Ji[j * 3 + k];
      g[c * 3 + k] = d1(gv, gd);
    }
#pragma unroll
  for (int c = 0; c < NC; c++) sr[c] = d1(0.0);
  QF::eval(prm, v, g, sr, f);
  if constexpr (FULL) {
#pragma unroll
    for (int c = 0; c < NC; c++) col[c] = sr[c].d * dJ;
  }
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      col[VOFF + c * 3 + j] =
          (f[c * 3 + 0].d * Ji[j * 3 + 0] + f[c * 3 + 1].d * Ji[j * 3 + 1] + f[c * 3 + 2].d * Ji[j * 3 + 2]) * dJ;
}

template <int N, int NC, int MODE, class QF, int LAYOUT>
__global__ void __launch_bounds__(256) k_element(const ElemLaunch P)
{
  constexpr int  NQ       = N * N * N;
  constexpr int  EB       = 256 / NQ;
  constexpr bool FULL     = (LAYOUT == LAYOUT_FULL_SYM);
  constexpr bool JVP      = (MODE == MODE_JVP);
  constexpr bool HAS_SRC  = JVP ? FULL : QF::HAS_SOURCE;
  constexpr bool NEED_VAL = JVP ? FULL : QF::USES_VALUE;
  constexpr int  M        = FULL ? 4 * NC : 3 * NC;
  constexpr int  DS       = M * (M + 1) / 2;
  constexpr int  R        = NC * NQ;  // one LDS region
  constexpr int  PER      = (6 + (HAS_SRC ? 1 : 0)) * R;
  constexpr int  TAB      = 4 * N * N + N;

  __shared__ double smem[EB * PER];
  __shared__ double stab[TAB];
  __shared__ double sxv[JVP ? 1 : EB * 24];

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;
  const int     li     = tid - el * NQ;
  const int     i0     = li % N;
  const int     i1     = (li / N) % N;
  const int     i2     = li / (N * N);
  const int64_t slot   = P.slot_begin + logical_block(P.xcd_remap) * EB + el;
  const bool    active = (el < EB) && (slot < P.slot_end);
  double*       S      = smem + (el < EB ? el : 0) * PER;

  for (int k = tid; k < TAB; k += blockDim.x) stab[k] = P.tables[k];
  const double* tB  = stab;
  const double* tG  = stab + N * N;
  const double* tBw = stab + 2 * N * N;
  const double* tGw = stab + 3 * N * N;
  const double* txq = stab + 4 * N * N;

  // ---- gather: the nodal loads are issued first ...
  int64_t node = 0;
  double  ug[NC];
  if (active) {
    node = P.map[size_t(slot) * NQ + li];
#pragma unroll
    for (int c = 0; c < NC; c++) ug[c] = P.u[lidx(node, c, NC, P.nnodes, P.ordering)];
  }
  // ... then the stored tangent for J.v.  Memory counters retire in order, so
  // waiting for the nodal values does not wait for the tangent stream, which
  // stays in flight through the interpolation stages.
  double Dr[JVP ? DS : 1];
  if constexpr (JVP) {
    if (active) {
      typedef double dbl2 __attribute__((ext_vector_type(2)));
      const double* Dp = P.D + size_t(slot) * tangent_slot_stride(DS, NQ);
#pragma unroll
      for (int k = 0; k + 1 < DS; k += 2) {
        const dbl2 v = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(Dp + tangent_offset(k, li, DS, NQ)));
        Dr[k]        = v.x;
        Dr[k + 1]    = v.y;
      }
      if constexpr (DS & 1) Dr[DS - 1] = __builtin_nontemporal_load(Dp + tangent_offset(DS - 1, li, DS, NQ));
    }
  }
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) S[c * NQ + li] = ug[c];
    if constexpr (!JVP) {
      for (int k = li; k < 24; k += NQ) sxv[el * 24 + k] = P.coords[size_t(slot) * 24 + k];
    }
  }
  lds_barrier();

  // ---- interpolate, stage x: (z, y, qx)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* u  = S + c * NQ + (i2 * N + i1) * N;
      double        s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int x = 0; x < N; x++) {
        s0 += tB[i0 * N + x] * u[x];
        s1 += tG[i0 * N + x] * u[x];
      }
      S[1 * R + c * NQ + li] = s0;
      S[2 * R + c * NQ + li] = s1;
    }
  }
  lds_barrier();

  // ---- stage y: (z, qy, qx)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* a10 = S + 1 * R + c * NQ;
      const double* a11 = S + 2 * R + c * NQ;
      double        s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int y = 0; y < N; y++) {
        const int    idx = (i2 * N + y) * N + i0;
        const double b = tB[i1 * N + y], g = tG[i1 * N + y];
        s0 += b * a10[idx];
        s1 += b * a11[idx];
        s2 += g * a10[idx];
      }
      S[3 * R + c * NQ + li] = s0;
      S[4 * R + c * NQ + li] = s1;
      S[5 * R + c * NQ + li] = s2;
    }
  }
  lds_barrier();

  // ---- stage z + quadrature point: (qz, qy, qx)
  if (active) {
    double val[NC], gp[3 * NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* a20 = S + 3 * R + c * NQ;
      const double* a21 = S + 4 * R + c * NQ;
      const double* a22 = S + 5 * R + c * NQ;
      double        v = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
#pragma unroll
      for (int z = 0; z < N; z++) {
        const int    idx = (z * N + i1) * N + i0;
        const double b = tB[i2 * N + z], g = tG[i2 * N + z];
        if constexpr (NEED_VAL) v += b * a20[idx];
        g0 += b * a21[idx];
        g1 += b * a22[idx];
        g2 += g * a20[idx];
      }
      val[c]        = v;
      gp[c * 3 + 0] = g0;
      gp[c * 3 + 1] = g1;
      gp[c * 3 + 2] = g2;
    }

    double src[NC], fp[3 * NC];
    if constexpr (JVP) {
      // chain rule with the stored symmetric tangent (fn/tensor.hpp:1791-1869)
      double in[M], out[M];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        if constexpr (FULL) in[c] = val[c];
#pragma unroll
        for (int d = 0; d < 3; d++) in[(FULL ? NC : 0) + c * 3 + d] = gp[c * 3 + d];
      }
#pragma unroll
      for (int r = 0; r < M; r++) out[r] = 0.0;
#pragma unroll
      for (int r = 0; r < M; r++)
#pragma unroll
        for (int s = r; s < M; s++) {
          const double Dk = Dr[packed_index<M>(r, s)];
          out[r] += Dk * in[s];
          if (s != r) out[s] += Dk * in[r];
        }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        src[c] = FULL ? out[c] : 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) fp[c * 3 + d] = out[(FULL ? NC : 0) + c * 3 + d];
      }
    } else {
      double J[9];
      trilinear_jacobian(sxv + el * 24, txq[i0], txq[i1], txq[i2], J);
      qpoint_eval<QF, NC, FULL>(P.params, J, val, gp, src, fp);
      if constexpr (MODE == MODE_LINEARIZE) {
        double Ji[9];
        inv3(J, Ji);
        const double dJ = det3(J);
        double*      Dp = P.D + size_t(slot) * tangent_slot_stride(DS, NQ);
#pragma unroll 1
        for (int s = 0; s < M; s++) {
          double col[M];
          qpoint_tangent_column<QF, NC, FULL>(P.params, Ji, dJ, val, gp, s, col);
          // keep the upper triangle: entry (r, s) for r <= s
          for (int r = 0; r <= s; r++)
            __builtin_nontemporal_store(col[r], Dp + tangent_offset(packed_index<M>(r, s), li, DS, NQ));
        }
      }
    }
    // flux -> regions 0..2 (u and stage-x partials are dead), source -> region 6
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
      for (int d = 0; d < 3; d++) S[(c * 3 + d) * NQ + li] = fp[c * 3 + d];
      if constexpr (HAS_SRC) S[6 * R + c * NQ + li] = src[c];
    }
  }
  lds_barrier();

  // ---- integrate, stage qx: (qz, qy, x)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* f0 = S + (c * 3 + 0) * NQ;
      const double* f1 = S + (c * 3 + 1) * NQ;
      const double* f2 = S + (c * 3 + 2) * NQ;
      const double* sr = S + 6 * R + c * NQ;
      double        s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int qx = 0; qx < N; qx++) {
        const int    q  = (i2 * N + i1) * N + qx;
        const double bw = tBw[qx * N + i0], gw = tGw[qx * N + i0];
        if constexpr (HAS_SRC) s0 += sr[q] * bw;
        s0 += f0[q] * gw;
        s1 += f1[q] * bw;
        s2 += f2[q] * bw;
      }
      S[3 * R + c * NQ + li] = s0;
      S[4 * R + c * NQ + li] = s1;
      S[5 * R + c * NQ + li] = s2;
    }
  }
  lds_barrier();

  // ---- stage qy: (qz, y, x)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* a20 = S + 3 * R + c * NQ;
      const double* a21 = S + 4 * R + c * NQ;
      const double* a22 = S + 5 * R + c * NQ;
      double        s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int qy = 0; qy < N; qy++) {
        const int    idx = (i2 * N + qy) * N + i0;
        const double bw = tBw[qy * N + i1], gw = tGw[qy * N + i1];
        s0 += a20[idx] * bw + a21[idx] * gw;
        s1 += a22[idx] * bw;
      }
      S[0 * R + c * NQ + li] = s0;
      S[1 * R + c * NQ + li] = s1;
    }
  }
  lds_barrier();

  // ---- stage qz + scatter-add: (z, y, x)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* a10 = S + 0 * R + c * NQ;
      const double* a11 = S + 1 * R + c * NQ;
      double        r   = 0.0;
#pragma unroll
      for (int qz = 0; qz < N; qz++) {
        const int idx = (qz * N + i1) * N + i0;
        r += a10[idx] * tBw[qz * N + i2] + a11[idx] * tGw[qz * N + i2];
      }
      scatter_add(P.r, lidx(node, c, NC, P.nnodes, P.ordering), r, P.scatter);
    }
  }
}

// Diagonal of the linearized operator from the stored tangent:
//   diag(c, j) = sum_q [w phi_j, w grad phi_j]_c . D(q) . [phi_j, grad phi_j]_c
// (the diagonal entries element_gradient_kernel would assemble,
//  fn/domain_integral_kernels.hpp:327-361).  Lane li owns node j = li.
template <int N, int NC, int LAYOUT>
__global__ void __launch_bounds__(256) k_diagonal(const ElemLaunch P)
{
  constexpr int  NQ   = N * N * N;
  constexpr int  EB   = 256 / NQ;
  constexpr bool FULL = (LAYOUT == LAYOUT_FULL_SYM);
  constexpr int  M    = FULL ? 4 * NC : 3 * NC;
  constexpr int  DS   = M * (M + 1) / 2;
  constexpr int  VO   = FULL ? NC : 0;
  constexpr int  TAB  = 4 * N * N + N;
  __shared__ double stab[TAB];

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;
  const int     li     = tid - el * NQ;
  const int     jx     = li % N, jy = (li / N) % N, jz = li / (N * N);
  const int64_t slot   = P.slot_begin + logical_block(P.xcd_remap) * EB + el;
  const bool    active = (el < EB) && (slot < P.slot_end);
  for (int k = tid; k < TAB; k += blockDim.x) stab[k] = P.tables[k];
  lds_barrier();
  if (!active) return;
  const double* tB  = stab;
  const double* tG  = stab + N * N;
  const double* tBw = stab + 2 * N * N;
  const double* tGw = stab + 3 * N * N;

  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) acc[c] = 0.0;
  const double* Dbase = P.D + size_t(slot) * tangent_slot_stride(DS, NQ);
  auto          Dat   = [&](int k, int q) { return Dbase[tangent_offset(k, q, DS, NQ)]; };
  for (int q = 0; q < NQ; q++) {
    const int    qx = q % N, qy = (q / N) % N, qz = q / (N * N);
    const double bx = tB[qx * N + jx], by = tB[qy * N + jy], bz = tB[qz * N + jz];
    const double gx = tG[qx * N + jx], gy = tG[qy * N + jy], gz = tG[qz * N + jz];
    const double wbx = tBw[qx * N + jx], wby = tBw[qy * N + jy], wbz = tBw[qz * N + jz];
    const double wgx = tGw[qx * N + jx], wgy = tGw[qy * N + jy], wgz = tGw[qz * N + jz];
    const double ug[3] = {gx * by * bz, bx * gy * bz, bx * by * gz};
    const double tg[3] = {wgx * wby * wbz, wbx * wgy * wbz, wbx * wby * wgz};
    const double uv = bx * by * bz, tv = wbx * wby * wbz;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      double a = 0.0;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const int r = VO + c * 3 + i, s = VO + c * 3 + j;
          const int k = r <= s ? packed_index<M>(r, s) : packed_index<M>(s, r);
          a += tg[i] * Dat(k, q) * ug[j];
        }
      if constexpr (FULL) {
        a += tv * Dat(packed_index<M>(c, c), q) * uv;
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const int s = VO + c * 3 + j;
          const int k = packed_index<M>(c, s);  // c < s always
          a += tv * Dat(k, q) * ug[j] + tg[j] * Dat(k, q) * uv;
        }
      }
      acc[c] += a;
    }
  }
  const int64_t node = P.map[size_t(slot) * NQ + li];
#pragma unroll
  for (int c = 0; c < NC; c++) scatter_add(P.r, lidx(node, c, NC, P.nnodes, P.ordering), acc[c], P.scatter);
}

template <int N, int NC, int MODE, class QF, int LAYOUT>
inline hipError_t launch_k(const ElemLaunch& P)
{
  constexpr int NQ = N * N * N;
  constexpr int EB = 256 / NQ;
  const int64_t n  = P.slot_end - P.slot_begin;
  if (n <= 0) return hipSuccess;
  const unsigned grid = unsigned((n + EB - 1) / EB);
  if constexpr (MODE == MODE_DIAG)
    hipLaunchKernelGGL((k_diagonal<N, NC, LAYOUT>), dim3(grid), dim3(256), 0, P.stream, P);
  else
    hipLaunchKernelGGL((k_element<N, NC, MODE, QF, LAYOUT>), dim3(grid), dim3(256), 0, P.stream, P);
  return hipGetLastError();
}

// Dispatch on (mode, qfunction, components, tangent layout) for one order.
template <int N>
inline int dispatch_order(const ElemLaunch& P, hipError_t* err)
{
  const bool full = (P.layout == LAYOUT_FULL_SYM);
  if (P.mode == MODE_JVP || P.mode == MODE_DIAG) {
    const bool diag = (P.mode == MODE_DIAG);
    if (P.nc == 1 && !full)
      *err = diag ? launch_k<N, 1, MODE_DIAG, QFNone<1>, LAYOUT_GRAD_SYM>(P)
                  : launch_k<N, 1, MODE_JVP, QFNone<1>, LAYOUT_GRAD_SYM>(P);
    else if (P.nc == 1 && full)
      *err = diag ? launch_k<N, 1, MODE_DIAG, QFNone<1>, LAYOUT_FULL_SYM>(P)
                  : launch_k<N, 1, MODE_JVP, QFNone<1>, LAYOUT_FULL_SYM>(P);
    else if (P.nc == 3 && !full)
      *err = diag ? launch_k<N, 3, MODE_DIAG, QFNone<3>, LAYOUT_GRAD_SYM>(P)
                  : launch_k<N, 3, MODE_JVP, QFNone<3>, LAYOUT_GRAD_SYM>(P);
    else if (P.nc == 3 && full)
      *err = diag ? launch_k<N, 3, MODE_DIAG, QFNone<3>, LAYOUT_FULL_SYM>(P)
                  : launch_k<N, 3, MODE_JVP, QFNone<3>, LAYOUT_FULL_SYM>(P);
    else
      return 0;
    return 1;
  }
  const bool lin = (P.mode == MODE_LINEARIZE);
#define SXK_QF(QFT, NCV, LAY)                                                                     \
  *err = lin ? launch_k<N, NCV, MODE_LINEARIZE, QFT, LAY>(P) : launch_k<N, NCV, MODE_RESIDUAL, QFT, LAY>(P); \
  return 1;
  switch (P.qf) {
    case QF_ID_HEAT:
      if (P.nc == 1) { SXK_QF(QFHeat, 1, LAYOUT_GRAD_SYM) }
      break;
    case QF_ID_MASS_DIFFUSION:
      if (P.nc == 1) { SXK_QF(QFMassDiffusion<1>, 1, LAYOUT_FULL_SYM) }
      if (P.nc == 3) { SXK_QF(QFMassDiffusion<3>, 3, LAYOUT_FULL_SYM) }
      break;
    case QF_ID_NEOHOOKEAN:
      if (P.nc == 3) { SXK_QF(QFNeoHookean, 3, LAYOUT_GRAD_SYM) }
      break;
    case QF_ID_LINEAR_ELASTIC:
      if (P.nc == 3) { SXK_QF(QFLinearElastic, 3, LAYOUT_GRAD_SYM) }
      break;
  }
#undef SXK_QF
  return 0;
}

}  // namespace sxk
