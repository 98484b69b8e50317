// sx_element.hpp -- fused element kernels for H1 hexahedra on gfx950.
//
// Per element, the pipeline of fn/domain_integral_kernels.hpp:155-304 runs in
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
// (6 + has_source) * nc * Q^3 values, reused across stages.
//
// The scatter-add is either a plain read-modify-write (one launch per element
// colour: no two elements of a colour share a node) or an FP64/FP32 atomic add
// (one launch for all elements).
//
// Precision: residual and linearization compute in FP64.  The stored tangent
// is kept in S = double, float or bf16; the J.v kernel computes in T = double
// or float and exchanges vectors in V = double or float (float arithmetic on
// double vectors is the operator of the mixed-precision solvers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sx_launch.hpp"
#include "sx_qfunctions.hpp"

namespace sxk {

// ---------------------------------------------------------------------------
// storage types of the tangent
// ---------------------------------------------------------------------------
struct bf16 {
  uint16_t bits;
};

__device__ __forceinline__ double to_real(double v, double) { return v; }
__device__ __forceinline__ float  to_real(float v, float) { return v; }
__device__ __forceinline__ double to_real(float v, double) { return double(v); }
__device__ __forceinline__ float  to_real(double v, float) { return float(v); }
__device__ __forceinline__ float  bf16_to_float(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

// FP64 -> storage, round to nearest even
__device__ __forceinline__ double   to_store(double v, double*) { return v; }
__device__ __forceinline__ float    to_store(double v, float*) { return float(v); }
__device__ __forceinline__ uint16_t to_store(double v, bf16*)
{
  const uint32_t u = __float_as_uint(float(v));
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));  // inf / NaN
  return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

template <typename S>
struct StoreTraits {
  using raw = S;
};
template <>
struct StoreTraits<bf16> {
  using raw = uint16_t;
};

// Per-slot tangent layout, in units of S: the DS packed entries of quadrature
// point q are grouped VEC = 16 / sizeof(S) at a time, [g][q][VEC], so that lane
// q reads 16 bytes per group.  The REM entries of the last, partial group
// follow entry-major so that each load is coalesced across the points: for
// 2-byte S in pairs, [RP][q][2] (lane q reads one 4-byte word per pair), then
// an odd last entry as [q]; for wider S one entry at a time, [REM][q].
template <typename S, int DS, int NQ>
struct TangentLayout {
  static constexpr int     VEC    = 16 / int(sizeof(S));
  static constexpr int     G      = DS / VEC;
  static constexpr int     REM    = DS - G * VEC;
  static constexpr int     RP     = sizeof(S) == 2 ? REM / 2 : 0;  // remainder pairs
  static constexpr int64_t stride = (int64_t(DS) * NQ * int64_t(sizeof(S)) + 15) / 16 * (16 / int64_t(sizeof(S)));
  __host__ __device__ static constexpr int offset(int k, int q)
  {
    return (k < G * VEC)          ? (k / VEC) * VEC * NQ + q * VEC + (k % VEC)
           : (k - G * VEC < 2 * RP) ? G * VEC * NQ + ((k - G * VEC) / 2) * 2 * NQ + q * 2 + (k - G * VEC) % 2
                                    : G * VEC * NQ + 2 * RP * NQ + (k - G * VEC - 2 * RP) * NQ + q;
  }
};

// packed index of the symmetric tangent: upper triangle, column-major, so a
// forward-mode pass seeded at input s produces entries s(s+1)/2 .. s(s+1)/2+s
__host__ __device__ constexpr int packed_index(int r, int s) { return r <= s ? s * (s + 1) / 2 + r : r * (r + 1) / 2 + s; }

// Shape of the stored tangent for a layout: M inputs/outputs per point
// (VAL: the field value is an input and the source an output), DS stored
// entries, and where entry D(r, s) = d out_r / d in_s lives.  Column s of a
// forward-mode pass holds entries idx(r, s) for r < rows(s), in increasing
// storage order.
template <int LAYOUT, int NC>
struct TanShape {
  static constexpr bool VAL = (LAYOUT == LAYOUT_FULL_SYM);
  static constexpr bool SYM = (LAYOUT != LAYOUT_GRAD_FULL);
  static constexpr int  M   = VAL ? 4 * NC : 3 * NC;
  static constexpr int  DS  = SYM ? M * (M + 1) / 2 : M * M;
  static constexpr int  VO  = VAL ? NC : 0;  // offset of the gradient inputs
  __host__ __device__ static constexpr int idx(int r, int s) { return SYM ? packed_index(r, s) : s * M + r; }
  __host__ __device__ static constexpr int rows(int s) { return SYM ? s + 1 : M; }
};

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

// Cross-lane reads within a wave (DPP): quad_bcast<x> gives every lane the
// value of lane x of its group of 4 lanes; row_ror<k> the value of the lane k
// places away (cyclically) within its row of 16 lanes.  With one p = 3 element
// per wave and lane = (i2 N + i1) N + i0, a quad holds one x line of nodes or
// points and a row one z layer, so the x and y contractions need no LDS.
// Every lane reads a valid source lane under these controls, so the moves
// carry no old value (no zeroing move ahead of each).
template <int X>
__device__ __forceinline__ float quad_bcast(float v)
{
  static_assert(X >= 0 && X < 4, "lane in a quad");
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), X * 0x55, 0xf, 0xf, true));
}
template <int K>
__device__ __forceinline__ float row_ror(float v)
{
  static_assert(K >= 1 && K < 16, "rotation within a row");
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + K, 0xf, 0xf, true));
}
template <int X>
__device__ __forceinline__ double quad_bcast(double v)
{
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), X * 0x55, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), X * 0x55, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
template <int K>
__device__ __forceinline__ double row_ror(double v)
{
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x120 + K, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x120 + K, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
template <int K>
__device__ __forceinline__ int row_ror_i(int v)
{
  static_assert(K >= 1 && K < 16, "rotation within a row");
  return __builtin_amdgcn_mov_dpp(v, 0x120 + K, 0xf, 0xf, true);
}

template <typename T>
__device__ __forceinline__ void scatter_add(T* r, int64_t o, T v, int atomic)
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

// Scatter-add of one element's NC x NQ nodal values in address order.  Lane li
// of the element holds acc[c] for its node li; the values and node ids go
// through LDS (sv: the element's NC * NQ values, sn: the workgroup's node ids,
// NQ per element), then pass t of the element's NQ lanes adds entries
// t NQ .. t NQ + NQ - 1 of the element's (node, component) list, node-major for
// byVDIM and component-major for byNODES: one wave instruction then covers
// whole rows of consecutive addresses (a structured mesh numbers the nodes of
// an element row consecutively) instead of one component at a 3-entry stride,
// which cuts the memory transactions of the adds about threefold.
// With merge, the nodes an element shares with the previous element of the
// workgroup across its i0 = 0 face (consecutive elements along x) are added
// into that element's values in LDS first and skipped in the passes: one add
// per shared node and workgroup instead of two.  Every lane of the workgroup
// calls it (it holds barriers); sv_prev is the previous element's sv.
template <int N, int NC, typename T, typename V>
__device__ __forceinline__ void scatter_element(V* r, const T* acc, T* sv, T* sv_prev, int* sn_wg, int el, int eb,
                                                int64_t node, int li, bool active, int64_t nnodes, int ordering,
                                                int atomic, bool merge)
{
  constexpr int NQ = N * N * N;
  int*          sn = sn_wg + (el < eb ? el : 0) * NQ;  // lanes past the workgroup's elements touch nothing
  if (el < eb) {
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) sv[ordering ? li * NC + c : c * NQ + li] = acc[c];
    }
    sn[li] = active ? int(node) : -1;
  }
  lds_barrier();
  if (merge) {
    const int i0 = li % N;
    if (active && el > 0 && i0 == 0) {
      const int jp = li + (N - 1);  // (N - 1, i1, i2) of the previous element
      if (sn_wg[(el - 1) * NQ + jp] == int(node)) {
#pragma unroll
        for (int c = 0; c < NC; c++) sv_prev[ordering ? jp * NC + c : c * NQ + jp] += acc[c];
        sn[li] = -1;
      }
    }
    lds_barrier();
  }
  if (active) {
#pragma unroll
    for (int t = 0; t < NC; t++) {
      const int f  = t * NQ + li;
      const int j  = ordering ? f / NC : li;
      const int c  = ordering ? f - j * NC : t;
      const int nd = sn[j];
      if (nd >= 0) scatter_add(r, lidx(nd, c, NC, nnodes, ordering), V(sv[f]), atomic);
    }
  }
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
// fn/finite_element.hpp:246-322: grad_x = grad_xi inv(J); source *= detJ;
// flux_parent[c][j] = detJ sum_k flux[c][k] inv(J)[j][k].
template <class QF, int NC>
__device__ __forceinline__ void qpoint_eval(const double* prm, const double* Ji, double dJ, const double* val,
                                            const double* gp, double* src, double* fp, double* st)
{
  double g[3 * NC], s[NC], f[3 * NC];
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
      g[c * 3 + k] = gp[c * 3 + 0] * Ji[0 * 3 + k] + gp[c * 3 + 1] * Ji[1 * 3 + k] + gp[c * 3 + 2] * Ji[2 * 3 + k];
#pragma unroll
  for (int c = 0; c < NC; c++) s[c] = 0.0;
  if constexpr (QF::STATE > 0)
    QF::eval(prm, val, g, s, f, st);
  else
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
                                                      const double* val, const double* gp, const double* st0,
                                                      int s, double* col)
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
  if constexpr (QF::STATE > 0) {
    // every pass starts from the point's state before this evaluation
    double st[QF::STATE];
#pragma unroll
    for (int k = 0; k < QF::STATE; k++) st[k] = st0[k];
    QF::eval(prm, v, g, sr, f, st);
  } else {
    QF::eval(prm, v, g, sr, f);
  }
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

// Directional derivative of the quadrature-point stage with respect to the
// shape: the Jacobian J = dX/dxi + dp/dxi moves by dJ = d(dp)/dxi, and
// inv(J), det(J), grad_x u = grad_xi u inv(J), the qfunction and the
// transformation of its source and flux are carried as one-direction duals
// (fn/shape_aware_functional.hpp:37-125 differentiated in p).  src and fp get
// the derivative parts.  The point's state enters as a copy (nothing is
// committed).
template <class QF, int NC>
__device__ __forceinline__ void qpoint_shape_dir(const double* prm, const double* J, const double* dJdir,
                                                 const double* val, const double* gp, const double* st0,
                                                 double* src, double* fp)
{
  d1 Jd[9], Jid[9];
#pragma unroll
  for (int k = 0; k < 9; k++) Jd[k] = d1(J[k], dJdir[k]);
  inv3(Jd, Jid);
  const d1 dJd = det3(Jd);
  d1       v[NC], g[3 * NC], s[NC], f[3 * NC];
#pragma unroll
  for (int c = 0; c < NC; c++) v[c] = d1(val[c]);
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int k = 0; k < 3; k++)
      g[c * 3 + k] = gp[c * 3 + 0] * Jid[0 * 3 + k] + gp[c * 3 + 1] * Jid[1 * 3 + k] + gp[c * 3 + 2] * Jid[2 * 3 + k];
#pragma unroll
  for (int c = 0; c < NC; c++) s[c] = d1(0.0);
  if constexpr (QF::STATE > 0) {
    double st[QF::STATE];
#pragma unroll
    for (int k = 0; k < QF::STATE; k++) st[k] = st0[k];
    QF::eval(prm, v, g, s, f, st);
  } else {
    QF::eval(prm, v, g, s, f);
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    src[c] = (s[c] * dJd).d;
#pragma unroll
    for (int j = 0; j < 3; j++)
      fp[c * 3 + j] = ((f[c * 3 + 0] * Jid[j * 3 + 0] + f[c * 3 + 1] * Jid[j * 3 + 1] + f[c * 3 + 2] * Jid[j * 3 + 2]) * dJd).d;
  }
}

// Loads the DS tangent entries of quadrature point q as T.
template <typename T, typename S, int DS, int NQ>
__device__ __forceinline__ void load_tangent(const void* base, int q, T* Dr)
{
  using L = TangentLayout<S, DS, NQ>;
  // native vector types per storage type: a 16-byte integer vector
  // reassembled into doubles gets split by the back end into a 16-byte load
  // plus a second, 8-byte load of the same address
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  typedef double       d2 __attribute__((ext_vector_type(2)));
  typedef float        f4 __attribute__((ext_vector_type(4)));
  const S* Dp = static_cast<const S*>(base);
#pragma unroll
  for (int g = 0; g < L::G; g++) {
    const S* gp = Dp + g * L::VEC * NQ + q * L::VEC;
    if constexpr (sizeof(S) == 8) {
      const d2 v    = __builtin_nontemporal_load(reinterpret_cast<const d2*>(gp));
      Dr[2 * g]     = T(v.x);
      Dr[2 * g + 1] = T(v.y);
    } else if constexpr (sizeof(S) == 4) {
      const f4 v    = __builtin_nontemporal_load(reinterpret_cast<const f4*>(gp));
      Dr[4 * g]     = T(v.x);
      Dr[4 * g + 1] = T(v.y);
      Dr[4 * g + 2] = T(v.z);
      Dr[4 * g + 3] = T(v.w);
    } else {
      const u4 v = __builtin_nontemporal_load(reinterpret_cast<const u4*>(gp));
      const unsigned int w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 4; i++) {
        Dr[8 * g + 2 * i]     = T(__uint_as_float(w[i] << 16));
        Dr[8 * g + 2 * i + 1] = T(__uint_as_float(w[i] & 0xffff0000u));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < L::REM; j++) {
    const auto raw = __builtin_nontemporal_load(
        reinterpret_cast<const typename StoreTraits<S>::raw*>(Dp + L::offset(L::G * L::VEC + j, q)));
    if constexpr (sizeof(S) == 2)
      Dr[L::G * L::VEC + j] = T(bf16_to_float(uint32_t(raw)));
    else
      Dr[L::G * L::VEC + j] = T(raw);
  }
}

// The DS tangent entries of one quadrature point held in registers for the
// J.v chain rule: as T, or, for a BF16 tangent, still packed two to a 32-bit
// register and widened at their one use (the same conversions, half the
// registers: 25 instead of 45 for the Neo-Hookean tangent, which lets more
// waves share a SIMD).  get(k) needs a compile-time k after unrolling.
template <typename T, typename S, int DS, int NQ>
struct TangentRegs {
  T d[DS];
  __device__ __forceinline__ void load(const void* base, int q) { load_tangent<T, S, DS, NQ>(base, q, d); }
  __device__ __forceinline__ T    get(int k) const { return d[k]; }
};
template <typename T, int DS, int NQ>
struct TangentRegs<T, bf16, DS, NQ> {
  using L = TangentLayout<bf16, DS, NQ>;
  static constexpr int K0 = L::G * L::VEC;  // first remainder entry
  static constexpr int NW = L::G * 4 + L::RP;
  uint32_t w[NW > 0 ? NW : 1];  // entries 2i, 2i + 1 as word i (low half first): the groups, then the remainder pairs
  float    last;                // an odd last entry, widened on load
  __device__ __forceinline__ void load(const void* base, int q)
  {
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const bf16* Dp = static_cast<const bf16*>(base);
#pragma unroll
    for (int g = 0; g < L::G; g++) {
      const u4 v   = __builtin_nontemporal_load(reinterpret_cast<const u4*>(Dp + g * L::VEC * NQ + q * L::VEC));
      w[4 * g]     = v.x;
      w[4 * g + 1] = v.y;
      w[4 * g + 2] = v.z;
      w[4 * g + 3] = v.w;
    }
#pragma unroll
    for (int p = 0; p < L::RP; p++)
      w[4 * L::G + p] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(Dp + L::offset(K0 + 2 * p, q)));
    if constexpr (L::REM % 2 == 1)
      last = bf16_to_float(
          uint32_t(__builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(Dp + L::offset(DS - 1, q)))));
    else
      last = 0.0f;
  }
  __device__ __forceinline__ T get(int k) const
  {
    if (L::REM % 2 == 1 && k == DS - 1) return T(last);
    const uint32_t x = w[k / 2];  // groups and remainder pairs alike: word k / 2
    return T(__uint_as_float((k & 1) ? (x & 0xffff0000u) : (x << 16)));
  }
};

// Stores tangent entry k (FP64 value) of quadrature point q: entries arrive in
// increasing k; whole 16-byte groups are stored at once from `pend`.
template <typename S, int DS, int NQ>
__device__ __forceinline__ void store_tangent_entry(void* base, int q, int k, double v,
                                                    typename StoreTraits<S>::raw* pend)
{
  using L   = TangentLayout<S, DS, NQ>;
  using raw = typename StoreTraits<S>::raw;
  raw* Dp   = static_cast<raw*>(base);
  if (k < L::G * L::VEC) {
    pend[k % L::VEC] = raw(to_store(v, static_cast<S*>(nullptr)));
    if (k % L::VEC == L::VEC - 1) {
      typedef unsigned int u4 __attribute__((ext_vector_type(4)));
      u4 w;
      __builtin_memcpy(&w, pend, 16);
      __builtin_nontemporal_store(w, reinterpret_cast<u4*>(Dp + (k / L::VEC) * L::VEC * NQ + q * L::VEC));
    }
  } else {
    __builtin_nontemporal_store(raw(to_store(v, static_cast<S*>(nullptr))), Dp + L::offset(k, q));
  }
}

// ---------------------------------------------------------------------------
// residual (MODE_RESIDUAL) and linearization (MODE_LINEARIZE, tangent in S)
// ---------------------------------------------------------------------------
template <int N, int NC, int MODE, class QF, int LAYOUT, typename S>
__global__ void __launch_bounds__(256) k_element(const ElemLaunch P)
{
  constexpr int  NQ      = N * N * N;
  constexpr int  EB      = 256 / NQ;
  using Shape            = TanShape<LAYOUT, NC>;
  constexpr bool FULL    = Shape::VAL;
  constexpr bool HAS_SRC = QF::HAS_SOURCE;
  constexpr int  M       = Shape::M;
  constexpr int  DS      = Shape::DS;
  constexpr int  NST     = QF::STATE;
  constexpr int  R       = NC * NQ;  // one LDS region
  constexpr int  PER     = (6 + (HAS_SRC ? 1 : 0)) * R;
  constexpr int  TAB     = 4 * N * N + N;
  using Lay              = TangentLayout<S, DS, NQ>;

  __shared__ double smem[EB * PER];
  __shared__ double stab[TAB];
  __shared__ double sxv[EB * 24];
  __shared__ int    snode[EB * NQ];

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;
  const int     li     = tid - el * NQ;
  const int     i0     = li % N;
  const int     i1     = (li / N) % N;
  const int     i2     = li / (N * N);
  const int64_t slot   = P.slot_begin + logical_block(P.xcd_remap) * EB + el;
  const bool    active = (el < EB) && (slot < P.slot_end);
  double*       S_     = smem + (el < EB ? el : 0) * PER;
  const double* u      = static_cast<const double*>(P.u);
  double*       r      = static_cast<double*>(P.r);

  for (int k = tid; k < TAB; k += blockDim.x) stab[k] = P.tables[k];
  const double* tB  = stab;
  const double* tG  = stab + N * N;
  const double* tBw = stab + 2 * N * N;
  const double* tGw = stab + 3 * N * N;
  const double* txq = stab + 4 * N * N;

  // ---- quadrature-point state (issued first, consumed at the qfunction)
  double  st[NST > 0 ? NST : 1];
  double* stp = nullptr;
  if constexpr (NST > 0) {
    if (active) {
      stp = P.qdata + size_t(P.qrow[slot]) * NST * NQ + li;
#pragma unroll
      for (int k = 0; k < NST; k++) st[k] = stp[k * NQ];
    }
  }

  // ---- shape displacement (ShapeAwareFunctional, fn/shape_aware_functional.hpp:37-50):
  // the parent gradient of p at this lane's quadrature point, added to the
  // geometric Jacobian below, so dx/dxi = dX/dxi + dp/dxi = (I + dp/dX) dX/dxi.
  // P.shape is the same for every lane, so the barriers inside are uniform.
  // parent gradient (at this lane's point) of a 3-component field on the
  // element's nodes, by the interpolation's sum factorization; uses LDS
  // regions 0..5 before the gather
  auto parent_grad3 = [&](const double* field, double* out9) {
    if (active) {
      const int64_t nd = P.map[size_t(slot) * NQ + li];
#pragma unroll
      for (int c = 0; c < 3; c++) S_[c * NQ + li] = field[lidx(nd, c, 3, P.nnodes, P.ordering)];
    }
    lds_barrier();
    if (active) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double* pp = S_ + c * NQ + (i2 * N + i1) * N;
        double        s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int x = 0; x < N; x++) {
          s0 += tB[i0 * N + x] * pp[x];
          s1 += tG[i0 * N + x] * pp[x];
        }
        S_[1 * R + c * NQ + li] = s0;
        S_[2 * R + c * NQ + li] = s1;
      }
    }
    lds_barrier();
    if (active) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
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
    lds_barrier();
    if (active) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double* a20 = S_ + 3 * R + c * NQ;
        const double* a21 = S_ + 4 * R + c * NQ;
        const double* a22 = S_ + 5 * R + c * NQ;
#pragma unroll
        for (int z = 0; z < N; z++) {
          const int    idx = (z * N + i1) * N + i0;
          const double b = tB[i2 * N + z], g = tG[i2 * N + z];
          out9[c * 3 + 0] += b * a21[idx];
          out9[c * 3 + 1] += b * a22[idx];
          out9[c * 3 + 2] += g * a20[idx];
        }
      }
    }
    lds_barrier();  // the next pass / the gather overwrites the regions read above
  };
  double gps[9], dgps[9];
#pragma unroll
  for (int k = 0; k < 9; k++) gps[k] = dgps[k] = 0.0;
  if constexpr (NC == 3) {
    if (P.shape != nullptr) parent_grad3(P.shape, gps);
    if constexpr (MODE == MODE_SHAPE_JVP) parent_grad3(P.dshape, dgps);
  }

  // ---- gather
  int64_t node = 0;
  if (active) {
    node = P.map[size_t(slot) * NQ + li];
#pragma unroll
    for (int c = 0; c < NC; c++) S_[c * NQ + li] = u[lidx(node, c, NC, P.nnodes, P.ordering)];
    for (int k = li; k < 24; k += NQ) sxv[el * 24 + k] = P.coords[size_t(slot) * 24 + k];
  }
  lds_barrier();

  // ---- interpolate, stage x: (z, y, qx)
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
  lds_barrier();

  // ---- stage y: (z, qy, qx)
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
  lds_barrier();

  // ---- stage z + quadrature point: (qz, qy, qx)
  if (active) {
    double val[NC], gp[3 * NC];
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
        if constexpr (QF::USES_VALUE) v += b * a20[idx];
        g0 += b * a21[idx];
        g1 += b * a22[idx];
        g2 += g * a20[idx];
      }
      val[c]        = v;
      gp[c * 3 + 0] = g0;
      gp[c * 3 + 1] = g1;
      gp[c * 3 + 2] = g2;
    }

    double J[9], Ji[9];
    trilinear_jacobian(sxv + el * 24, txq[i0], txq[i1], txq[i2], J);
#pragma unroll
    for (int k = 0; k < 9; k++) J[k] += gps[k];
    inv3(J, Ji);
    const double dJ = det3(J);
    double       src[NC], fp[3 * NC];
    if constexpr (MODE == MODE_SHAPE_JVP) {
      qpoint_shape_dir<QF, NC>(P.params, J, dgps, val, gp, st, src, fp);
    } else {
    if constexpr (MODE == MODE_LINEARIZE) {
      // columns s = 0..M-1 give entries idx(r, s), r < rows(s), in increasing
      // storage order; every column starts from the state before this call
      void*                          Dp = static_cast<char*>(P.D) + size_t(slot) * Lay::stride * sizeof(S);
      typename StoreTraits<S>::raw pend[Lay::VEC];
#pragma unroll
      for (int s = 0; s < M; s++) {
        double col[M];
        qpoint_tangent_column<QF, NC, FULL>(P.params, Ji, dJ, val, gp, st, s, col);
#pragma unroll
        for (int rr = 0; rr < Shape::rows(s); rr++)
          store_tangent_entry<S, DS, NQ>(Dp, li, Shape::idx(rr, s), col[rr], pend);
      }
    }
    qpoint_eval<QF, NC>(P.params, Ji, dJ, val, gp, src, fp, st);
    if constexpr (NST > 0) {
      // batch_apply_qf (fn/domain_integral_kernels.hpp:125-150): commit the
      // point's updated state only when asked to
      if (P.update_state) {
#pragma unroll
        for (int k = 0; k < NST; k++) stp[k * NQ] = st[k];
      }
    }
    }
    // flux -> regions 0..2 (u and stage-x partials are dead), source -> region 6
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
      for (int d = 0; d < 3; d++) S_[(c * 3 + d) * NQ + li] = fp[c * 3 + d];
      if constexpr (HAS_SRC) S_[6 * R + c * NQ + li] = src[c];
    }
  }
  lds_barrier();

  // ---- integrate, stage qx: (qz, qy, x)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* f0 = S_ + (c * 3 + 0) * NQ;
      const double* f1 = S_ + (c * 3 + 1) * NQ;
      const double* f2 = S_ + (c * 3 + 2) * NQ;
      const double* sr = S_ + 6 * R + c * NQ;
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
      S_[3 * R + c * NQ + li] = s0;
      S_[4 * R + c * NQ + li] = s1;
      S_[5 * R + c * NQ + li] = s2;
    }
  }
  lds_barrier();

  // ---- stage qy: (qz, y, x)
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* a20 = S_ + 3 * R + c * NQ;
      const double* a21 = S_ + 4 * R + c * NQ;
      const double* a22 = S_ + 5 * R + c * NQ;
      double        s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int qy = 0; qy < N; qy++) {
        const int    idx = (i2 * N + qy) * N + i0;
        const double bw = tBw[qy * N + i1], gw = tGw[qy * N + i1];
        s0 += a20[idx] * bw + a21[idx] * gw;
        s1 += a22[idx] * bw;
      }
      S_[0 * R + c * NQ + li] = s0;
      S_[1 * R + c * NQ + li] = s1;
    }
  }
  lds_barrier();

  // ---- stage qz + scatter-add: (z, y, x)
  double acc[NC];
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const double* a10 = S_ + 0 * R + c * NQ;
      const double* a11 = S_ + 1 * R + c * NQ;
      acc[c]            = 0.0;
#pragma unroll
      for (int qz = 0; qz < N; qz++) {
        const int idx = (qz * N + i1) * N + i0;
        acc[c] += a10[idx] * tBw[qz * N + i2] + a11[idx] * tGw[qz * N + i2];
      }
    }
  }
  if (P.ordered) {
    scatter_element<N, NC, double, double>(r, acc, S_ + 2 * R, smem + (el > 0 && el < EB ? el - 1 : 0) * PER + 2 * R,
                                           snode, el, EB, node, li, active, P.nnodes, P.ordering, P.scatter,
                                           P.ordered == 2);
  } else if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) scatter_add(r, lidx(node, c, NC, P.nnodes, P.ordering), acc[c], P.scatter);
  }
}

// ---------------------------------------------------------------------------
// J.v with the stored tangent (fn/domain_integral_kernels.hpp:223-304): vectors
// and arithmetic in T, tangent stored in S
// ---------------------------------------------------------------------------
// Lanes per workgroup of k_jvp: 256, except p = 4 in FP32 arithmetic, where
// one element (125 lanes) per 128-lane workgroup measured faster (FP32 tangent
// 0.580 -> 0.574 ms, BF16 0.505 -> 0.485 ms at 48^3; FP64 and the other orders
// are faster at 256).
template <int N, typename T>
constexpr int jvp_group_lanes()
{
  return (N == 5 && sizeof(T) == 4) ? 128 : 256;
}

template <int N, int NC, int LAYOUT, typename T, typename S, typename V>
__global__ void __launch_bounds__((jvp_group_lanes<N, T>())) k_jvp(const ElemLaunch P)
{
  constexpr int  NQ   = N * N * N;
  constexpr int  EB   = jvp_group_lanes<N, T>() / NQ;
  using Shape         = TanShape<LAYOUT, NC>;
  constexpr bool FULL = Shape::VAL;
  constexpr int  M    = Shape::M;
  constexpr int  DS   = Shape::DS;
  constexpr int  R    = NC * NQ;
  constexpr int  PER  = (6 + (FULL ? 1 : 0)) * R;
  constexpr int  TAB  = 4 * N * N;
  using Lay           = TangentLayout<S, DS, NQ>;

  __shared__ T   smem[EB * PER];
  __shared__ T   stab[TAB];
  __shared__ int snode[EB * NQ];

  // a Krylov loop that already converged: the whole grid leaves (the flag
  // was written by an earlier kernel on the stream, so every block reads the same)
  if (P.stop != nullptr && *P.stop != 0.0) return;

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;
  const int     li     = tid - el * NQ;
  const int     i0     = li % N;
  const int     i1     = (li / N) % N;
  const int     i2     = li / (N * N);
  const int64_t slot   = P.slot_begin + logical_block(P.xcd_remap) * EB + el;
  const bool    active = (el < EB) && (slot < P.slot_end);
  T*            S_     = smem + (el < EB ? el : 0) * PER;
  const V*      u      = static_cast<const V*>(P.u);
  V*            r      = static_cast<V*>(P.r);

  for (int k = tid; k < TAB; k += blockDim.x) stab[k] = T(P.tables[k]);
  const T* tB  = stab;
  const T* tG  = stab + N * N;
  const T* tBw = stab + 2 * N * N;
  const T* tGw = stab + 3 * N * N;

  // ---- gather loads first, then the tangent stream: memory counters retire
  // in order, so waiting for the nodal values does not wait for the tangent,
  // which stays in flight through the interpolation stages
  int64_t node = 0;
  T       ug[NC];
  if (active) {
    node = P.map[size_t(slot) * NQ + li];
#pragma unroll
    for (int c = 0; c < NC; c++) ug[c] = T(u[lidx(node, c, NC, P.nnodes, P.ordering)]);
  }
  TangentRegs<T, S, DS, NQ> Dr;
  if (active) Dr.load(static_cast<const char*>(P.D) + size_t(slot) * Lay::stride * sizeof(S), li);
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) S_[c * NQ + li] = ug[c];
  }
  lds_barrier();

  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T* uu = S_ + c * NQ + (i2 * N + i1) * N;
      T        s0 = T(0), s1 = T(0);
#pragma unroll
      for (int x = 0; x < N; x++) {
        s0 += tB[i0 * N + x] * uu[x];
        s1 += tG[i0 * N + x] * uu[x];
      }
      S_[1 * R + c * NQ + li] = s0;
      S_[2 * R + c * NQ + li] = s1;
    }
  }
  lds_barrier();

  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T* a10 = S_ + 1 * R + c * NQ;
      const T* a11 = S_ + 2 * R + c * NQ;
      T        s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
      for (int y = 0; y < N; y++) {
        const int idx = (i2 * N + y) * N + i0;
        const T   b = tB[i1 * N + y], g = tG[i1 * N + y];
        s0 += b * a10[idx];
        s1 += b * a11[idx];
        s2 += g * a10[idx];
      }
      S_[3 * R + c * NQ + li] = s0;
      S_[4 * R + c * NQ + li] = s1;
      S_[5 * R + c * NQ + li] = s2;
    }
  }
  lds_barrier();

  if (active) {
    T in[M], out[M];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T* a20 = S_ + 3 * R + c * NQ;
      const T* a21 = S_ + 4 * R + c * NQ;
      const T* a22 = S_ + 5 * R + c * NQ;
      T        v = T(0), g0 = T(0), g1 = T(0), g2 = T(0);
#pragma unroll
      for (int z = 0; z < N; z++) {
        const int idx = (z * N + i1) * N + i0;
        const T   b = tB[i2 * N + z], g = tG[i2 * N + z];
        if constexpr (FULL) v += b * a20[idx];
        g0 += b * a21[idx];
        g1 += b * a22[idx];
        g2 += g * a20[idx];
      }
      if constexpr (FULL) in[c] = v;
      in[(FULL ? NC : 0) + c * 3 + 0] = g0;
      in[(FULL ? NC : 0) + c * 3 + 1] = g1;
      in[(FULL ? NC : 0) + c * 3 + 2] = g2;
    }
    // chain rule with the stored symmetric tangent (fn/tensor.hpp:1791-1869)
#pragma unroll
    for (int k = 0; k < M; k++) out[k] = T(0);
#pragma unroll
    for (int s = 0; s < M; s++)
#pragma unroll
      for (int rr = 0; rr < Shape::rows(s); rr++) {
        const T Dk = Dr.get(Shape::idx(rr, s));
        out[rr] += Dk * in[s];
        if (Shape::SYM && rr != s) out[s] += Dk * in[rr];
      }
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
      for (int d = 0; d < 3; d++) S_[(c * 3 + d) * NQ + li] = out[(FULL ? NC : 0) + c * 3 + d];
      if constexpr (FULL) S_[6 * R + c * NQ + li] = out[c];
    }
  }
  lds_barrier();

  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T* f0 = S_ + (c * 3 + 0) * NQ;
      const T* f1 = S_ + (c * 3 + 1) * NQ;
      const T* f2 = S_ + (c * 3 + 2) * NQ;
      const T* sr = S_ + 6 * R + c * NQ;
      T        s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
      for (int qx = 0; qx < N; qx++) {
        const int q  = (i2 * N + i1) * N + qx;
        const T   bw = tBw[qx * N + i0], gw = tGw[qx * N + i0];
        if constexpr (FULL) s0 += sr[q] * bw;
        s0 += f0[q] * gw;
        s1 += f1[q] * bw;
        s2 += f2[q] * bw;
      }
      S_[3 * R + c * NQ + li] = s0;
      S_[4 * R + c * NQ + li] = s1;
      S_[5 * R + c * NQ + li] = s2;
    }
  }
  lds_barrier();

  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T* a20 = S_ + 3 * R + c * NQ;
      const T* a21 = S_ + 4 * R + c * NQ;
      const T* a22 = S_ + 5 * R + c * NQ;
      T        s0 = T(0), s1 = T(0);
#pragma unroll
      for (int qy = 0; qy < N; qy++) {
        const int idx = (i2 * N + qy) * N + i0;
        const T   bw = tBw[qy * N + i1], gw = tGw[qy * N + i1];
        s0 += a20[idx] * bw + a21[idx] * gw;
        s1 += a22[idx] * bw;
      }
      S_[0 * R + c * NQ + li] = s0;
      S_[1 * R + c * NQ + li] = s1;
    }
  }
  lds_barrier();

  T acc[NC];
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const T* a10 = S_ + 0 * R + c * NQ;
      const T* a11 = S_ + 1 * R + c * NQ;
      acc[c]       = T(0);
#pragma unroll
      for (int qz = 0; qz < N; qz++) {
        const int idx = (qz * N + i1) * N + i0;
        acc[c] += a10[idx] * tBw[qz * N + i2] + a11[idx] * tGw[qz * N + i2];
      }
    }
  }
  if (P.ordered) {
    scatter_element<N, NC, T, V>(r, acc, S_ + 2 * R, smem + (el > 0 && el < EB ? el - 1 : 0) * PER + 2 * R, snode,
                                 el, EB, node, li, active, P.nnodes, P.ordering, P.scatter, P.ordered == 2);
  } else if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) scatter_add(r, lidx(node, c, NC, P.nnodes, P.ordering), V(acc[c]), P.scatter);
  }
}

// J.v for p = 3 (N = 4), arithmetic in T: the contractions along x and y
// move values between lanes of the element's wave (quad_bcast, row_ror)
// instead of through LDS; only the z contractions and the scatter use LDS.
// Vectors V (float, or double for FP32 arithmetic on FP64 vectors), tangent S.
// The same stages, in the same order of operations per term, as k_jvp.
template <typename T, int NC, int LAYOUT, typename S, typename V, int EB>
__global__ void __launch_bounds__(64 * EB) k_jvp_xy(const ElemLaunch P)
{
  constexpr int  N    = 4;
  constexpr int  NQ   = N * N * N;
  using Shape         = TanShape<LAYOUT, NC>;
  constexpr bool FULL = Shape::VAL;
  constexpr int  M    = Shape::M;
  constexpr int  DS   = Shape::DS;
  constexpr int  R    = NC * NQ;
  constexpr int  PER  = 3 * R;  // the y-interpolation outputs, later the y-integration outputs and the scatter
  constexpr int  TAB  = 4 * N * N;
  using Lay           = TangentLayout<S, DS, NQ>;

  __shared__ T   smem[EB * PER];
  __shared__ T   stab[TAB];
  __shared__ int snode[EB * NQ];

  if (P.stop != nullptr && *P.stop != 0.0) return;

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;  // one element per wave
  const int     li     = tid - el * NQ;
  const int     i0     = li & 3;
  const int     i1     = (li >> 2) & 3;
  const int     i2     = li >> 4;
  const int64_t slot   = P.slot_begin + logical_block(P.xcd_remap) * EB + el;
  const bool    active = (el < EB) && (slot < P.slot_end);  // the same for every lane of a wave
  T*            S_     = smem + (el < EB ? el : 0) * PER;
  const V*      u      = static_cast<const V*>(P.u);
  V*            r      = static_cast<V*>(P.r);

  for (int k = tid; k < TAB; k += blockDim.x) stab[k] = T(P.tables[k]);
  const T* tB  = stab;
  const T* tG  = stab + N * N;
  const T* tBw = stab + 2 * N * N;
  const T* tGw = stab + 3 * N * N;

  int* const sn_el = snode + (el < EB ? el : 0) * NQ;
  T          ug[NC];
  if (active) {
    const int64_t node = P.map[size_t(slot) * NQ + li];
#pragma unroll
    for (int c = 0; c < NC; c++) ug[c] = T(u[lidx(node, c, NC, P.nnodes, P.ordering)]);
    sn_el[li] = int(node);  // read back for the scatter: no register holds it through the element work
  }
  TangentRegs<T, S, DS, NQ> Dr;
  if (active) Dr.load(static_cast<const char*>(P.D) + size_t(slot) * Lay::stride * sizeof(S), li);
  // the y index of the lane a rotation by 4 m within the row reads (m = 0..3)
  // the y index of the lane a rotation by 4 m within the row reads is
  // (i1 + m ydir) mod 4, ydir found by rotating the lane's own index once
  const int ydir = ((row_ror_i<4>(li & 15) >> 2) - i1) & 3;
  auto      ys   = [&](int m) { return (i1 + m * ydir) & 3; };
  lds_barrier();  // the tables

  T acc[NC];
  {
    // ---- interpolation: x through the quad, y through the row
    T b0[NC], b1[NC], b2[NC];
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const T v[4] = {quad_bcast<0>(ug[c]), quad_bcast<1>(ug[c]), quad_bcast<2>(ug[c]), quad_bcast<3>(ug[c])};
        T       s0 = T(0), s1 = T(0);
#pragma unroll
        for (int x = 0; x < N; x++) {
          s0 += tB[i0 * N + x] * v[x];
          s1 += tG[i0 * N + x] * v[x];
        }
        const T a0[4] = {s0, row_ror<4>(s0), row_ror<8>(s0), row_ror<12>(s0)};
        const T a1[4] = {s1, row_ror<4>(s1), row_ror<8>(s1), row_ror<12>(s1)};
        T       t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const T b = tB[i1 * N + ys(m)], g = tG[i1 * N + ys(m)];
          t0 += b * a0[m];
          t1 += b * a1[m];
          t2 += g * a0[m];
        }
        b0[c] = t0;
        b1[c] = t1;
        b2[c] = t2;
      }
#pragma unroll
      for (int c = 0; c < NC; c++) {
        S_[0 * R + c * NQ + li] = b0[c];
        S_[1 * R + c * NQ + li] = b1[c];
        S_[2 * R + c * NQ + li] = b2[c];
      }
    }
    lds_barrier();

    // ---- z through LDS, then the chain rule at the point
    T out[M];
    if (active) {
      T in[M];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const T* a20 = S_ + 0 * R + c * NQ;
        const T* a21 = S_ + 1 * R + c * NQ;
        const T* a22 = S_ + 2 * R + c * NQ;
        T        v = T(0), g0 = T(0), g1 = T(0), g2 = T(0);
#pragma unroll
        for (int z = 0; z < N; z++) {
          const int idx = (z * N + i1) * N + i0;
          const T   b = tB[i2 * N + z], g = tG[i2 * N + z];
          if constexpr (FULL) v += b * a20[idx];
          g0 += b * a21[idx];
          g1 += b * a22[idx];
          g2 += g * a20[idx];
        }
        if constexpr (FULL) in[c] = v;
        in[(FULL ? NC : 0) + c * 3 + 0] = g0;
        in[(FULL ? NC : 0) + c * 3 + 1] = g1;
        in[(FULL ? NC : 0) + c * 3 + 2] = g2;
      }
#pragma unroll
      for (int k = 0; k < M; k++) out[k] = T(0);
#pragma unroll
      for (int s = 0; s < M; s++)
#pragma unroll
        for (int rr = 0; rr < Shape::rows(s); rr++) {
          const T Dk = Dr.get(Shape::idx(rr, s));
          out[rr] += Dk * in[s];
          if (Shape::SYM && rr != s) out[s] += Dk * in[rr];
        }
    }

    // ---- integration: x through the quad, y through the row, from registers
    T d0[NC], d1[NC];
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const T f0 = out[(FULL ? NC : 0) + c * 3 + 0], f1 = out[(FULL ? NC : 0) + c * 3 + 1];
        const T f2 = out[(FULL ? NC : 0) + c * 3 + 2];
        const T q0[4] = {quad_bcast<0>(f0), quad_bcast<1>(f0), quad_bcast<2>(f0), quad_bcast<3>(f0)};
        const T q1[4] = {quad_bcast<0>(f1), quad_bcast<1>(f1), quad_bcast<2>(f1), quad_bcast<3>(f1)};
        const T q2[4] = {quad_bcast<0>(f2), quad_bcast<1>(f2), quad_bcast<2>(f2), quad_bcast<3>(f2)};
        T       sr[4] = {T(0), T(0), T(0), T(0)};
        if constexpr (FULL) {
          sr[0] = quad_bcast<0>(out[c]);
          sr[1] = quad_bcast<1>(out[c]);
          sr[2] = quad_bcast<2>(out[c]);
          sr[3] = quad_bcast<3>(out[c]);
        }
        T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
        for (int qx = 0; qx < N; qx++) {
          const T bw = tBw[qx * N + i0], gw = tGw[qx * N + i0];
          if constexpr (FULL) s0 += sr[qx] * bw;
          s0 += q0[qx] * gw;
          s1 += q1[qx] * bw;
          s2 += q2[qx] * bw;
        }
        const T a20[4] = {s0, row_ror<4>(s0), row_ror<8>(s0), row_ror<12>(s0)};
        const T a21[4] = {s1, row_ror<4>(s1), row_ror<8>(s1), row_ror<12>(s1)};
        const T a22[4] = {s2, row_ror<4>(s2), row_ror<8>(s2), row_ror<12>(s2)};
        T       t0 = T(0), t1 = T(0);
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const T bw = tBw[ys(m) * N + i1], gw = tGw[ys(m) * N + i1];
          t0 += a20[m] * bw + a21[m] * gw;
          t1 += a22[m] * bw;
        }
        d0[c] = t0;
        d1[c] = t1;
      }
    }
    lds_barrier();  // every lane has read the interpolation slots
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        S_[0 * R + c * NQ + li] = d0[c];
        S_[1 * R + c * NQ + li] = d1[c];
      }
    }
    lds_barrier();

    // ---- z integration through LDS
    if (active) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const T* a10 = S_ + 0 * R + c * NQ;
        const T* a11 = S_ + 1 * R + c * NQ;
        acc[c]       = T(0);
#pragma unroll
        for (int qz = 0; qz < N; qz++) {
          const int idx = (qz * N + i1) * N + i0;
          acc[c] += a10[idx] * tBw[qz * N + i2] + a11[idx] * tGw[qz * N + i2];
        }
      }
    }
  }
  // slot 2 is free here: the z integration reads slots 0 and 1 only
  const int64_t node = active ? int64_t(sn_el[li]) : 0;
  if (P.ordered) {
    scatter_element<N, NC, T, V>(r, acc, S_ + 2 * R, smem + (el > 0 && el < EB ? el - 1 : 0) * PER + 2 * R, snode,
                                 el, EB, node, li, active, P.nnodes, P.ordering, P.scatter, P.ordered == 2);
  } else if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) scatter_add(r, lidx(node, c, NC, P.nnodes, P.ordering), V(acc[c]), P.scatter);
  }
}

// Diagonal of the linearized operator from the stored tangent (FP64 out):
//   diag(c, j) = sum_q [w phi_j, w grad phi_j]_c . D(q) . [phi_j, grad phi_j]_c
// (the diagonal entries element_gradient_kernel would assemble,
//  fn/domain_integral_kernels.hpp:327-361).
// Every term is a product of one 1D factor per direction, so the sum over q
// factorizes: with phi_j = B(qx, jx) B(qy, jy) B(qz, jz), the term of tangent
// entry (test grad a, trial grad b) carries in direction d the factor
// (w G)(q_d, j_d) if a = d else (w B), times G(q_d, j_d) if b = d else B.  The
// products w B B, w G B (= w B G) and w G G are the three 1D tables.  Entries
// (a, b) and (b, a) share their tables and are summed first; the value rows of
// FULL_SYM add the (value, value) term and the (value, grad) pairs.  Per
// component this leaves NF fields per point, each contracted over qx, qy, qz in
// turn through LDS (lane = point, then (jx, qy, qz), (jx, jy, qz), node).
template <int LAYOUT, int NC>
struct DiagFields {
  using Shape                = TanShape<LAYOUT, NC>;
  static constexpr bool FULL = Shape::VAL;
  static constexpr int  NF   = FULL ? 10 : 6;  // per component: 6 gradient pairs (+ value, 3 value-gradient pairs)
  // gradient pair f < 6: (a, b) = (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
  __host__ __device__ static constexpr int pa(int f) { return f < 3 ? f : (f == 5 ? 1 : 0); }
  __host__ __device__ static constexpr int pb(int f) { return f < 3 ? f : (f == 3 ? 1 : 2); }
  // table (0: w B B, 1: w G B, 2: w G G) of field f in direction d
  __host__ __device__ static constexpr int table(int f, int d)
  {
    return f < 6 ? int(pa(f) == d) + int(pb(f) == d) : (f == 6 ? 0 : int(f - 7 == d));
  }
};

template <int N, int NC, int LAYOUT, typename S>
__global__ void __launch_bounds__(256) k_diagonal(const ElemLaunch P)
{
  constexpr int NQ = N * N * N;
  constexpr int EB = 256 / NQ;
  using Shape      = TanShape<LAYOUT, NC>;
  using F          = DiagFields<LAYOUT, NC>;
  constexpr int NF = F::NF;
  constexpr int VO = Shape::VO;
  constexpr int NN = N * N;
  using Lay        = TangentLayout<S, Shape::DS, NQ>;
  static_assert(NF >= NC, "the scatter stages NC x NQ values in the element's field slots");
  __shared__ double stab[3 * NN];  // [w B B | w G B | w G G][q * N + j]
  __shared__ double sf[EB * NF * NQ];
  __shared__ int    snode[EB * NQ];

  const int     tid    = threadIdx.x;
  const int     el     = tid / NQ;
  const int     li     = tid - el * NQ;
  const int     a0     = li % N, a1 = (li / N) % N, a2 = li / NN;
  const int64_t slot   = P.slot_begin + logical_block(P.xcd_remap) * EB + el;
  const bool    active = (el < EB) && (slot < P.slot_end);  // uniform over the element's lanes
  double*       sv     = sf + (el < EB ? el : 0) * NF * NQ;
  for (int k = tid; k < NN; k += blockDim.x) {
    const double b = P.tables[k], g = P.tables[NN + k], bw = P.tables[2 * NN + k], gw = P.tables[3 * NN + k];
    stab[k]          = bw * b;
    stab[NN + k]     = gw * b;
    stab[2 * NN + k] = gw * g;
  }

  using raw        = typename StoreTraits<S>::raw;
  const raw* Dbase = reinterpret_cast<const raw*>(static_cast<const char*>(P.D) +
                                                  size_t(active ? slot : 0) * Lay::stride * sizeof(S));
  auto       Dat   = [&](int k) -> double {
    const raw v = Dbase[Lay::offset(k, li)];
    if constexpr (sizeof(S) == 2)
      return double(bf16_to_float(uint32_t(v)));
    else
      return double(v);
  };
  // the point's fields, all components, loaded up front
  double fv[NC][NF];
  if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int g0 = VO + c * 3;
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const int ra = g0 + F::pa(f), rb = g0 + F::pb(f);
        fv[c][f]     = ra == rb ? Dat(Shape::idx(ra, rb)) : Dat(Shape::idx(ra, rb)) + Dat(Shape::idx(rb, ra));
      }
      if constexpr (F::FULL) {
        fv[c][6] = Dat(Shape::idx(c, c));
#pragma unroll
        for (int d = 0; d < 3; d++) fv[c][7 + d] = Dat(Shape::idx(c, g0 + d)) + Dat(Shape::idx(g0 + d, c));
      }
    }
  }
  const double* tx = stab + 0;  // offset by table * NN below
  double        acc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    acc[c] = 0.0;
    __syncthreads();  // the tables (c = 0); the previous component's z reads (c > 0)
    if (active) {
#pragma unroll
      for (int f = 0; f < NF; f++) sv[f * NQ + li] = fv[c][f];
    }
    __syncthreads();
    double t[NF];
    if (active) {  // x: lane (jx, qy, qz)
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double* X = tx + F::table(f, 0) * NN;
        double        r = 0.0;
#pragma unroll
        for (int q = 0; q < N; q++) r += X[q * N + a0] * sv[f * NQ + (a2 * N + a1) * N + q];
        t[f] = r;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int f = 0; f < NF; f++) sv[f * NQ + li] = t[f];
    }
    __syncthreads();
    if (active) {  // y: lane (jx, jy, qz)
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double* Y = tx + F::table(f, 1) * NN;
        double        r = 0.0;
#pragma unroll
        for (int q = 0; q < N; q++) r += Y[q * N + a1] * sv[f * NQ + (a2 * N + q) * N + a0];
        t[f] = r;
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int f = 0; f < NF; f++) sv[f * NQ + li] = t[f];
    }
    __syncthreads();
    if (active) {  // z: lane = node (jx, jy, jz)
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double* Z = tx + F::table(f, 2) * NN;
#pragma unroll
        for (int q = 0; q < N; q++) acc[c] += Z[q * N + a2] * sv[f * NQ + (q * N + a1) * N + a0];
      }
    }
  }
  // scatter in address order (as the J.v's), through the element's field slots
  const int64_t node = active ? int64_t(P.map[size_t(slot) * NQ + li]) : 0;
  double*       out  = static_cast<double*>(P.r);
  __syncthreads();  // the last z stage has read the slots
  if (P.ordered) {
    scatter_element<N, NC, double, double>(out, acc, sv, sf + (el > 0 && el < EB ? el - 1 : 0) * NF * NQ, snode, el,
                                           EB, node, li, active, P.nnodes, P.ordering, P.scatter, P.ordered == 2);
  } else if (active) {
#pragma unroll
    for (int c = 0; c < NC; c++) scatter_add(out, lidx(node, c, NC, P.nnodes, P.ordering), acc[c], P.scatter);
  }
}

// Element gradient assembled into CSR values (element_gradient_kernel,
// fn/domain_integral_kernels.hpp:327-361, and Gradient::assemble,
// fn/functional.hpp:569-667, which scatters K_e through lookup tables):
//   K_e[(c,i)][(c',j)] = sum_q [w phi_i, w grad phi_i]_c . D(q) . [phi_j, grad phi_j]_c'
// One workgroup per element: the element's tangent is staged in LDS as FP64
// ([q][DS]), lane i owns test node i and loops over trial nodes j and
// quadrature points; each entry is added at its CSR position, found by binary
// search in the row's sorted columns.  Plain read-modify-write when the launch
// covers one element colour (deterministic), atomics otherwise.
template <int N, int NC, int LAYOUT, typename S, bool STAGE>
__global__ void __launch_bounds__(128) k_assemble(const ElemLaunch P)
{
  constexpr int NQ = N * N * N;
  using Shape      = TanShape<LAYOUT, NC>;
  constexpr bool VAL = Shape::VAL;
  constexpr int  DS  = Shape::DS;
  constexpr int  VO  = Shape::VO;
  constexpr int  TAB = 4 * N * N + N;
  using Lay          = TangentLayout<S, DS, NQ>;
  __shared__ double stab[TAB];
  extern __shared__ double sD[];  // [q][DS]

  const int     li   = threadIdx.x;
  const int64_t slot = P.slot_begin + blockIdx.x;
  for (int k = li; k < TAB; k += blockDim.x) stab[k] = P.tables[k];
  using raw       = typename StoreTraits<S>::raw;
  const raw* Dsrc = reinterpret_cast<const raw*>(static_cast<const char*>(P.D) + size_t(slot) * Lay::stride * sizeof(S));
  auto       Dat  = [&](int k, int q) -> double {
    const raw v = Dsrc[Lay::offset(k, q)];
    if constexpr (sizeof(S) == 2)
      return double(bf16_to_float(uint32_t(v)));
    else
      return double(v);
  };
  if constexpr (STAGE)
    for (int t = li; t < NQ * DS; t += blockDim.x) sD[t] = Dat(t % DS, t / DS);
  __syncthreads();
  if (li >= NQ) return;
  const double* tB  = stab;
  const double* tG  = stab + N * N;
  const double* tBw = stab + 2 * N * N;
  const double* tGw = stab + 3 * N * N;
  const int     ix = li % N, iy = (li / N) % N, iz = li / (N * N);
  const int64_t node_i = P.map[size_t(slot) * NQ + li];

  for (int jn = 0; jn < NQ; jn++) {
    const int jx = jn % N, jy = (jn / N) % N, jz = jn / (N * N);
    double    acc[NC][NC];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int c2 = 0; c2 < NC; c2++) acc[c][c2] = 0.0;
    for (int q = 0; q < NQ; q++) {
      const int qx = q % N, qy = (q / N) % N, qz = q / (N * N);
      // tangent entries of point q: staged in LDS, or read through the caches
      auto Dq = [&](int k) -> double {
        if constexpr (STAGE)
          return sD[q * DS + k];
        else
          return Dat(k, q);
      };
      const double  ug[3] = {tG[qx * N + jx] * tB[qy * N + jy] * tB[qz * N + jz],
                             tB[qx * N + jx] * tG[qy * N + jy] * tB[qz * N + jz],
                             tB[qx * N + jx] * tB[qy * N + jy] * tG[qz * N + jz]};
      const double  tg[3] = {tGw[qx * N + ix] * tBw[qy * N + iy] * tBw[qz * N + iz],
                             tBw[qx * N + ix] * tGw[qy * N + iy] * tBw[qz * N + iz],
                             tBw[qx * N + ix] * tBw[qy * N + iy] * tGw[qz * N + iz]};
      const double  uv = tB[qx * N + jx] * tB[qy * N + jy] * tB[qz * N + jz];
      const double  tv = tBw[qx * N + ix] * tBw[qy * N + iy] * tBw[qz * N + iz];
#pragma unroll
      for (int c = 0; c < NC; c++)
#pragma unroll
        for (int c2 = 0; c2 < NC; c2++) {
          double a = 0.0;
#pragma unroll
          for (int x = 0; x < 3; x++)
#pragma unroll
            for (int y = 0; y < 3; y++) a += tg[x] * Dq(Shape::idx(VO + c * 3 + x, VO + c2 * 3 + y)) * ug[y];
          if constexpr (VAL) {
            a += tv * Dq(Shape::idx(c, c2)) * uv;
#pragma unroll
            for (int y = 0; y < 3; y++)
              a += tv * Dq(Shape::idx(c, VO + c2 * 3 + y)) * ug[y] + tg[y] * Dq(Shape::idx(VO + c * 3 + y, c2)) * uv;
          }
          acc[c][c2] += a;
        }
    }
    const int64_t node_j = P.map[size_t(slot) * NQ + jn];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int64_t row = lidx(node_i, c, NC, P.nnodes, P.ordering);
#pragma unroll
      for (int c2 = 0; c2 < NC; c2++) {
        const int32_t colv = int32_t(lidx(node_j, c2, NC, P.nnodes, P.ordering));
        int64_t       lo = P.csr_row[row], hi = P.csr_row[row + 1];  // the column is in [lo, hi)
        while (hi - lo > 1) {
          const int64_t mid = (lo + hi) >> 1;
          if (P.csr_col[mid] <= colv)
            lo = mid;
          else
            hi = mid;
        }
        scatter_add(P.csr_val, lo, acc[c][c2], P.scatter);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launch + dispatch
// ---------------------------------------------------------------------------
template <int N, typename K>
inline hipError_t launch_grid(const ElemLaunch& P, K kernel)
{
  constexpr int EB = 256 / (N * N * N);
  const int64_t n  = P.slot_end - P.slot_begin;
  if (n <= 0) return hipSuccess;
  const unsigned grid = unsigned((n + EB - 1) / EB);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, P.stream, P);
  return hipGetLastError();
}

// k_jvp_xy: EB elements (waves) per workgroup.  With a BF16 tangent 2 per
// workgroup measured 1.5 % faster than 4 (fewer waves wait at each barrier); with
// an FP32 tangent 4 measured 3 % faster than 2 (64^3 p = 3 Neo-Hookean).
template <typename S>
constexpr int xy_elems_per_group()
{
  return sizeof(S) == 2 ? 2 : 4;
}
template <int EB, typename K>
inline hipError_t launch_grid_xy(const ElemLaunch& P, K kernel)
{
  const int64_t n = P.slot_end - P.slot_begin;
  if (n <= 0) return hipSuccess;
  const unsigned grid = unsigned((n + EB - 1) / EB);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * EB), 0, P.stream, P);
  return hipGetLastError();
}

template <int N, typename T, typename K>
inline hipError_t launch_grid_jvp(const ElemLaunch& P, K kernel)
{
  constexpr int WG = jvp_group_lanes<N, T>();
  constexpr int EB = WG / (N * N * N);
  const int64_t n  = P.slot_end - P.slot_begin;
  if (n <= 0) return hipSuccess;
  const unsigned grid = unsigned((n + EB - 1) / EB);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(WG), 0, P.stream, P);
  return hipGetLastError();
}

template <int N, int NC, int LAYOUT, typename S>
inline int dispatch_gradient(const ElemLaunch& P, hipError_t* err)
{
  if (P.mode == MODE_ASSEMBLE) {
    const int64_t n = P.slot_end - P.slot_begin;
    if (n <= 0) return 1;
    // the element's tangent is staged in LDS when it fits in 48 KB
    constexpr int    NQ    = N * N * N;
    constexpr size_t bytes = size_t(NQ) * TanShape<LAYOUT, NC>::DS * sizeof(double);
    constexpr bool   stage = bytes <= 48 * 1024;
    const unsigned   threads = unsigned((NQ + 63) / 64 * 64);
    hipLaunchKernelGGL((k_assemble<N, NC, LAYOUT, S, stage>), dim3(unsigned(n)), dim3(threads), stage ? bytes : 0,
                       P.stream, P);
    *err = hipGetLastError();
    return 1;
  }
  if (P.mode == MODE_DIAG) {
    *err = launch_grid<N>(P, k_diagonal<N, NC, LAYOUT, S>);
    return 1;
  }
  if (P.vprec == PREC_F64) {
    *err = launch_grid_jvp<N, double>(P, k_jvp<N, NC, LAYOUT, double, S, double>);
    return 1;
  }
  if (P.vprec == PREC_F32) {
    if constexpr (N == 4)
      *err = launch_grid_xy<xy_elems_per_group<S>()>(P, k_jvp_xy<float, NC, LAYOUT, S, float, xy_elems_per_group<S>()>);
    else
      *err = launch_grid_jvp<N, float>(P, k_jvp<N, NC, LAYOUT, float, S, float>);
    return 1;
  }
  if (P.vprec == PREC_MIXED) {
    if constexpr (N == 4)
      *err = launch_grid_xy<xy_elems_per_group<S>()>(P, k_jvp_xy<float, NC, LAYOUT, S, double, xy_elems_per_group<S>()>);
    else
      *err = launch_grid_jvp<N, float>(P, k_jvp<N, NC, LAYOUT, float, S, double>);
    return 1;
  }
  return 0;
}

template <int N, int NC, int LAYOUT>
inline int dispatch_gradient_prec(const ElemLaunch& P, hipError_t* err)
{
  switch (P.tprec) {
    case PREC_F64: return dispatch_gradient<N, NC, LAYOUT, double>(P, err);
    case PREC_F32: return dispatch_gradient<N, NC, LAYOUT, float>(P, err);
    case PREC_BF16: return dispatch_gradient<N, NC, LAYOUT, bf16>(P, err);
  }
  return 0;
}

template <int N, int NC, class QF, int LAYOUT>
inline int dispatch_eval(const ElemLaunch& P, hipError_t* err)
{
  if (P.mode == MODE_RESIDUAL) {
    *err = launch_grid<N>(P, k_element<N, NC, MODE_RESIDUAL, QF, LAYOUT, double>);
    return 1;
  }
  if (P.mode == MODE_SHAPE_JVP) {
    if constexpr (NC == 3) {
      *err = launch_grid<N>(P, k_element<N, NC, MODE_SHAPE_JVP, QF, LAYOUT, double>);
      return 1;
    }
    return 0;
  }
  switch (P.tprec) {
    case PREC_F64: *err = launch_grid<N>(P, k_element<N, NC, MODE_LINEARIZE, QF, LAYOUT, double>); return 1;
    case PREC_F32: *err = launch_grid<N>(P, k_element<N, NC, MODE_LINEARIZE, QF, LAYOUT, float>); return 1;
    case PREC_BF16: *err = launch_grid<N>(P, k_element<N, NC, MODE_LINEARIZE, QF, LAYOUT, bf16>); return 1;
  }
  return 0;
}

// Dispatch on (mode, qfunction, components, tangent layout, precisions) for one order.
template <int N>
inline int dispatch_order(const ElemLaunch& P, hipError_t* err)
{
  if (P.mode == MODE_JVP || P.mode == MODE_DIAG || P.mode == MODE_ASSEMBLE) {
    if (P.nc == 1) {
      if (P.layout == LAYOUT_GRAD_SYM) return dispatch_gradient_prec<N, 1, LAYOUT_GRAD_SYM>(P, err);
      if (P.layout == LAYOUT_FULL_SYM) return dispatch_gradient_prec<N, 1, LAYOUT_FULL_SYM>(P, err);
    }
    if (P.nc == 3) {
      if (P.layout == LAYOUT_GRAD_SYM) return dispatch_gradient_prec<N, 3, LAYOUT_GRAD_SYM>(P, err);
      if (P.layout == LAYOUT_FULL_SYM) return dispatch_gradient_prec<N, 3, LAYOUT_FULL_SYM>(P, err);
      if (P.layout == LAYOUT_GRAD_FULL) return dispatch_gradient_prec<N, 3, LAYOUT_GRAD_FULL>(P, err);
    }
    return 0;
  }
  switch (P.qf) {
    case QF_ID_HEAT:
      if (P.nc == 1) return dispatch_eval<N, 1, QFHeat, LAYOUT_GRAD_SYM>(P, err);
      break;
    case QF_ID_MASS_DIFFUSION:
      if (P.nc == 1) return dispatch_eval<N, 1, QFMassDiffusion<1>, LAYOUT_FULL_SYM>(P, err);
      if (P.nc == 3) return dispatch_eval<N, 3, QFMassDiffusion<3>, LAYOUT_FULL_SYM>(P, err);
      break;
    case QF_ID_NEOHOOKEAN:
      if (P.nc == 3) return dispatch_eval<N, 3, QFNeoHookean, LAYOUT_GRAD_SYM>(P, err);
      break;
    case QF_ID_LINEAR_ELASTIC:
      if (P.nc == 3) return dispatch_eval<N, 3, QFLinearElastic, LAYOUT_GRAD_SYM>(P, err);
      break;
    case QF_ID_J2:
      if (P.nc == 3) return dispatch_eval<N, 3, QFJ2, LAYOUT_GRAD_FULL>(P, err);
      break;
    case QF_ID_QOI_VALUE:
      // residual only: the QoI gradient is never linearized
      if (P.mode != MODE_RESIDUAL) return 0;
      if (P.nc == 1) {
        *err = launch_grid<N>(P, k_element<N, 1, MODE_RESIDUAL, QFQoiValue<1>, LAYOUT_GRAD_SYM, double>);
        return 1;
      }
      if (P.nc == 3) {
        *err = launch_grid<N>(P, k_element<N, 3, MODE_RESIDUAL, QFQoiValue<3>, LAYOUT_GRAD_SYM, double>);
        return 1;
      }
      break;
  }
  return 0;
}

}  // namespace sxk
