// sx_launch.hpp -- launch descriptors shared by the host library and the
// per-order kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sxk {

enum Mode { MODE_RESIDUAL = 0, MODE_LINEARIZE = 1, MODE_JVP = 2, MODE_DIAG = 3 };

// Stored-tangent layouts (symmetric, upper triangle packed):
//   GRAD_SYM: d(parent flux)/d(parent grad), M = 3 nc   (heat: 6, Neo-Hookean: 45)
//   FULL_SYM: d[source, parent flux]/d[value, parent grad], M = 4 nc
enum Layout { LAYOUT_GRAD_SYM = 0, LAYOUT_FULL_SYM = 1 };

enum QFId { QF_ID_HEAT = 1, QF_ID_MASS_DIFFUSION = 2, QF_ID_NEOHOOKEAN = 3, QF_ID_LINEAR_ELASTIC = 4 };

struct ElemLaunch {
  int           order, nc, qf, mode, layout, ordering;
  int           scatter;    // 0: plain read-modify-write (colour launches), 1: FP64 atomics
  int           xcd_remap;  // 1: give each XCD a contiguous range of element blocks
  const int*    map;     // [slot][npe] global node ids
  const double* coords;  // [slot][8 corners][3], corner = a + 2b + 4c
  int64_t       slot_begin, slot_end;
  int64_t       nnodes;
  const double* u;       // input L-vector (trial)
  double*       r;       // output L-vector (test), accumulated into
  double*       D;       // stored tangent [slot][DS][Q^3]
  const double* tables;  // B, G, Bw, Gw [Q][N] and GL points [Q]
  double        params[8];
  double        t;
  hipStream_t   stream;
};

inline int tangent_m(int nc, int layout) { return layout == LAYOUT_FULL_SYM ? 4 * nc : 3 * nc; }
inline int tangent_ds(int nc, int layout)
{
  int m = tangent_m(nc, layout);
  return m * (m + 1) / 2;
}

// Tangent storage per slot (element): the DS packed entries of every
// quadrature point q, grouped in pairs so that lane q reads 16 bytes at once:
//   entry k < 2*floor(DS/2):  [(k/2)][q][k%2]
//   last entry when DS odd:   [floor(DS/2)][q]  (after the pairs)
// The slot stride is DS*NQ rounded up to even (16-byte alignment).
__host__ __device__ constexpr int64_t tangent_slot_stride(int ds, int nq) { return (int64_t(ds) * nq + 1) & ~int64_t(1); }
__host__ __device__ constexpr int tangent_offset(int k, int q, int ds, int nq)
{
  return (k < (ds & ~1)) ? (k >> 1) * 2 * nq + q * 2 + (k & 1) : (ds >> 1) * 2 * nq + q;
}

// defined in sx_kernels_p{1,2,3,4}.hip; return 0 when the combination is not compiled
int launch_element_p1(const ElemLaunch& P, hipError_t* err);
int launch_element_p2(const ElemLaunch& P, hipError_t* err);
int launch_element_p3(const ElemLaunch& P, hipError_t* err);
int launch_element_p4(const ElemLaunch& P, hipError_t* err);

inline int launch_element(const ElemLaunch& P, hipError_t* err)
{
  switch (P.order) {
    case 1: return launch_element_p1(P, err);
    case 2: return launch_element_p2(P, err);
    case 3: return launch_element_p3(P, err);
    case 4: return launch_element_p4(P, err);
  }
  return 0;
}

}  // namespace sxk
