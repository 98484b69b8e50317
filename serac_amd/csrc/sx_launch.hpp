// sx_launch.hpp -- launch descriptors shared by the host library and the
// per-order kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sxk {

// MODE_SHAPE_JVP: the derivative of the residual with respect to the shape
// displacement in direction `dshape`, recomputed by forward-mode duals
enum Mode { MODE_RESIDUAL = 0, MODE_LINEARIZE = 1, MODE_JVP = 2, MODE_DIAG = 3, MODE_ASSEMBLE = 4, MODE_SHAPE_JVP = 5 };

// Stored-tangent layouts:
//   GRAD_SYM:  d(parent flux)/d(parent grad), M = 3 nc, symmetric, upper
//              triangle packed (heat: 6 entries, Neo-Hookean: 45)
//   FULL_SYM:  d[source, parent flux]/d[value, parent grad], M = 4 nc, symmetric
//   GRAD_FULL: d(parent flux)/d(parent grad), M = 3 nc, all M^2 entries
//              (J2 plasticity: 81)
enum Layout { LAYOUT_GRAD_SYM = 0, LAYOUT_FULL_SYM = 1, LAYOUT_GRAD_FULL = 2 };

// precisions of the stored tangent (tprec) and of J.v vectors (vprec);
// PREC_MIXED (J.v only): FP64 vectors, FP32 element arithmetic
enum Precision { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16 = 2, PREC_MIXED = 3 };

enum QFId {
  QF_ID_HEAT           = 1,
  QF_ID_MASS_DIFFUSION = 2,
  QF_ID_NEOHOOKEAN     = 3,
  QF_ID_LINEAR_ELASTIC = 4,
  QF_ID_J2             = 5,
  QF_ID_QOI_VALUE      = 6  // internal: the gradient of the QoI int u_c dV
};

struct ElemLaunch {
  int           order, nc, qf, mode, layout, ordering;
  int           tprec;      // Precision of the stored tangent
  int           vprec;      // Precision of the J.v vectors (residual/linearize: FP64)
  int           scatter;    // 0: plain read-modify-write (colour launches), 1: atomics
  int           xcd_remap;  // 1: give each XCD a contiguous range of element blocks
  int           ordered;    // 1: scatter each element's values in address order (through LDS)
  const int*    map;        // [slot][npe] global node ids
  const double* coords;     // [slot][8 corners][3], corner = a + 2b + 4c
  int64_t       slot_begin, slot_end;
  int64_t       nnodes;
  const void*   u;       // input L-vector (trial), FP64 or (J.v) vprec
  void*         r;       // output L-vector (test), accumulated into
  void*         D;       // stored tangent, per-slot layout of sx_element.hpp TangentLayout
  const double* tables;  // B, G, Bw, Gw [Q][N] and GL points [Q]
  double*       qdata;         // quadrature-point state [row][k][q] (residual/linearize)
  const int*    qrow;          // [slot] -> state row (the element's position in its domain)
  int           update_state;  // write the updated state back
  const int64_t* csr_row;      // assembly: CSR row offsets (test DOFs)
  const int32_t* csr_col;      // assembly: sorted column ids per row (trial DOFs)
  double*        csr_val;      // assembly: values, accumulated into
  const double*  shape;        // residual/linearize: shape displacement p (3-component L-vector of the
                               // test space) or null; the geometry becomes X + p
  const double*  dshape;       // MODE_SHAPE_JVP: the direction of the shape displacement
  const double*  stop;         // J.v: when non-null and *stop != 0 the kernel returns at once
  double        params[16];
  double        t;
  hipStream_t   stream;
};

inline int tangent_m(int nc, int layout) { return layout == LAYOUT_FULL_SYM ? 4 * nc : 3 * nc; }
inline int tangent_ds(int nc, int layout)
{
  int m = tangent_m(nc, layout);
  return layout == LAYOUT_GRAD_FULL ? m * m : m * (m + 1) / 2;
}
inline int precision_bytes(int prec) { return prec == PREC_F64 ? 8 : prec == PREC_F32 ? 4 : 2; }
// bytes of tangent per slot: DS * NQ entries rounded up to 16 bytes
inline int64_t tangent_slot_bytes(int ds, int nq, int prec)
{
  return (int64_t(ds) * nq * precision_bytes(prec) + 15) / 16 * 16;
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
