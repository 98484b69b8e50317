// sx_boundary.hpp -- launch descriptor of the boundary-face kernels in
// sx_boundary.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sxk {

enum FaceMode { FMODE_RESIDUAL = 0, FMODE_LINEARIZE = 1, FMODE_JVP = 2, FMODE_DIAG = 3, FMODE_ASSEMBLE = 4 };

// boundary qfunction ids (the values of SX_BQF_* in serac_amd.h)
enum BQFId { BQF_ID_TRACTION = 101, BQF_ID_PRESSURE = 102, BQF_ID_FLUX = 103 };

struct FaceLaunch {
  int            order, nc, qf, mode, ordering;
  int            vprec;          // J.v vectors: 0 FP64, 1 FP32
  const int*     fmap;           // [slot][Q*Q] node ids, xi1 fastest
  const double*  corners;        // [slot][4][3], corner a + 2b at (xi1, xi2) = (a, b)
  int64_t        slot_begin, slot_end;
  int64_t        nnodes;
  const void*    u;              // input L-vector
  void*          r;              // output L-vector, accumulated into
  double*        D;              // tangent [slot][nc * 3nc][Q*Q], FP64
  const double*  tables;         // B, G, Bw, Gw [Q][N] and GL points [Q]
  const int64_t* csr_row;        // assembly
  const int32_t* csr_col;
  double*        csr_val;
  double         params[16];
  const double*  shape;          // residual/linearize: shape displacement (3-component L-vector) or null
  hipStream_t    stream;
};

// returns 0 when the (order, qfunction, components) combination is not compiled
int launch_face(const FaceLaunch& P, hipError_t* err);

}  // namespace sxk
