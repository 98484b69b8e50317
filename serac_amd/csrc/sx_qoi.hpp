// sx_qoi.hpp -- launch interface of the quantity-of-interest value kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sxk {

// ids of the compiled QoI integrands (the values of SX_QOI_* in serac_amd.h)
enum QoiId { QOI_ID_MEASURE = 1, QOI_ID_POSITION = 2, QOI_ID_VALUE = 3, QOI_ID_NEOHOOKEAN_ENERGY = 4, QOI_ID_INNER = 5 };

struct QoiLaunch {
  int           order, nc, qoi, ordering;
  const int*    map;      // [slot][npe] global node ids (faces: [slot][Q*Q])
  const double* coords;   // [slot][8 corners][3] (faces: [slot][4][3])
  int64_t       nslots;
  int64_t       nnodes;
  const double* u;        // FP64 L-vector of the trial space
  const double* u2;       // QOI_ID_INNER: the second trial argument (same space layout)
  const double* tables;   // B, G, Bw, Gw [Q][N] and GL points [Q]
  double        params[16];
  double*       partial;  // [slot]: the element's sum of q detJ w
  hipStream_t   stream;
};

hipError_t launch_qoi(const QoiLaunch& P);
// boundary faces: map = [slot][Q*Q] face nodes (b Q + a), coords = [slot][4][3]
// corners (a + 2 b); measure, position and value QoIs
hipError_t launch_qoi_face(const QoiLaunch& P);

}  // namespace sxk
