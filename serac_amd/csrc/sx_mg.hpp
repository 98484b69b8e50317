// sx_mg.hpp -- kernels of the geometric multigrid V-cycle on the vertex grids
// of a structured (cuboid) mesh (sx_mg.hip).
//
// Level 0 is the order-p nodal grid (matrix-free operator); level 1 is the
// mesh's vertex grid and each further level halves the intervals.  Vectors on
// the vertex levels are [z][y][x][3] in FP64.  A level operator is a 27-point
// stencil of 3x3 blocks stored structure-of-arrays: S[(o * 9 + co * 3 + ci) *
// nv + v] couples output component co of vertex v with input component ci of
// vertex v + o, o = (ox, oy, oz) in {-1, 0, 1}^3 numbered (oz + 1) * 9 + (oy +
// 1) * 3 + (ox + 1).  `mask` marks the constrained entries of a level (one
// byte per vertex component): their rows and columns of S are zero and the
// level's vectors hold zero there.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sx_unet.hpp"

namespace sxm {

struct Grid3 {
  int n[3];  // vertices per direction
};

// level 0 <-> 1: rc = P^T r on the vertex grid (zero where mask is set), in
// three passes, one direction each, through scratch t1 and t2 of
// fe_restrict_scratch(G, 0) and (G, 1) doubles; z += P ec on the nodal grid
// (P: the order-p interpolation of the vertex field)
int64_t    fe_restrict_scratch(const sxu::FEGrid& G, int pass);
hipError_t fe_restrict(const sxu::FEGrid& G, const double* r, double* rc, const uint8_t* mask, double* t1, double* t2,
                       hipStream_t s);
hipError_t fe_prolong_add(const sxu::FEGrid& G, const double* ec, double* z, hipStream_t s);
// vertex level l <-> l + 1 (trilinear): rc = P^T rf (zero where maskc is set);
// zf += P ec (skipping entries where maskf is set)
hipError_t grid_restrict(const double* rf, Grid3 f, double* rc, Grid3 c, const uint8_t* maskc, hipStream_t s);
hipError_t grid_prolong_add(const double* ec, Grid3 c, double* zf, Grid3 f, const uint8_t* maskf, hipStream_t s);
// y = S x
hipError_t stencil_apply(const double* S, Grid3 g, const double* x, double* y, hipStream_t s);
// probing: e = 1 on component ci of the vertices whose index is c (mod 3) in
// every direction (and not masked), else 0;  extract writes y (= P^T A P e on
// the coarse grid) into the stencil entries that colour determines
hipError_t probe_fill(double* e, Grid3 g, int cx, int cy, int cz, int ci, const uint8_t* mask, hipStream_t s);
hipError_t probe_extract(const double* y, Grid3 g, int cx, int cy, int cz, int ci, const uint8_t* mask, double* S,
                         hipStream_t s);
// dinv = 1 / (the stencil's diagonal), 0 where masked
hipError_t stencil_dinv(const double* S, Grid3 g, const uint8_t* mask, double* dinv, hipStream_t s);
// dinv = 1 / diag, 0 on the listed rows
hipError_t diag_inverse(const double* diag, int64_t n, const int32_t* rows, int64_t nr, double* dinv, hipStream_t s);
// mask[rows] = 1 in a nodal-grid sized mask, then the vertex mask: a vertex
// component is constrained when its own node's component is
hipError_t vertex_mask(const sxu::FEGrid& G, const uint8_t* node_mask, uint8_t* vmask, hipStream_t s);
hipError_t rows_mask(const int32_t* rows, int64_t nr, uint8_t* mask, hipStream_t s);
// vmask of the coarse level: the constraint of the coinciding fine vertex
hipError_t coarsen_mask(const uint8_t* mf, Grid3 f, uint8_t* mc, Grid3 c, hipStream_t s);

// Chebyshev smoothing (x0 = 0 or not):
//   cheb_init:  d = dinv b / theta; x = d (x_add = 0) or x += d; r = b (copy_r)
//   cheb_step:  r -= t; d = c1 d + c2 dinv r; x += d
//   residual:   r = b - t;  sub_inplace: r -= t
hipError_t cheb_init(const double* b, const double* dinv, double theta, double* x, double* d, double* r, int x_add,
                     int copy_r, int64_t n, hipStream_t s);
hipError_t cheb_step(double* r, const double* t, double* d, const double* dinv, double c1, double c2, double* x,
                     int64_t n, hipStream_t s);
hipError_t residual(const double* b, const double* t, double* r, int64_t n, hipStream_t s);
hipError_t sub_inplace(double* r, const double* t, int64_t n, hipStream_t s);
// zeroes the masked entries of x
hipError_t apply_mask(double* x, const uint8_t* mask, int64_t n, hipStream_t s);
// x = A b for a dense row-major n x n A
hipError_t dense_mv(const double* A, int n, const double* b, double* x, hipStream_t s);
// PCG half-step without the Jacobi solve: den = sum(part_den), alpha =
// scal[0] / den; x += alpha d; r -= alpha Ad; Ad = 0; block 0 records
// scal[1] = den, scal[2] = alpha, scal[5] = scal[0]; nothing once stop[0] is set
hipError_t pcg_step_xr(double* x, double* r, double* Ad, const double* d, double* scal, const double* part_den,
                       int nb, int64_t n, const double* stop, hipStream_t s);
// vertex-level Galerkin operator from the stored GRAD_SYM tangent of one
// 3-component domain integral (FP64, or FP32 with tprec_f32), accumulated into
// S (atomics); map: [slot][(order + 1)^3] node ids of the lexicographic nodal
// grid gn; tables: the integral's B, G, Bw, Gw and GL points
hipError_t q1_galerkin(int order, int tprec_f32, const int* map, const void* D, int64_t nslots, const double* tables,
                      Grid3 gn, Grid3 gv, double* S, hipStream_t s);
// a BF16 copy of a stored tangent (ds = 45 or 81 entries per point, FP64 or
// FP32 source with src_f32) in the BF16 layout; dst holds
// bf16_tangent_elems(order, ds, nslots) entries
int64_t    bf16_tangent_elems(int order, int ds, int64_t nslots);
hipError_t tangent_to_bf16(int order, int ds, int src_f32, const void* src, uint16_t* dst, int64_t nslots,
                           hipStream_t s);
// Sc = P^T Sf P for the trilinear transfer from coarse grid c to fine grid f
// (writes every coarse entry; constrained coarse rows/columns still need
// stencil_mask)
hipError_t stencil_rap(const double* Sf, Grid3 f, double* Sc, Grid3 c, hipStream_t s);
// zero the stencil entries with a constrained row or column
hipError_t stencil_mask(double* S, Grid3 g, const uint8_t* mask, hipStream_t s);
// *flag = 1 (atomic max) when a listed constrained DOF has an unconstrained
// vertex component whose hat is nonzero at it
hipError_t vertex_compatible(const sxu::FEGrid& G, const int32_t* rows, int64_t nr, const uint8_t* vmask, int* flag,
                             hipStream_t s);
// FP32 level-0 vectors for the V-cycle's smoothing (the same operations, FP32
// storage and arithmetic; transfers accumulate in FP64); to_f32 zeroes the
// masked entries
hipError_t fe_restrict(const sxu::FEGrid& G, const float* r, double* rc, const uint8_t* mask, double* t1, double* t2,
                       hipStream_t s);
hipError_t fe_prolong_add(const sxu::FEGrid& G, const double* ec, float* z, hipStream_t s);
hipError_t cheb_init(const float* b, const float* dinv, double theta, float* x, float* d, float* r, int x_add,
                     int copy_r, int64_t n, hipStream_t s);
hipError_t cheb_step(float* r, const float* t, float* d, const float* dinv, double c1, double c2, float* x, int64_t n,
                     hipStream_t s);
hipError_t residual(const float* b, const float* t, float* r, int64_t n, hipStream_t s);
hipError_t sub_inplace(float* r, const float* t, int64_t n, hipStream_t s);
hipError_t apply_mask(float* x, const uint8_t* mask, int64_t n, hipStream_t s);
hipError_t to_f32(const double* x, float* y, const uint8_t* mask, int64_t n, hipStream_t s);
hipError_t to_f64(const float* x, double* y, int64_t n, hipStream_t s);
// power iteration helpers: v = dinv t / |dinv t| (norms on the host)
hipError_t mul_dinv(const double* dinv, const double* t, double* v, int64_t n, hipStream_t s);
hipError_t scale(double* v, double a, int64_t n, hipStream_t s);
hipError_t seed_field(double* v, int64_t n, uint64_t seed, const uint8_t* mask, hipStream_t s);

}  // namespace sxm
