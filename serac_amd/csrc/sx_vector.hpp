// sx_vector.hpp -- host launchers of the vector kernels in sx_vector.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sxk {

hipError_t vec_fill(double* x, double v, int64_t n, hipStream_t s);
hipError_t vec_fill_f32(float* x, float v, int64_t n, hipStream_t s);
hipError_t vec_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s);
// number of partial sums the reductions use (size of the `part` scratch)
int reduction_blocks();
// scal[slot] = (a, b); op 1 then sets scal[2] = scal[0]/scal[1]
hipError_t vec_dot_dev(const double* a, const double* b, int64_t n, double* part, double* scal, int slot, int op,
                       hipStream_t s);
// Jacobi-PCG in two kernels per iteration, each reducing the previous
// kernel's reduction_blocks() partials itself (in the fixed order of
// k_reduce_final's sum):
//   dot_partial: part = partials of (a, b)
//   cg_step_x:   alpha = scal[0] / sum(part_den); x += alpha d; r -= alpha Ad;
//                z = r/diag; Ad = 0; part_rz = partials of (r, z);
//                scal[1] = den, scal[2] = alpha, scal[5] = scal[0]
//   cg_step_d:   betanom = sum(part_rz); d = z + betanom / scal[5] d;
//                scal[0] = scal[3] = betanom, scal[4] = beta; sets stop =
//                {code, it} with code 1 (betanom <= goal: converged), 2
//                (betanom < 0: preconditioner not positive definite) or 3
//                (betanom not a number)
// Both leave everything alone once stop[0] is set (stop: device {code, iteration}).
hipError_t dot_partial(const double* a, const double* b, int64_t n, double* part, hipStream_t s);
hipError_t cg_step_x(double* x, double* r, double* z, double* Ad, const double* d, const double* diag, double* scal,
                     const double* part_den, double* part_rz, int64_t n, const double* stop, hipStream_t s);
hipError_t cg_step_d(double* d, const double* z, double* scal, const double* part_rz, int64_t n, double goal, int it,
                     double* stop, hipStream_t s);
// z = r/diag; scal[slot] = (r, z)
hipError_t jacobi_dot(const double* r, const double* diag, double* z, int64_t n, double* part, double* scal,
                      int slot, hipStream_t s);
// GMRES building blocks.  multi_dot: out[j] (+)= (V_j, w) for j < k, V_j at
// V + j ld; part needs k * multi_dot_blocks() doubles.  multi_axpy:
// y += sign * sum_j c[j] V_j (c in device memory).  scale_copy: y = a x.
// jacobi_apply: z = x / diag.
int        multi_dot_blocks();
hipError_t multi_dot(const double* V, int64_t ld, int k, const double* w, int64_t n, double* part, double* out,
                     bool accumulate, hipStream_t s);
hipError_t multi_axpy(const double* V, int64_t ld, int k, const double* c_dev, double sign, double* y, int64_t n,
                      hipStream_t s);
hipError_t scale_copy(double a, const double* x, double* y, int64_t n, hipStream_t s);
hipError_t jacobi_apply(const double* x, const double* diag, double* z, int64_t n, hipStream_t s);
// GMRES(m) least-squares state on the device, gmres_state_doubles(m) doubles:
//   H [(m + 1) x m, column-major] | cs [m] | sn [m] | g [m + 1] | st [8]
// st: 0 |g_{j+1}| (the residual estimate), 1 1/h_{j+1,j} (0 at breakdown),
//     2 columns done this cycle, 3 stop (0 running, 1 converged or breakdown,
//     3 non-finite estimate), 4 goal.
// gmres_cycle_start: g = beta e_0, st = {beta, 0, 0, 0, goal}.
// gmres_column: column j of H from the two Gram-Schmidt passes hh[0..j] +
//   hh[m+1..m+1+j] and |w|^2 = hh[2(m+1)]; the previous rotations, the one that
//   zeroes H(j+1, j), g; sets st.  Leaves everything alone once st[3] is set.
// gmres_scale_inv: y *= st[1] (normalises the new Arnoldi vector).
// gmres_solve_y: y = H^-1 g over the st[2] columns done (upper triangular).
size_t     gmres_state_doubles(int m);
hipError_t gmres_cycle_start(double* gs, int m, double beta, double goal, hipStream_t s);
hipError_t gmres_column(const double* hh, int j, int m, double* gs, hipStream_t s);
hipError_t gmres_scale_inv(double* y, int64_t n, int m, const double* gs, hipStream_t s);
hipError_t gmres_solve_y(const double* gs, int m, double* y, hipStream_t s);
hipError_t set_rows(double* x, const int32_t* rows, int64_t nr, double v, hipStream_t s);
// x[rows[i]] = vals[i]
hipError_t set_rows_values(double* x, const int32_t* rows, const double* vals, int64_t nr, hipStream_t s);
// y = a x + b y  (elementwise, FP64)
hipError_t vec_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);
hipError_t copy_rows(double* dst, const double* src, const int32_t* rows, int64_t nr, hipStream_t s);
hipError_t gather_e(const int* map, const double* L, double* E, int64_t ne, int npe, int nc, int64_t nnodes,
                    int ordering, hipStream_t s);
hipError_t scatter_e(const int64_t* offs, const int32_t* ents, const double* E, double* L, int64_t nnodes, int npe,
                     int nc, int ordering, hipStream_t s);

// E-vector layouts: dir 0 copies the reference layout [e][c][j] into the
// component-major [c][e][j], dir 1 back
hipError_t e_transpose(const double* in, double* out, int64_t ne, int nc, int npe, int dir, hipStream_t s);

// y = A x, A in CSR (int64 row offsets, int32 columns)
hipError_t csr_spmv(const int64_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                    int64_t nrows, hipStream_t s);

}  // namespace sxk
