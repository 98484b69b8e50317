// sx_vector.hip -- L/E-vector kernels: fills, AXPYs, deterministic FP64
// reductions, E-vector gather / scatter-add, essential-row edits and the fused
// Jacobi-PCG update kernels.  All are HBM-bound streaming kernels; grids are
// sized for 256 CUs with grid-stride loops.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sx_vector.hpp"

namespace sxk {

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(int64_t n)
{
  int64_t g = (n + kThreads - 1) / kThreads;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return unsigned(g);
}

__global__ void __launch_bounds__(kThreads) k_fill(double* x, double v, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) x[i] = v;
}

__global__ void __launch_bounds__(kThreads) k_fill_f32(float* x, float v, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) x[i] = v;
}

__global__ void __launch_bounds__(kThreads) k_axpy(double a, const double* x, double* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] += a * x[i];
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
  // wave64 shuffle tree, then the 4 wave partials in LDS (fixed order)
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return s;
}

__global__ void __launch_bounds__(kThreads) k_dot_partial(const double* a, const double* b, int64_t n, double* part)
{
  __shared__ double sh[4];
  double            s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    s += a[i] * b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Final reduction of nb partials into scal[slot]; op 1 then sets
// scal[2] = scal[0] / scal[1]
__global__ void __launch_bounds__(kThreads) k_reduce_final(const double* part, int nb, double* scal, int slot, int op)
{
  __shared__ double sh[4];
  double            s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    scal[slot] = s;
    if (op == 1) scal[2] = scal[0] / scal[1];
  }
}

// Sum of the nb reduction partials, the same fixed order as k_reduce_final
// (every block of a kernel gets the same bits)
__device__ __forceinline__ double sum_partials(const double* part, int nb, double* sh, double* bc)
{
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) *bc = s;
  __syncthreads();
  return *bc;
}

// First half of a Jacobi-PCG iteration: den = sum of part_den (the partials
// of (d, Ad)), alpha = scal[0] / den; x += alpha d; r -= alpha Ad; z = r /
// diag; Ad = 0 for the next J.v; partials of (r, z) into part_rz.  Block 0
// records scal[1] = den, scal[2] = alpha and scal[5] = scal[0].
__global__ void __launch_bounds__(kThreads) k_cg_step_x(double* x, double* r, double* z, double* Ad, const double* d,
                                                         const double* diag, double* scal, const double* part_den,
                                                         double* part_rz, int nb, int64_t n, const double* stop)
{
  __shared__ double sh[4];
  __shared__ double bc;
  if (stop[0] != 0.0) return;  // the whole block leaves before any barrier
  const double den   = sum_partials(part_den, nb, sh, &bc);
  const double nom   = scal[0];
  const double alpha = nom / den;
  double       s     = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    x[i] += alpha * d[i];
    const double ri = r[i] - alpha * Ad[i];
    Ad[i]           = 0.0;
    r[i]            = ri;
    const double zi = ri / diag[i];
    z[i]            = zi;
    s += ri * zi;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = s;
    if (blockIdx.x == 0) {
      scal[1] = den;
      scal[2] = alpha;
      scal[5] = nom;
    }
  }
}

// Second half: betanom = sum of part_rz, beta = betanom / scal[5]; d = z +
// beta d.  Block 0 records scal[0] = scal[3] = betanom, scal[4] = beta and the
// stop flag {code, it}: code 1 when betanom is at most goal, 2 when betanom is
// negative (the preconditioner is not positive definite: mfem's CGSolver stops
// unconverged there), 3 when it is not a number.  Block 0 writes the flag while
// other blocks of this launch may be starting, so each block reads it once,
// through LDS, and leaves as a whole; a block that leaves skips only the d
// update, which no later step reads.
__global__ void __launch_bounds__(kThreads) k_cg_step_d(double* d, const double* z, double* scal,
                                                         const double* part_rz, int nb, int64_t n, double goal, int it,
                                                         double* stop)
{
  __shared__ double sh[4];
  __shared__ double bc;
  __shared__ int    stopped;
  if (threadIdx.x == 0) stopped = __hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0.0;
  __syncthreads();
  if (stopped) return;  // block-uniform
  const double betanom = sum_partials(part_rz, nb, sh, &bc);
  const double beta    = betanom / scal[5];
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    d[i] = z[i] + beta * d[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[0] = betanom;
    scal[3] = betanom;
    scal[4] = beta;
    const double code = betanom != betanom ? 3.0 : betanom < 0.0 ? 2.0 : betanom <= goal ? 1.0 : 0.0;
    if (code != 0.0) {
      stop[1] = double(it);
      __hip_atomic_store(stop, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// z = r / diag; partial (r, z)
__global__ void __launch_bounds__(kThreads) k_jacobi(const double* r, const double* diag, double* z, int64_t n,
                                                      double* part)
{
  __shared__ double sh[4];
  double            s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const double zi = r[i] / diag[i];
    z[i]            = zi;
    s += r[i] * zi;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// part[j * gridDim.x + bx] = partial (V_j, w) over block bx's grid-stride
// range; grid = (nb, k), V = k vectors of length n at stride ld
__global__ void __launch_bounds__(kThreads) k_multi_dot_partial(const double* V, int64_t ld, const double* w, int64_t n,
                                                                double* part)
{
  __shared__ double sh[4];
  const double*     v = V + int64_t(blockIdx.y) * ld;
  double            s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    s += v[i] * w[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[int64_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// out[j] (+)= sum of row j of part[k][nb], one block per row, fixed order
__global__ void __launch_bounds__(kThreads) k_reduce_rows(const double* part, int nb, double* out, int accumulate)
{
  __shared__ double sh[4];
  const double*     row = part + int64_t(blockIdx.x) * nb;
  double            s   = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) s += row[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = accumulate ? out[blockIdx.x] + s : s;
}

// y += sign * sum_j c[j] V_j, coefficients read from device memory
__global__ void __launch_bounds__(kThreads) k_multi_axpy(const double* V, int64_t ld, int k, const double* c,
                                                         double sign, double* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    double s = 0.0;
    for (int j = 0; j < k; j++) s += c[j] * V[int64_t(j) * ld + i];
    y[i] += sign * s;
  }
}

// y = a x
__global__ void __launch_bounds__(kThreads) k_scale_copy(double a, const double* x, double* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] = a * x[i];
}

// ---- GMRES least-squares state (layout in sx_vector.hpp) ----
__device__ __forceinline__ double* gm_st(double* gs, int m) { return gs + size_t(m + 1) * m + 2 * size_t(m) + m + 1; }
__device__ __forceinline__ const double* gm_st(const double* gs, int m)
{
  return gs + size_t(m + 1) * m + 2 * size_t(m) + m + 1;
}

__global__ void k_gmres_cycle_start(double* gs, int m, double beta, double goal)
{
  double* g  = gs + size_t(m + 1) * m + 2 * size_t(m);
  double* st = gm_st(gs, m);
  for (int i = threadIdx.x; i <= m; i += blockDim.x) g[i] = i == 0 ? beta : 0.0;
  if (threadIdx.x == 0) {
    st[0] = beta;
    st[1] = 0.0;
    st[2] = 0.0;
    st[3] = 0.0;
    st[4] = goal;
  }
}

// one thread: the rotations of a column are sequential
__global__ void k_gmres_column(const double* hh, int j, int m, double* gs)
{
  double* H  = gs;
  double* cs = gs + size_t(m + 1) * m;
  double* sn = cs + m;
  double* g  = sn + m;
  double* st = gm_st(gs, m);
  if (st[3] != 0.0) return;
  double* col = H + size_t(j) * (m + 1);
  for (int i = 0; i <= j; i++) col[i] = hh[i] + hh[m + 1 + i];
  const double hnext = sqrt(fmax(hh[2 * (m + 1)], 0.0));
  col[j + 1]         = hnext;
  for (int i = 0; i < j; i++) {
    const double a = col[i], c = col[i + 1];
    col[i]         = cs[i] * a + sn[i] * c;
    col[i + 1]     = -sn[i] * a + cs[i] * c;
  }
  const double a = col[j], c = col[j + 1], den = hypot(a, c);
  cs[j]            = den > 0.0 ? a / den : 1.0;
  sn[j]            = den > 0.0 ? c / den : 0.0;
  col[j]           = den;
  col[j + 1]       = 0.0;
  g[j + 1]         = -sn[j] * g[j];
  g[j]             = cs[j] * g[j];
  const double est = fabs(g[j + 1]);
  st[0]            = est;
  st[1]            = hnext > 0.0 ? 1.0 / hnext : 0.0;
  st[2]            = double(j + 1);
  if (!(est == est) || isinf(est))
    st[3] = 3.0;
  else if (est <= st[4] || hnext == 0.0)
    st[3] = 1.0;
}

__global__ void __launch_bounds__(kThreads) k_gmres_scale_inv(double* y, int64_t n, const double* st)
{
  const double a = st[1];
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) y[i] *= a;
}

// one thread: back substitution over the columns done
__global__ void k_gmres_solve_y(const double* gs, int m, double* y)
{
  const double* H    = gs;
  const double* g    = gs + size_t(m + 1) * m + 2 * size_t(m);
  const double* st   = gm_st(gs, m);
  const int     jend = int(st[2]);
  for (int i = jend - 1; i >= 0; i--) {
    double acc = g[i];
    for (int k = i + 1; k < jend; k++) acc -= H[size_t(k) * (m + 1) + i] * y[k];
    const double d = H[size_t(i) * (m + 1) + i];
    y[i]           = d != 0.0 ? acc / d : 0.0;
  }
}

// z = x / diag (Jacobi application without a reduction)
__global__ void __launch_bounds__(kThreads) k_div(const double* x, const double* diag, double* z, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    z[i] = x[i] / diag[i];
}

__global__ void __launch_bounds__(kThreads) k_set_rows(double* x, const int32_t* rows, int64_t nr, double v)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < nr; i += int64_t(gridDim.x) * kThreads)
    x[rows[i]] = v;
}

__global__ void __launch_bounds__(kThreads) k_set_rows_values(double* x, const int32_t* rows, const double* vals,
                                                               int64_t nr)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < nr; i += int64_t(gridDim.x) * kThreads)
    x[rows[i]] = vals[i];
}

__global__ void __launch_bounds__(kThreads) k_axpby(double a, const double* x, double b, double* y, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] = a * x[i] + b * y[i];
}

__global__ void __launch_bounds__(kThreads) k_copy_rows(double* dst, const double* src, const int32_t* rows,
                                                         int64_t nr)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < nr; i += int64_t(gridDim.x) * kThreads)
    dst[rows[i]] = src[rows[i]];
}

// E[(e*nc + c)*npe + j] = L[lidx(map[e*npe + j], c)]
__global__ void __launch_bounds__(kThreads) k_gather_e(const int* map, const double* L, double* E, int64_t ne,
                                                        int npe, int nc, int64_t nnodes, int ordering)
{
  const int64_t n = ne * nc * npe;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const int64_t j    = i % npe;
    const int64_t ec   = i / npe;
    const int     c    = int(ec % nc);
    const int64_t e    = ec / nc;
    const int64_t node = map[e * npe + j];
    E[i]               = L[ordering ? node * nc + c : int64_t(c) * nnodes + node];
  }
}

// node-centric scatter-add: for each (node, c) add its element entries in
// increasing element order, the order of the reference's serial loop
// (fn/element_restriction.cpp:456-467), so the result is bit-identical.
__global__ void __launch_bounds__(kThreads) k_scatter_e(const int64_t* offs, const int32_t* ents, const double* E,
                                                         double* L, int64_t nnodes, int npe, int nc, int ordering)
{
  const int64_t n = nnodes * nc;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const int64_t node = ordering ? i / nc : i % nnodes;
    const int     c    = ordering ? int(i % nc) : int(i / nnodes);
    const int64_t li   = ordering ? node * nc + c : int64_t(c) * nnodes + node;
    double        acc  = L[li];
    for (int64_t k = offs[node]; k < offs[node + 1]; k++) {
      const int64_t ej = ents[k];
      const int64_t e = ej / npe, j = ej % npe;
      acc += E[(e * nc + c) * npe + j];
    }
    L[li] = acc;
  }
}

// y = A x for CSR A: one wavefront per row; lanes stride over the row and
// their partial sums are combined in a fixed tree, so results are
// reproducible run to run.
__global__ void __launch_bounds__(kThreads) k_csr_spmv(const int64_t* rp, const int32_t* col, const double* val,
                                                       const double* x, double* y, int64_t nrows)
{
  const int     lane  = threadIdx.x & 63;
  const int64_t wpb   = kThreads / 64;
  for (int64_t row = int64_t(blockIdx.x) * wpb + (threadIdx.x >> 6); row < nrows; row += int64_t(gridDim.x) * wpb) {
    double s = 0.0;
    for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += 64) s += val[k] * x[col[k]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[row] = s;
  }
}

}  // namespace

hipError_t vec_fill(double* x, double v, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kThreads), 0, s, x, v, n);
  return hipGetLastError();
}

hipError_t vec_fill_f32(float* x, float v, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(kThreads), 0, s, x, v, n);
  return hipGetLastError();
}

hipError_t vec_axpy(double a, const double* x, double* y, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(kThreads), 0, s, a, x, y, n);
  return hipGetLastError();
}

int reduction_blocks() { return 1024; }

hipError_t vec_dot_dev(const double* a, const double* b, int64_t n, double* part, double* scal, int slot, int op,
                       hipStream_t s)
{
  const int nb = reduction_blocks();
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kThreads), 0, s, a, b, n, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, s, part, nb, scal, slot, op);
  return hipGetLastError();
}

hipError_t dot_partial(const double* a, const double* b, int64_t n, double* part, hipStream_t s)
{
  hipLaunchKernelGGL(k_dot_partial, dim3(reduction_blocks()), dim3(kThreads), 0, s, a, b, n, part);
  return hipGetLastError();
}

hipError_t cg_step_x(double* x, double* r, double* z, double* Ad, const double* d, const double* diag, double* scal,
                     const double* part_den, double* part_rz, int64_t n, const double* stop, hipStream_t s)
{
  const int nb = reduction_blocks();
  hipLaunchKernelGGL(k_cg_step_x, dim3(nb), dim3(kThreads), 0, s, x, r, z, Ad, d, diag, scal, part_den, part_rz, nb,
                     n, stop);
  return hipGetLastError();
}

hipError_t cg_step_d(double* d, const double* z, double* scal, const double* part_rz, int64_t n, double goal, int it,
                     double* stop, hipStream_t s)
{
  const int nb = reduction_blocks();
  hipLaunchKernelGGL(k_cg_step_d, dim3(nb), dim3(kThreads), 0, s, d, z, scal, part_rz, nb, n, goal, it, stop);
  return hipGetLastError();
}

hipError_t jacobi_dot(const double* r, const double* diag, double* z, int64_t n, double* part, double* scal,
                      int slot, hipStream_t s)
{
  const int nb = reduction_blocks();
  hipLaunchKernelGGL(k_jacobi, dim3(nb), dim3(kThreads), 0, s, r, diag, z, n, part);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kThreads), 0, s, part, nb, scal, slot, 0);
  return hipGetLastError();
}

int multi_dot_blocks() { return 256; }

hipError_t multi_dot(const double* V, int64_t ld, int k, const double* w, int64_t n, double* part, double* out,
                     bool accumulate, hipStream_t s)
{
  if (k <= 0) return hipSuccess;
  const int nb = multi_dot_blocks();
  hipLaunchKernelGGL(k_multi_dot_partial, dim3(unsigned(nb), unsigned(k)), dim3(kThreads), 0, s, V, ld, w, n, part);
  hipLaunchKernelGGL(k_reduce_rows, dim3(unsigned(k)), dim3(kThreads), 0, s, part, nb, out, int(accumulate));
  return hipGetLastError();
}

hipError_t multi_axpy(const double* V, int64_t ld, int k, const double* c_dev, double sign, double* y, int64_t n,
                      hipStream_t s)
{
  if (k <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kThreads), 0, s, V, ld, k, c_dev, sign, y, n);
  return hipGetLastError();
}

hipError_t scale_copy(double a, const double* x, double* y, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(n)), dim3(kThreads), 0, s, a, x, y, n);
  return hipGetLastError();
}

size_t gmres_state_doubles(int m) { return size_t(m + 1) * m + 2 * size_t(m) + m + 1 + 8; }

hipError_t gmres_cycle_start(double* gs, int m, double beta, double goal, hipStream_t s)
{
  hipLaunchKernelGGL(k_gmres_cycle_start, dim3(1), dim3(kThreads), 0, s, gs, m, beta, goal);
  return hipGetLastError();
}

hipError_t gmres_column(const double* hh, int j, int m, double* gs, hipStream_t s)
{
  hipLaunchKernelGGL(k_gmres_column, dim3(1), dim3(1), 0, s, hh, j, m, gs);
  return hipGetLastError();
}

hipError_t gmres_scale_inv(double* y, int64_t n, int m, const double* gs, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  const double* st = gs + size_t(m + 1) * m + 2 * size_t(m) + m + 1;
  hipLaunchKernelGGL(k_gmres_scale_inv, dim3(grid_for(n)), dim3(kThreads), 0, s, y, n, st);
  return hipGetLastError();
}

hipError_t gmres_solve_y(const double* gs, int m, double* y, hipStream_t s)
{
  hipLaunchKernelGGL(k_gmres_solve_y, dim3(1), dim3(1), 0, s, gs, m, y);
  return hipGetLastError();
}

hipError_t jacobi_apply(const double* x, const double* diag, double* z, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_div, dim3(grid_for(n)), dim3(kThreads), 0, s, x, diag, z, n);
  return hipGetLastError();
}

hipError_t set_rows(double* x, const int32_t* rows, int64_t nr, double v, hipStream_t s)
{
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_set_rows, dim3(grid_for(nr)), dim3(kThreads), 0, s, x, rows, nr, v);
  return hipGetLastError();
}

hipError_t set_rows_values(double* x, const int32_t* rows, const double* vals, int64_t nr, hipStream_t s)
{
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_set_rows_values, dim3(grid_for(nr)), dim3(kThreads), 0, s, x, rows, vals, nr);
  return hipGetLastError();
}

hipError_t vec_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kThreads), 0, s, a, x, b, y, n);
  return hipGetLastError();
}

hipError_t copy_rows(double* dst, const double* src, const int32_t* rows, int64_t nr, hipStream_t s)
{
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_rows, dim3(grid_for(nr)), dim3(kThreads), 0, s, dst, src, rows, nr);
  return hipGetLastError();
}

hipError_t gather_e(const int* map, const double* L, double* E, int64_t ne, int npe, int nc, int64_t nnodes,
                    int ordering, hipStream_t s)
{
  const int64_t n = ne * nc * npe;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_e, dim3(grid_for(n)), dim3(kThreads), 0, s, map, L, E, ne, npe, nc, nnodes, ordering);
  return hipGetLastError();
}

hipError_t scatter_e(const int64_t* offs, const int32_t* ents, const double* E, double* L, int64_t nnodes, int npe,
                     int nc, int ordering, hipStream_t s)
{
  const int64_t n = nnodes * nc;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_e, dim3(grid_for(n)), dim3(kThreads), 0, s, offs, ents, E, L, nnodes, npe, nc,
                     ordering);
  return hipGetLastError();
}

__global__ void __launch_bounds__(kThreads) k_e_transpose(const double* in, double* out, int64_t ne, int nc, int npe,
                                                           int dir)
{
  const int64_t n = ne * nc * npe;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    // i indexes the reference layout [e][c][j]
    const int64_t j = i % npe, ec = i / npe, c = ec % nc, e = ec / nc;
    const int64_t o = (c * ne + e) * npe + j;  // component-major [c][e][j]
    if (dir == 0)
      out[o] = in[i];
    else
      out[i] = in[o];
  }
}

hipError_t e_transpose(const double* in, double* out, int64_t ne, int nc, int npe, int dir, hipStream_t s)
{
  const int64_t n = ne * nc * npe;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_e_transpose, dim3(grid_for(n)), dim3(kThreads), 0, s, in, out, ne, nc, npe, dir);
  return hipGetLastError();
}

hipError_t csr_spmv(const int64_t* rp, const int32_t* col, const double* val, const double* x, double* y,
                    int64_t nrows, hipStream_t s)
{
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_spmv, dim3(grid_for(nrows * 64)), dim3(kThreads), 0, s, rp, col, val, x, y, nrows);
  return hipGetLastError();
}

}  // namespace sxk
