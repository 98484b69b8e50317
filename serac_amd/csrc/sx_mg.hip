// sx_mg.hip -- kernels of the geometric multigrid V-cycle on the vertex grids
// of a structured mesh: grid transfers, the 27-point block-stencil operator of
// the coarse levels, Galerkin probing, Chebyshev smoothing updates and the
// dense coarsest solve.  All FP64.
#include <hip/hip_runtime.h>

#include "sx_element.hpp"
#include "sx_mg.hpp"

namespace sxm {
namespace {

constexpr int kThreads = 256;

inline unsigned blocks_for(int64_t n, unsigned cap = 8192)
{
  const int64_t b = (n + kThreads - 1) / kThreads;
  return unsigned(b < 1 ? 1 : (b > cap ? cap : b));
}

__device__ __forceinline__ int64_t nverts(const Grid3& g) { return int64_t(g.n[0]) * g.n[1] * g.n[2]; }

__device__ __forceinline__ void coords(const Grid3& g, int64_t v, int& ix, int& iy, int& iz)
{
  ix = int(v % g.n[0]);
  iy = int((v / g.n[0]) % g.n[1]);
  iz = int(v / (int64_t(g.n[0]) * g.n[1]));
}

// weight of the vertex at offset 0 for a node d = j - p i away from it
__device__ __forceinline__ double fe_weight(const sxu::FEGrid& G, int d)
{
  return d >= 0 ? 1.0 - G.xi[d] : G.xi[G.p + d];
}

// One direction of the separable restriction P^T = Rz Ry Rx: out has the
// dims of in except nv[dir] along dir, out(i) = sum_d w(d) in(p i + d).  The
// last pass (dir 2) applies the vertex mask.
template <typename TI>
__global__ void __launch_bounds__(kThreads) k_fe_restrict_dir(sxu::FEGrid G, const TI* in, Grid3 din, double* out,
                                                              Grid3 dout, int dir, const uint8_t* mask)
{
  const int64_t n      = int64_t(dout.n[0]) * dout.n[1] * dout.n[2] * 3;
  const int64_t stride = dir == 0 ? 3 : (dir == 1 ? int64_t(din.n[0]) * 3 : int64_t(din.n[0]) * din.n[1] * 3);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kThreads) {
    if (mask && mask[t]) {
      out[t] = 0.0;
      continue;
    }
    const int c = int(t % 3);
    int       o[3];
    coords(dout, t / 3, o[0], o[1], o[2]);
    const int i = o[dir];
    o[dir]      = G.p * i;
    // the input entry at the vertex's own node, then offsets along dir
    const TI* base = in + ((int64_t(o[2]) * din.n[1] + o[1]) * din.n[0] + o[0]) * 3 + c;
    double    s    = 0.0;
    for (int d = -G.p; d <= G.p; d++) {
      const int j = G.p * i + d;
      if (j < 0 || j >= din.n[dir]) continue;
      s += fe_weight(G, d) * double(base[int64_t(d) * stride]);
    }
    out[t] = s;
  }
}

template <typename TO>
__global__ void __launch_bounds__(kThreads) k_fe_prolong_add(sxu::FEGrid G, const double* ec, TO* z)
{
  const int64_t nn = int64_t(G.nn[0]) * G.nn[1] * G.nn[2];
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nn * 3; t += int64_t(gridDim.x) * kThreads) {
    const int     c = int(t % 3);
    const int64_t j = t / 3;
    const int     jv[3] = {int(j % G.nn[0]), int((j / G.nn[0]) % G.nn[1]), int(j / (int64_t(G.nn[0]) * G.nn[1]))};
    int           e[3], k[3];
    for (int d = 0; d < 3; d++) {
      e[d] = jv[d] / G.p;
      k[d] = jv[d] - G.p * e[d];
    }
    double s = 0.0;
    for (int b = 0; b < 8; b++) {
      const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
      if ((bx && k[0] == 0) || (by && k[1] == 0) || (bz && k[2] == 0)) continue;
      const double w = (bx ? G.xi[k[0]] : 1.0 - G.xi[k[0]]) * (by ? G.xi[k[1]] : 1.0 - G.xi[k[1]]) *
                       (bz ? G.xi[k[2]] : 1.0 - G.xi[k[2]]);
      const int64_t v = (int64_t(e[2] + bz) * G.nv[1] + (e[1] + by)) * G.nv[0] + (e[0] + bx);
      s += w * ec[v * 3 + c];
    }
    z[t] = TO(double(z[t]) + s);
  }
}

__global__ void __launch_bounds__(kThreads) k_grid_restrict(const double* rf, Grid3 f, double* rc, Grid3 c,
                                                            const uint8_t* maskc)
{
  const int64_t nc = nverts(c);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nc * 3; t += int64_t(gridDim.x) * kThreads) {
    if (maskc[t]) {
      rc[t] = 0.0;
      continue;
    }
    const int comp = int(t % 3);
    int       ix, iy, iz;
    coords(c, t / 3, ix, iy, iz);
    double s = 0.0;
    for (int oz = -1; oz <= 1; oz++) {
      const int jz = 2 * iz + oz;
      if (jz < 0 || jz >= f.n[2]) continue;
      const double wz = oz ? 0.5 : 1.0;
      for (int oy = -1; oy <= 1; oy++) {
        const int jy = 2 * iy + oy;
        if (jy < 0 || jy >= f.n[1]) continue;
        const double wzy = wz * (oy ? 0.5 : 1.0);
        for (int ox = -1; ox <= 1; ox++) {
          const int jx = 2 * ix + ox;
          if (jx < 0 || jx >= f.n[0]) continue;
          s += wzy * (ox ? 0.5 : 1.0) * rf[((int64_t(jz) * f.n[1] + jy) * f.n[0] + jx) * 3 + comp];
        }
      }
    }
    rc[t] = s;
  }
}

__global__ void __launch_bounds__(kThreads) k_grid_prolong_add(const double* ec, Grid3 c, double* zf, Grid3 f,
                                                               const uint8_t* maskf)
{
  const int64_t nf = nverts(f);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nf * 3; t += int64_t(gridDim.x) * kThreads) {
    if (maskf[t]) continue;
    const int comp = int(t % 3);
    int       iv[3];
    coords(f, t / 3, iv[0], iv[1], iv[2]);
    // per direction: the coarse vertices (one or two) and their weights
    int    lo[3], cnt[3];
    double w[3];
    for (int d = 0; d < 3; d++) {
      lo[d]  = iv[d] >> 1;
      cnt[d] = (iv[d] & 1) ? 2 : 1;
      w[d]   = (iv[d] & 1) ? 0.5 : 1.0;
    }
    double s = 0.0;
    for (int bz = 0; bz < cnt[2]; bz++)
      for (int by = 0; by < cnt[1]; by++)
        for (int bx = 0; bx < cnt[0]; bx++) {
          const int64_t v = (int64_t(lo[2] + bz) * c.n[1] + (lo[1] + by)) * c.n[0] + (lo[0] + bx);
          s += w[0] * w[1] * w[2] * ec[v * 3 + comp];
        }
    zf[t] += s;
  }
}

__global__ void __launch_bounds__(kThreads) k_stencil_apply(const double* S, Grid3 g, const double* x, double* y)
{
  const int64_t nv = nverts(g);
  for (int64_t v = int64_t(blockIdx.x) * kThreads + threadIdx.x; v < nv; v += int64_t(gridDim.x) * kThreads) {
    int ix, iy, iz;
    coords(g, v, ix, iy, iz);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int oz = -1; oz <= 1; oz++) {
      const int jz = iz + oz;
      if (jz < 0 || jz >= g.n[2]) continue;
      for (int oy = -1; oy <= 1; oy++) {
        const int jy = iy + oy;
        if (jy < 0 || jy >= g.n[1]) continue;
        for (int ox = -1; ox <= 1; ox++) {
          const int jx = ix + ox;
          if (jx < 0 || jx >= g.n[0]) continue;
          const int64_t j  = (int64_t(jz) * g.n[1] + jy) * g.n[0] + jx;
          const double  x0 = x[j * 3], x1 = x[j * 3 + 1], x2 = x[j * 3 + 2];
          const double* s  = S + int64_t((oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)) * 9 * nv + v;
          a0 += s[0] * x0 + s[nv] * x1 + s[2 * nv] * x2;
          a1 += s[3 * nv] * x0 + s[4 * nv] * x1 + s[5 * nv] * x2;
          a2 += s[6 * nv] * x0 + s[7 * nv] * x1 + s[8 * nv] * x2;
        }
      }
    }
    y[v * 3]     = a0;
    y[v * 3 + 1] = a1;
    y[v * 3 + 2] = a2;
  }
}

__global__ void __launch_bounds__(kThreads) k_probe_fill(double* e, Grid3 g, int cx, int cy, int cz, int ci,
                                                         const uint8_t* mask)
{
  const int64_t nv = nverts(g);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nv * 3; t += int64_t(gridDim.x) * kThreads) {
    int ix, iy, iz;
    coords(g, t / 3, ix, iy, iz);
    const bool on = int(t % 3) == ci && ix % 3 == cx && iy % 3 == cy && iz % 3 == cz && !mask[t];
    e[t]          = on ? 1.0 : 0.0;
  }
}

// the offset o in {-1, 0, 1} with (i + o) = c (mod 3)
__device__ __forceinline__ int colour_offset(int i, int c)
{
  const int o = ((c - i) % 3 + 3) % 3;
  return o == 2 ? -1 : o;
}

__global__ void __launch_bounds__(kThreads) k_probe_extract(const double* y, Grid3 g, int cx, int cy, int cz, int ci,
                                                            const uint8_t* mask, double* S)
{
  const int64_t nv = nverts(g);
  for (int64_t v = int64_t(blockIdx.x) * kThreads + threadIdx.x; v < nv; v += int64_t(gridDim.x) * kThreads) {
    int ix, iy, iz;
    coords(g, v, ix, iy, iz);
    const int ox = colour_offset(ix, cx), oy = colour_offset(iy, cy), oz = colour_offset(iz, cz);
    const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
    if (jx < 0 || jx >= g.n[0] || jy < 0 || jy >= g.n[1] || jz < 0 || jz >= g.n[2]) continue;
    const int64_t j = (int64_t(jz) * g.n[1] + jy) * g.n[0] + jx;
    if (mask[j * 3 + ci]) continue;
    const int o = (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1);
    for (int co = 0; co < 3; co++)
      if (!mask[v * 3 + co]) S[(int64_t(o) * 9 + co * 3 + ci) * nv + v] = y[v * 3 + co];
  }
}

__global__ void __launch_bounds__(kThreads) k_stencil_dinv(const double* S, Grid3 g, const uint8_t* mask,
                                                           double* dinv)
{
  const int64_t nv = nverts(g);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nv * 3; t += int64_t(gridDim.x) * kThreads) {
    const int64_t v = t / 3;
    const int     c = int(t % 3);
    dinv[t]         = mask[t] ? 0.0 : 1.0 / S[(int64_t(13) * 9 + c * 3 + c) * nv + v];
  }
}

__global__ void __launch_bounds__(kThreads) k_diag_inverse(const double* diag, int64_t n, double* dinv)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    dinv[i] = 1.0 / diag[i];
}

__global__ void __launch_bounds__(kThreads) k_set_rows0(double* x, const int32_t* rows, int64_t nr)
{
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < nr; k += int64_t(gridDim.x) * kThreads)
    x[rows[k]] = 0.0;
}

__global__ void __launch_bounds__(kThreads) k_rows_mask(const int32_t* rows, int64_t nr, uint8_t* mask)
{
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < nr; k += int64_t(gridDim.x) * kThreads)
    mask[rows[k]] = 1;
}

__global__ void __launch_bounds__(kThreads) k_vertex_mask(sxu::FEGrid G, const uint8_t* node_mask, uint8_t* vmask)
{
  const int64_t nv = int64_t(G.nv[0]) * G.nv[1] * G.nv[2];
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nv * 3; t += int64_t(gridDim.x) * kThreads) {
    const int64_t i    = t / 3;
    const int     ix   = int(i % G.nv[0]), iy = int((i / G.nv[0]) % G.nv[1]), iz = int(i / (int64_t(G.nv[0]) * G.nv[1]));
    const int64_t node = (int64_t(G.p) * iz * G.nn[1] + int64_t(G.p) * iy) * G.nn[0] + int64_t(G.p) * ix;
    vmask[t]           = node_mask[node * 3 + t % 3];
  }
}

__global__ void __launch_bounds__(kThreads) k_coarsen_mask(const uint8_t* mf, Grid3 f, uint8_t* mc, Grid3 c)
{
  const int64_t nc = nverts(c);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nc * 3; t += int64_t(gridDim.x) * kThreads) {
    int ix, iy, iz;
    coords(c, t / 3, ix, iy, iz);
    const int64_t j = (int64_t(2 * iz) * f.n[1] + 2 * iy) * f.n[0] + 2 * ix;
    mc[t]           = mf[j * 3 + t % 3];
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_cheb_init(const T* b, const T* dinv, double inv_theta, T* x, T* d,
                                                        T* r, int x_add, int copy_r, int64_t n)
{
  const T it = T(inv_theta);
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const T bi = b[i];
    const T di = dinv[i] * bi * it;
    d[i]       = di;
    x[i]       = x_add ? x[i] + di : di;
    if (copy_r) r[i] = bi;
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_cheb_step(T* r, const T* t, T* d, const T* dinv, double c1, double c2,
                                                        T* x, int64_t n)
{
  const T a = T(c1), b = T(c2);
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const T ri = r[i] - t[i];
    r[i]       = ri;
    const T di = a * d[i] + b * dinv[i] * ri;
    d[i]       = di;
    x[i] += di;
  }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_residual(const T* b, const T* t, T* r, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    r[i] = b[i] - t[i];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_sub_inplace(T* r, const T* t, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    r[i] -= t[i];
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_apply_mask(T* x, const uint8_t* mask, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    if (mask[i]) x[i] = T(0);
}

// y = T2(x), zero where mask is set (mask may be null)
template <typename T1, typename T2>
__global__ void __launch_bounds__(kThreads) k_convert(const T1* x, T2* y, const uint8_t* mask, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    y[i] = (mask && mask[i]) ? T2(0) : T2(x[i]);
}

// one wave per row
__global__ void __launch_bounds__(kThreads) k_dense_mv(const double* A, int n, const double* b, double* x)
{
  const int row  = int(blockIdx.x) * (kThreads / 64) + int(threadIdx.x >> 6);
  const int lane = int(threadIdx.x & 63);
  if (row >= n) return;  // whole waves leave; no barrier follows
  double s = 0.0;
  for (int k = lane; k < n; k += 64) s += A[int64_t(row) * n + k] * b[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[row] = s;
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return s;
}

__global__ void __launch_bounds__(kThreads) k_pcg_step_xr(double* x, double* r, double* Ad, const double* d,
                                                          double* scal, const double* part_den, int nb, int64_t n,
                                                          const double* stop)
{
  __shared__ double sh[4];
  __shared__ double bc;
  if (stop[0] != 0.0) return;  // the whole block leaves before any barrier
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) s += part_den[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) bc = s;
  __syncthreads();
  const double den = bc, nom = scal[0], alpha = nom / den;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    x[i] += alpha * d[i];
    r[i] -= alpha * Ad[i];
    Ad[i] = 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[1] = den;
    scal[2] = alpha;
    scal[5] = nom;
  }
}

__global__ void __launch_bounds__(kThreads) k_mul_dinv(const double* dinv, const double* t, double* v, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    v[i] = dinv[i] * t[i];
}

__global__ void __launch_bounds__(kThreads) k_scale(double* v, double a, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    v[i] *= a;
}

// a fixed pseudo-random field in [-1, 1) (splitmix64 of the index and seed)
__global__ void __launch_bounds__(kThreads) k_seed_field(double* v, int64_t n, uint64_t seed, const uint8_t* mask)
{
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    uint64_t z = uint64_t(i) + seed * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z          = z ^ (z >> 31);
    const double u = double(z >> 11) * (1.0 / 9007199254740992.0);
    v[i]           = (mask && mask[i]) ? 0.0 : 2.0 * u - 1.0;
  }
}

// Vertex-level Galerkin operator P^T J P assembled element by element from the
// stored GRAD_SYM tangent (9 x 9, packed upper triangle, d(parent flux) /
// d(parent grad) at each quadrature point; the quadrature weights are not in
// it).  The trilinear hat of corner a lies in the order-p space, so its block
// is exact: K[(a,ci),(b,cj)] = sum_q w_q sum_{d,k} g_a,d(q) D_q[(ci,d),(cj,k)]
// g_b,k(q), with g the parent gradients of the hats.  One workgroup per
// element; quadrature points in chunks of QC: the chunk's tangent and
// w D G products in LDS, each thread accumulating up to three of the 576
// entries, then FP64 atomic adds into the stencil.
template <int N, typename S>
__global__ void __launch_bounds__(kThreads) k_q1_galerkin(const int* map, const void* D, int64_t nslots,
                                                          const double* tables, int p, Grid3 gn, Grid3 gv, double* St)
{
  constexpr int NQ = N * N * N, DS = 45, QC = 4, NE = 576;
  using L          = sxk::TangentLayout<S, DS, NQ>;
  __shared__ double g[NQ * 24];  // [q][corner][dir]
  __shared__ double w[NQ];
  __shared__ double Dl[QC * DS];
  __shared__ double DG[QC * 216];  // [qq][(ci,d)][(b,cj)]
  __shared__ int    vpos[8][3];
  __shared__ int    vid[8];
  const int         tid = threadIdx.x;
  const double*     pts = tables + 4 * N * N;
  for (int e = tid; e < NQ; e += kThreads) {
    const int qx = e % N, qy = (e / N) % N, qz = e / (N * N);
    double    wx = 0.0, wy = 0.0, wz = 0.0;
    for (int i = 0; i < N; i++) {
      wx += tables[2 * N * N + qx * N + i];
      wy += tables[2 * N * N + qy * N + i];
      wz += tables[2 * N * N + qz * N + i];
    }
    w[e]              = wx * wy * wz;
    const double x[3] = {pts[qx], pts[qy], pts[qz]};
    for (int a = 0; a < 8; a++) {
      const int    b[3] = {a & 1, (a >> 1) & 1, a >> 2};
      double       f[3], df[3];
      for (int d = 0; d < 3; d++) {
        f[d]  = b[d] ? x[d] : 1.0 - x[d];
        df[d] = b[d] ? 1.0 : -1.0;
      }
      g[e * 24 + a * 3 + 0] = df[0] * f[1] * f[2];
      g[e * 24 + a * 3 + 1] = f[0] * df[1] * f[2];
      g[e * 24 + a * 3 + 2] = f[0] * f[1] * df[2];
    }
  }
  const int64_t nv = int64_t(gv.n[0]) * gv.n[1] * gv.n[2];
  for (int64_t slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
    if (tid < 8) {
      const int li   = ((tid >> 2) * (N - 1) * N + ((tid >> 1) & 1) * (N - 1)) * N + (tid & 1) * (N - 1);
      const int node = map[slot * NQ + li];
      const int jx = node % gn.n[0], jy = (node / gn.n[0]) % gn.n[1], jz = node / (gn.n[0] * gn.n[1]);
      vpos[tid][0] = jx / p;
      vpos[tid][1] = jy / p;
      vpos[tid][2] = jz / p;
      vid[tid]     = (jz / p * gv.n[1] + jy / p) * gv.n[0] + jx / p;
    }
    double     acc[3] = {0.0, 0.0, 0.0};
    const S*   base   = static_cast<const S*>(D) + slot * L::stride;
    for (int q0 = 0; q0 < NQ; q0 += QC) {
      __syncthreads();  // the previous chunk's DG is consumed (and the tables are written)
      for (int e = tid; e < QC * DS; e += kThreads) {
        const int qq = e / DS, k = e % DS;
        Dl[e]        = q0 + qq < NQ ? double(base[L::offset(k, q0 + qq)]) : 0.0;
      }
      __syncthreads();
      for (int e = tid; e < QC * 216; e += kThreads) {
        const int qq = e / 216, r = e % 216, row = r / 24, col = r % 24, b = col / 3, cj = col % 3;
        double    s = 0.0;
        if (q0 + qq < NQ) {
          const double* gb = g + (q0 + qq) * 24 + b * 3;
          for (int k = 0; k < 3; k++) s += Dl[qq * DS + sxk::packed_index(row, cj * 3 + k)] * gb[k];
          s *= w[q0 + qq];
        }
        DG[e] = s;
      }
      __syncthreads();
      for (int m = 0; m < 3; m++) {
        const int e = tid + m * kThreads;
        if (e >= NE) break;
        const int a = e / 72, ci = (e / 24) % 3, col = e % 24;
        double    s = 0.0;
        for (int qq = 0; qq < QC && q0 + qq < NQ; qq++) {
          const double* ga = g + (q0 + qq) * 24 + a * 3;
          const double* dg = DG + qq * 216 + (ci * 3) * 24 + col;
          s += ga[0] * dg[0] + ga[1] * dg[24] + ga[2] * dg[48];
        }
        acc[m] += s;
      }
    }
    for (int m = 0; m < 3; m++) {
      const int e = tid + m * kThreads;
      if (e >= NE) break;
      const int a = e / 72, ci = (e / 24) % 3, b = (e % 24) / 3, cj = e % 3;
      const int o = (vpos[b][2] - vpos[a][2] + 1) * 9 + (vpos[b][1] - vpos[a][1] + 1) * 3 + (vpos[b][0] - vpos[a][0] + 1);
      atomicAdd(St + (int64_t(o) * 9 + ci * 3 + cj) * nv + vid[a], acc[m]);
    }
    __syncthreads();  // vpos / vid of this element are read before the next element writes them
  }
}

// A BF16 copy of a stored tangent (FP64 or FP32 source) in the BF16 layout:
// entry k of quadrature point q of each slot, rounded to nearest even.
template <typename SRC, int DS, int NQ>
__global__ void __launch_bounds__(kThreads) k_tangent_to_bf16(const SRC* src, uint16_t* dst, int64_t nslots)
{
  using Ls          = sxk::TangentLayout<SRC, DS, NQ>;
  using Lb          = sxk::TangentLayout<sxk::bf16, DS, NQ>;
  const int64_t n   = nslots * DS * NQ;
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < n; t += int64_t(gridDim.x) * kThreads) {
    const int64_t slot = t / (DS * NQ);
    const int     e = int(t % (DS * NQ)), q = e % NQ, k = e / NQ;
    const double  v = double(src[slot * Ls::stride + Ls::offset(k, q)]);
    dst[slot * Lb::stride + Lb::offset(k, q)] = sxk::to_store(v, static_cast<sxk::bf16*>(nullptr));
  }
}

// Galerkin product of a vertex-level stencil with trilinear transfers:
// Sc[I][O] = sum_i sum_j w_I(i) Sf[i][j - i] w_{I+O}(j) over fine vertices i
// within one of 2I and j within one of 2(I + O) (weights 1 on the coinciding
// vertex, 1/2 beside it), one thread per (coarse vertex, offset) and its 3x3
// block.  Constrained fine rows and columns of Sf are zero; the coarse ones
// are zeroed afterwards (stencil_mask).
__global__ void __launch_bounds__(kThreads) k_stencil_rap(const double* Sf, Grid3 f, double* Sc, Grid3 c)
{
  const int64_t nvf = nverts(f), nvc = nverts(c);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nvc * 27; t += int64_t(gridDim.x) * kThreads) {
    const int64_t I = t % nvc;
    const int     O = int(t / nvc);
    int           Ic[3];
    coords(c, I, Ic[0], Ic[1], Ic[2]);
    const int Od[3] = {O % 3 - 1, (O / 3) % 3 - 1, O / 9 - 1};
    // per direction: the (i, j) pairs with |i - 2I| <= 1, |j - 2(I + O)| <= 1,
    // |j - i| <= 1, inside the fine grid, and their weight products
    int    pi[3][9], pj[3][9], np[3];
    double pw[3][9];
    bool   empty = false;
    for (int d = 0; d < 3; d++) {
      np[d]        = 0;
      const int J  = Ic[d] + Od[d];
      if (J < 0 || J >= c.n[d]) {
        empty = true;
        break;
      }
      for (int a = -1; a <= 1; a++) {
        const int i = 2 * Ic[d] + a;
        if (i < 0 || i >= f.n[d]) continue;
        for (int b = -1; b <= 1; b++) {
          const int j = 2 * J + b;
          if (j < 0 || j >= f.n[d] || j - i < -1 || j - i > 1) continue;
          pi[d][np[d]] = i;
          pj[d][np[d]] = j;
          pw[d][np[d]] = (a ? 0.5 : 1.0) * (b ? 0.5 : 1.0);
          np[d]++;
        }
      }
    }
    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!empty) {
      for (int kz = 0; kz < np[2]; kz++)
        for (int ky = 0; ky < np[1]; ky++)
          for (int kx = 0; kx < np[0]; kx++) {
            const double  w  = pw[0][kx] * pw[1][ky] * pw[2][kz];
            const int64_t i  = (int64_t(pi[2][kz]) * f.n[1] + pi[1][ky]) * f.n[0] + pi[0][kx];
            const int     o  = (pj[2][kz] - pi[2][kz] + 1) * 9 + (pj[1][ky] - pi[1][ky] + 1) * 3 + (pj[0][kx] - pi[0][kx] + 1);
            const double* sf = Sf + int64_t(o) * 9 * nvf + i;
            for (int k = 0; k < 9; k++) acc[k] += w * sf[int64_t(k) * nvf];
          }
    }
    for (int k = 0; k < 9; k++) Sc[(int64_t(O) * 9 + k) * nvc + I] = acc[k];
  }
}

// zero the stencil entries whose row or column is constrained
__global__ void __launch_bounds__(kThreads) k_stencil_mask(double* S, Grid3 g, const uint8_t* mask)
{
  const int64_t nv = nverts(g);
  for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < nv * 243; t += int64_t(gridDim.x) * kThreads) {
    const int64_t v = t % nv;
    const int     k = int(t / nv), o = k / 9, co = (k / 3) % 3, ci = k % 3;
    int           ix, iy, iz;
    coords(g, v, ix, iy, iz);
    const int jx = ix + o % 3 - 1, jy = iy + (o / 3) % 3 - 1, jz = iz + o / 9 - 1;
    const bool out = jx < 0 || jx >= g.n[0] || jy < 0 || jy >= g.n[1] || jz < 0 || jz >= g.n[2];
    if (out || mask[v * 3 + co] || mask[((int64_t(jz) * g.n[1] + jy) * g.n[0] + jx) * 3 + ci]) S[t] = 0.0;
  }
}

// flag[0] = 1 when a constrained node component has an unconstrained vertex
// whose hat is nonzero there (then P_0 of the unconstrained vertex space does
// not vanish on the constraints and the element-wise operator differs from
// the probed Z P^T J P Z)
__global__ void __launch_bounds__(kThreads) k_vertex_compatible(sxu::FEGrid G, const int32_t* rows, int64_t nr,
                                                                const uint8_t* vmask, int* flag)
{
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < nr; k += int64_t(gridDim.x) * kThreads) {
    const int64_t dof = rows[k], node = dof / 3;
    const int     c = int(dof % 3);
    const int     j[3] = {int(node % G.nn[0]), int((node / G.nn[0]) % G.nn[1]), int(node / (int64_t(G.nn[0]) * G.nn[1]))};
    int           lo[3], cnt[3];
    for (int d = 0; d < 3; d++) {
      lo[d]  = j[d] / G.p;
      cnt[d] = (j[d] % G.p) ? 2 : 1;
    }
    bool bad = false;
    for (int bz = 0; bz < cnt[2]; bz++)
      for (int by = 0; by < cnt[1]; by++)
        for (int bx = 0; bx < cnt[0]; bx++) {
          const int64_t v = (int64_t(lo[2] + bz) * G.nv[1] + (lo[1] + by)) * G.nv[0] + (lo[0] + bx);
          bad             = bad || !vmask[v * 3 + c];
        }
    if (bad) atomicMax(flag, 1);
  }
}

}  // namespace

int64_t fe_restrict_scratch(const sxu::FEGrid& G, int pass)
{
  return pass == 0 ? int64_t(G.nn[2]) * G.nn[1] * G.nv[0] * 3 : int64_t(G.nn[2]) * G.nv[1] * G.nv[0] * 3;
}

hipError_t fe_restrict(const sxu::FEGrid& G, const double* r, double* rc, const uint8_t* mask, double* t1, double* t2,
                       hipStream_t s)
{
  const Grid3 d0{{G.nn[0], G.nn[1], G.nn[2]}}, d1{{G.nv[0], G.nn[1], G.nn[2]}}, d2{{G.nv[0], G.nv[1], G.nn[2]}},
      d3{{G.nv[0], G.nv[1], G.nv[2]}};
  const uint8_t* none = nullptr;
  hipLaunchKernelGGL(k_fe_restrict_dir<double>, dim3(blocks_for(fe_restrict_scratch(G, 0))), dim3(kThreads), 0, s, G, r, d0,
                     t1, d1, 0, none);
  hipLaunchKernelGGL(k_fe_restrict_dir<double>, dim3(blocks_for(fe_restrict_scratch(G, 1))), dim3(kThreads), 0, s, G, t1, d1,
                     t2, d2, 1, none);
  hipLaunchKernelGGL(k_fe_restrict_dir<double>, dim3(blocks_for(int64_t(G.nv[0]) * G.nv[1] * G.nv[2] * 3)), dim3(kThreads), 0,
                     s, G, t2, d2, rc, d3, 2, mask);
  return hipGetLastError();
}

hipError_t fe_prolong_add(const sxu::FEGrid& G, const double* ec, double* z, hipStream_t s)
{
  const int64_t n = int64_t(G.nn[0]) * G.nn[1] * G.nn[2] * 3;
  hipLaunchKernelGGL(k_fe_prolong_add<double>, dim3(blocks_for(n)), dim3(kThreads), 0, s, G, ec, z);
  return hipGetLastError();
}

hipError_t grid_restrict(const double* rf, Grid3 f, double* rc, Grid3 c, const uint8_t* maskc, hipStream_t s)
{
  const int64_t n = int64_t(c.n[0]) * c.n[1] * c.n[2] * 3;
  hipLaunchKernelGGL(k_grid_restrict, dim3(blocks_for(n)), dim3(kThreads), 0, s, rf, f, rc, c, maskc);
  return hipGetLastError();
}

hipError_t grid_prolong_add(const double* ec, Grid3 c, double* zf, Grid3 f, const uint8_t* maskf, hipStream_t s)
{
  const int64_t n = int64_t(f.n[0]) * f.n[1] * f.n[2] * 3;
  hipLaunchKernelGGL(k_grid_prolong_add, dim3(blocks_for(n)), dim3(kThreads), 0, s, ec, c, zf, f, maskf);
  return hipGetLastError();
}

hipError_t stencil_apply(const double* S, Grid3 g, const double* x, double* y, hipStream_t s)
{
  const int64_t n = int64_t(g.n[0]) * g.n[1] * g.n[2];
  hipLaunchKernelGGL(k_stencil_apply, dim3(blocks_for(n)), dim3(kThreads), 0, s, S, g, x, y);
  return hipGetLastError();
}

hipError_t probe_fill(double* e, Grid3 g, int cx, int cy, int cz, int ci, const uint8_t* mask, hipStream_t s)
{
  const int64_t n = int64_t(g.n[0]) * g.n[1] * g.n[2] * 3;
  hipLaunchKernelGGL(k_probe_fill, dim3(blocks_for(n)), dim3(kThreads), 0, s, e, g, cx, cy, cz, ci, mask);
  return hipGetLastError();
}

hipError_t probe_extract(const double* y, Grid3 g, int cx, int cy, int cz, int ci, const uint8_t* mask, double* S,
                         hipStream_t s)
{
  const int64_t n = int64_t(g.n[0]) * g.n[1] * g.n[2];
  hipLaunchKernelGGL(k_probe_extract, dim3(blocks_for(n)), dim3(kThreads), 0, s, y, g, cx, cy, cz, ci, mask, S);
  return hipGetLastError();
}

hipError_t stencil_dinv(const double* S, Grid3 g, const uint8_t* mask, double* dinv, hipStream_t s)
{
  const int64_t n = int64_t(g.n[0]) * g.n[1] * g.n[2] * 3;
  hipLaunchKernelGGL(k_stencil_dinv, dim3(blocks_for(n)), dim3(kThreads), 0, s, S, g, mask, dinv);
  return hipGetLastError();
}

hipError_t diag_inverse(const double* diag, int64_t n, const int32_t* rows, int64_t nr, double* dinv, hipStream_t s)
{
  hipLaunchKernelGGL(k_diag_inverse, dim3(blocks_for(n)), dim3(kThreads), 0, s, diag, n, dinv);
  if (nr > 0) hipLaunchKernelGGL(k_set_rows0, dim3(blocks_for(nr)), dim3(kThreads), 0, s, dinv, rows, nr);
  return hipGetLastError();
}

hipError_t vertex_mask(const sxu::FEGrid& G, const uint8_t* node_mask, uint8_t* vmask, hipStream_t s)
{
  const int64_t n = int64_t(G.nv[0]) * G.nv[1] * G.nv[2] * 3;
  hipLaunchKernelGGL(k_vertex_mask, dim3(blocks_for(n)), dim3(kThreads), 0, s, G, node_mask, vmask);
  return hipGetLastError();
}

hipError_t rows_mask(const int32_t* rows, int64_t nr, uint8_t* mask, hipStream_t s)
{
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rows_mask, dim3(blocks_for(nr)), dim3(kThreads), 0, s, rows, nr, mask);
  return hipGetLastError();
}

hipError_t coarsen_mask(const uint8_t* mf, Grid3 f, uint8_t* mc, Grid3 c, hipStream_t s)
{
  const int64_t n = int64_t(c.n[0]) * c.n[1] * c.n[2] * 3;
  hipLaunchKernelGGL(k_coarsen_mask, dim3(blocks_for(n)), dim3(kThreads), 0, s, mf, f, mc, c);
  return hipGetLastError();
}

hipError_t cheb_init(const double* b, const double* dinv, double theta, double* x, double* d, double* r, int x_add,
                     int copy_r, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_cheb_init<double>, dim3(blocks_for(n)), dim3(kThreads), 0, s, b, dinv, 1.0 / theta, x, d, r, x_add,
                     copy_r, n);
  return hipGetLastError();
}

hipError_t cheb_step(double* r, const double* t, double* d, const double* dinv, double c1, double c2, double* x,
                     int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_cheb_step<double>, dim3(blocks_for(n)), dim3(kThreads), 0, s, r, t, d, dinv, c1, c2, x, n);
  return hipGetLastError();
}

hipError_t residual(const double* b, const double* t, double* r, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_residual<double>, dim3(blocks_for(n)), dim3(kThreads), 0, s, b, t, r, n);
  return hipGetLastError();
}

hipError_t sub_inplace(double* r, const double* t, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_sub_inplace<double>, dim3(blocks_for(n)), dim3(kThreads), 0, s, r, t, n);
  return hipGetLastError();
}

hipError_t apply_mask(double* x, const uint8_t* mask, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_apply_mask<double>, dim3(blocks_for(n)), dim3(kThreads), 0, s, x, mask, n);
  return hipGetLastError();
}

hipError_t dense_mv(const double* A, int n, const double* b, double* x, hipStream_t s)
{
  const unsigned rows_per_block = kThreads / 64;
  hipLaunchKernelGGL(k_dense_mv, dim3((unsigned(n) + rows_per_block - 1) / rows_per_block), dim3(kThreads), 0, s, A,
                     n, b, x);
  return hipGetLastError();
}

hipError_t pcg_step_xr(double* x, double* r, double* Ad, const double* d, double* scal, const double* part_den,
                       int nb, int64_t n, const double* stop, hipStream_t s)
{
  hipLaunchKernelGGL(k_pcg_step_xr, dim3(unsigned(nb)), dim3(kThreads), 0, s, x, r, Ad, d, scal, part_den, nb, n,
                     stop);
  return hipGetLastError();
}

hipError_t mul_dinv(const double* dinv, const double* t, double* v, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_mul_dinv, dim3(blocks_for(n)), dim3(kThreads), 0, s, dinv, t, v, n);
  return hipGetLastError();
}

hipError_t scale(double* v, double a, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(kThreads), 0, s, v, a, n);
  return hipGetLastError();
}

hipError_t seed_field(double* v, int64_t n, uint64_t seed, const uint8_t* mask, hipStream_t s)
{
  hipLaunchKernelGGL(k_seed_field, dim3(blocks_for(n)), dim3(kThreads), 0, s, v, n, seed, mask);
  return hipGetLastError();
}

hipError_t q1_galerkin(int order, int tprec_f32, const int* map, const void* D, int64_t nslots, const double* tables,
                      Grid3 gn, Grid3 gv, double* S, hipStream_t s)
{
  const unsigned grid = unsigned(nslots < 16384 ? (nslots < 1 ? 1 : nslots) : 16384);
#define SXM_Q1(NN)                                                                                                  \
  if (order + 1 == NN) {                                                                                            \
    if (tprec_f32)                                                                                                  \
      hipLaunchKernelGGL((k_q1_galerkin<NN, float>), dim3(grid), dim3(kThreads), 0, s, map, D, nslots, tables,      \
                         order, gn, gv, S);                                                                         \
    else                                                                                                            \
      hipLaunchKernelGGL((k_q1_galerkin<NN, double>), dim3(grid), dim3(kThreads), 0, s, map, D, nslots, tables,     \
                         order, gn, gv, S);                                                                         \
    return hipGetLastError();                                                                                       \
  }
  SXM_Q1(2)
  SXM_Q1(3)
  SXM_Q1(4)
  SXM_Q1(5)
#undef SXM_Q1
  return hipErrorInvalidValue;
}

int64_t bf16_tangent_elems(int order, int ds, int64_t nslots)
{
  const int64_t nq = int64_t(order + 1) * (order + 1) * (order + 1);
  return nslots * ((ds * nq * 2 + 15) / 16 * 8);
}

hipError_t tangent_to_bf16(int order, int ds, int src_f32, const void* src, uint16_t* dst, int64_t nslots,
                           hipStream_t s)
{
  const int64_t n    = nslots * ds * int64_t(order + 1) * (order + 1) * (order + 1);
  const unsigned grid = blocks_for(n);
#define SXM_BF(NN, DSV)                                                                                             \
  if (order + 1 == NN && ds == DSV) {                                                                               \
    if (src_f32)                                                                                                    \
      hipLaunchKernelGGL((k_tangent_to_bf16<float, DSV, NN * NN * NN>), dim3(grid), dim3(kThreads), 0, s,          \
                         static_cast<const float*>(src), dst, nslots);                                              \
    else                                                                                                            \
      hipLaunchKernelGGL((k_tangent_to_bf16<double, DSV, NN * NN * NN>), dim3(grid), dim3(kThreads), 0, s,         \
                         static_cast<const double*>(src), dst, nslots);                                             \
    return hipGetLastError();                                                                                       \
  }
  SXM_BF(2, 45)
  SXM_BF(3, 45)
  SXM_BF(4, 45)
  SXM_BF(5, 45)
  SXM_BF(2, 81)
  SXM_BF(3, 81)
  SXM_BF(4, 81)
  SXM_BF(5, 81)
#undef SXM_BF
  return hipErrorInvalidValue;
}

hipError_t stencil_rap(const double* Sf, Grid3 f, double* Sc, Grid3 c, hipStream_t s)
{
  const int64_t n = int64_t(c.n[0]) * c.n[1] * c.n[2] * 27;
  hipLaunchKernelGGL(k_stencil_rap, dim3(blocks_for(n)), dim3(kThreads), 0, s, Sf, f, Sc, c);
  return hipGetLastError();
}

hipError_t stencil_mask(double* S, Grid3 g, const uint8_t* mask, hipStream_t s)
{
  const int64_t n = int64_t(g.n[0]) * g.n[1] * g.n[2] * 243;
  hipLaunchKernelGGL(k_stencil_mask, dim3(blocks_for(n)), dim3(kThreads), 0, s, S, g, mask);
  return hipGetLastError();
}

hipError_t vertex_compatible(const sxu::FEGrid& G, const int32_t* rows, int64_t nr, const uint8_t* vmask, int* flag,
                             hipStream_t s)
{
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vertex_compatible, dim3(blocks_for(nr)), dim3(kThreads), 0, s, G, rows, nr, vmask, flag);
  return hipGetLastError();
}

// FP32 level-0 vectors
hipError_t fe_restrict(const sxu::FEGrid& G, const float* r, double* rc, const uint8_t* mask, double* t1, double* t2,
                       hipStream_t s)
{
  const Grid3 d0{{G.nn[0], G.nn[1], G.nn[2]}}, d1{{G.nv[0], G.nn[1], G.nn[2]}}, d2{{G.nv[0], G.nv[1], G.nn[2]}},
      d3{{G.nv[0], G.nv[1], G.nv[2]}};
  const uint8_t* none = nullptr;
  hipLaunchKernelGGL(k_fe_restrict_dir<float>, dim3(blocks_for(fe_restrict_scratch(G, 0))), dim3(kThreads), 0, s, G, r,
                     d0, t1, d1, 0, none);
  hipLaunchKernelGGL(k_fe_restrict_dir<double>, dim3(blocks_for(fe_restrict_scratch(G, 1))), dim3(kThreads), 0, s, G,
                     static_cast<const double*>(t1), d1, t2, d2, 1, none);
  hipLaunchKernelGGL(k_fe_restrict_dir<double>, dim3(blocks_for(int64_t(G.nv[0]) * G.nv[1] * G.nv[2] * 3)),
                     dim3(kThreads), 0, s, G, static_cast<const double*>(t2), d2, rc, d3, 2, mask);
  return hipGetLastError();
}
hipError_t fe_prolong_add(const sxu::FEGrid& G, const double* ec, float* z, hipStream_t s)
{
  const int64_t n = int64_t(G.nn[0]) * G.nn[1] * G.nn[2] * 3;
  hipLaunchKernelGGL(k_fe_prolong_add<float>, dim3(blocks_for(n)), dim3(kThreads), 0, s, G, ec, z);
  return hipGetLastError();
}
hipError_t cheb_init(const float* b, const float* dinv, double theta, float* x, float* d, float* r, int x_add,
                     int copy_r, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_cheb_init<float>, dim3(blocks_for(n)), dim3(kThreads), 0, s, b, dinv, 1.0 / theta, x, d, r,
                     x_add, copy_r, n);
  return hipGetLastError();
}
hipError_t cheb_step(float* r, const float* t, float* d, const float* dinv, double c1, double c2, float* x, int64_t n,
                     hipStream_t s)
{
  hipLaunchKernelGGL(k_cheb_step<float>, dim3(blocks_for(n)), dim3(kThreads), 0, s, r, t, d, dinv, c1, c2, x, n);
  return hipGetLastError();
}
hipError_t residual(const float* b, const float* t, float* r, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_residual<float>, dim3(blocks_for(n)), dim3(kThreads), 0, s, b, t, r, n);
  return hipGetLastError();
}
hipError_t sub_inplace(float* r, const float* t, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_sub_inplace<float>, dim3(blocks_for(n)), dim3(kThreads), 0, s, r, t, n);
  return hipGetLastError();
}
hipError_t apply_mask(float* x, const uint8_t* mask, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_apply_mask<float>, dim3(blocks_for(n)), dim3(kThreads), 0, s, x, mask, n);
  return hipGetLastError();
}
hipError_t to_f32(const double* x, float* y, const uint8_t* mask, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks_for(n)), dim3(kThreads), 0, s, x, y, mask, n);
  return hipGetLastError();
}
hipError_t to_f64(const float* x, double* y, int64_t n, hipStream_t s)
{
  const uint8_t* none = nullptr;
  hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks_for(n)), dim3(kThreads), 0, s, x, y, none, n);
  return hipGetLastError();
}

}  // namespace sxm
