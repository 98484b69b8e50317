// serac_oracle.hpp -- CPU FP64 restatement of serac's matrix-free residual / J.v path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
// oracle/liboracle.so and used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  The product library (serac_amd/) never
// includes, links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the LLNL/serac checkout; fn/ = src/serac/numerics/functional/).
//
// Pinning (how each part is checked independently of this restatement) is
// listed per function as "pinned by:"; see tests/test_oracle_*.py.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------
// 1D Gauss-Legendre rules on [0,1].
// Restates fn/polynomials.hpp:42-52 (nodes) and :271-283 (weights).  The
// literal decimals are the reference's (data, kept alike for 1e-12 parity).
// pinned by: exact integration of x^k, k <= 2q-1 (test_oracle_basis.py).
// ---------------------------------------------------------------------------
inline void gl_rule(int q, double* x, double* w)
{
  static const double X[8][8] = {
      {0.50000000000000000},
      {0.2113248654051871, 0.7886751345948129},
      {0.1127016653792583, 0.500000000000000, 0.887298334620742},
      {0.0694318442029737, 0.330009478207572, 0.669990521792428, 0.930568155797026},
      {0.0469100770306680, 0.230765344947158, 0.500000000000000, 0.769234655052842, 0.953089922969332},
      {0.0337652428984240, 0.169395306766868, 0.380690406958402, 0.619309593041598, 0.830604693233132,
       0.966234757101576},
      {0.0254460438286207, 0.129234407200303, 0.297077424311301, 0.500000000000000, 0.702922575688699,
       0.87076559279970, 0.97455395617138},
      {0.0198550717512319, 0.101666761293187, 0.237233795041836, 0.408282678752175, 0.591717321247825,
       0.76276620495816, 0.89833323870681, 0.98014492824877}};
  static const double W[8][8] = {
      {1.000000000000000},
      {0.500000000000000, 0.500000000000000},
      {0.277777777777778, 0.444444444444444, 0.277777777777778},
      {0.173927422568727, 0.326072577431273, 0.326072577431273, 0.173927422568727},
      {0.118463442528095, 0.239314335249683, 0.284444444444444, 0.239314335249683, 0.118463442528095},
      {0.085662246189585, 0.180380786524069, 0.233956967286346, 0.233956967286346, 0.180380786524069,
       0.085662246189585},
      {0.0647424830844348, 0.139852695744638, 0.190915025252559, 0.208979591836735, 0.190915025252559,
       0.139852695744638, 0.0647424830844348},
      {0.0506142681451881, 0.111190517226687, 0.156853322938944, 0.181341891689181, 0.181341891689181,
       0.156853322938944, 0.111190517226687, 0.0506142681451881}};
  if (q < 1 || q > 8) throw std::runtime_error("gl_rule: q must be in 1..8");
  for (int i = 0; i < q; i++) {
    x[i] = X[q - 1][i];
    w[i] = W[q - 1][i];
  }
}

// ---------------------------------------------------------------------------
// Gauss-Lobatto nodes on [0,1] for n = 2..5 points.
// n <= 4 restates fn/polynomials.hpp:27-33 (the reference's n=4 basis uses the
// exact nodes (5 -+ sqrt5)/10 through sqrt5, :506-508).  n = 5 (p = 4) is not
// in the reference (static_assert n <= 4, :496); its interior nodes are the
// roots of P4'(2x-1): 1/2 -+ sqrt(21)/14.
// pinned by: Kronecker property of the basis at these nodes; for n = 5 also the
// Lobatto rule built on them integrates x^k exactly for k <= 2n-3.
// ---------------------------------------------------------------------------
inline void gll_nodes(int n, double* x)
{
  if (n == 1) {
    x[0] = 0.5;
    return;
  }
  if (n == 2) {
    x[0] = 0.0;
    x[1] = 1.0;
    return;
  }
  if (n == 3) {
    x[0] = 0.0;
    x[1] = 0.5;
    x[2] = 1.0;
    return;
  }
  if (n == 4) {
    const double s5 = std::sqrt(5.0);
    x[0]            = 0.0;
    x[1]            = (5.0 - s5) / 10.0;
    x[2]            = (5.0 + s5) / 10.0;
    x[3]            = 1.0;
    return;
  }
  if (n == 5) {
    const double a = std::sqrt(21.0) / 14.0;
    x[0]           = 0.0;
    x[1]           = 0.5 - a;
    x[2]           = 0.5;
    x[3]           = 0.5 + a;
    x[4]           = 1.0;
    return;
  }
  throw std::runtime_error("gll_nodes: n must be in 1..5");
}

// Lagrange interpolating polynomials through the GLL nodes, and derivatives.
// Restates GaussLobattoInterpolation / GaussLobattoInterpolationDerivative
// (fn/polynomials.hpp:493-543) in product form, valid for every n (the
// reference's closed forms stop at n = 4).
// pinned by: B(x_j) = delta_ij, sum_i B_i = 1, sum_i G_i = 0, and agreement with
// the reference's closed forms for n <= 4 typed independently in the test.
inline void gll_basis(int n, double x, double* B)
{
  double nodes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  gll_nodes(n, nodes);
  for (int i = 0; i < n; i++) {
    double v = 1.0;
    for (int j = 0; j < n; j++)
      if (j != i) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    B[i] = v;
  }
}

inline void gll_dbasis(int n, double x, double* G)
{
  double nodes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  gll_nodes(n, nodes);
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int k = 0; k < n; k++) {
      if (k == i) continue;
      double v = 1.0 / (nodes[i] - nodes[k]);
      for (int j = 0; j < n; j++)
        if (j != i && j != k) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
      s += v;
    }
    G[i] = s;
  }
}

// 1D interpolation matrices at the GL points: B(q,i), G(q,i), optionally
// weighted by the 1D GL weight of row q (fn/detail/hexahedron_H1.inl:96-130).
struct Basis1D {
  int                 n = 0, q = 0;
  std::vector<double> B, G, Bw, Gw, xq, wq;  // [q][n] row-major
  Basis1D() = default;
  Basis1D(int n_, int q_) : n(n_), q(q_), B(q_ * n_), G(q_ * n_), Bw(q_ * n_), Gw(q_ * n_), xq(q_), wq(q_)
  {
    gl_rule(q, xq.data(), wq.data());
    for (int a = 0; a < q; a++) {
      gll_basis(n, xq[a], &B[a * n]);
      gll_dbasis(n, xq[a], &G[a * n]);
      for (int i = 0; i < n; i++) {
        Bw[a * n + i] = B[a * n + i] * wq[a];
        Gw[a * n + i] = G[a * n + i] * wq[a];
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Counter-based RNG (SplitMix64 finalizer), shared bit-for-bit with the
// product library's mesh generator and with the numpy field generators.
//   h = mix(seed); h = mix(h ^ stream); h = mix(h ^ idx); u = (h >> 11) * 2^-53
// ---------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double rng_u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
  uint64_t h = mix64(seed);
  h          = mix64(h ^ stream);
  h          = mix64(h ^ idx);
  return double(h >> 11) * (1.0 / 9007199254740992.0);
}
inline double rng_uniform(uint64_t seed, uint64_t stream, uint64_t idx, double a, double b)
{
  return a + (b - a) * rng_u01(seed, stream, idx);
}

// ---------------------------------------------------------------------------
// Hex mesh with trilinear geometry.  Element vertices use mfem's hexahedron
// vertex order (the order of data/meshes/*.mesh):
//   v0(0,0,0) v1(1,0,0) v2(1,1,0) v3(0,1,0) v4(0,0,1) v5(1,0,1) v6(1,1,1) v7(0,1,1)
// LEX2V maps a lexicographic corner a + 2b + 4c to that vertex slot.
// The structured generator restates buildCuboidMesh -> MakeCartesian3D
// (src/serac/mesh/mesh_utils.cpp:149-153): vertex (i,j,k) at (i/nx,j/ny,k/nz)*L,
// lexicographic vertex and element ids (mfem's own ordering is parity unpinned).
// ---------------------------------------------------------------------------
static const int LEX2V[8] = {0, 1, 3, 2, 4, 5, 7, 6};

struct Mesh {
  int                 nv = 0, ne = 0;
  std::vector<double> X;    // [nv][3]
  std::vector<int>    hex;  // [ne][8] in mfem vertex order
  bool                structured = false;
  int                 nx = 0, ny = 0, nz = 0;
  double              lx = 1, ly = 1, lz = 1;
};

inline Mesh cuboid(int nx, int ny, int nz, double lx, double ly, double lz)
{
  Mesh m;
  m.structured = true;
  m.nx = nx, m.ny = ny, m.nz = nz, m.lx = lx, m.ly = ly, m.lz = lz;
  m.nv = (nx + 1) * (ny + 1) * (nz + 1);
  m.ne = nx * ny * nz;
  m.X.resize(size_t(m.nv) * 3);
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        size_t v       = size_t(i) + size_t(nx + 1) * (size_t(j) + size_t(ny + 1) * size_t(k));
        m.X[3 * v + 0] = lx * double(i) / double(nx);
        m.X[3 * v + 1] = ly * double(j) / double(ny);
        m.X[3 * v + 2] = lz * double(k) / double(nz);
      }
  m.hex.resize(size_t(m.ne) * 8);
  auto vid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
  for (int ez = 0; ez < nz; ez++)
    for (int ey = 0; ey < ny; ey++)
      for (int ex = 0; ex < nx; ex++) {
        size_t e = size_t(ex) + size_t(nx) * (size_t(ey) + size_t(ny) * size_t(ez));
        for (int c = 0; c < 8; c++) {
          int a = c & 1, b = (c >> 1) & 1, d = (c >> 2) & 1;
          m.hex[8 * e + LEX2V[c]] = vid(ex + a, ey + b, ez + d);
        }
      }
  return m;
}

// Interior vertices of a structured mesh move by h_d * U[-amp, amp] per
// coordinate d, with RNG stream 1000 + d and index = vertex id.
inline void perturb(Mesh& m, uint64_t seed, double amp)
{
  if (!m.structured) throw std::runtime_error("perturb: structured meshes only");
  double h[3] = {m.lx / m.nx, m.ly / m.ny, m.lz / m.nz};
  for (int k = 1; k < m.nz; k++)
    for (int j = 1; j < m.ny; j++)
      for (int i = 1; i < m.nx; i++) {
        size_t v = size_t(i) + size_t(m.nx + 1) * (size_t(j) + size_t(m.ny + 1) * size_t(k));
        for (int d = 0; d < 3; d++) m.X[3 * v + d] += h[d] * rng_uniform(seed, 1000 + d, v, -amp, amp);
      }
}

// ---------------------------------------------------------------------------
// H1 space on hexes: per-element node lists in lexicographic local order
// (x fastest), the `dof_info` of fn/element_restriction.cpp:216-277.
// Structured meshes number the global (p*nx+1)x(p*ny+1)x(p*nz+1) node grid
// lexicographically.  Unstructured meshes number nodes by first appearance
// (elements in order, local nodes lexicographic) of a canonical topological
// key (vertex / edge / face / interior), which makes shared nodes agree
// between neighbours whatever their relative orientation.  mfem's own global
// numbering is parity unpinned (needs mfem).
// ---------------------------------------------------------------------------
struct Space {
  int              p = 1, nc = 1, ordering = 1;  // ordering: 0 byNODES, 1 byVDIM
  int              ne = 0, npe = 0;
  int64_t          nnodes = 0;
  std::vector<int> emap;  // [ne][npe] global node ids
  int64_t          ndof() const { return nnodes * nc; }
  // L index of (node, comp), fn/element_restriction.cpp:425-432
  int64_t lidx(int64_t node, int c) const { return ordering == 1 ? node * nc + c : int64_t(c) * nnodes + node; }
};

using NodeKey = std::array<int64_t, 6>;

inline NodeKey topo_key(const Mesh& m, int e, int p, int i, int j, int k)
{
  int  ijk[3] = {i, j, k};
  int  nfree  = 0;
  int  fr[3];
  for (int d = 0; d < 3; d++)
    if (ijk[d] != 0 && ijk[d] != p) fr[nfree++] = d;
  auto corner = [&](int a, int b, int c) { return int64_t(m.hex[8 * size_t(e) + LEX2V[a + 2 * b + 4 * c]]); };
  auto cfix   = [&](int d) { return ijk[d] == 0 ? 0 : 1; };
  if (nfree == 0) return {0, corner(cfix(0), cfix(1), cfix(2)), 0, 0, 0, 0};
  if (nfree == 3) return {3, e, i, j, k, 0};
  if (nfree == 1) {
    int d   = fr[0];
    int c0[3], c1[3];
    for (int q = 0; q < 3; q++) c0[q] = c1[q] = (q == d) ? 0 : cfix(q);
    c1[d]      = 1;
    int64_t gA = corner(c0[0], c0[1], c0[2]), gB = corner(c1[0], c1[1], c1[2]);
    int     t  = ijk[d];
    if (gA < gB) return {1, gA, gB, t, 0, 0};
    return {1, gB, gA, p - t, 0, 0};
  }
  // face: two free directions d1 < d2 with parameters s, t
  int d1 = fr[0], d2 = fr[1];
  int s = ijk[d1], t = ijk[d2];
  int64_t g[2][2];
  for (int u = 0; u < 2; u++)
    for (int v = 0; v < 2; v++) {
      int c[3];
      for (int q = 0; q < 3; q++) c[q] = cfix(q);
      c[d1]   = u;
      c[d2]   = v;
      g[u][v] = corner(c[0], c[1], c[2]);
    }
  int u0 = 0, v0 = 0;
  for (int u = 0; u < 2; u++)
    for (int v = 0; v < 2; v++)
      if (g[u][v] < g[u0][v0]) u0 = u, v0 = v;
  int64_t nu = g[1 - u0][v0], nv = g[u0][1 - v0];
  int     su = (u0 == 0) ? s : p - s;
  int     tv = (v0 == 0) ? t : p - t;
  if (nu < nv) return {2, g[u0][v0], nu, nv, su, tv};
  return {2, g[u0][v0], nv, nu, tv, su};
}

inline Space make_space(const Mesh& m, int p, int nc, int ordering)
{
  Space s;
  s.p = p, s.nc = nc, s.ordering = ordering, s.ne = m.ne;
  int n = p + 1;
  s.npe = n * n * n;
  s.emap.resize(size_t(m.ne) * s.npe);
  if (m.structured) {
    int64_t gx = int64_t(p) * m.nx + 1, gy = int64_t(p) * m.ny + 1, gz = int64_t(p) * m.nz + 1;
    s.nnodes = gx * gy * gz;
    for (int ez = 0; ez < m.nz; ez++)
      for (int ey = 0; ey < m.ny; ey++)
        for (int ex = 0; ex < m.nx; ex++) {
          size_t e = size_t(ex) + size_t(m.nx) * (size_t(ey) + size_t(m.ny) * size_t(ez));
          for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
              for (int i = 0; i < n; i++) {
                int64_t X = int64_t(p) * ex + i, Y = int64_t(p) * ey + j, Z = int64_t(p) * ez + k;
                s.emap[e * s.npe + size_t(i + n * (j + n * k))] = int(X + gx * (Y + gy * Z));
              }
        }
    return s;
  }
  std::map<NodeKey, int> ids;
  int                    next = 0;
  for (int e = 0; e < m.ne; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          NodeKey key = topo_key(m, e, p, i, j, k);
          auto    it  = ids.find(key);
          int     id;
          if (it == ids.end()) {
            id = next++;
            ids.emplace(key, id);
          } else {
            id = it->second;
          }
          s.emap[size_t(e) * s.npe + size_t(i + n * (j + n * k))] = id;
        }
  s.nnodes = next;
  return s;
}

// Trilinear map of element e at parent point (xi, eta, zeta) in [0,1]^3.
inline void trilinear(const Mesh& m, int e, const double xi[3], double x[3])
{
  double N[2][3];
  for (int d = 0; d < 3; d++) {
    N[0][d] = 1.0 - xi[d];
    N[1][d] = xi[d];
  }
  x[0] = x[1] = x[2] = 0.0;
  for (int c = 0; c < 8; c++) {
    int    a = c & 1, b = (c >> 1) & 1, d = (c >> 2) & 1;
    double w = N[a][0] * N[b][1] * N[d][2];
    int    v = m.hex[8 * size_t(e) + LEX2V[c]];
    for (int q = 0; q < 3; q++) x[q] += w * m.X[3 * size_t(v) + q];
  }
}

// Physical coordinates of every node (GLL points of each cell's trilinear map).
inline std::vector<double> node_coords(const Mesh& m, const Space& s)
{
  int                 n = s.p + 1;
  double              g[8];
  std::vector<double> xyz(size_t(s.nnodes) * 3, 0.0);
  gll_nodes(n, g);
  for (int e = 0; e < m.ne; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          double xi[3] = {g[i], g[j], g[k]}, x[3];
          trilinear(m, e, xi, x);
          int64_t node = s.emap[size_t(e) * s.npe + size_t(i + n * (j + n * k))];
          for (int q = 0; q < 3; q++) xyz[3 * node + q] = x[q];
        }
  return xyz;
}

// Gather / ScatterAdd, fn/element_restriction.cpp:443-467.
// E layout: [e][comp][lexicographic node] (:448).
inline void gather(const Space& s, const double* L, double* E)
{
  for (int e = 0; e < s.ne; e++)
    for (int c = 0; c < s.nc; c++)
      for (int j = 0; j < s.npe; j++)
        E[(size_t(e) * s.nc + c) * s.npe + j] = L[s.lidx(s.emap[size_t(e) * s.npe + j], c)];
}
inline void scatter_add(const Space& s, const double* E, double* L)
{
  for (int e = 0; e < s.ne; e++)
    for (int c = 0; c < s.nc; c++)
      for (int j = 0; j < s.npe; j++)
        L[s.lidx(s.emap[size_t(e) * s.npe + j], c)] += E[(size_t(e) * s.nc + c) * s.npe + j];
}

// ---------------------------------------------------------------------------
// Sum-factorized interpolate / integrate, fn/detail/hexahedron_H1.inl:169-280.
// X: [c][z][y][x] nodal values (n per dim).  Outputs per quadrature point
// q = (qz*Q + qy)*Q + qx (:255): value[c][q], grad[c][d][q] (parent coords).
// pinned by: equality with the full-basis evaluation below (shape_functions,
// :39-84), adjointness <Iu, w> = <u, I^T w>, and polynomial reproduction.
// ---------------------------------------------------------------------------
template <typename T>
inline void interpolate(const Basis1D& b, int nc, const T* X, T* value, T* grad)
{
  const int n = b.n, Q = b.q;
  thread_local std::vector<T> A10, A11, A20, A21, A22;  // scratch reused across elements
  A10.resize(n * n * Q), A11.resize(n * n * Q), A20.resize(n * Q * Q), A21.resize(n * Q * Q), A22.resize(n * Q * Q);
  const int nq = Q * Q * Q;
  for (int c = 0; c < nc; c++) {
    const T* Xc = X + size_t(c) * n * n * n;
    // contract x: A1*(z,y,qx)
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int qx = 0; qx < Q; qx++) {
          T s0 = T(0), s1 = T(0);
          for (int x = 0; x < n; x++) {
            s0 = s0 + b.B[qx * n + x] * Xc[(z * n + y) * n + x];
            s1 = s1 + b.G[qx * n + x] * Xc[(z * n + y) * n + x];
          }
          A10[(z * n + y) * Q + qx] = s0;
          A11[(z * n + y) * Q + qx] = s1;
        }
    // contract y: A2*(z,qy,qx)
    for (int z = 0; z < n; z++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++) {
          T s0 = T(0), s1 = T(0), s2 = T(0);
          for (int y = 0; y < n; y++) {
            s0 = s0 + b.B[qy * n + y] * A10[(z * n + y) * Q + qx];
            s1 = s1 + b.B[qy * n + y] * A11[(z * n + y) * Q + qx];
            s2 = s2 + b.G[qy * n + y] * A10[(z * n + y) * Q + qx];
          }
          A20[(z * Q + qy) * Q + qx] = s0;
          A21[(z * Q + qy) * Q + qx] = s1;
          A22[(z * Q + qy) * Q + qx] = s2;
        }
    // contract z
    for (int qz = 0; qz < Q; qz++)
      for (int qy = 0; qy < Q; qy++)
        for (int qx = 0; qx < Q; qx++) {
          T v = T(0), g0 = T(0), g1 = T(0), g2 = T(0);
          for (int z = 0; z < n; z++) {
            v  = v + b.B[qz * n + z] * A20[(z * Q + qy) * Q + qx];
            g0 = g0 + b.B[qz * n + z] * A21[(z * Q + qy) * Q + qx];
            g1 = g1 + b.B[qz * n + z] * A22[(z * Q + qy) * Q + qx];
            g2 = g2 + b.G[qz * n + z] * A20[(z * Q + qy) * Q + qx];
          }
          int q                        = (qz * Q + qy) * Q + qx;
          value[size_t(c) * nq + q]    = v;
          grad[(size_t(c) * 3 + 0) * nq + q] = g0;
          grad[(size_t(c) * 3 + 1) * nq + q] = g1;
          grad[(size_t(c) * 3 + 2) * nq + q] = g2;
        }
  }
}

// source[c][q] (may be null), flux[c][d][q] (may be null); r[c][node] +=.
inline void integrate(const Basis1D& b, int nc, const double* source, const double* flux, double* r)
{
  const int n = b.n, Q = b.q, nq = Q * Q * Q;
  thread_local std::vector<double> A20, A21, A22, A10, A11;  // scratch reused across elements
  A20.resize(Q * Q * n), A21.resize(Q * Q * n), A22.resize(Q * Q * n), A10.resize(Q * n * n), A11.resize(Q * n * n);
  for (int c = 0; c < nc; c++) {
    auto S = [&](int q) { return source ? source[size_t(c) * nq + q] : 0.0; };
    auto F = [&](int d, int q) { return flux ? flux[(size_t(c) * 3 + d) * nq + q] : 0.0; };
    for (int qz = 0; qz < Q; qz++)
      for (int qy = 0; qy < Q; qy++)
        for (int x = 0; x < n; x++) {
          double s0 = 0, s1 = 0, s2 = 0;
          for (int qx = 0; qx < Q; qx++) {
            int q = (qz * Q + qy) * Q + qx;
            s0 += S(q) * b.Bw[qx * n + x] + F(0, q) * b.Gw[qx * n + x];
            s1 += F(1, q) * b.Bw[qx * n + x];
            s2 += F(2, q) * b.Bw[qx * n + x];
          }
          A20[(qz * Q + qy) * n + x] = s0;
          A21[(qz * Q + qy) * n + x] = s1;
          A22[(qz * Q + qy) * n + x] = s2;
        }
    for (int qz = 0; qz < Q; qz++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          double s0 = 0, s1 = 0;
          for (int qy = 0; qy < Q; qy++) {
            s0 += A20[(qz * Q + qy) * n + x] * b.Bw[qy * n + y] + A21[(qz * Q + qy) * n + x] * b.Gw[qy * n + y];
            s1 += A22[(qz * Q + qy) * n + x] * b.Bw[qy * n + y];
          }
          A10[(qz * n + y) * n + x] = s0;
          A11[(qz * n + y) * n + x] = s1;
        }
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          double s = 0;
          for (int qz = 0; qz < Q; qz++)
            s += A10[(qz * n + y) * n + x] * b.Bw[qz * n + z] + A11[(qz * n + y) * n + x] * b.Gw[qz * n + z];
          r[(size_t(c) * n + z) * n * n + y * n + x] += s;
        }
  }
}

// Full-basis (brute force) variants: phi_j(q) = B(qx,jx)B(qy,jy)B(qz,jz) etc.,
// the shape_functions / shape_function_gradients of hexahedron_H1.inl:39-84
// evaluated at the tensor GL points.  Used only to pin the sum factorization.
inline void interpolate_brute(const Basis1D& b, int nc, const double* X, double* value, double* grad)
{
  const int n = b.n, Q = b.q, nq = Q * Q * Q, nd = n * n * n;
  for (int c = 0; c < nc; c++)
    for (int q = 0; q < nq; q++) {
      int    qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
      double v = 0, g[3] = {0, 0, 0};
      for (int j = 0; j < nd; j++) {
        int    jx = j % n, jy = (j / n) % n, jz = j / (n * n);
        double u  = X[size_t(c) * nd + j];
        double bx = b.B[qx * n + jx], by = b.B[qy * n + jy], bz = b.B[qz * n + jz];
        double gx = b.G[qx * n + jx], gy = b.G[qy * n + jy], gz = b.G[qz * n + jz];
        v += bx * by * bz * u;
        g[0] += gx * by * bz * u;
        g[1] += bx * gy * bz * u;
        g[2] += bx * by * gz * u;
      }
      value[size_t(c) * nq + q] = v;
      for (int d = 0; d < 3; d++) grad[(size_t(c) * 3 + d) * nq + q] = g[d];
    }
}

inline void integrate_brute(const Basis1D& b, int nc, const double* source, const double* flux, double* r)
{
  const int n = b.n, Q = b.q, nq = Q * Q * Q, nd = n * n * n;
  for (int c = 0; c < nc; c++)
    for (int j = 0; j < nd; j++) {
      int    jx = j % n, jy = (j / n) % n, jz = j / (n * n);
      double s  = 0;
      for (int q = 0; q < nq; q++) {
        int    qx = q % Q, qy = (q / Q) % Q, qz = q / (Q * Q);
        double w  = b.wq[qx] * b.wq[qy] * b.wq[qz];
        double bx = b.B[qx * n + jx], by = b.B[qy * n + jy], bz = b.B[qz * n + jz];
        double gx = b.G[qx * n + jx], gy = b.G[qy * n + jy], gz = b.G[qz * n + jz];
        double t  = 0;
        if (source) t += source[size_t(c) * nq + q] * bx * by * bz;
        if (flux) {
          t += flux[(size_t(c) * 3 + 0) * nq + q] * gx * by * bz;
          t += flux[(size_t(c) * 3 + 1) * nq + q] * bx * gy * bz;
          t += flux[(size_t(c) * 3 + 2) * nq + q] * bx * by * gz;
        }
        s += w * t;
      }
      r[size_t(c) * nd + j] += s;
    }
}

// ---------------------------------------------------------------------------
// Geometric factors, fn/geometric_factors.cpp:17-58: interpolate the order-1
// mesh nodes of each element at the Q^3 GL points.
//   X[e][i][q] = x_i,   J[e][j][i][q] = d x_i / d xi_j.
// pinned by: constant J on affine meshes; sum_q detJ w = volume of the cube
// even with perturbed interior vertices (trilinear cells tile it).
// ---------------------------------------------------------------------------
inline void geometric_factors(const Mesh& m, int Q, std::vector<double>& Xq, std::vector<double>& Jq)
{
  Basis1D b(2, Q);
  int     nq = Q * Q * Q;
  Xq.assign(size_t(m.ne) * 3 * nq, 0.0);
  Jq.assign(size_t(m.ne) * 9 * nq, 0.0);
  std::vector<double> Xe(24), val(3 * nq), grd(9 * nq);
  for (int e = 0; e < m.ne; e++) {
    for (int c = 0; c < 8; c++) {
      int v = m.hex[8 * size_t(e) + LEX2V[c]];
      for (int i = 0; i < 3; i++) Xe[i * 8 + c] = m.X[3 * size_t(v) + i];
    }
    interpolate(b, 3, Xe.data(), val.data(), grd.data());
    for (int q = 0; q < nq; q++)
      for (int i = 0; i < 3; i++) {
        Xq[(size_t(e) * 3 + i) * nq + q] = val[i * nq + q];
        for (int j = 0; j < 3; j++) Jq[((size_t(e) * 3 + j) * 3 + i) * nq + q] = grd[(i * 3 + j) * nq + q];
      }
  }
}

// ---------------------------------------------------------------------------
// Forward-mode dual numbers with an N-wide gradient (fn/dual.hpp:28-147,
// 352-372).  Each gradient slot is propagated by the same arithmetic as the
// value, so slot k equals a single-direction derivative seeded at input k.
// ---------------------------------------------------------------------------
template <int N>
struct dual {
  double                value = 0;
  std::array<double, N> gradient{};
  dual() = default;
  dual(double v) : value(v) { gradient.fill(0.0); }
};
template <int N>
inline dual<N> operator+(const dual<N>& a, const dual<N>& b)
{
  dual<N> r;
  r.value = a.value + b.value;
  for (int i = 0; i < N; i++) r.gradient[i] = a.gradient[i] + b.gradient[i];
  return r;
}
template <int N>
inline dual<N> operator-(const dual<N>& a, const dual<N>& b)
{
  dual<N> r;
  r.value = a.value - b.value;
  for (int i = 0; i < N; i++) r.gradient[i] = a.gradient[i] - b.gradient[i];
  return r;
}
template <int N>
inline dual<N> operator-(const dual<N>& a)
{
  dual<N> r;
  r.value = -a.value;
  for (int i = 0; i < N; i++) r.gradient[i] = -a.gradient[i];
  return r;
}
template <int N>
inline dual<N> operator*(const dual<N>& a, const dual<N>& b)
{
  dual<N> r;
  r.value = a.value * b.value;
  for (int i = 0; i < N; i++) r.gradient[i] = a.value * b.gradient[i] + a.gradient[i] * b.value;
  return r;
}
template <int N>
inline dual<N> operator*(double a, const dual<N>& b)
{
  dual<N> r;
  r.value = a * b.value;
  for (int i = 0; i < N; i++) r.gradient[i] = a * b.gradient[i];
  return r;
}
template <int N>
inline dual<N> operator*(const dual<N>& b, double a)
{
  return a * b;
}
template <int N>
inline dual<N> operator+(const dual<N>& a, double b)
{
  dual<N> r = a;
  r.value += b;
  return r;
}
template <int N>
inline dual<N> operator+(double b, const dual<N>& a)
{
  return a + b;
}
template <int N>
inline dual<N> operator-(const dual<N>& a, double b)
{
  dual<N> r = a;
  r.value -= b;
  return r;
}
template <int N>
inline dual<N> operator/(const dual<N>& a, const dual<N>& b)
{
  dual<N> r;
  r.value = a.value / b.value;
  for (int i = 0; i < N; i++) r.gradient[i] = (a.gradient[i] * b.value - a.value * b.gradient[i]) / (b.value * b.value);
  return r;
}
template <int N>
inline dual<N> operator/(const dual<N>& a, double b)
{
  dual<N> r;
  r.value = a.value / b;
  for (int i = 0; i < N; i++) r.gradient[i] = a.gradient[i] / b;
  return r;
}
template <int N>
inline dual<N> operator/(double a, const dual<N>& b)
{
  dual<N> r;
  r.value = a / b.value;
  for (int i = 0; i < N; i++) r.gradient[i] = -(a / (b.value * b.value)) * b.gradient[i];
  return r;
}
template <int N>
inline dual<N> log1p(const dual<N>& a)
{
  dual<N> r;
  r.value = std::log1p(a.value);
  for (int i = 0; i < N; i++) r.gradient[i] = a.gradient[i] / (1.0 + a.value);
  return r;
}
inline double log1p(double a) { return std::log1p(a); }
template <int N>
inline dual<N> sqrt(const dual<N>& a)
{
  dual<N> r;
  r.value = std::sqrt(a.value);
  for (int i = 0; i < N; i++) r.gradient[i] = a.gradient[i] / (2.0 * r.value);
  return r;
}
inline double value_of(double a) { return a; }
// r.gradient = -q.gradient / df_dx: the derivative of the root x(q) of
// f(x, q) = 0 when df/dq = 1 (fn/tuple_tensor_dual_functions.hpp:661-666)
inline void set_root_dual(double&, double, double) {}
template <int N>
inline void set_root_dual(dual<N>& r, const dual<N>& q, double df_dx)
{
  for (int i = 0; i < N; i++) r.gradient[i] = -q.gradient[i] / df_dx;
}
template <int N>
inline double value_of(const dual<N>& a)
{
  return a.value;
}

// 3x3 helpers, fn/tensor.hpp:1242-1290 (det, detApIm1), :1610-1630 (inv).
template <typename T>
struct M3 {
  T a[3][3];
};
template <typename T>
inline T det3(const M3<T>& A)
{
  return A.a[0][0] * A.a[1][1] * A.a[2][2] + A.a[0][1] * A.a[1][2] * A.a[2][0] + A.a[0][2] * A.a[1][0] * A.a[2][1] -
         A.a[0][0] * A.a[1][2] * A.a[2][1] - A.a[0][1] * A.a[1][0] * A.a[2][2] - A.a[0][2] * A.a[1][1] * A.a[2][0];
}
// det(A + I) - 1 without cancellation: tr(A) + I2(A) + det(A)
template <typename T>
inline T detApIm1(const M3<T>& A)
{
  auto a = [&](int i, int j) -> const T& { return A.a[i][j]; };
  return a(0, 0) + a(1, 1) + a(2, 2) - a(0, 1) * a(1, 0) * (1.0 + a(2, 2)) + a(0, 0) * a(1, 1) * (1.0 + a(2, 2)) -
         a(0, 2) * a(2, 0) * (1.0 + a(1, 1)) - a(1, 2) * a(2, 1) * (1.0 + a(0, 0)) + a(0, 0) * a(2, 2) +
         a(1, 1) * a(2, 2) + a(0, 1) * a(1, 2) * a(2, 0) + a(0, 2) * a(1, 0) * a(2, 1);
}
template <typename T>
inline M3<T> inv3(const M3<T>& A)
{
  T     inv_det = 1.0 / det3(A);
  M3<T> R;
  auto  a   = [&](int i, int j) -> const T& { return A.a[i][j]; };
  R.a[0][0] = (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) * inv_det;
  R.a[0][1] = (a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2)) * inv_det;
  R.a[0][2] = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) * inv_det;
  R.a[1][0] = (a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2)) * inv_det;
  R.a[1][1] = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) * inv_det;
  R.a[1][2] = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) * inv_det;
  R.a[2][0] = (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0)) * inv_det;
  R.a[2][1] = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) * inv_det;
  R.a[2][2] = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) * inv_det;
  return R;
}

// ---------------------------------------------------------------------------
// Quadrature-point functions.  Each takes physical-space inputs (value[c],
// grad[c][3]) and returns (source[c], flux[c][3]) in physical space, like the
// integrands the reference calls at fn/domain_integral_kernels.hpp:106-150.
// ---------------------------------------------------------------------------
enum QFKind { QF_HEAT = 1, QF_MASS_DIFFUSION = 2, QF_NEOHOOKEAN = 3, QF_LINEAR_ELASTIC = 4, QF_J2_LINEAR = 5 };

struct QFParams {
  int    kind = QF_HEAT;
  double p[16] = {0};
};

// Heat: ThermalMaterialIntegrand(LinearIsotropicConductor) + setSource
// (src/serac/physics/heat_transfer.hpp:381-406, :493-503;
//  physics/materials/thermal_material.hpp:50-55), steady state:
//   source = -s,  flux = -(-k grad T) = k grad T.     p = {k, s}
// MassDiffusion: {a u, b grad u} (fn/tests/functional_comparisons.cpp:32-33,
//   the mass + diffusion comparison).                   p = {a, b}
// NeoHookean: MaterialStressFunctor<NeoHookean>, geometric nonlinearity on
//   (physics/solid_mechanics.hpp:834-852, materials/solid_material.hpp:112-122):
//   sigma = (lambda log1p(J-1) I + G (H H^T + H^T + H)) / J, lambda = K - 2G/3,
//   flux = sigma . inv(I+H)^T . det(I+H), source = density * 0 (quasi-static).
//                                                       p = {density, K, G}
// LinearElastic: sigma = lambda tr(eps) I + 2 mu eps, flux = sigma.  p = {lambda, mu}
// state size in doubles per quadrature point (0: stateless)
inline int qf_state_size(int kind) { return kind == QF_J2_LINEAR ? 10 : 0; }

// flux = stress . inv(F)^T . det(F), F = I + H with geometric nonlinearity,
// else F = I (MaterialStressFunctor, phy/solid_mechanics.hpp:843-849)
template <typename T>
inline void stress_to_flux(const M3<T>& sigma, const T* H, bool geom_nonlin, T* flux)
{
  if (!geom_nonlin) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) flux[i * 3 + j] = sigma.a[i][j];
    return;
  }
  M3<T> F;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) F.a[i][j] = H[i * 3 + j];
  for (int i = 0; i < 3; i++) F.a[i][i] = F.a[i][i] + 1.0;
  M3<T> Fi = inv3(F);
  T     dF = det3(F);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      // (sigma . Fi^T)_ij = sum_k sigma_ik Fi_jk
      T s = sigma.a[i][0] * Fi.a[j][0] + sigma.a[i][1] * Fi.a[j][1] + sigma.a[i][2] * Fi.a[j][2];
      flux[i * 3 + j] = s * dF;
    }
}

// solve_scalar_equation (fn/tuple_tensor_dual_functions.hpp:573-672) on the
// J2 consistency residual f(x) = q - a x - (sigma_y + Hi (eqps_old + x)),
// a = 3G + Hk: Newton safeguarded by bisection inside [lo, hi], on values.
// df_dx is the gradient f(make_dual(x)) carries, -a - Hi.
struct ScalarSolve {
  double x = 0, df_dx = 0;
  bool   converged  = false;
  int    iterations = 0;
};
inline ScalarSolve j2_scalar_solve(double q, double a, double sy, double Hi, double e0, double x0, double lo,
                                   double hi, double xtol, double rtol, int max_iter)
{
  auto         f    = [&](double x) { return q - a * x - (sy + Hi * (e0 + x)); };
  const double dfdx = -a - Hi;
  double       fl = f(lo), fh = f(hi);
  if (fl * fh > 0) throw std::runtime_error("solve_scalar_equation: root not bracketed by input bounds");
  ScalarSolve r;
  double      x = 0.0;
  if (fl == 0) {
    x           = lo;
    r.converged = true;
  } else if (fh == 0) {
    x           = hi;
    r.converged = true;
  }
  double df = dfdx;
  if (!r.converged) {
    double xl = lo, xh = hi;
    if (fl > 0) {
      xl = hi;
      xh = lo;
    }
    if (x0 < lo || x0 > hi) x0 = 0.5 * (lo + hi);
    x              = x0;
    double dx_old  = std::abs(hi - lo);
    double dx      = dx_old;
    double fval    = f(x);
    while (!r.converged) {
      if (r.iterations == max_iter) break;
      if ((x - xh) * df - fval > 0 || (x - xl) * df - fval < 0 || std::abs(2. * fval) > std::abs(dx_old * df)) {
        dx_old      = dx;
        dx          = 0.5 * (xh - xl);
        x           = xl + dx;
        r.converged = (x == xl);
      } else {
        dx_old        = dx;
        dx            = fval / df;
        const double t = x;
        x -= dx;
        r.converged = (x == t);
      }
      fval        = f(x);
      df          = dfdx;
      r.converged = r.converged || (std::abs(dx) < xtol) || (std::abs(fval) < rtol);
      if (fval < 0)
        xl = x;
      else
        xh = x;
      ++r.iterations;
    }
  }
  r.x     = x;
  r.df_dx = df;
  return r;
}

// J2SmallStrain<LinearHardening> (phy/materials/solid_material.hpp:132-148,
// 199-258).  H = du/dX (3x3 row-major); state = {plastic strain [3][3]
// row-major, accumulated plastic strain}, updated in place when the return
// map runs.  p = {E, nu, sigma_y, Hi, Hk, ...}.
template <typename T>
inline M3<T> j2_small_strain(const double* p, const T* H, double* state)
{
  const double   E = p[0], nu = p[1], sy = p[2], Hi = p[3], Hk = p[4];
  constexpr double tol = 1e-10;
  const double   K = E / (3.0 * (1.0 - 2.0 * nu));
  const double   G = 0.5 * E / (1.0 + nu);
  T              el[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) el[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]) - state[i * 3 + j];
  T tr_el = T(0.0);
  for (int i = 0; i < 3; i++) tr_el = tr_el + el[i * 4];
  const T pr = K * tr_el;
  T       s[9];
  for (int k = 0; k < 9; k++) s[k] = el[k];
  for (int i = 0; i < 3; i++) s[i * 4] = s[i * 4] - tr_el / 3.0;
  for (int k = 0; k < 9; k++) s[k] = (2.0 * G) * s[k];
  T eta[9];
  for (int k = 0; k < 9; k++) eta[k] = s[k] - ((2.0 / 3.0) * Hk) * state[k];
  T sq = T(0.0);
  for (int k = 0; k < 9; k++) sq = sq + eta[k] * eta[k];
  using std::sqrt;
  const T      q   = std::sqrt(1.5) * sqrt(sq);
  const double e0  = state[9];
  const double a   = 3.0 * G + Hk;
  const double qv  = value_of(q);
  if (qv - a * 0.0 - (sy + Hi * (e0 + 0.0)) > tol * sy) {
    const double upper = (qv - (sy + Hi * e0)) / a;
    ScalarSolve  r     = j2_scalar_solve(qv, a, sy, Hi, e0, 0.0, 0.0, upper, 0.0, tol * sy, 25);
    // dual root: x' = -df / df_dx with df = q' (f(x, q) carries q's gradient)
    T delta = T(r.x);
    set_root_dual(delta, q, r.df_dx);
    T Np[9];
    for (int k = 0; k < 9; k++) Np[k] = (1.5 * eta[k]) / q;
    for (int k = 0; k < 9; k++) s[k] = s[k] - ((2.0 * G) * delta) * Np[k];
    state[9] += r.x;
    for (int k = 0; k < 9; k++) state[k] += r.x * value_of(Np[k]);
  }
  M3<T> sigma;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sigma.a[i][j] = (i == j) ? s[i * 3 + j] + pr : s[i * 3 + j];
  return sigma;
}

template <typename T>
inline void qf_eval(const QFParams& P, int nc, const T* value, const T* grad, T* source, T* flux,
                    double* state = nullptr)
{
  switch (P.kind) {
    case QF_HEAT: {
      source[0] = T(-P.p[1]);
      for (int d = 0; d < 3; d++) flux[d] = P.p[0] * grad[d];
      return;
    }
    case QF_MASS_DIFFUSION: {
      for (int c = 0; c < nc; c++) {
        source[c] = P.p[0] * value[c];
        for (int d = 0; d < 3; d++) flux[c * 3 + d] = P.p[1] * grad[c * 3 + d];
      }
      return;
    }
    case QF_NEOHOOKEAN: {
      const double K = P.p[1], G = P.p[2];
      const double lambda = K - (2.0 / 3.0) * G;
      M3<T>        H;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H.a[i][j] = grad[i * 3 + j];
      T     Jm1 = detApIm1(H);
      T     J   = Jm1 + 1.0;
      T     L   = lambda * log1p(Jm1);
      M3<T> sigma;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          T hht = H.a[i][0] * H.a[j][0] + H.a[i][1] * H.a[j][1] + H.a[i][2] * H.a[j][2];
          T bmi = hht + H.a[j][i] + H.a[i][j];
          T v   = G * bmi;
          if (i == j) v = v + L;
          sigma.a[i][j] = v / J;
        }
      M3<T> F = H;
      for (int i = 0; i < 3; i++) F.a[i][i] = F.a[i][i] + 1.0;
      M3<T> Fi = inv3(F);
      T     dF = det3(F);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          // (sigma . Fi^T)_ij = sum_k sigma_ik Fi_jk
          T s = sigma.a[i][0] * Fi.a[j][0] + sigma.a[i][1] * Fi.a[j][1] + sigma.a[i][2] * Fi.a[j][2];
          flux[i * 3 + j] = s * dF;
        }
      for (int c = 0; c < 3; c++) source[c] = T(0.0);
      return;
    }
    case QF_J2_LINEAR: {
      // MaterialStressFunctor<J2SmallStrain<LinearHardening>>, quasi-static;
      // p = {E, nu, sigma_y, Hi, Hk, density, geom_nonlin}
      if (!state) throw std::runtime_error("qf_eval: J2 needs quadrature-point state");
      M3<T> sigma = j2_small_strain(P.p, grad, state);
      stress_to_flux(sigma, grad, P.p[6] != 0.0, flux);
      for (int c = 0; c < 3; c++) source[c] = T(0.0);
      return;
    }
    case QF_LINEAR_ELASTIC: {
      const double lam = P.p[0], mu = P.p[1];
      T            tr  = grad[0] + grad[4] + grad[8];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          T v = mu * (grad[i * 3 + j] + grad[j * 3 + i]);
          if (i == j) v = v + lam * tr;
          flux[i * 3 + j] = v;
        }
      for (int c = 0; c < 3; c++) source[c] = T(0.0);
      return;
    }
  }
  throw std::runtime_error("qf_eval: unknown qfunction");
}

// The integrands of two trial arguments (DependsOn<0, 1>): the solid material
// with its acceleration, MaterialStressFunctor (phy/solid_mechanics.hpp:834-852:
// source density * d2u/dt2, the flux of the quasi-static material), and the
// thermal material with its rate, ThermalMaterialIntegrand
// (phy/heat_transfer.hpp:380-406: source heat_capacity * du/dt, plus the
// constant source the heat integrand here carries).  rate_coefficient gives
// that density or volumetric heat capacity from the parameter list; 0 marks a
// qfunction without a two-argument form.
inline double rate_coefficient(const QFParams& P)
{
  switch (P.kind) {
    case QF_NEOHOOKEAN: return P.p[0];      // {density, K, G}
    case QF_J2_LINEAR: return P.p[5];       // {E, nu, sigma_y, Hi, Hk, density, geom}
    case QF_LINEAR_ELASTIC: return P.p[2];  // {lambda, mu, density}
    case QF_HEAT: return P.p[2];            // {k, s, rho c_p}
  }
  return 0.0;
}

// qf_eval with the value of the second argument (rate) at the point
template <typename T>
inline void qf_eval2(const QFParams& P, int nc, const T* value, const T* grad, const double* rate, T* source,
                     T* flux, double* state = nullptr)
{
  qf_eval(P, nc, value, grad, source, flux, state);
  const double c = rate_coefficient(P);
  for (int i = 0; i < nc; i++) source[i] = source[i] + c * rate[i];
}

// ---------------------------------------------------------------------------
// Boundary qfunctions (the integrands AddBoundaryIntegral receives,
// fn/boundary_integral_kernels.hpp:106-130): reference position X, dX/dxi
// (3 x 2, dX[i][k] = d X_i / d xi_k), field value u[nc] and face-parent
// derivative du[nc][2] -> source[nc] (the caller multiplies by
// |dX/dxi1 x dX/dxi2|).
//   TRACTION: setTraction (phy/solid_mechanics.hpp:1023-1037), traction(X, n) =
//     A n + b, source = -1 * (A n_hat + b).              p = {A (9), b (3)}
//   PRESSURE: setPressure (:1071-1101), source = p n / |dX/dxi1 x dX/dxi2| with
//     n = dx/dxi1 x dx/dxi2, x = X + u if geom_nonlin else X. p = {p, geom}
//   FLUX: setFluxBCs (phy/heat_transfer.hpp:536-552), flux = q0 + h (T - T_inf).
//                                                        p = {q0, h, T_inf}
// ---------------------------------------------------------------------------
enum BQFKind { BQF_TRACTION = 101, BQF_PRESSURE = 102, BQF_FLUX = 103 };
inline int bqf_components(int kind) { return kind == BQF_FLUX ? 1 : 3; }

template <typename T>
inline void bqf_eval(const QFParams& P, const double* X, const double dX[3][2], const T* u, const T* du, T* src)
{
  (void)X;
  switch (P.kind) {
    case BQF_TRACTION: {
      double n[3] = {dX[1][0] * dX[2][1] - dX[2][0] * dX[1][1], dX[2][0] * dX[0][1] - dX[0][0] * dX[2][1],
                     dX[0][0] * dX[1][1] - dX[1][0] * dX[0][1]};
      double nn   = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int i = 0; i < 3; i++) n[i] = n[i] / nn;
      for (int i = 0; i < 3; i++)
        src[i] = T(-1.0 * (P.p[i * 3 + 0] * n[0] + P.p[i * 3 + 1] * n[1] + P.p[i * 3 + 2] * n[2] + P.p[9 + i]));
      return;
    }
    case BQF_PRESSURE: {
      const bool g = P.p[1] != 0.0;
      T          a[3], b[3];
      for (int i = 0; i < 3; i++) {
        a[i] = g ? T(dX[i][0]) + du[i * 2 + 0] : T(dX[i][0]);
        b[i] = g ? T(dX[i][1]) + du[i * 2 + 1] : T(dX[i][1]);
      }
      T      n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      double r[3] = {dX[1][0] * dX[2][1] - dX[2][0] * dX[1][1], dX[2][0] * dX[0][1] - dX[0][0] * dX[2][1],
                     dX[0][0] * dX[1][1] - dX[1][0] * dX[0][1]};
      double nr   = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      for (int i = 0; i < 3; i++) src[i] = P.p[0] * (n[i] / nr);
      return;
    }
    case BQF_FLUX: {
      src[0] = P.p[0] + P.p[1] * (u[0] - P.p[2]);
      return;
    }
  }
  throw std::runtime_error("bqf_eval: unknown boundary qfunction");
}

// Which derivative blocks the reference's compile-time `zero` sparsity keeps
// (fn/tensor.hpp:123-127): [source|flux] x [value|grad].
inline void qf_blocks(int kind, bool blk[2][2])
{
  blk[0][0] = blk[0][1] = blk[1][0] = blk[1][1] = false;
  switch (kind) {
    case QF_HEAT: blk[1][1] = true; break;
    case QF_MASS_DIFFUSION: blk[0][0] = blk[1][1] = true; break;
    case QF_NEOHOOKEAN: blk[1][1] = true; break;
    case QF_LINEAR_ELASTIC: blk[1][1] = true; break;
    case QF_J2_LINEAR: blk[1][1] = true; break;
  }
}
inline bool qf_has_source(int kind) { return kind == QF_HEAT || kind == QF_MASS_DIFFUSION; }

}  // namespace oracle
