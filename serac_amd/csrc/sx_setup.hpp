// sx_setup.hpp -- host-side set-up of the matrix-free path: 1D rules and
// bases, meshes, H1 node numbering, element colouring.  Set-up runs once per
// mesh/space on the host, as in the reference; only the per-application path
// runs on the GPU.
#pragma once

#include <cmath>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace sxh {

// Gauss-Legendre nodes/weights on [0,1] for q = 1..8: the reference's literal
// table (fn/polynomials.hpp:42-52, 271-283), kept identical so that the
// quadrature is bit-for-bit the reference's.
inline void gauss_legendre(int q, double* x, double* w)
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
  if (q < 1 || q > 8) throw std::invalid_argument("Gauss-Legendre rule needs 1..8 points");
  for (int i = 0; i < q; i++) x[i] = X[q - 1][i], w[i] = W[q - 1][i];
}

// Gauss-Lobatto points on [0,1]: 0, interior roots of P'_{n-1}(2x-1), 1.
inline std::vector<double> lobatto_points(int n)
{
  switch (n) {
    case 2: return {0.0, 1.0};
    case 3: return {0.0, 0.5, 1.0};
    case 4: return {0.0, (5.0 - std::sqrt(5.0)) / 10.0, (5.0 + std::sqrt(5.0)) / 10.0, 1.0};
    case 5: return {0.0, 0.5 - std::sqrt(21.0) / 14.0, 0.5, 0.5 + std::sqrt(21.0) / 14.0, 1.0};
  }
  throw std::invalid_argument("Gauss-Lobatto points defined for 2..5 nodes");
}

// Values and derivatives of the Lagrange polynomials through `pts` at x.
inline void lagrange(const std::vector<double>& pts, double x, double* phi, double* dphi)
{
  const int n = int(pts.size());
  for (int i = 0; i < n; i++) {
    double p = 1.0, dp = 0.0;
    for (int j = 0; j < n; j++) {
      if (j == i) continue;
      const double a = (x - pts[j]) / (pts[i] - pts[j]);
      dp             = dp * a + p / (pts[i] - pts[j]);
      p *= a;
    }
    phi[i]  = p;
    dphi[i] = dp;
  }
}

// Device table for N = Q = p+1: B, G, Bw, Gw as [q][n], then GL points [q].
inline std::vector<double> element_tables(int N)
{
  const int           Q = N;
  std::vector<double> t(size_t(4 * N * N + N));
  double              xq[8], wq[8], phi[8], dphi[8];
  gauss_legendre(Q, xq, wq);
  auto pts = lobatto_points(N);
  for (int q = 0; q < Q; q++) {
    lagrange(pts, xq[q], phi, dphi);
    for (int i = 0; i < N; i++) {
      t[0 * N * N + q * N + i] = phi[i];
      t[1 * N * N + q * N + i] = dphi[i];
      t[2 * N * N + q * N + i] = phi[i] * wq[q];
      t[3 * N * N + q * N + i] = dphi[i] * wq[q];
    }
    t[4 * N * N + q] = xq[q];
  }
  return t;
}

// SplitMix64-finalizer counter RNG:
//   h = f(seed); h = f(h ^ stream); h = f(h ^ idx); u = (h >> 11) 2^-53
inline uint64_t splitmix(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double counter_uniform(uint64_t seed, uint64_t stream, uint64_t idx)
{
  uint64_t h = splitmix(splitmix(splitmix(seed) ^ stream) ^ idx);
  return double(h >> 11) * 0x1.0p-53;
}

// mfem hexahedron vertex slot of the lexicographic corner a + 2b + 4c
constexpr int kCornerToVertex[8] = {0, 1, 3, 2, 4, 5, 7, 6};

struct HexMesh {
  int64_t             nv = 0, ne = 0;
  std::vector<double> X;         // [nv][3]
  std::vector<int>    hex;       // [ne][8], mfem order
  std::vector<int>    attr;      // [ne] element attributes (1 when not given)
  std::vector<int>    bdr;       // [nb][4] boundary quadrilaterals (vertex ids)
  std::vector<int>    bdr_attr;  // [nb] boundary attributes
  bool                structured = false;
  int                 n[3]       = {0, 0, 0};
  double              len[3]     = {1, 1, 1};

  int64_t vertex_id(int64_t i, int64_t j, int64_t k) const { return i + (n[0] + 1) * (j + (n[1] + 1) * k); }
  int     corner(int64_t e, int c) const { return hex[size_t(e) * 8 + kCornerToVertex[c]]; }
};

inline HexMesh make_cuboid(int nx, int ny, int nz, double lx, double ly, double lz)
{
  if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("cuboid mesh needs at least one element per direction");
  HexMesh m;
  m.structured = true;
  m.n[0] = nx, m.n[1] = ny, m.n[2] = nz;
  m.len[0] = lx, m.len[1] = ly, m.len[2] = lz;
  m.nv = int64_t(nx + 1) * (ny + 1) * (nz + 1);
  m.ne = int64_t(nx) * ny * nz;
  m.X.resize(size_t(m.nv) * 3);
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        const int64_t v = m.vertex_id(i, j, k);
        m.X[3 * v + 0]  = lx * double(i) / double(nx);
        m.X[3 * v + 1]  = ly * double(j) / double(ny);
        m.X[3 * v + 2]  = lz * double(k) / double(nz);
      }
  m.hex.resize(size_t(m.ne) * 8);
  int64_t e = 0;
  for (int ez = 0; ez < nz; ez++)
    for (int ey = 0; ey < ny; ey++)
      for (int ex = 0; ex < nx; ex++, e++)
        for (int c = 0; c < 8; c++)
          m.hex[size_t(e) * 8 + kCornerToVertex[c]] =
              int(m.vertex_id(ex + (c & 1), ey + ((c >> 1) & 1), ez + ((c >> 2) & 1)));
  m.attr.assign(size_t(m.ne), 1);
  // boundary quadrilaterals with mfem's Cartesian attributes: 1 z = 0, 2 y = 0,
  // 3 x = max, 4 y = max, 5 x = 0, 6 z = max (vertices counter-clockwise seen
  // from outside)
  auto quad = [&](int attr, int64_t a, int64_t b, int64_t c, int64_t d) {
    m.bdr.insert(m.bdr.end(), {int(a), int(b), int(c), int(d)});
    m.bdr_attr.push_back(attr);
  };
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      quad(1, m.vertex_id(i, j, 0), m.vertex_id(i, j + 1, 0), m.vertex_id(i + 1, j + 1, 0), m.vertex_id(i + 1, j, 0));
      quad(6, m.vertex_id(i, j, nz), m.vertex_id(i + 1, j, nz), m.vertex_id(i + 1, j + 1, nz),
           m.vertex_id(i, j + 1, nz));
    }
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nx; i++) {
      quad(2, m.vertex_id(i, 0, k), m.vertex_id(i + 1, 0, k), m.vertex_id(i + 1, 0, k + 1), m.vertex_id(i, 0, k + 1));
      quad(4, m.vertex_id(i, ny, k), m.vertex_id(i, ny, k + 1), m.vertex_id(i + 1, ny, k + 1),
           m.vertex_id(i + 1, ny, k));
    }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      quad(5, m.vertex_id(0, j, k), m.vertex_id(0, j, k + 1), m.vertex_id(0, j + 1, k + 1), m.vertex_id(0, j + 1, k));
      quad(3, m.vertex_id(nx, j, k), m.vertex_id(nx, j + 1, k), m.vertex_id(nx, j + 1, k + 1),
           m.vertex_id(nx, j, k + 1));
    }
  return m;
}

// Unsupported content of an otherwise well-formed mesh file.
struct mesh_unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// MFEM mesh v1.0 text, the format of the reference's data/meshes/*.mesh and of
// mfem::Mesh(std::istream&) behind buildMeshFromFile
// (src/serac/mesh/mesh_utils.cpp:19-43): sections `dimension`, `elements`
// (`attr geom v...`), `boundary` and `vertices` (count, space dimension, one
// coordinate row per vertex); `#` starts a comment.  Only 3D meshes of
// hexahedra (geometry 5) with quadrilateral boundary elements (geometry 3)
// and straight vertex geometry are accepted; curved meshes (a `nodes`
// section) and other element types raise mesh_unsupported.
inline HexMesh parse_mfem_mesh(const std::string& text)
{
  std::vector<std::string> lines;
  {
    std::istringstream in(text);
    std::string        line;
    while (std::getline(in, line)) {
      const size_t h = line.find('#');
      if (h != std::string::npos) line.resize(h);
      const size_t b = line.find_first_not_of(" \t\r");
      if (b == std::string::npos) continue;
      const size_t e = line.find_last_not_of(" \t\r");
      lines.push_back(line.substr(b, e - b + 1));
    }
  }
  auto bad = [](const std::string& what) { return std::invalid_argument("MFEM mesh: " + what); };
  if (lines.empty() || lines[0].rfind("MFEM mesh v1.0", 0) != 0) throw bad("expected the header 'MFEM mesh v1.0'");
  auto ints = [&](size_t i, size_t want) {
    if (i >= lines.size()) throw bad("truncated file");
    std::istringstream  in(lines[i]);
    std::vector<double> v;
    double              x;
    while (in >> x) v.push_back(x);
    if (v.size() != want) throw bad("line '" + lines[i] + "' has " + std::to_string(v.size()) + " numbers, expected " +
                                    std::to_string(want));
    return v;
  };
  auto count = [&](size_t i) {
    if (i >= lines.size()) throw bad("truncated file");
    const double v = ints(i, 1)[0];
    if (v < 0 || v != std::floor(v)) throw bad("bad count '" + lines[i] + "'");
    return int64_t(v);
  };
  HexMesh m;
  bool    have_dim = false, have_elems = false, have_verts = false;
  size_t  i        = 1;
  while (i < lines.size()) {
    const std::string& key = lines[i];
    if (key == "dimension") {
      if (count(i + 1) != 3) throw mesh_unsupported("MFEM mesh: only 3D meshes are supported");
      have_dim = true;
      i += 2;
    } else if (key == "elements" || key == "boundary") {
      const bool    elem = key == "elements";
      const int64_t n    = count(i + 1);
      for (int64_t k = 0; k < n; k++) {
        const size_t li = i + 2 + size_t(k);
        if (li >= lines.size()) throw bad("truncated " + key + " section");
        std::istringstream in(lines[li]);
        int                attr = 0, geom = -1;
        in >> attr >> geom;
        if (elem && geom != 5) throw mesh_unsupported("MFEM mesh: only hexahedra (geometry 5) are supported");
        if (!elem && geom != 3)
          throw mesh_unsupported("MFEM mesh: boundary elements of a hex mesh must be quadrilaterals (geometry 3)");
        const int nvtx = elem ? 8 : 4;
        const auto v   = ints(li, size_t(2 + nvtx));
        for (int c = 0; c < nvtx; c++) (elem ? m.hex : m.bdr).push_back(int(v[size_t(2 + c)]));
        (elem ? m.attr : m.bdr_attr).push_back(attr);
      }
      if (elem) {
        m.ne       = n;
        have_elems = true;
      }
      i += 2 + size_t(n);
    } else if (key == "vertices") {
      m.nv = count(i + 1);
      if (i + 2 < lines.size() && lines[i + 2] == "nodes")
        throw mesh_unsupported("MFEM mesh: curved meshes (a nodes section) are not supported");
      if (count(i + 2) != 3) throw mesh_unsupported("MFEM mesh: vertices must have 3 coordinates");
      for (int64_t k = 0; k < m.nv; k++) {
        const auto x = ints(i + 3 + size_t(k), 3);
        m.X.insert(m.X.end(), x.begin(), x.end());
      }
      have_verts = true;
      i += 3 + size_t(m.nv);
    } else if (key == "nodes") {
      throw mesh_unsupported("MFEM mesh: curved meshes (a nodes section) are not supported");
    } else if (key == "mfem_serial_mesh_end" || key == "mfem_mesh_end") {
      break;
    } else {
      throw bad("unknown section '" + key + "'");
    }
  }
  if (!have_dim || !have_elems || !have_verts) throw bad("missing dimension, elements or vertices section");
  if (m.ne == 0 || m.nv == 0) throw bad("empty mesh");
  for (int v : m.hex)
    if (v < 0 || v >= m.nv) throw bad("element vertex index out of range");
  for (int v : m.bdr)
    if (v < 0 || v >= m.nv) throw bad("boundary vertex index out of range");
  return m;
}

inline void perturb_interior(HexMesh& m, uint64_t seed, double amp)
{
  if (!m.structured) throw std::invalid_argument("perturbation is defined for cuboid meshes");
  for (int k = 1; k < m.n[2]; k++)
    for (int j = 1; j < m.n[1]; j++)
      for (int i = 1; i < m.n[0]; i++) {
        const int64_t v = m.vertex_id(i, j, k);
        for (int d = 0; d < 3; d++) {
          const double h = m.len[d] / m.n[d];
          m.X[3 * v + d] += h * (-amp + 2.0 * amp * counter_uniform(seed, 1000 + d, uint64_t(v)));
        }
      }
}

// Node numbering.  Structured meshes: the global (p n_x + 1) x ... node grid,
// x fastest.  Unstructured meshes: nodes keyed by the entity they sit on
// (vertex, edge, face, interior) in an orientation-free parametrisation, and
// numbered in order of first appearance (elements in order, local nodes
// lexicographic).
struct NodeKeyHash {
  size_t operator()(const std::array<int64_t, 6>& k) const
  {
    uint64_t h = 0x12345678ull;
    for (auto v : k) h = splitmix(h ^ uint64_t(v));
    return size_t(h);
  }
};

inline std::array<int64_t, 6> entity_key(const HexMesh& m, int64_t e, int p, const int ijk[3])
{
  int free_dirs[3], nfree = 0, side[3];
  for (int d = 0; d < 3; d++) {
    side[d] = ijk[d] == 0 ? 0 : 1;
    if (ijk[d] != 0 && ijk[d] != p) free_dirs[nfree++] = d;
  }
  auto vtx = [&](const int c[3]) { return int64_t(m.corner(e, c[0] + 2 * c[1] + 4 * c[2])); };
  if (nfree == 0) return {0, vtx(side), 0, 0, 0, 0};
  if (nfree == 3) return {3, e, ijk[0], ijk[1], ijk[2], 0};
  if (nfree == 1) {
    const int d  = free_dirs[0];
    int       lo[3] = {side[0], side[1], side[2]}, hi[3] = {side[0], side[1], side[2]};
    lo[d] = 0, hi[d] = 1;
    const int64_t a = vtx(lo), b = vtx(hi);
    return a < b ? std::array<int64_t, 6>{1, a, b, ijk[d], 0, 0} : std::array<int64_t, 6>{1, b, a, p - ijk[d], 0, 0};
  }
  // face: corners g[u][v] along the two free directions
  const int d1 = free_dirs[0], d2 = free_dirs[1];
  int64_t   g[2][2];
  for (int u = 0; u < 2; u++)
    for (int v = 0; v < 2; v++) {
      int c[3] = {side[0], side[1], side[2]};
      c[d1] = u, c[d2] = v;
      g[u][v] = vtx(c);
    }
  int ou = 0, ov = 0;
  for (int u = 0; u < 2; u++)
    for (int v = 0; v < 2; v++)
      if (g[u][v] < g[ou][ov]) ou = u, ov = v;
  const int64_t nu = g[1 - ou][ov], nv = g[ou][1 - ov];
  const int64_t su = ou == 0 ? ijk[d1] : p - ijk[d1];
  const int64_t sv = ov == 0 ? ijk[d2] : p - ijk[d2];
  return nu < nv ? std::array<int64_t, 6>{2, g[ou][ov], nu, nv, su, sv}
                 : std::array<int64_t, 6>{2, g[ou][ov], nv, nu, sv, su};
}

inline int64_t number_nodes(const HexMesh& m, int p, std::vector<int>& emap)
{
  const int n = p + 1, npe = n * n * n;
  emap.resize(size_t(m.ne) * npe);
  if (m.structured) {
    const int64_t gx = int64_t(p) * m.n[0] + 1, gy = int64_t(p) * m.n[1] + 1, gz = int64_t(p) * m.n[2] + 1;
    if (gx * gy * gz > INT32_MAX) throw std::invalid_argument("more than 2^31 nodes");
    int64_t e = 0;
    for (int ez = 0; ez < m.n[2]; ez++)
      for (int ey = 0; ey < m.n[1]; ey++)
        for (int ex = 0; ex < m.n[0]; ex++, e++) {
          int* out = &emap[size_t(e) * npe];
          for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++) {
              const int64_t row = int64_t(p) * ex + gx * (int64_t(p) * ey + j + gy * (int64_t(p) * ez + k));
              for (int i = 0; i < n; i++) *out++ = int(row + i);
            }
        }
    return gx * gy * gz;
  }
  std::unordered_map<std::array<int64_t, 6>, int, NodeKeyHash> ids;
  int                                                           next = 0;
  for (int64_t e = 0; e < m.ne; e++) {
    int* out = &emap[size_t(e) * npe];
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          const int ijk[3] = {i, j, k};
          auto      ins    = ids.emplace(entity_key(m, e, p, ijk), next);
          if (ins.second) next++;
          *out++ = ins.first->second;
        }
  }
  return next;
}

// Colouring such that no two elements of a colour share a node.
inline int color_elements(const HexMesh& m, int p, const std::vector<int>& emap, int64_t nnodes,
                          std::vector<int>& color)
{
  color.assign(size_t(m.ne), 0);
  if (m.structured) {
    int64_t e = 0;
    for (int ez = 0; ez < m.n[2]; ez++)
      for (int ey = 0; ey < m.n[1]; ey++)
        for (int ex = 0; ex < m.n[0]; ex++, e++) color[size_t(e)] = (ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2);
    return 8;
  }
  const int            npe = (p + 1) * (p + 1) * (p + 1);
  std::vector<int64_t> off(size_t(nnodes) + 1, 0);
  for (auto v : emap) off[size_t(v) + 1]++;
  for (size_t i = 0; i < size_t(nnodes); i++) off[i + 1] += off[i];
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  std::vector<int>     elems(emap.size());
  for (int64_t e = 0; e < m.ne; e++)
    for (int j = 0; j < npe; j++) {
      const int v                 = emap[size_t(e) * npe + j];
      elems[size_t(pos[size_t(v)]++)] = int(e);
    }
  int               ncolors = 0;
  std::vector<char> used;
  for (int64_t e = 0; e < m.ne; e++) {
    used.assign(size_t(ncolors) + 1, 0);
    for (int j = 0; j < npe; j++) {
      const int v = emap[size_t(e) * npe + j];
      for (int64_t k = off[size_t(v)]; k < off[size_t(v) + 1]; k++) {
        const int o = elems[size_t(k)];
        if (o < e) used[size_t(color[size_t(o)])] = 1;
      }
    }
    int c = 0;
    while (used[size_t(c)]) c++;
    color[size_t(e)] = c;
    if (c + 1 > ncolors) ncolors = c + 1;
  }
  return ncolors;
}

}  // namespace sxh
