// sx_capi.cpp -- the C-ABI of include/serac_amd.h: meshes, spaces, forms,
// gradient operators and the Newton / Jacobi-CG solvers.  Host code drives
// the HIP kernels of sx_kernels_p*.hip and sx_vector.hip on one stream per
// form; nothing here computes on the CPU per application.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/serac_amd.h"
#include "sx_boundary.hpp"
#include "sx_launch.hpp"
#include "sx_setup.hpp"
#include "sx_qoi.hpp"
#include "sx_mg.hpp"
#include "sx_unet.hpp"
#include "sx_vector.hpp"

namespace {

struct sx_failure : std::runtime_error {
  int code;
  sx_failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

thread_local std::string g_last_error;

}  // namespace

// the message sx_last_error() returns, set by other translation units of the library
namespace sxh {
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace sxh

namespace {

#define SX_HIP(call)                                                                                         \
  do {                                                                                                       \
    hipError_t sx_e_ = (call);                                                                               \
    if (sx_e_ != hipSuccess)                                                                                 \
      throw sx_failure(SX_ERR_HIP, std::string(#call) + " failed: " + hipGetErrorString(sx_e_));            \
  } while (0)

#define SX_REQUIRE(cond, code, msg)          \
  do {                                       \
    if (!(cond)) throw sx_failure(code, msg); \
  } while (0)

template <typename F>
int api(F&& f)
{
  try {
    f();
    return SX_OK;
  } catch (const sx_failure& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return SX_ERR_ARG;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SX_ERR_ARG;
  }
}

// owning device buffer
struct DeviceBuffer {
  void*  ptr   = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void release()
  {
    if (ptr) (void)hipFree(ptr);
    ptr   = nullptr;
    bytes = 0;
  }
  void allocate(size_t n)
  {
    if (n == bytes && ptr) return;
    release();
    if (n == 0) return;
    SX_HIP(hipMalloc(&ptr, n));
    bytes = n;
  }
  template <typename T>
  void upload(const std::vector<T>& v)
  {
    allocate(v.size() * sizeof(T));
    if (!v.empty()) SX_HIP(hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  template <typename T>
  T* as() const
  {
    return static_cast<T*>(ptr);
  }
};

struct Runtime {
  std::once_flag                 once;
  int                            device = 0;
  std::unique_ptr<DeviceBuffer> tables[6];  // index N = nodes per direction (2..5)
  std::mutex                     mu;
};
Runtime& rt()
{
  static Runtime r;
  return r;
}

void ensure_device()
{
  static bool ok = false;
  if (ok) return;
  int n = 0;
  SX_HIP(hipGetDeviceCount(&n));
  SX_REQUIRE(n > 0, SX_ERR_HIP, "no HIP device visible");
  SX_HIP(hipSetDevice(rt().device));
  ok = true;
}

}  // namespace

// Scratch of the allocation-free entry points (sx_cg_solve_noalloc): every
// GPU the library runs on holds its own copy until the process ends.  Host
// code never uses the array's name as an address: the device address of the
// current GPU's copy comes from hipGetSymbolAddress (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[SX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex    g_scratch_mu;
void*         g_scratch_dev[kMaxScratchDevices] = {};

// The device address of the current GPU's scratch array, looked up once per
// device (sx_init does it for the device it selects).
void* noalloc_scratch()
{
  int dev = 0;
  SX_HIP(hipGetDevice(&dev));
  SX_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, SX_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    SX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SX_REQUIRE(p != nullptr, SX_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch array");
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

const double* element_tables(int N)
{
  ensure_device();
  std::lock_guard<std::mutex> lk(rt().mu);
  auto&                       t = rt().tables[N];
  if (!t) {
    t = std::make_unique<DeviceBuffer>();
    t->upload(sxh::element_tables(N));
  }
  return t->as<const double>();
}

}  // namespace

// ---------------------------------------------------------------------------
// handle types
// ---------------------------------------------------------------------------
struct sx_mesh {
  sxh::HexMesh m;
};

struct sx_space {
  sx_mesh*         mesh = nullptr;  // borrowed, like the reference's FE space pointers
  int              order = 1, nc = 1, ordering = SX_ORDERING_BYVDIM, npe = 8;
  int64_t          nnodes = 0;
  std::vector<int> emap;    // [ne][npe]
  std::vector<int> color;   // [ne]
  int              ncolors = 0;
  DeviceBuffer     d_emap;  // element order (E-vector gather)
  DeviceBuffer     d_offs, d_ents;  // node -> (e*npe + j), increasing e
  // DOF-level sparsity of forms on this space (built on first assembly)
  std::vector<int64_t> csr_row;
  std::vector<int32_t> csr_col;
  // numbering: native = the library's own (lexicographic node grid on cuboid
  // meshes); otherwise node_of_own[own node] = the caller's node id
  bool                 native = true;
  std::vector<int32_t> node_of_own;
  int64_t          ndof() const { return nnodes * nc; }
  int64_t          node_id(int64_t own) const { return native ? own : int64_t(node_of_own[size_t(own)]); }
};

// Assembled linearization (Gradient::assemble, fn/functional.hpp:569-667): CSR
// on the device, int64 row offsets, sorted int32 columns per row.
struct sx_csr {
  int64_t      nrows = 0, ncols = 0, nnz = 0;
  DeviceBuffer row, col, val;
  hipStream_t  stream = nullptr;
};

// QuadratureData<State> (fn/quadrature_data.hpp:111-167): `size` doubles per
// quadrature point of `nelem` elements; on the device as [elem][k][q] so that
// the lanes of an element (one per point) read consecutive doubles.
struct sx_qdata {
  int64_t      nelem = 0;
  int          nq = 0, size = 0;
  DeviceBuffer d;
};

struct Integral {
  int                  qf = 0;
  double               params[SX_MAX_PARAMS] = {0};
  std::vector<int>     depends_on;
  int64_t              nslots = 0;
  std::vector<int64_t> color_off;  // [ncolors + 1] slot ranges
  DeviceBuffer         d_map;      // [slot][npe]
  DeviceBuffer         d_coords;   // [slot][24]
  DeviceBuffer         d_tangent;  // per-slot TangentLayout (sx_element.hpp)
  DeviceBuffer         d_tangent_lo;  // BF16 copy for a multigrid smoother (set up with the hierarchy)
  int                  layout  = sxk::LAYOUT_GRAD_SYM;
  int                  tprec   = sxk::PREC_F64;  // precision the tangent was stored in
  bool                 has_tangent = false;
  sx_qdata*            qdata = nullptr;  // borrowed; must outlive the form
  DeviceBuffer         d_qrow;           // [slot] -> element position in the domain
  std::vector<int>     slot_pos;         // [slot] -> element position in the domain (host)
  DeviceBuffer         d_emap_e;         // E-vector entry points: [slot][npe] -> pos * npe + j
  // a two-argument integral (DependsOn<a, b>) is held as the integrand in its
  // first argument plus, as a companion, the mass integral of its second
  bool                 companion     = false;  // the mass part of a two-argument integral
  bool                 has_companion = false;  // the integrand part of one
};

// AddBoundaryIntegral (fn/functional.hpp:328-354): faces of the selected
// boundary, each as its element's node trace in outward (xi1, xi2) order.
struct BoundaryIntegral {
  int                  qf = 0, nc = 1;
  double               params[SX_MAX_PARAMS] = {0};
  std::vector<int>     depends_on;
  int64_t              nslots = 0;
  std::vector<int64_t> color_off;  // [ncolors + 1] slot ranges
  DeviceBuffer         d_fmap;     // [slot][Q*Q]
  DeviceBuffer         d_corners;  // [slot][4][3]
  DeviceBuffer         d_tangent;  // [slot][nc * 3nc][Q*Q], FP64
  bool                 has_tangent = false;
};

// The learned coarse-grid preconditioner of a structured vector space:
//   M(r) = B(r) + gain s P N(P^T r / s),  s = rms(P^T r)
// with B = D^-1 (Jacobi) or, with a multigrid attached, the V-cycle; P the
// vertex (Q1) -> order-p interpolation and N the U-Net; plus the bank of
// (coarse residual, coarse correction) pairs it trains on.
struct sx_learned {
  const sx_space*                            space = nullptr;
  sx_unet*                                   net   = nullptr;
  sxu::FEGrid                                G{};
  int64_t                                    nv = 0;  // vertex nodes
  DeviceBuffer                               rc, xin, y, ss;
  std::vector<std::unique_ptr<DeviceBuffer>> pr, pt;  // pairs: P^T r_k and the vertex-mean correction
  std::vector<char>                          ok;      // the pair's target is complete
  int                                        max_pairs = 256, stride = 50, window = 50, stall = 20;
  int                                        next = 0;  // ring position once the bank is full
  int64_t                                    fallbacks = 0;
  double                                     gain      = 1.0;  // target scale: N learns t / (s gain)
  // with a multigrid attached: M(r) = MG(r) + gain s P N(P^T r / s), the
  // network trained on the V-cycle's remaining error (borrowed, same space)
  sx_multigrid*                              mg = nullptr;
  ~sx_learned()
  {
    if (net) sx_unet_destroy(net);
  }
};

struct sx_grad {
  sx_form* form = nullptr;
  int      wrt  = 0;
  bool     valid = false;
  bool     shape = false;  // the derivative in the shape displacement (SX_WRT_SHAPE)
};

// Geometric multigrid preconditioner of a structured 3-component space: a
// V-cycle over the order-p nodal grid (level 0: the form's matrix-free J_bc)
// and the mesh's vertex grids (level 1, then halving the intervals), with
// Galerkin coarse operators P^T A P found by probing, Chebyshev-Jacobi
// smoothing on every level but the coarsest, and a dense solve there.
struct MgLevel {
  sxm::Grid3          g{};
  int64_t             n = 0;  // vector entries
  DeviceBuffer        S, dinv, mask, x, b, r, d, t;
  double              lmax = 0.0, theta = 1.0;
  std::vector<double> c1, c2;  // Chebyshev steps 2..degree
};

struct sx_multigrid {
  const sx_space*      space = nullptr;
  sxu::FEGrid          G{};
  std::vector<MgLevel> lv;
  DeviceBuffer         coarse_inv, part, scal, diag;
  DeviceBuffer         rt1, rt2;        // scratch of the separable level-0 restriction
  DeviceBuffer         ess;             // essential DOFs of the last sx_mg_setup
  int64_t              ness   = 0;
  sx_grad*             set_g  = nullptr;  // the gradient of the last sx_mg_setup
  int                  ncoarse = 0;
  int                  fine_degree = 2, coarse_degree = 2, power_iters = 12;
  double               range = 12.0, safety = 1.15;
  int64_t              setups = 0;
  int                  galerkin = 0;  // level 1: 0 element-wise from the tangent where possible, 1 probing
  int                  smoother_bf16 = 0;  // level-0 applications with BF16 tangent copies
  bool                 lo_active     = false;
  int                  smoother_f32  = 0;  // level-0 smoothing on FP32 vectors
  DeviceBuffer         fb, fx, fr, fd, ft, fdinv;  // its FP32 level-0 vectors
  int64_t              element_setups = 0;
  DeviceBuffer         flag;
  int                  reuse  = 1;  // linearizations one hierarchy serves inside a Newton call
  int                  uses   = INT32_MAX;  // linear solves since the last setup (INT32_MAX: set up at the next)
};

struct sx_form {
  const sx_space*                        test = nullptr;
  std::vector<const sx_space*>           trials;
  std::vector<std::unique_ptr<Integral>> integrals;
  std::vector<std::unique_ptr<BoundaryIntegral>> bintegrals;
  std::vector<std::unique_ptr<sx_grad>>  grads;
  hipStream_t                            stream = nullptr;
  int                                    scatter   = SX_SCATTER_COLORED;
  int                                    xcd_remap = 0;  // hardware block order
  int                                    scatter_order = 2;  // address order, shared faces merged in LDS first
  int                                    tprec     = sxk::PREC_F64;
  int                                    update_qdata = 0;  // updateQdata(bool), fn/functional.hpp:523
  int                                    jvp_arith    = sxk::PREC_F64;  // element arithmetic of FP64 J.v
  int                                    linsolver    = SX_LINSOLVER_CG;  // linear solver of the Newton loops
  int                                    gmres_restart = 50;
  int                                    line_search   = 0;  // max_line_search_iterations (0: plain Newton)
  int                                    min_newton    = 0;  // min_iterations of the Newton loops
  int                                    print_level   = 0;  // > 0: the Newton loops report each iteration on stderr
  sx_learned*                            learned      = nullptr;        // borrowed
  sx_multigrid*                          mg           = nullptr;        // borrowed: multigrid-PCG in the Newton loops
  const double*                          shape        = nullptr;        // borrowed: shape displacement (device)
  int                                    learn_mode   = 0;              // SX_PRECOND_*
  // differentiate_wrt on the shape (SX_WRT_SHAPE): the gradient, and the
  // trial inputs and shape displacement of its linearization point
  std::unique_ptr<sx_grad>                   shape_grad;
  std::vector<std::unique_ptr<DeviceBuffer>> shape_u;
  DeviceBuffer                               shape_p;
  // solver scratch
  DeviceBuffer part, scal;
  ~sx_form()
  {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// set while the library adds an integrand of its own (the QoI gradients),
// which the public entry points do not accept
thread_local bool g_internal_qf = false;

int qf_components(int qf)
{
  if (qf == sxk::QF_ID_QOI_VALUE) return g_internal_qf ? -1 : 0;
  switch (qf) {
    case SX_QF_HEAT: return 1;
    case SX_QF_NEOHOOKEAN:
    case SX_QF_LINEAR_ELASTIC:
    case SX_QF_J2_LINEAR: return 3;
    case SX_QF_MASS_DIFFUSION: return -1;  // scalar or vector
  }
  return 0;
}
int qf_layout(int qf)
{
  if (qf == SX_QF_MASS_DIFFUSION) return sxk::LAYOUT_FULL_SYM;
  if (qf == SX_QF_J2_LINEAR) return sxk::LAYOUT_GRAD_FULL;
  return sxk::LAYOUT_GRAD_SYM;
}
// doubles of quadrature-point state the qfunction carries
int qf_state_size(int qf) { return qf == SX_QF_J2_LINEAR ? 10 : 0; }

sxk::ElemLaunch base_launch(const sx_form* f, const Integral& I, int mode)
{
  const sx_space* s = f->test;
  sxk::ElemLaunch P{};
  P.order    = s->order;
  P.nc       = s->nc;
  P.qf       = I.qf;
  P.mode     = mode;
  P.layout   = I.layout;
  P.ordering = s->ordering;
  P.scatter   = f->scatter;
  P.xcd_remap = f->xcd_remap;
  P.ordered   = f->scatter_order;
  P.map      = I.d_map.as<const int>();
  P.coords   = I.d_coords.as<const double>();
  P.nnodes   = s->nnodes;
  P.D        = I.d_tangent.ptr;
  P.tprec    = I.tprec;
  P.vprec    = sxk::PREC_F64;
  P.tables   = element_tables(s->order + 1);
  P.shape    = f->shape;
  for (int k = 0; k < SX_MAX_PARAMS; k++) P.params[k] = I.params[k];
  if (I.qdata) {
    P.qdata        = I.qdata->d.as<double>();
    P.qrow         = I.d_qrow.as<const int>();
    P.update_state = f->update_qdata;
  }
  P.stream = f->stream;
  return P;
}

// one launch per colour; returns the number of launches.  lo: J.v with the
// BF16 copy of the tangent (d_tangent_lo)
int run_integral(const sx_form* f, const Integral& I, int mode, const void* u, void* r, hipEvent_t* ev = nullptr,
                 int vprec = sxk::PREC_F64, bool lo = false, const double* stop = nullptr)
{
  sxk::ElemLaunch P = base_launch(f, I, mode);
  P.stop            = stop;
  P.u               = u;
  P.r               = r;
  P.vprec           = vprec;
  if (lo) {
    P.D     = I.d_tangent_lo.ptr;
    P.tprec = sxk::PREC_BF16;
  }
  int launches      = 0;
  for (size_t c = 0; c + 1 < I.color_off.size(); c++) {
    P.slot_begin = I.color_off[c];
    P.slot_end   = I.color_off[c + 1];
    if (P.slot_end <= P.slot_begin) continue;
    if (ev) SX_HIP(hipEventRecord(ev[2 * launches], f->stream));
    hipError_t err = hipSuccess;
    SX_REQUIRE(sxk::launch_element(P, &err), SX_ERR_UNSUPPORTED,
               "element kernel not compiled for order " + std::to_string(P.order) + ", qfunction " +
                   std::to_string(P.qf) + ", " + std::to_string(P.nc) + " components");
    SX_HIP(err);
    if (ev) SX_HIP(hipEventRecord(ev[2 * launches + 1], f->stream));
    launches++;
  }
  return launches;
}

// Sparsity of the space's forms: DOF (node a, component c) couples with
// (node b, c') for every node b sharing an element with a; rows and columns
// follow the space's ordering, columns sorted within a row.
void build_csr_pattern(sx_space* s)
{
  if (!s->csr_row.empty()) return;
  const int64_t nn = s->nnodes, npe = s->npe, ne = int64_t(s->emap.size()) / npe;
  const int     nc = s->nc;
  std::vector<int64_t> eptr(size_t(nn) + 1, 0);
  for (int v : s->emap) eptr[size_t(v) + 1]++;
  for (int64_t a = 0; a < nn; a++) eptr[size_t(a) + 1] += eptr[size_t(a)];
  std::vector<int32_t> elist(static_cast<size_t>(eptr[size_t(nn)]));
  {
    std::vector<int64_t> fill(eptr.begin(), eptr.end() - 1);
    for (int64_t e = 0; e < ne; e++)
      for (int64_t j = 0; j < npe; j++) elist[size_t(fill[size_t(s->emap[size_t(e * npe + j)])]++)] = int32_t(e);
  }
  std::vector<int64_t> nptr(size_t(nn) + 1, 0);
  std::vector<int32_t> adj, buf;
  std::vector<int64_t> mark(size_t(nn), -1);
  for (int64_t a = 0; a < nn; a++) {
    buf.clear();
    for (int64_t k = eptr[size_t(a)]; k < eptr[size_t(a) + 1]; k++) {
      const int64_t e = elist[size_t(k)];
      for (int64_t j = 0; j < npe; j++) {
        const int b = s->emap[size_t(e * npe + j)];
        if (mark[size_t(b)] != a) {
          mark[size_t(b)] = a;
          buf.push_back(b);
        }
      }
    }
    std::sort(buf.begin(), buf.end());
    adj.insert(adj.end(), buf.begin(), buf.end());
    nptr[size_t(a) + 1] = int64_t(adj.size());
  }
  auto& rp = s->csr_row;
  auto& cl = s->csr_col;
  rp.assign(size_t(nn * nc) + 1, 0);
  cl.reserve(size_t(nptr[size_t(nn)]) * size_t(nc) * size_t(nc));
  int64_t row = 0;
  auto    emit_row = [&](int64_t a) {
    if (s->ordering == SX_ORDERING_BYVDIM) {
      for (int64_t k = nptr[size_t(a)]; k < nptr[size_t(a) + 1]; k++)
        for (int c2 = 0; c2 < nc; c2++) cl.push_back(int32_t(adj[size_t(k)] * nc + c2));
    } else {
      for (int c2 = 0; c2 < nc; c2++)
        for (int64_t k = nptr[size_t(a)]; k < nptr[size_t(a) + 1]; k++) cl.push_back(int32_t(c2 * nn + adj[size_t(k)]));
    }
    rp[size_t(++row)] = int64_t(cl.size());
  };
  if (s->ordering == SX_ORDERING_BYVDIM) {
    for (int64_t a = 0; a < nn; a++)
      for (int c = 0; c < nc; c++) emit_row(a);
  } else {
    for (int c = 0; c < nc; c++)
      for (int64_t a = 0; a < nn; a++) emit_row(a);
  }
}

// ---- boundary faces ----
// Local face lf of a hex (parent axes x, y, z): the fixed axis and the
// (xi1, xi2) parametrization whose dX/dxi1 x dX/dxi2 points out of a
// positively oriented element.  Node (a, b) of the face, a along xi1, is the
// element's lexicographic node (i0, i1, i2) below (n = nodes per direction).
inline void face_node(int lf, int a, int b, int n, int& i0, int& i1, int& i2)
{
  const int m = n - 1;
  switch (lf) {
    case 0: i0 = 0, i1 = b, i2 = a; break;  // x = 0: (z, y)
    case 1: i0 = m, i1 = a, i2 = b; break;  // x = 1: (y, z)
    case 2: i0 = a, i1 = 0, i2 = b; break;  // y = 0: (x, z)
    case 3: i0 = b, i1 = m, i2 = a; break;  // y = 1: (z, x)
    case 4: i0 = b, i1 = a, i2 = 0; break;  // z = 0: (y, x)
    default: i0 = a, i1 = b, i2 = m; break; // z = 1: (x, y)
  }
}

struct BoundaryFace {
  int e, lf;
};

struct FaceKey {
  std::array<int, 4> v;  // sorted corner vertices
  int                e, lf;
};

// every (element, local face) keyed by its sorted corner vertices
std::vector<FaceKey> element_face_keys(const sxh::HexMesh& m)
{
  std::vector<FaceKey> keys;
  keys.reserve(size_t(m.ne) * 6);
  for (int64_t e = 0; e < m.ne; e++)
    for (int lf = 0; lf < 6; lf++) {
      FaceKey k;
      for (int c = 0; c < 4; c++) {
        int i0, i1, i2;
        face_node(lf, c & 1, c >> 1, 2, i0, i1, i2);
        k.v[size_t(c)] = m.corner(e, i0 + 2 * i1 + 4 * i2);
      }
      std::sort(k.v.begin(), k.v.end());
      k.e  = int(e);
      k.lf = lf;
      keys.push_back(k);
    }
  return keys;
}

// the faces of the boundary elements whose attribute a has bit a - 1 set in
// attr_mask, in boundary-element order (the selection of
// Domain::ofBoundaryElements by attribute and of the essential boundary
// attribute sets).  A boundary element must be a face of the mesh; with
// exterior_only it must belong to exactly one element, whose face it is.
std::vector<BoundaryFace> attribute_faces(const sxh::HexMesh& m, uint64_t attr_mask, bool exterior_only)
{
  std::vector<FaceKey> keys = element_face_keys(m);
  std::sort(keys.begin(), keys.end(), [](const FaceKey& a, const FaceKey& b) {
    return a.v != b.v ? a.v < b.v : (a.e != b.e ? a.e < b.e : a.lf < b.lf);
  });
  std::vector<BoundaryFace> out;
  const size_t              nb = m.bdr_attr.size();
  for (size_t q = 0; q < nb; q++) {
    const int a = m.bdr_attr[q];
    if (a < 1 || a > 64 || !((attr_mask >> (a - 1)) & 1)) continue;
    FaceKey k{};
    for (int c = 0; c < 4; c++) k.v[size_t(c)] = m.bdr[4 * q + size_t(c)];
    std::sort(k.v.begin(), k.v.end());
    auto lo = std::lower_bound(keys.begin(), keys.end(), k, [](const FaceKey& x, const FaceKey& y) { return x.v < y.v; });
    SX_REQUIRE(lo != keys.end() && lo->v == k.v, SX_ERR_ARG,
               "boundary element " + std::to_string(q) + " is not a face of the mesh");
    const bool shared = (lo + 1) != keys.end() && (lo + 1)->v == k.v;
    SX_REQUIRE(!(exterior_only && shared), SX_ERR_UNSUPPORTED,
               "boundary element " + std::to_string(q) + " is an interior face; boundary integrals need exterior faces");
    out.push_back({lo->e, lo->lf});
  }
  return out;
}

// faces that belong to exactly one element, in (element, local face) order,
// restricted by face_mask to the bounding-box planes (bit 0 x = min, 1 x = max,
// 2 y = min, ...); face_mask 0 selects the whole boundary
std::vector<BoundaryFace> boundary_faces(const sxh::HexMesh& m, int face_mask)
{
  const std::vector<FaceKey> keys = element_face_keys(m);
  std::vector<size_t>        order(keys.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return keys[a].v < keys[b].v; });
  std::vector<char> single(keys.size(), 0);
  for (size_t i = 0; i < order.size();) {
    size_t j = i;
    while (j < order.size() && keys[order[j]].v == keys[order[i]].v) j++;
    if (j - i == 1) single[order[i]] = 1;
    i = j;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t v = 0; v < m.nv; v++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], m.X[size_t(3 * v + d)]);
      hi[d] = std::max(hi[d], m.X[size_t(3 * v + d)]);
    }
  std::vector<BoundaryFace> out;
  for (size_t i = 0; i < keys.size(); i++) {
    if (!single[i]) continue;
    bool take = (face_mask == 0);
    for (int bit = 0; bit < 6 && !take; bit++) {
      if (!(face_mask >> bit & 1)) continue;
      const int    d = bit / 2;
      const double plane = (bit & 1) ? hi[d] : lo[d], tol = 1e-12 * std::max(hi[d] - lo[d], 1.0);
      bool         on = true;
      for (int c = 0; c < 4; c++) on = on && std::fabs(m.X[size_t(3 * keys[i].v[size_t(c)] + d)] - plane) <= tol;
      take = on;
    }
    if (take) out.push_back({keys[i].e, keys[i].lf});
  }
  return out;
}

// The exterior faces (faces of exactly one element) in (element, local face)
// order, each with the attribute of the boundary element on it (-1 when the
// mesh has none there).
struct ExteriorFace {
  FaceKey key;
  int     attr;
};

std::vector<ExteriorFace> exterior_faces(const sxh::HexMesh& m)
{
  std::vector<FaceKey> keys = element_face_keys(m);
  std::vector<size_t>  order(keys.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return keys[a].v < keys[b].v; });
  std::vector<char> single(keys.size(), 0);
  for (size_t i = 0; i < order.size();) {
    size_t j = i;
    while (j < order.size() && keys[order[j]].v == keys[order[i]].v) j++;
    if (j - i == 1) single[order[i]] = 1;
    i = j;
  }
  std::map<std::array<int, 4>, int> battr;
  for (size_t q = 0; q < m.bdr_attr.size(); q++) {
    std::array<int, 4> v;
    for (int c = 0; c < 4; c++) v[size_t(c)] = m.bdr[4 * q + size_t(c)];
    std::sort(v.begin(), v.end());
    battr.emplace(v, m.bdr_attr[q]);
  }
  std::vector<ExteriorFace> out;
  for (size_t i = 0; i < keys.size(); i++) {
    if (!single[i]) continue;
    const auto it = battr.find(keys[i].v);
    out.push_back({keys[i], it == battr.end() ? -1 : it->second});
  }
  return out;
}

// Domain::ofBoundaryElements(mesh, predicate): the exterior faces the
// predicate selects, given their corner coordinates and attribute
std::vector<BoundaryFace> predicate_faces(const sxh::HexMesh& m, sx_face_predicate pred, void* user)
{
  std::vector<BoundaryFace> out;
  for (const ExteriorFace& f : exterior_faces(m)) {
    double xyz[12];
    for (int k = 0; k < 4; k++) {
      int i0, i1, i2;
      face_node(f.key.lf, k & 1, k >> 1, 2, i0, i1, i2);
      const int v = m.corner(f.key.e, i0 + 2 * i1 + 4 * i2);
      for (int d = 0; d < 3; d++) xyz[3 * k + d] = m.X[size_t(3 * v + d)];
    }
    if (pred(xyz, f.attr, user)) out.push_back({f.key.e, f.key.lf});
  }
  return out;
}

// the exterior faces with the given vertex quadruples, in list order
std::vector<BoundaryFace> listed_faces(const sxh::HexMesh& m, const int32_t* fv, int64_t nf)
{
  std::map<std::array<int, 4>, BoundaryFace> ext;
  for (const ExteriorFace& f : exterior_faces(m)) ext.emplace(f.key.v, BoundaryFace{f.key.e, f.key.lf});
  std::vector<BoundaryFace> out;
  for (int64_t i = 0; i < nf; i++) {
    std::array<int, 4> v;
    for (int c = 0; c < 4; c++) v[size_t(c)] = fv[4 * i + c];
    std::sort(v.begin(), v.end());
    const auto it = ext.find(v);
    SX_REQUIRE(it != ext.end(), SX_ERR_ARG, "face " + std::to_string(i) + " is not an exterior face of the mesh");
    out.push_back(it->second);
  }
  return out;
}

bool is_boundary_qf(int qf) { return qf == SX_BQF_TRACTION || qf == SX_BQF_PRESSURE || qf == SX_BQF_FLUX; }
int  bqf_components(int qf) { return qf == SX_BQF_FLUX ? 1 : 3; }

sxk::FaceLaunch face_launch(const sx_form* f, const BoundaryIntegral& B, int mode)
{
  const sx_space* s = f->test;
  sxk::FaceLaunch P{};
  P.order    = s->order;
  P.nc       = s->nc;
  P.qf       = B.qf;
  P.mode     = mode;
  P.ordering = s->ordering;
  P.fmap     = B.d_fmap.as<const int>();
  P.corners  = B.d_corners.as<const double>();
  P.nnodes   = s->nnodes;
  P.D        = B.d_tangent.as<double>();
  P.tables   = element_tables(s->order + 1);
  P.shape    = B.nc == 3 ? f->shape : nullptr;
  for (int k = 0; k < SX_MAX_PARAMS; k++) P.params[k] = B.params[k];
  P.stream = f->stream;
  return P;
}

// one launch per face colour; returns the number of launches
int run_boundary(const sx_form* f, const BoundaryIntegral& B, int mode, const void* u, void* r, hipEvent_t* ev,
                 int vprec, const sx_csr* A = nullptr)
{
  sxk::FaceLaunch P = face_launch(f, B, mode);
  P.u               = u;
  P.r               = r;
  P.vprec           = vprec == sxk::PREC_F32 ? 1 : 0;
  if (A) {
    P.csr_row = A->row.as<const int64_t>();
    P.csr_col = A->col.as<const int32_t>();
    P.csr_val = A->val.as<double>();
  }
  int launches = 0;
  for (size_t c = 0; c + 1 < B.color_off.size(); c++) {
    P.slot_begin = B.color_off[c];
    P.slot_end   = B.color_off[c + 1];
    if (P.slot_end <= P.slot_begin) continue;
    if (ev) SX_HIP(hipEventRecord(ev[2 * launches], f->stream));
    hipError_t err = hipSuccess;
    SX_REQUIRE(sxk::launch_face(P, &err), SX_ERR_UNSUPPORTED,
               "boundary kernel not compiled for order " + std::to_string(P.order) + ", qfunction " +
                   std::to_string(P.qf));
    SX_HIP(err);
    if (ev) SX_HIP(hipEventRecord(ev[2 * launches + 1], f->stream));
    launches++;
  }
  return launches;
}

bool depends(const BoundaryIntegral& B, int trial)
{
  for (int d : B.depends_on)
    if (d == trial) return true;
  return false;
}

bool depends(const Integral& I, int trial)
{
  return std::find(I.depends_on.begin(), I.depends_on.end(), trial) != I.depends_on.end();
}

// a linearized gradient in a trial argument (the shape gradient only applies)
void require_trial_grad(const sx_grad* g, const char* what)
{
  SX_REQUIRE(!g->shape, SX_ERR_UNSUPPORTED, "the shape gradient supports sx_grad_apply only");
  SX_REQUIRE(g->valid, SX_ERR_STATE, what);
}

// df = dR/dp dx: the shape derivative of the domain integrals at the
// linearization point saved by sx_form_eval(..., SX_WRT_SHAPE, ...)
void apply_shape_gradient(sx_grad* g, const double* dx, double* df)
{
  sx_form* f = g->form;
  SX_HIP(sxk::vec_fill(df, 0.0, f->test->ndof(), f->stream));
  for (auto& I : f->integrals) {
    sxk::ElemLaunch P = base_launch(f, *I, sxk::MODE_SHAPE_JVP);
    P.u               = f->shape_u[size_t(I->depends_on[0])]->as<const double>();
    P.shape           = f->shape_p.as<const double>();
    P.dshape          = dx;
    P.r               = df;
    for (size_t c = 0; c + 1 < I->color_off.size(); c++) {
      P.slot_begin = I->color_off[c];
      P.slot_end   = I->color_off[c + 1];
      if (P.slot_end <= P.slot_begin) continue;
      hipError_t err = hipSuccess;
      SX_REQUIRE(sxk::launch_element(P, &err), SX_ERR_UNSUPPORTED,
                 "shape derivative not compiled for order " + std::to_string(P.order) + ", qfunction " +
                     std::to_string(P.qf));
      SX_HIP(err);
    }
  }
}

// df = J dx; out_zeroed: df already holds zeros (the fill is skipped); lo:
// the domain integrals use their BF16 tangent copies
// stop: a device flag that, once non-zero, makes the J.v kernels return at
// once (a Krylov loop that converged inside a queued chunk of iterations)
void apply_gradient(sx_grad* g, const void* dx, void* df, hipEvent_t* ev = nullptr, int* nlaunch = nullptr,
                    int vprec = sxk::PREC_F64, bool out_zeroed = false, bool lo = false, const double* stop = nullptr)
{
  SX_REQUIRE(g && g->form, SX_ERR_ARG, "null gradient");
  SX_REQUIRE(g->valid, SX_ERR_STATE, "gradient applied before a linearization (call sx_form_eval with wrt)");
  SX_REQUIRE(dx && df, SX_ERR_ARG, "null vector");
  if (g->shape) {
    SX_REQUIRE(vprec == sxk::PREC_F64 && !ev, SX_ERR_UNSUPPORTED, "the shape gradient applies to FP64 vectors");
    apply_shape_gradient(g, static_cast<const double*>(dx), static_cast<double*>(df));
    if (nlaunch) *nlaunch = 0;
    return;
  }
  SX_REQUIRE(vprec == sxk::PREC_F64 || vprec == sxk::PREC_F32, SX_ERR_ARG, "J.v vectors are FP64 or FP32");
  sx_form* f = g->form;
  if (out_zeroed) {
  } else if (vprec == sxk::PREC_F64)
    SX_HIP(sxk::vec_fill(static_cast<double*>(df), 0.0, f->test->ndof(), f->stream));
  else
    SX_HIP(sxk::vec_fill_f32(static_cast<float*>(df), 0.0f, f->test->ndof(), f->stream));
  // FP64 vectors with FP32 element arithmetic when the form asks for it
  const int kprec = (vprec == sxk::PREC_F64 && f->jvp_arith == sxk::PREC_F32) ? sxk::PREC_MIXED : vprec;
  int       n     = 0;
  for (auto& I : f->integrals)
    if (depends(*I, g->wrt) && I->has_tangent)
      n += run_integral(f, *I, sxk::MODE_JVP, dx, df, ev ? ev + 2 * n : nullptr, kprec,
                        lo && I->d_tangent_lo.ptr != nullptr, stop);
  for (auto& B : f->bintegrals)
    if (depends(*B, g->wrt) && B->has_tangent)
      n += run_boundary(f, *B, sxk::FMODE_JVP, dx, df, ev ? ev + 2 * n : nullptr, vprec);
  if (nlaunch) *nlaunch = n;
}

}  // namespace

// ---------------------------------------------------------------------------
// runtime helpers
// ---------------------------------------------------------------------------
extern "C" {

const char* sx_last_error(void) { return g_last_error.c_str(); }

int sx_init(int device)
{
  return api([&] {
    int n = 0;
    SX_HIP(hipGetDeviceCount(&n));
    SX_REQUIRE(device >= 0 && device < n, SX_ERR_ARG, "device index out of range");
    rt().device = device;
    SX_HIP(hipSetDevice(device));
    (void)noalloc_scratch();
  });
}

int sx_device_name(char* buf, size_t n)
{
  return api([&] {
    ensure_device();
    hipDeviceProp_t prop;
    SX_HIP(hipGetDeviceProperties(&prop, rt().device));
    std::snprintf(buf, n, "%s %s", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName);
  });
}

int sx_malloc(void** ptr, size_t bytes)
{
  return api([&] {
    ensure_device();
    SX_REQUIRE(ptr, SX_ERR_ARG, "null output pointer");
    SX_HIP(hipMalloc(ptr, bytes));
  });
}
int sx_free(void* ptr)
{
  return api([&] {
    if (ptr) SX_HIP(hipFree(ptr));
  });
}
int sx_memcpy_htod(void* dst, const void* src, size_t bytes)
{
  return api([&] { SX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int sx_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
  return api([&] { SX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int sx_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
  return api([&] { SX_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int sx_memset_zero(void* ptr, size_t bytes)
{
  return api([&] {
    SX_REQUIRE(bytes % sizeof(double) == 0, SX_ERR_ARG, "sx_memset_zero expects a whole number of doubles");
    SX_HIP(sxk::vec_fill(static_cast<double*>(ptr), 0.0, int64_t(bytes / sizeof(double)), nullptr));
    SX_HIP(hipDeviceSynchronize());
  });
}
int sx_device_synchronize(void)
{
  return api([&] { SX_HIP(hipDeviceSynchronize()); });
}

// ---------------------------------------------------------------------------
// meshes
// ---------------------------------------------------------------------------
int sx_mesh_create_cuboid(int nx, int ny, int nz, double lx, double ly, double lz, sx_mesh** out)
{
  return api([&] {
    SX_REQUIRE(out, SX_ERR_ARG, "null output");
    auto m = std::make_unique<sx_mesh>();
    m->m   = sxh::make_cuboid(nx, ny, nz, lx, ly, lz);
    *out   = m.release();
  });
}

int sx_mesh_create(int nverts, const double* coords, int nelem, const int* hex_vertices, sx_mesh** out)
{
  return api([&] {
    SX_REQUIRE(out && coords && hex_vertices && nverts > 0 && nelem > 0, SX_ERR_ARG, "invalid mesh arrays");
    auto m  = std::make_unique<sx_mesh>();
    m->m.nv = nverts;
    m->m.ne = nelem;
    m->m.X.assign(coords, coords + size_t(nverts) * 3);
    m->m.hex.assign(hex_vertices, hex_vertices + size_t(nelem) * 8);
    m->m.attr.assign(size_t(nelem), 1);
    for (int v : m->m.hex) SX_REQUIRE(v >= 0 && v < nverts, SX_ERR_ARG, "hex vertex index out of range");
    *out = m.release();
  });
}

int sx_mesh_read_mfem(const char* text, int64_t nbytes, sx_mesh** out)
{
  return api([&] {
    SX_REQUIRE(text && out && nbytes >= 0, SX_ERR_ARG, "null argument");
    auto m = std::make_unique<sx_mesh>();
    try {
      m->m = sxh::parse_mfem_mesh(std::string(text, size_t(nbytes)));
    } catch (const sxh::mesh_unsupported& e) {
      throw sx_failure(SX_ERR_UNSUPPORTED, e.what());
    }
    *out = m.release();
  });
}

int sx_mesh_read_mfem_file(const char* path, sx_mesh** out)
{
  return api([&] {
    SX_REQUIRE(path && out, SX_ERR_ARG, "null argument");
    std::FILE* fp = std::fopen(path, "rb");
    SX_REQUIRE(fp, SX_ERR_ARG, std::string("cannot open mesh file '") + path + "'");
    std::string text;
    char        buf[65536];
    size_t      got;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, got);
    std::fclose(fp);
    const int rc = sx_mesh_read_mfem(text.data(), int64_t(text.size()), out);
    if (rc != SX_OK) throw sx_failure(rc, g_last_error);
  });
}

int sx_mesh_attributes(const sx_mesh* mesh, int32_t* attrs)
{
  return api([&] {
    SX_REQUIRE(mesh && attrs, SX_ERR_ARG, "null argument");
    for (int64_t e = 0; e < mesh->m.ne; e++) attrs[e] = mesh->m.attr.empty() ? 1 : mesh->m.attr[size_t(e)];
  });
}

int sx_mesh_boundary_info(const sx_mesh* mesh, int64_t* nbdr)
{
  return api([&] {
    SX_REQUIRE(mesh && nbdr, SX_ERR_ARG, "null argument");
    *nbdr = int64_t(mesh->m.bdr_attr.size());
  });
}

int sx_mesh_boundary(const sx_mesh* mesh, int32_t* quads, int32_t* attrs)
{
  return api([&] {
    SX_REQUIRE(mesh, SX_ERR_ARG, "null mesh");
    if (quads) std::copy(mesh->m.bdr.begin(), mesh->m.bdr.end(), quads);
    if (attrs) std::copy(mesh->m.bdr_attr.begin(), mesh->m.bdr_attr.end(), attrs);
  });
}

int sx_mesh_perturb(sx_mesh* mesh, uint64_t seed, double amp)
{
  return api([&] {
    SX_REQUIRE(mesh, SX_ERR_ARG, "null mesh");
    SX_REQUIRE(amp >= 0.0 && amp < 0.5, SX_ERR_ARG, "perturbation amplitude must be in [0, 0.5)");
    sxh::perturb_interior(mesh->m, seed, amp);
  });
}

int sx_mesh_info(const sx_mesh* mesh, int64_t* nverts, int64_t* nelem)
{
  return api([&] {
    SX_REQUIRE(mesh, SX_ERR_ARG, "null mesh");
    if (nverts) *nverts = mesh->m.nv;
    if (nelem) *nelem = mesh->m.ne;
  });
}

int sx_mesh_vertices(const sx_mesh* mesh, double* coords_out)
{
  return api([&] {
    SX_REQUIRE(mesh && coords_out, SX_ERR_ARG, "null argument");
    std::memcpy(coords_out, mesh->m.X.data(), mesh->m.X.size() * sizeof(double));
  });
}

int sx_mesh_destroy(sx_mesh* mesh)
{
  delete mesh;
  return SX_OK;
}

// ---------------------------------------------------------------------------
// spaces
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

std::unique_ptr<sx_space> new_space(sx_mesh* mesh, int order, int components, int ordering)
{
  SX_REQUIRE(mesh, SX_ERR_ARG, "null argument");
  SX_REQUIRE(order >= 1 && order <= 4, SX_ERR_UNSUPPORTED, "H1 hex order must be 1..4");
  SX_REQUIRE(components == 1 || components == 3, SX_ERR_UNSUPPORTED, "components must be 1 or 3");
  SX_REQUIRE(ordering == SX_ORDERING_BYNODES || ordering == SX_ORDERING_BYVDIM, SX_ERR_ARG, "bad ordering");
  ensure_device();
  auto s      = std::make_unique<sx_space>();
  s->mesh     = mesh;
  s->order    = order;
  s->nc       = components;
  s->ordering = ordering;
  s->npe      = (order + 1) * (order + 1) * (order + 1);
  return s;
}

// colouring and the device copies of a space whose emap / nnodes are set
void finish_space(sx_space* s)
{
  SX_REQUIRE(s->ndof() < INT32_MAX, SX_ERR_UNSUPPORTED, "more than 2^31 DOFs");
  s->ncolors = sxh::color_elements(s->mesh->m, s->order, s->emap, s->nnodes, s->color);
  s->d_emap.upload(s->emap);
  // node -> element entries in increasing element order
  std::vector<int64_t> offs(size_t(s->nnodes) + 1, 0);
  for (int v : s->emap) offs[size_t(v) + 1]++;
  for (size_t i = 0; i < size_t(s->nnodes); i++) offs[i + 1] += offs[i];
  std::vector<int64_t> pos(offs.begin(), offs.end() - 1);
  std::vector<int32_t> ents(s->emap.size());
  for (size_t k = 0; k < s->emap.size(); k++) ents[size_t(pos[size_t(s->emap[k])]++)] = int32_t(k);
  s->d_offs.upload(offs);
  s->d_ents.upload(ents);
}

}  // namespace

extern "C" {

int sx_space_create_h1_hex(sx_mesh* mesh, int order, int components, int ordering, sx_space** out)
{
  return api([&] {
    SX_REQUIRE(out, SX_ERR_ARG, "null argument");
    auto s    = new_space(mesh, order, components, ordering);
    s->nnodes = sxh::number_nodes(mesh->m, order, s->emap);
    finish_space(s.get());
    *out = s.release();
  });
}

int sx_space_create_h1_hex_map(sx_mesh* mesh, int order, int components, int ordering, int64_t nnodes,
                               const int32_t* elem_nodes, sx_space** out)
{
  return api([&] {
    SX_REQUIRE(out && elem_nodes, SX_ERR_ARG, "null argument");
    auto               s = new_space(mesh, order, components, ordering);
    std::vector<int>   own;
    const int64_t      nown = sxh::number_nodes(mesh->m, order, own);
    SX_REQUIRE(nnodes == nown, SX_ERR_ARG,
               "the map numbers " + std::to_string(nnodes) + " nodes; the order-" + std::to_string(order) +
                   " conforming space of this mesh has " + std::to_string(nown));
    // the caller's map must relabel the mesh's conforming nodes one to one:
    // the same caller node wherever the library sees the same node, and no
    // caller node on two library nodes
    std::vector<int32_t> ext_of_own(size_t(nown), -1), own_of_ext(size_t(nnodes), -1);
    for (size_t k = 0; k < own.size(); k++) {
      const int32_t a = own[k], b = elem_nodes[k];
      const int64_t e = int64_t(k) / s->npe, j = int64_t(k) % s->npe;
      SX_REQUIRE(b >= 0 && b < nnodes, SX_ERR_ARG,
                 "element " + std::to_string(e) + " node " + std::to_string(j) + ": node id " + std::to_string(b) +
                     " out of range [0, " + std::to_string(nnodes) + ")");
      if (ext_of_own[size_t(a)] < 0 && own_of_ext[size_t(b)] < 0) {
        ext_of_own[size_t(a)] = b;
        own_of_ext[size_t(b)] = a;
      }
      SX_REQUIRE(ext_of_own[size_t(a)] == b && own_of_ext[size_t(b)] == a, SX_ERR_ARG,
                 "element " + std::to_string(e) + " node " + std::to_string(j) +
                     ": the map does not number the mesh's shared nodes consistently (node id " +
                     std::to_string(b) + ")");
    }
    s->nnodes = nnodes;
    s->emap.assign(elem_nodes, elem_nodes + own.size());
    s->native = true;
    for (int64_t a = 0; a < nown; a++) s->native = s->native && ext_of_own[size_t(a)] == a;
    if (!s->native) s->node_of_own.assign(ext_of_own.begin(), ext_of_own.end());
    finish_space(s.get());
    *out = s.release();
  });
}

int sx_space_info(const sx_space* s, int64_t* nnodes, int64_t* ndofs, int64_t* nelem, int* nodes_per_elem)
{
  return api([&] {
    SX_REQUIRE(s, SX_ERR_ARG, "null space");
    if (nnodes) *nnodes = s->nnodes;
    if (ndofs) *ndofs = s->ndof();
    if (nelem) *nelem = s->mesh->m.ne;
    if (nodes_per_elem) *nodes_per_elem = s->npe;
  });
}

int sx_space_element_nodes(const sx_space* s, int32_t* map_out)
{
  return api([&] {
    SX_REQUIRE(s && map_out, SX_ERR_ARG, "null argument");
    std::memcpy(map_out, s->emap.data(), s->emap.size() * sizeof(int32_t));
  });
}

int sx_space_node_coords(const sx_space* s, double* xyz)
{
  return api([&] {
    SX_REQUIRE(s && xyz, SX_ERR_ARG, "null argument");
    const auto& m = s->mesh->m;
    const int   n = s->order + 1;
    auto        g = sxh::lobatto_points(n);
    for (int64_t e = 0; e < m.ne; e++)
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            const double w[3][2] = {{1 - g[i], g[i]}, {1 - g[j], g[j]}, {1 - g[k], g[k]}};
            double       x[3]    = {0, 0, 0};
            for (int c = 0; c < 8; c++) {
              const double wt = w[0][c & 1] * w[1][(c >> 1) & 1] * w[2][(c >> 2) & 1];
              const int    v  = m.corner(e, c);
              for (int d = 0; d < 3; d++) x[d] += wt * m.X[3 * size_t(v) + d];
            }
            const int64_t node = s->emap[size_t(e) * s->npe + size_t(i + n * (j + n * k))];
            for (int d = 0; d < 3; d++) xyz[3 * node + d] = x[d];
          }
  });
}

int sx_space_boundary_dofs(const sx_space* s, int face_mask, int comp_mask, int32_t* dofs_out, int64_t cap,
                           int64_t* count)
{
  return api([&] {
    SX_REQUIRE(s && count, SX_ERR_ARG, "null argument");
    const auto& m = s->mesh->m;
    SX_REQUIRE(m.structured, SX_ERR_UNSUPPORTED, "boundary selection by face needs a cuboid mesh");
    const int64_t        p = s->order;
    const int64_t        g[3] = {p * m.n[0] + 1, p * m.n[1] + 1, p * m.n[2] + 1};
    std::vector<int32_t> dofs;
    for (int64_t k = 0; k < g[2]; k++)
      for (int64_t j = 0; j < g[1]; j++)
        for (int64_t i = 0; i < g[0]; i++) {
          const int64_t idx[3] = {i, j, k};
          bool          on     = false;
          for (int d = 0; d < 3; d++) {
            if ((face_mask >> (2 * d)) & 1) on = on || idx[d] == 0;
            if ((face_mask >> (2 * d + 1)) & 1) on = on || idx[d] == g[d] - 1;
          }
          if (!on) continue;
          const int64_t node = s->node_id(i + g[0] * (j + g[1] * k));
          for (int c = 0; c < s->nc; c++)
            if ((comp_mask >> c) & 1)
              dofs.push_back(int32_t(s->ordering ? node * s->nc + c : int64_t(c) * s->nnodes + node));
        }
    std::sort(dofs.begin(), dofs.end());
    *count = int64_t(dofs.size());
    if (dofs_out) {
      SX_REQUIRE(cap >= int64_t(dofs.size()), SX_ERR_ARG, "boundary DOF buffer too small");
      std::memcpy(dofs_out, dofs.data(), dofs.size() * sizeof(int32_t));
    }
  });
}

int sx_space_boundary_dofs_attr(const sx_space* s, uint64_t attr_mask, int comp_mask, int32_t* dofs_out, int64_t cap,
                                int64_t* count)
{
  return api([&] {
    SX_REQUIRE(s && count, SX_ERR_ARG, "null argument");
    const int                 n     = s->order + 1;
    std::vector<BoundaryFace> faces = attribute_faces(s->mesh->m, attr_mask, false);
    std::vector<int32_t>      dofs;
    for (const auto& fc : faces)
      for (int b = 0; b < n; b++)
        for (int a = 0; a < n; a++) {
          int i0, i1, i2;
          face_node(fc.lf, a, b, n, i0, i1, i2);
          const int64_t node = s->emap[size_t(fc.e) * size_t(s->npe) + size_t((i2 * n + i1) * n + i0)];
          for (int c = 0; c < s->nc; c++)
            if ((comp_mask >> c) & 1)
              dofs.push_back(int32_t(s->ordering ? node * s->nc + c : int64_t(c) * s->nnodes + node));
        }
    std::sort(dofs.begin(), dofs.end());
    dofs.erase(std::unique(dofs.begin(), dofs.end()), dofs.end());
    *count = int64_t(dofs.size());
    if (dofs_out) {
      SX_REQUIRE(cap >= int64_t(dofs.size()), SX_ERR_ARG, "boundary DOF buffer too small");
      std::memcpy(dofs_out, dofs.data(), dofs.size() * sizeof(int32_t));
    }
  });
}

int sx_space_gather(const sx_space* s, const double* L, double* E)
{
  return api([&] {
    SX_REQUIRE(s && L && E, SX_ERR_ARG, "null argument");
    SX_HIP(sxk::gather_e(s->d_emap.as<const int>(), L, E, s->mesh->m.ne, s->npe, s->nc, s->nnodes, s->ordering,
                         nullptr));
    SX_HIP(hipDeviceSynchronize());
  });
}

int sx_space_scatter_add(const sx_space* s, const double* E, double* L)
{
  return api([&] {
    SX_REQUIRE(s && L && E, SX_ERR_ARG, "null argument");
    SX_HIP(sxk::scatter_e(s->d_offs.as<const int64_t>(), s->d_ents.as<const int32_t>(), E, L, s->nnodes, s->npe,
                          s->nc, s->ordering, nullptr));
    SX_HIP(hipDeviceSynchronize());
  });
}

int sx_space_destroy(sx_space* s)
{
  delete s;
  return SX_OK;
}

// ---------------------------------------------------------------------------
// forms
// ---------------------------------------------------------------------------
int sx_form_create(const sx_space* test, const sx_space* const* trials, int ntrials, sx_form** out)
{
  return api([&] {
    SX_REQUIRE(test && out && ntrials >= 1 && trials, SX_ERR_ARG, "form needs a test space and >= 1 trial space");
    ensure_device();
    auto f  = std::make_unique<sx_form>();
    f->test = test;
    for (int i = 0; i < ntrials; i++) {
      SX_REQUIRE(trials[i], SX_ERR_ARG, "null trial space");
      f->trials.push_back(trials[i]);
      auto g  = std::make_unique<sx_grad>();
      g->form = f.get();
      g->wrt  = i;
      f->grads.push_back(std::move(g));
    }
    SX_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
    f->part.allocate(size_t(sxk::reduction_blocks()) * sizeof(double));
    f->scal.allocate(16 * sizeof(double));
    *out = f.release();
  });
}

int sx_form_add_domain_integral(sx_form* f, int qfunc_id, const double* params, int nparams, const int* depends_on,
                                int ndep, const int* elems, int nelem)
{
  return sx_form_add_domain_integral_qdata(f, qfunc_id, params, nparams, depends_on, ndep, elems, nelem, nullptr);
}

namespace {

// One domain integral of qfunction qfunc_id in trial argument `trial`
// (companion: the mass part of a two-argument integral).
void add_domain_integral_one(sx_form* f, int qfunc_id, const double* params, int nparams, int trial, const int* elems,
                             int nelem, sx_qdata* qdata, bool companion)
{
  SX_REQUIRE(f, SX_ERR_ARG, "null form");
  SX_REQUIRE(nparams >= 0 && nparams <= SX_MAX_PARAMS && (nparams == 0 || params), SX_ERR_ARG,
             "qfunction parameters: at most 16 doubles");
  const int nc_qf = qf_components(qfunc_id);
  SX_REQUIRE(nc_qf != 0, SX_ERR_UNSUPPORTED, "unknown qfunction id " + std::to_string(qfunc_id));
  SX_REQUIRE(trial >= 0 && trial < int(f->trials.size()), SX_ERR_ARG, "DependsOn index out of range");
  const sx_space* ts = f->trials[size_t(trial)];
  const sx_space* vs = f->test;
  SX_REQUIRE(ts->mesh == vs->mesh, SX_ERR_UNSUPPORTED, "test and trial spaces must share a mesh");
  SX_REQUIRE(ts->order == vs->order && ts->nc == vs->nc && ts->ordering == vs->ordering, SX_ERR_UNSUPPORTED,
             "test and trial spaces must have the same order, components and ordering");
  SX_REQUIRE(ts == vs || (ts->native && vs->native) || ts->emap == vs->emap, SX_ERR_UNSUPPORTED,
             "test and trial spaces must number their nodes alike");
  SX_REQUIRE(nc_qf < 0 || nc_qf == vs->nc, SX_ERR_ARG, "qfunction does not match the number of components");
  const int ssize = qf_state_size(qfunc_id);
  SX_REQUIRE(ssize == 0 || qdata, SX_ERR_ARG, "this qfunction needs quadrature-point state (sx_qdata)");
  SX_REQUIRE(ssize > 0 || !qdata, SX_ERR_ARG, "this qfunction carries no quadrature-point state");
  const int nq = (vs->order + 1) * (vs->order + 1) * (vs->order + 1);

  auto I = std::make_unique<Integral>();
  I->qf  = qfunc_id;
  for (int k = 0; k < nparams; k++) I->params[k] = params[k];
  I->depends_on = {trial};
  I->layout     = qf_layout(qfunc_id);

  const auto& m = vs->mesh->m;
  std::vector<int> domain;
  if (elems && nelem > 0) {
    domain.assign(elems, elems + nelem);
    for (int e : domain) SX_REQUIRE(e >= 0 && e < m.ne, SX_ERR_ARG, "domain element out of range");
  } else {
    domain.resize(size_t(m.ne));
    for (int64_t e = 0; e < m.ne; e++) domain[size_t(e)] = int(e);
  }
  if (qdata) {
    SX_REQUIRE(qdata->nelem == int64_t(domain.size()) && qdata->nq == nq && qdata->size == ssize, SX_ERR_ARG,
               "quadrature data shape (" + std::to_string(qdata->nelem) + " x " + std::to_string(qdata->nq) +
                   " x " + std::to_string(qdata->size) + ") does not match the domain (" +
                   std::to_string(domain.size()) + " x " + std::to_string(nq) + " x " + std::to_string(ssize) +
                   ")");
  }
  // slots: positions in the domain sorted by element colour (stable); with
  // atomic scatter, one range in domain order
  std::vector<int> pos;
  pos.reserve(domain.size());
  if (f->scatter == SX_SCATTER_ATOMIC) {
    for (size_t i = 0; i < domain.size(); i++) pos.push_back(int(i));
    I->color_off = {0, int64_t(pos.size())};
  } else {
    I->color_off.assign(size_t(vs->ncolors) + 1, 0);
    for (int c = 0; c < vs->ncolors; c++) {
      for (size_t i = 0; i < domain.size(); i++)
        if (vs->color[size_t(domain[i])] == c) pos.push_back(int(i));
      I->color_off[size_t(c) + 1] = int64_t(pos.size());
    }
  }
  std::vector<int> slots(pos.size());
  for (size_t k = 0; k < pos.size(); k++) slots[k] = domain[size_t(pos[k])];
  I->nslots   = int64_t(slots.size());
  I->slot_pos = pos;
  if (qdata) {
    I->qdata = qdata;
    I->d_qrow.upload(pos);
  }
  std::vector<int>    map(size_t(I->nslots) * vs->npe);
  std::vector<double> xyz(size_t(I->nslots) * 24);
  for (int64_t k = 0; k < I->nslots; k++) {
    const int e = slots[size_t(k)];
    std::memcpy(&map[size_t(k) * vs->npe], &vs->emap[size_t(e) * vs->npe], sizeof(int) * vs->npe);
    for (int c = 0; c < 8; c++) {
      const int v = m.corner(e, c);
      for (int d = 0; d < 3; d++) xyz[size_t(k) * 24 + c * 3 + d] = m.X[3 * size_t(v) + d];
    }
  }
  I->d_map.upload(map);
  I->d_coords.upload(xyz);
  I->companion = companion;
  f->integrals.push_back(std::move(I));
}

// The coefficient of the second argument's value in the source of a
// two-argument integrand: MaterialStressFunctor's density * acceleration
// (phy/solid_mechanics.hpp:834-852) and ThermalMaterialIntegrand's
// heat_capacity * rate (phy/heat_transfer.hpp:380-406).  Which parameter holds
// it depends on the qfunction; 0 for qfunctions without a two-argument form.
bool rate_coefficient(int qf, const double* params, int nparams, double* coef)
{
  int k = -1;
  switch (qf) {
    case SX_QF_NEOHOOKEAN: k = 0; break;      // {density, K, G}
    case SX_QF_J2_LINEAR: k = 5; break;       // {E, nu, sigma_y, Hi, Hk, density, geom}
    case SX_QF_LINEAR_ELASTIC: k = 2; break;  // {lambda, mu, density}
    case SX_QF_HEAT: k = 2; break;            // {k, s, rho c_p}
    default: return false;
  }
  *coef = k < nparams ? params[k] : 0.0;
  return true;
}

}  // namespace

int sx_form_add_domain_integral_qdata(sx_form* f, int qfunc_id, const double* params, int nparams,
                                      const int* depends_on, int ndep, const int* elems, int nelem, sx_qdata* qdata)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    SX_REQUIRE(depends_on && (ndep == 1 || ndep == 2), SX_ERR_UNSUPPORTED,
               "compiled qfunctions take one trial argument (DependsOn<i>), or two for the integrands with a rate");
    SX_REQUIRE(nparams >= 0 && nparams <= SX_MAX_PARAMS && (nparams == 0 || params), SX_ERR_ARG,
               "qfunction parameters: at most 16 doubles");
    double coef = 0.0;
    if (ndep == 2) {
      SX_REQUIRE(rate_coefficient(qfunc_id, params, nparams, &coef), SX_ERR_UNSUPPORTED,
                 "qfunction " + std::to_string(qfunc_id) + " has no two-argument form (DependsOn<a, b>)");
      SX_REQUIRE(depends_on[1] >= 0 && depends_on[1] < int(f->trials.size()), SX_ERR_ARG,
                 "DependsOn index out of range");
      SX_REQUIRE(depends_on[0] != depends_on[1], SX_ERR_ARG, "DependsOn<a, b> needs two different arguments");
    }
    const size_t first = f->integrals.size();
    add_domain_integral_one(f, qfunc_id, params, nparams, depends_on[0], elems, nelem, qdata, false);
    if (ndep == 2) {
      // the second argument enters the source linearly, as coef * value: its
      // part is the mass integral {coef u, 0} in that argument
      const double mp[2] = {coef, 0.0};
      try {
        add_domain_integral_one(f, SX_QF_MASS_DIFFUSION, mp, 2, depends_on[1], elems, nelem, nullptr, true);
      } catch (...) {
        f->integrals.resize(first);
        throw;
      }
      f->integrals[first]->has_companion = true;
    }
  });
}

namespace {

// selects the faces of a boundary integral on the form's test-space mesh
using FaceSelector = std::function<std::vector<BoundaryFace>(const sxh::HexMesh&)>;

int add_boundary_integral(sx_form* f, int qfunc_id, const double* params, int nparams, const int* depends_on,
                          int ndep, const FaceSelector& select)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    SX_REQUIRE(nparams >= 0 && nparams <= SX_MAX_PARAMS && (nparams == 0 || params), SX_ERR_ARG,
               "qfunction parameters: at most 16 doubles");
    SX_REQUIRE(is_boundary_qf(qfunc_id), SX_ERR_UNSUPPORTED, "unknown boundary qfunction id " + std::to_string(qfunc_id));
    SX_REQUIRE(ndep == 1 && depends_on, SX_ERR_UNSUPPORTED,
               "compiled boundary qfunctions take exactly one trial argument (DependsOn<i>)");
    const int trial = depends_on[0];
    SX_REQUIRE(trial >= 0 && trial < int(f->trials.size()), SX_ERR_ARG, "DependsOn index out of range");
    const sx_space* vs = f->test;
    SX_REQUIRE(f->trials[size_t(trial)] == vs || (f->trials[size_t(trial)]->mesh == vs->mesh &&
                                                  f->trials[size_t(trial)]->order == vs->order &&
                                                  f->trials[size_t(trial)]->nc == vs->nc &&
                                                  f->trials[size_t(trial)]->ordering == vs->ordering),
               SX_ERR_UNSUPPORTED, "test and trial spaces must match");
    SX_REQUIRE(bqf_components(qfunc_id) == vs->nc, SX_ERR_ARG, "qfunction does not match the number of components");
    auto B = std::make_unique<BoundaryIntegral>();
    B->qf  = qfunc_id;
    B->nc  = vs->nc;
    for (int k = 0; k < nparams; k++) B->params[k] = params[k];
    B->depends_on = {trial};
    const auto& m = vs->mesh->m;
    const int   n = vs->order + 1, nf = n * n;
    std::vector<BoundaryFace> faces = select(m);
    // greedy colouring: faces of one colour share no node
    std::vector<int> fnodes(faces.size() * size_t(nf));
    for (size_t i = 0; i < faces.size(); i++)
      for (int b = 0; b < n; b++)
        for (int a = 0; a < n; a++) {
          int i0, i1, i2;
          face_node(faces[i].lf, a, b, n, i0, i1, i2);
          fnodes[i * size_t(nf) + size_t(b * n + a)] =
              vs->emap[size_t(faces[i].e) * size_t(vs->npe) + size_t((i2 * n + i1) * n + i0)];
        }
    std::vector<int>              color(faces.size(), -1);
    std::vector<std::vector<int>> node_colors(size_t(vs->nnodes));
    int                           ncolors = 0;
    for (size_t i = 0; i < faces.size(); i++) {
      int c = 0;
      for (;; c++) {
        bool used = false;
        for (int k = 0; k < nf && !used; k++)
          for (int x : node_colors[size_t(fnodes[i * size_t(nf) + size_t(k)])]) used = used || x == c;
        if (!used) break;
      }
      color[i] = c;
      ncolors  = std::max(ncolors, c + 1);
      for (int k = 0; k < nf; k++) node_colors[size_t(fnodes[i * size_t(nf) + size_t(k)])].push_back(c);
    }
    std::vector<int>    map;
    std::vector<double> corners;
    B->color_off.assign(size_t(ncolors) + 1, 0);
    for (int c = 0; c < ncolors; c++) {
      for (size_t i = 0; i < faces.size(); i++) {
        if (color[i] != c) continue;
        map.insert(map.end(), fnodes.begin() + long(i * size_t(nf)), fnodes.begin() + long((i + 1) * size_t(nf)));
        for (int k = 0; k < 4; k++) {
          int i0, i1, i2;
          face_node(faces[i].lf, k & 1, k >> 1, 2, i0, i1, i2);
          const int v = m.corner(faces[i].e, i0 + 2 * i1 + 4 * i2);
          for (int d = 0; d < 3; d++) corners.push_back(m.X[size_t(3 * v + d)]);
        }
      }
      B->color_off[size_t(c) + 1] = int64_t(map.size() / size_t(nf));
    }
    B->nslots = int64_t(faces.size());
    B->d_fmap.upload(map);
    B->d_corners.upload(corners);
    f->bintegrals.push_back(std::move(B));
  });
}

}  // namespace

int sx_form_add_boundary_integral(sx_form* f, int qfunc_id, const double* params, int nparams,
                                  const int* depends_on, int ndep, int face_mask)
{
  if (!(face_mask >= 0 && face_mask < 64)) {
    g_last_error = "face_mask selects among 6 planes";
    return SX_ERR_ARG;
  }
  return add_boundary_integral(f, qfunc_id, params, nparams, depends_on, ndep,
                               [face_mask](const sxh::HexMesh& m) { return boundary_faces(m, face_mask); });
}

int sx_form_add_boundary_integral_attr(sx_form* f, int qfunc_id, const double* params, int nparams,
                                       const int* depends_on, int ndep, uint64_t attr_mask)
{
  return add_boundary_integral(f, qfunc_id, params, nparams, depends_on, ndep, [attr_mask](const sxh::HexMesh& m) {
    return attribute_faces(m, attr_mask, true);
  });
}

int sx_form_add_boundary_integral_pred(sx_form* f, int qfunc_id, const double* params, int nparams,
                                       const int* depends_on, int ndep, sx_face_predicate pred, void* user)
{
  if (!pred) {
    g_last_error = "null face predicate";
    return SX_ERR_ARG;
  }
  return add_boundary_integral(f, qfunc_id, params, nparams, depends_on, ndep,
                               [pred, user](const sxh::HexMesh& m) { return predicate_faces(m, pred, user); });
}

int sx_form_add_boundary_integral_faces(sx_form* f, int qfunc_id, const double* params, int nparams,
                                        const int* depends_on, int ndep, const int32_t* face_vertices, int64_t nfaces)
{
  if (nfaces < 0 || (nfaces > 0 && !face_vertices)) {
    g_last_error = "null face list";
    return SX_ERR_ARG;
  }
  return add_boundary_integral(f, qfunc_id, params, nparams, depends_on, ndep,
                               [face_vertices, nfaces](const sxh::HexMesh& m) {
                                 return listed_faces(m, face_vertices, nfaces);
                               });
}

int sx_form_update_qdata(sx_form* f, int flag)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    f->update_qdata = flag ? 1 : 0;
  });
}

int sx_qdata_create(int64_t nelem, int nq, int size, const double* init, sx_qdata** out)
{
  return api([&] {
    SX_REQUIRE(out && nelem >= 0 && nq > 0 && size > 0, SX_ERR_ARG, "bad quadrature data shape");
    SX_REQUIRE(init, SX_ERR_ARG, "null initial state");
    auto q   = std::make_unique<sx_qdata>();
    q->nelem = nelem;
    q->nq    = nq;
    q->size  = size;
    std::vector<double> h(size_t(nelem) * size_t(size) * size_t(nq));
    for (int64_t e = 0; e < nelem; e++)
      for (int k = 0; k < size; k++)
        std::fill_n(&h[(size_t(e) * size_t(size) + size_t(k)) * size_t(nq)], nq, init[k]);
    q->d.upload(h);
    *out = q.release();
  });
}

int sx_qdata_info(const sx_qdata* q, int64_t* out)
{
  return api([&] {
    SX_REQUIRE(q && out, SX_ERR_ARG, "null argument");
    out[0] = q->nelem;
    out[1] = q->nq;
    out[2] = q->size;
  });
}

int sx_qdata_get(const sx_qdata* q, double* host)
{
  return api([&] {
    SX_REQUIRE(q && host, SX_ERR_ARG, "null argument");
    const size_t        n = size_t(q->nelem) * size_t(q->size) * size_t(q->nq);
    std::vector<double> h(n);
    SX_HIP(hipDeviceSynchronize());
    SX_HIP(hipMemcpy(h.data(), q->d.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < q->nelem; e++)
      for (int k = 0; k < q->size; k++)
        for (int p = 0; p < q->nq; p++)
          host[(size_t(e) * size_t(q->nq) + size_t(p)) * size_t(q->size) + size_t(k)] =
              h[(size_t(e) * size_t(q->size) + size_t(k)) * size_t(q->nq) + size_t(p)];
  });
}

int sx_qdata_set(sx_qdata* q, const double* host)
{
  return api([&] {
    SX_REQUIRE(q && host, SX_ERR_ARG, "null argument");
    const size_t        n = size_t(q->nelem) * size_t(q->size) * size_t(q->nq);
    std::vector<double> h(n);
    for (int64_t e = 0; e < q->nelem; e++)
      for (int k = 0; k < q->size; k++)
        for (int p = 0; p < q->nq; p++)
          h[(size_t(e) * size_t(q->size) + size_t(k)) * size_t(q->nq) + size_t(p)] =
              host[(size_t(e) * size_t(q->nq) + size_t(p)) * size_t(q->size) + size_t(k)];
    SX_HIP(hipDeviceSynchronize());
    SX_HIP(hipMemcpy(q->d.ptr, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
  });
}

int sx_qdata_destroy(sx_qdata* q)
{
  delete q;
  return SX_OK;
}

int sx_form_eval(sx_form* f, double t, const double* const* inputs, double* out, int wrt, sx_grad** grad_out)
{
  return api([&] {
    SX_REQUIRE(f && inputs && out, SX_ERR_ARG, "null argument");
    SX_REQUIRE((wrt >= -1 || wrt == SX_WRT_SHAPE) && wrt < int(f->trials.size()), SX_ERR_ARG, "wrt out of range");
    for (size_t i = 0; i < f->trials.size(); i++) SX_REQUIRE(inputs[i], SX_ERR_ARG, "null input vector");
    if (wrt == SX_WRT_SHAPE) {
      SX_REQUIRE(f->test->nc == 3, SX_ERR_UNSUPPORTED, "shape derivatives need a 3-component test space");
      SX_REQUIRE(f->bintegrals.empty(), SX_ERR_UNSUPPORTED,
                 "shape derivatives of boundary integrals are not built (domain integrals only)");
    }
    SX_HIP(sxk::vec_fill(out, 0.0, f->test->ndof(), f->stream));
    const int nq = (f->test->order + 1) * (f->test->order + 1) * (f->test->order + 1);
    for (auto& I : f->integrals) {
      const bool lin = (wrt >= 0) && depends(*I, wrt);
      if (lin) {
        const size_t bytes =
            size_t(I->nslots) * size_t(sxk::tangent_slot_bytes(sxk::tangent_ds(f->test->nc, I->layout), nq, f->tprec));
        I->d_tangent.allocate(bytes);
        I->tprec = f->tprec;
        I->has_tangent = true;
      }
      run_integral(f, *I, lin ? sxk::MODE_LINEARIZE : sxk::MODE_RESIDUAL, inputs[I->depends_on[0]], out);
    }
    for (auto& B : f->bintegrals) {
      const bool lin = (wrt >= 0) && depends(*B, wrt);
      if (lin) {
        const int q2 = (f->test->order + 1) * (f->test->order + 1);
        B->d_tangent.allocate(size_t(B->nslots) * size_t(B->nc) * size_t(3 * B->nc) * size_t(q2) * sizeof(double));
        B->has_tangent = true;
      }
      run_boundary(f, *B, lin ? sxk::FMODE_LINEARIZE : sxk::FMODE_RESIDUAL, inputs[B->depends_on[0]], out, nullptr,
                   sxk::PREC_F64);
    }
    (void)t;
    if (wrt >= 0) {
      f->grads[size_t(wrt)]->valid = true;
      if (grad_out) *grad_out = f->grads[size_t(wrt)].get();
    }
    if (wrt == SX_WRT_SHAPE) {
      // keep the linearization point: the trial inputs and the shape displacement
      if (!f->shape_grad) {
        f->shape_grad        = std::make_unique<sx_grad>();
        f->shape_grad->form  = f;
        f->shape_grad->wrt   = SX_WRT_SHAPE;
        f->shape_grad->shape = true;
      }
      f->shape_u.resize(f->trials.size());
      for (size_t i = 0; i < f->trials.size(); i++) {
        if (!f->shape_u[i]) f->shape_u[i] = std::make_unique<DeviceBuffer>();
        const size_t bytes = size_t(f->trials[i]->ndof()) * sizeof(double);
        f->shape_u[i]->allocate(bytes);
        SX_HIP(hipMemcpyAsync(f->shape_u[i]->ptr, inputs[i], bytes, hipMemcpyDeviceToDevice, f->stream));
      }
      const size_t pbytes = size_t(f->test->ndof()) * sizeof(double);
      f->shape_p.allocate(pbytes);
      if (f->shape)
        SX_HIP(hipMemcpyAsync(f->shape_p.ptr, f->shape, pbytes, hipMemcpyDeviceToDevice, f->stream));
      else
        SX_HIP(sxk::vec_fill(f->shape_p.as<double>(), 0.0, f->test->ndof(), f->stream));
      f->shape_grad->valid = true;
      if (grad_out) *grad_out = f->shape_grad.get();
    }
  });
}

int sx_form_set_option(sx_form* f, int option, int64_t value)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    switch (option) {
      case SX_OPT_SCATTER:
        SX_REQUIRE(value == SX_SCATTER_COLORED || value == SX_SCATTER_ATOMIC, SX_ERR_ARG, "bad scatter mode");
        SX_REQUIRE(f->integrals.empty(), SX_ERR_STATE, "scatter mode must be set before adding integrals");
        f->scatter = int(value);
        return;
      case SX_OPT_XCD_REMAP:
        f->xcd_remap = value ? 1 : 0;
        return;
      case SX_OPT_TANGENT_PRECISION:
        SX_REQUIRE(value == SX_PREC_F64 || value == SX_PREC_F32 || value == SX_PREC_BF16, SX_ERR_ARG,
                   "bad tangent precision");
        f->tprec = int(value);
        return;
      case SX_OPT_JVP_ARITHMETIC:
        SX_REQUIRE(value == SX_PREC_F64 || value == SX_PREC_F32, SX_ERR_ARG, "J.v arithmetic is FP64 or FP32");
        f->jvp_arith = int(value);
        return;
      case SX_OPT_LINEAR_SOLVER:
        SX_REQUIRE(value == SX_LINSOLVER_CG || value == SX_LINSOLVER_GMRES, SX_ERR_ARG, "linear solver is CG or GMRES");
        f->linsolver = int(value);
        return;
      case SX_OPT_GMRES_RESTART:
        SX_REQUIRE(value >= 1 && value <= 1000, SX_ERR_ARG, "GMRES restart length must be in 1..1000");
        f->gmres_restart = int(value);
        return;
      case SX_OPT_SCATTER_ORDER:
        SX_REQUIRE(value >= 0 && value <= 2, SX_ERR_ARG, "scatter order is 0, 1 or 2");
        f->scatter_order = int(value);
        return;
      case SX_OPT_LINE_SEARCH:
        SX_REQUIRE(value >= 0 && value <= 64, SX_ERR_ARG, "line search iterations must be in 0..64");
        f->line_search = int(value);
        return;
      case SX_OPT_MIN_NEWTON_ITERATIONS:
        SX_REQUIRE(value >= 0 && value <= 1000, SX_ERR_ARG, "minimum Newton iterations must be in 0..1000");
        f->min_newton = int(value);
        return;
      case SX_OPT_PRINT_LEVEL:
        SX_REQUIRE(value >= 0, SX_ERR_ARG, "print level must be >= 0");
        f->print_level = int(value);
        return;
    }
    throw sx_failure(SX_ERR_ARG, "unknown form option " + std::to_string(option));
  });
}

int sx_form_synchronize(sx_form* f)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    SX_HIP(hipStreamSynchronize(f->stream));
  });
}

int sx_form_stream(sx_form* f, void** hip_stream)
{
  return api([&] {
    SX_REQUIRE(f && hip_stream, SX_ERR_ARG, "null argument");
    *hip_stream = f->stream;
  });
}

int sx_form_destroy(sx_form* f)
{
  if (f && f->stream) (void)hipStreamSynchronize(f->stream);
  delete f;
  return SX_OK;
}

// ---------------------------------------------------------------------------
// gradient operator
// ---------------------------------------------------------------------------
int sx_grad_apply(sx_grad* g, const double* dx, double* df)
{
  return api([&] { apply_gradient(g, dx, df); });
}

}  // extern "C"

namespace {

// Domain integral k of a form run on E-vectors ([element][component][node] over
// the integral's domain positions): the element kernels with a map that gives
// every (element, node) its own entry, so there is no gather sharing and no
// scatter sum.  The kernels address an entry as component-major
// (c * ne * npe + pos * npe + j); the reference layout is transposed in and out.
// the k-th domain integral as the caller added them (a two-argument
// integral's companion is not counted)
Integral* user_integral(const sx_form* f, int k)
{
  if (!f || k < 0) return nullptr;
  for (const auto& I : f->integrals)
    if (!I->companion && k-- == 0) return I.get();
  return nullptr;
}

Integral& e_integral(sx_form* f, int k)
{
  SX_REQUIRE(f, SX_ERR_ARG, "null form");
  Integral* Ip = user_integral(f, k);
  SX_REQUIRE(Ip, SX_ERR_ARG, "domain integral index out of range");
  SX_REQUIRE(!f->shape, SX_ERR_UNSUPPORTED, "E-vector evaluation of a shape-aware form is not built");
  SX_REQUIRE(!Ip->has_companion, SX_ERR_UNSUPPORTED,
             "E-vector evaluation of a two-argument integral (DependsOn<a, b>) is not built");
  Integral& I = *Ip;
  if (!I.d_emap_e.ptr) {
    const int        npe = f->test->npe;
    std::vector<int> m(size_t(I.nslots) * size_t(npe));
    for (int64_t sl = 0; sl < I.nslots; sl++)
      for (int j = 0; j < npe; j++) m[size_t(sl) * size_t(npe) + size_t(j)] = I.slot_pos[size_t(sl)] * npe + j;
    I.d_emap_e.upload(m);
  }
  return I;
}

void run_integral_e(sx_form* f, Integral& I, int mode, const double* in_E, double* out_E, const double* stop = nullptr)
{
  const int     nc = f->test->nc, npe = f->test->npe;
  const int64_t n  = I.nslots * int64_t(npe) * nc;
  hipStream_t   s  = f->stream;
  DeviceBuffer  ti, to;
  const double* in  = in_E;
  double*       out = out_E;
  if (nc > 1) {
    ti.allocate(size_t(n) * 8);
    to.allocate(size_t(n) * 8);
    SX_HIP(sxk::e_transpose(in_E, ti.as<double>(), I.nslots, nc, npe, 0, s));
    in  = ti.as<const double>();
    out = to.as<double>();
  }
  SX_HIP(sxk::vec_fill(out, 0.0, n, s));
  sxk::ElemLaunch P = base_launch(f, I, mode);
  P.map             = I.d_emap_e.as<const int>();
  P.nnodes          = I.nslots * int64_t(npe);
  P.ordering        = SX_ORDERING_BYNODES;
  P.scatter         = SX_SCATTER_COLORED;  // every entry has one writer
  P.u               = in;
  P.r               = out;
  P.stop            = stop;
  P.slot_begin      = 0;
  P.slot_end        = I.nslots;
  if (P.slot_end > 0) {
    hipError_t err = hipSuccess;
    SX_REQUIRE(sxk::launch_element(P, &err), SX_ERR_UNSUPPORTED, "element kernel not compiled for this integral");
    SX_HIP(err);
  }
  if (nc > 1) SX_HIP(sxk::e_transpose(out, out_E, I.nslots, nc, npe, 1, s));
  SX_HIP(hipStreamSynchronize(s));  // the transposition buffers are released here
}

}  // namespace

extern "C" {

int sx_form_integral_info(const sx_form* f, int k, int64_t* nelem, int* nodes_per_elem, int* ncomp)
{
  return api([&] {
    const Integral* I = user_integral(f, k);
    SX_REQUIRE(I, SX_ERR_ARG, "domain integral index out of range");
    if (nelem) *nelem = I->nslots;
    if (nodes_per_elem) *nodes_per_elem = f->test->npe;
    if (ncomp) *ncomp = f->test->nc;
  });
}

int sx_form_eval_e(sx_form* f, int k, double t, const double* u_E, double* r_E, int linearize)
{
  return api([&] {
    SX_REQUIRE(u_E && r_E, SX_ERR_ARG, "null vector");
    Integral& I = e_integral(f, k);
    if (linearize) {
      const int    nq    = f->test->npe;
      const size_t bytes = size_t(I.nslots) *
                           size_t(sxk::tangent_slot_bytes(sxk::tangent_ds(f->test->nc, I.layout), nq, f->tprec));
      I.d_tangent.allocate(bytes);
      I.tprec       = f->tprec;
      I.has_tangent = true;
    }
    (void)t;
    run_integral_e(f, I, linearize ? sxk::MODE_LINEARIZE : sxk::MODE_RESIDUAL, u_E, r_E);
  });
}

int sx_grad_apply_e(sx_grad* g, int k, const double* du_E, double* dr_E)
{
  return api([&] {
    SX_REQUIRE(g && g->form && du_E && dr_E, SX_ERR_ARG, "null argument");
    SX_REQUIRE(!g->shape, SX_ERR_UNSUPPORTED, "the shape gradient has no E-vector form");
    Integral& I = e_integral(g->form, k);
    SX_REQUIRE(depends(I, g->wrt), SX_ERR_ARG, "the integral does not depend on the gradient's trial argument");
    SX_REQUIRE(I.has_tangent, SX_ERR_STATE, "E-vector J.v before a linearization of the integral");
    run_integral_e(g->form, I, sxk::MODE_JVP, du_E, dr_E);
  });
}

int sx_grad_diagonal(sx_grad* g, double* diag)
{
  return api([&] {
    SX_REQUIRE(g && g->form && diag, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "diagonal requested before a linearization");
    sx_form* f = g->form;
    SX_HIP(sxk::vec_fill(diag, 0.0, f->test->ndof(), f->stream));
    for (auto& I : f->integrals)
      if (depends(*I, g->wrt) && I->has_tangent) run_integral(f, *I, sxk::MODE_DIAG, nullptr, diag);
    for (auto& B : f->bintegrals)
      if (depends(*B, g->wrt) && B->has_tangent)
        run_boundary(f, *B, sxk::FMODE_DIAG, nullptr, diag, nullptr, sxk::PREC_F64);
  });
}

int sx_grad_assemble(sx_grad* g, sx_csr** out)
{
  return api([&] {
    SX_REQUIRE(g && g->form && out, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "assembly requested before a linearization");
    sx_form*  f  = g->form;
    sx_space* ts = const_cast<sx_space*>(f->test);
    SX_REQUIRE(ts->ndof() < (int64_t(1) << 31), SX_ERR_UNSUPPORTED, "CSR columns are int32");
    build_csr_pattern(ts);
    auto A    = std::make_unique<sx_csr>();
    A->nrows  = ts->ndof();
    A->ncols  = f->trials[size_t(g->wrt)]->ndof();
    A->nnz    = int64_t(ts->csr_col.size());
    A->stream = f->stream;
    A->row.upload(ts->csr_row);
    A->col.upload(ts->csr_col);
    A->val.allocate(size_t(std::max<int64_t>(A->nnz, 1)) * sizeof(double));
    SX_HIP(sxk::vec_fill(A->val.as<double>(), 0.0, A->nnz, f->stream));
    for (auto& I : f->integrals) {
      if (!(depends(*I, g->wrt) && I->has_tangent)) continue;
      sxk::ElemLaunch P = base_launch(f, *I, sxk::MODE_ASSEMBLE);
      P.csr_row         = A->row.as<const int64_t>();
      P.csr_col         = A->col.as<const int32_t>();
      P.csr_val         = A->val.as<double>();
      for (size_t c = 0; c + 1 < I->color_off.size(); c++) {
        P.slot_begin = I->color_off[c];
        P.slot_end   = I->color_off[c + 1];
        if (P.slot_end <= P.slot_begin) continue;
        hipError_t err = hipSuccess;
        SX_REQUIRE(sxk::launch_element(P, &err), SX_ERR_UNSUPPORTED, "assembly kernel not compiled for this form");
        SX_HIP(err);
      }
    }
    for (auto& B : f->bintegrals)
      if (depends(*B, g->wrt) && B->has_tangent)
        run_boundary(f, *B, sxk::FMODE_ASSEMBLE, nullptr, nullptr, nullptr, sxk::PREC_F64, A.get());
    SX_HIP(hipStreamSynchronize(f->stream));
    *out = A.release();
  });
}

int sx_csr_info(const sx_csr* A, int64_t* nrows, int64_t* ncols, int64_t* nnz)
{
  return api([&] {
    SX_REQUIRE(A, SX_ERR_ARG, "null matrix");
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = A->nnz;
  });
}

int sx_csr_arrays(const sx_csr* A, const int64_t** row_ptr, const int32_t** col, const double** val)
{
  return api([&] {
    SX_REQUIRE(A, SX_ERR_ARG, "null matrix");
    if (row_ptr) *row_ptr = A->row.as<const int64_t>();
    if (col) *col = A->col.as<const int32_t>();
    if (val) *val = A->val.as<const double>();
  });
}

int sx_csr_download(const sx_csr* A, int64_t* row_ptr, int32_t* col, double* val)
{
  return api([&] {
    SX_REQUIRE(A, SX_ERR_ARG, "null matrix");
    SX_HIP(hipStreamSynchronize(A->stream));
    if (row_ptr) SX_HIP(hipMemcpy(row_ptr, A->row.ptr, size_t(A->nrows + 1) * 8, hipMemcpyDeviceToHost));
    if (col && A->nnz) SX_HIP(hipMemcpy(col, A->col.ptr, size_t(A->nnz) * 4, hipMemcpyDeviceToHost));
    if (val && A->nnz) SX_HIP(hipMemcpy(val, A->val.ptr, size_t(A->nnz) * 8, hipMemcpyDeviceToHost));
  });
}

int sx_csr_spmv(const sx_csr* A, const double* x, double* y)
{
  return api([&] {
    SX_REQUIRE(A && x && y, SX_ERR_ARG, "null argument");
    SX_HIP(sxk::csr_spmv(A->row.as<const int64_t>(), A->col.as<const int32_t>(), A->val.as<const double>(), x, y,
                         A->nrows, A->stream));
    SX_HIP(hipStreamSynchronize(A->stream));
  });
}

int sx_csr_destroy(sx_csr* A)
{
  delete A;
  return SX_OK;
}

int sx_grad_tangent_bytes(const sx_grad* g, int64_t* bytes)
{
  return api([&] {
    SX_REQUIRE(g && bytes, SX_ERR_ARG, "null argument");
    int64_t b = 0;
    for (auto& I : g->form->integrals)
      if (depends(*I, g->wrt) && I->has_tangent) b += int64_t(I->d_tangent.bytes);
    *bytes = b;
  });
}

int sx_grad_apply_f32(sx_grad* g, const float* dx, float* df)
{
  return api([&] { apply_gradient(g, dx, df, nullptr, nullptr, sxk::PREC_F32); });
}

int sx_grad_apply_timed(sx_grad* g, int vec_precision, const void* dx, void* df, int nrep, double* ms_total,
                        double* ms_kernels, int* launches_per_apply)
{
  return api([&] {
    SX_REQUIRE(g && g->form && nrep >= 1, SX_ERR_ARG, "bad argument");
    sx_form* f   = g->form;
    int      per = 0;
    apply_gradient(g, dx, df, nullptr, &per, vec_precision);  // count launches
    const int               nev = 2 * per * nrep;
    std::vector<hipEvent_t> ev(size_t(nev) + 2);
    for (auto& e : ev) SX_HIP(hipEventCreate(&e));
    SX_HIP(hipStreamSynchronize(f->stream));
    SX_HIP(hipEventRecord(ev[size_t(nev)], f->stream));
    for (int r = 0; r < nrep; r++) apply_gradient(g, dx, df, ev.data() + size_t(2 * per * r), nullptr, vec_precision);
    SX_HIP(hipEventRecord(ev[size_t(nev) + 1], f->stream));
    SX_HIP(hipEventSynchronize(ev[size_t(nev) + 1]));
    float total = 0.0f;
    SX_HIP(hipEventElapsedTime(&total, ev[size_t(nev)], ev[size_t(nev) + 1]));
    double kern = 0.0;
    for (int k = 0; k < per * nrep; k++) {
      float t = 0.0f;
      SX_HIP(hipEventElapsedTime(&t, ev[size_t(2 * k)], ev[size_t(2 * k + 1)]));
      kern += t;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (ms_total) *ms_total = total;
    if (ms_kernels) *ms_kernels = kern;
    if (launches_per_apply) *launches_per_apply = per;
  });
}

// ---------------------------------------------------------------------------
// vectors
// ---------------------------------------------------------------------------
int sx_vec_dot(const double* a, const double* b, int64_t n, double* result)
{
  return api([&] {
    SX_REQUIRE(a && b && result, SX_ERR_ARG, "null argument");
    ensure_device();
    DeviceBuffer part, scal;
    part.allocate(size_t(sxk::reduction_blocks()) * sizeof(double));
    scal.allocate(8 * sizeof(double));
    SX_HIP(sxk::vec_dot_dev(a, b, n, part.as<double>(), scal.as<double>(), 0, 0, nullptr));
    SX_HIP(hipMemcpy(result, scal.ptr, sizeof(double), hipMemcpyDeviceToHost));
  });
}

int sx_vec_axpy(double alpha, const double* x, double* y, int64_t n)
{
  return api([&] {
    SX_REQUIRE(x && y, SX_ERR_ARG, "null argument");
    SX_HIP(sxk::vec_axpy(alpha, x, y, n, nullptr));
    SX_HIP(hipDeviceSynchronize());
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
namespace {

double read_scalar(sx_form* f, int slot)
{
  double v = 0.0;
  SX_HIP(hipMemcpyAsync(&v, f->scal.as<double>() + slot, sizeof(double), hipMemcpyDeviceToHost, f->stream));
  SX_HIP(hipStreamSynchronize(f->stream));
  return v;
}

struct Workspace {
  DeviceBuffer r, z, d, Ad, diag, ess, t, part2;
  DeviceBuffer V, hb, gpart;  // GMRES basis, coefficients, partial dots
  DeviceBuffer Zb;            // flexible GMRES: the preconditioned directions
  DeviceBuffer gs;            // GMRES least-squares state (sxk::gmres_state_doubles)
};

// Checks every essential DOF against [0, ndof) before the list reaches the
// device, where the row kernels index with it unchecked, then uploads it.
void upload_ess(DeviceBuffer& dst, const int32_t* ess, int64_t ness, int64_t ndof)
{
  SX_REQUIRE(ness >= 0 && (ness == 0 || ess), SX_ERR_ARG, "null essential DOF list");
  for (int64_t k = 0; k < ness; k++)
    SX_REQUIRE(ess[k] >= 0 && ess[k] < ndof, SX_ERR_ARG,
               "essential DOF " + std::to_string(ess[k]) + " out of range [0, " + std::to_string(ndof) + ")");
  dst.upload(std::vector<int32_t>(ess, ess + ness));
}

// Jacobi-PCG on J_bc (rows and columns of ess eliminated) following the
// preconditioned CG of mfem::CGSolver, as configured by
// num/equation_solver.cpp:916-919, 1018-1021.
using CgHook = std::function<void(int it, const double* x, const double* r)>;

void compute_diag(sx_grad* g, const int32_t* ess_dev, int64_t ness, double* diag)
{
  sx_form*      f = g->form;
  const int64_t n = f->test->ndof();
  SX_HIP(sxk::vec_fill(diag, 0.0, n, f->stream));
  for (auto& I : f->integrals)
    if (depends(*I, g->wrt) && I->has_tangent) run_integral(f, *I, sxk::MODE_DIAG, nullptr, diag);
  for (auto& B : f->bintegrals)
    if (depends(*B, g->wrt) && B->has_tangent)
      run_boundary(f, *B, sxk::FMODE_DIAG, nullptr, diag, nullptr, sxk::PREC_F64);
  SX_HIP(sxk::set_rows(diag, ess_dev, ness, 1.0, f->stream));
}

// The device vectors of one Jacobi-CG solve: five of ndof doubles and the
// reduction_blocks() partials of (r, z).
struct CgVectors {
  double *r, *z, *d, *Ad, *diag, *part_rz;
};

CgVectors cg_vectors(Workspace& w, int64_t n)
{
  w.r.allocate(size_t(n) * 8);
  w.z.allocate(size_t(n) * 8);
  w.d.allocate(size_t(n) * 8);
  w.Ad.allocate(size_t(n) * 8);
  w.diag.allocate(size_t(n) * 8);
  w.part2.allocate(size_t(sxk::reduction_blocks()) * 8);
  return {w.r.as<double>(), w.z.as<double>(), w.d.as<double>(), w.Ad.as<double>(), w.diag.as<double>(),
          w.part2.as<double>()};
}

void cg_core(sx_grad* g, const int32_t* ess_dev, int64_t ness, const double* b, double* x, double rel,
             double abs_tol, int max_iter, const CgVectors& v, sx_solve_report* rep, const CgHook* hook)
{
  sx_form*      f = g->form;
  const int64_t n = f->test->ndof();
  hipStream_t   s = f->stream;
  double *r = v.r, *z = v.z, *d = v.d, *Ad = v.Ad, *diag = v.diag;
  double *part = f->part.as<double>(), *scal = f->scal.as<double>();

  // Jacobi diagonal with identity rows for essential DOFs
  compute_diag(g, ess_dev, ness, diag);

  SX_HIP(sxk::vec_fill(x, 0.0, n, s));
  SX_HIP(hipMemcpyAsync(r, b, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  SX_HIP(sxk::jacobi_dot(r, diag, z, n, part, scal, 0, s));  // scal[0] = (r, z)
  SX_HIP(hipMemcpyAsync(d, z, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  const double nom0 = read_scalar(f, 0);
  SX_REQUIRE(std::isfinite(nom0), SX_ERR_NOCONV, "CG: the initial residual is not finite");
  rep->initial_norm = std::sqrt(std::max(nom0, 0.0));
  const double goal = std::max(nom0 * rel * rel, abs_tol * abs_tol);
  rep->final_norm   = rep->initial_norm;
  rep->converged    = nom0 >= 0.0 && nom0 <= goal;
  rep->linear_iterations = 0;
  if (hook) (*hook)(0, x, r);
  // (r, D^-1 r) < 0: D is not positive definite; mfem's CGSolver returns
  // unconverged with x = 0
  if (rep->converged || nom0 < 0.0) return;

  auto apply_bc = [&](const double* in, double* outv) {
    apply_gradient(g, in, outv);
    SX_HIP(sxk::copy_rows(outv, in, ess_dev, ness, s));
  };
  // One iteration is five launches: the J.v (into the Ad that cg_step_x
  // zeroed), the essential rows, the partials of (d, Ad), then cg_step_x and
  // cg_step_d, which reduce the partials themselves.  The convergence test
  // runs on the device (scal[8] = stop flag, scal[9] = the iteration that set
  // it): once (r, z) <= goal the step kernels leave x, r, z, d and the scalars
  // alone, so the host issues iterations in chunks and reads the state once
  // per chunk.  x, the iteration count and the final norm are those of a test
  // after every iteration; the J.v launches queued behind the converged
  // iteration (fewer than a chunk) are wasted.  A hook sees every iteration,
  // so with one the chunks are single iterations.
  double* part_rz = v.part_rz;
  double* stop    = scal + 8;
  SX_HIP(sxk::vec_fill(stop, 0.0, 2, s));
  apply_bc(d, Ad);
  SX_HIP(sxk::dot_partial(d, Ad, n, part, s));
  int it = 1;
  while (it <= max_iter) {
    const int chunk = hook ? 1 : std::min(std::max(it / 4, 1), std::min(32, max_iter - it + 1));
    for (int j = 0; j < chunk; j++, it++) {
      SX_HIP(sxk::cg_step_x(x, r, z, Ad, d, diag, scal, part, part_rz, n, stop, s));
      if (hook) (*hook)(it, x, r);
      SX_HIP(sxk::cg_step_d(d, z, scal, part_rz, n, goal, it, stop, s));
      apply_gradient(g, d, Ad, nullptr, nullptr, sxk::PREC_F64, true, false, stop);
      SX_HIP(sxk::copy_rows(Ad, d, ess_dev, ness, s));
      SX_HIP(sxk::dot_partial(d, Ad, n, part, s));
    }
    double st[7];  // scal[3..9]
    SX_HIP(hipMemcpyAsync(st, scal + 3, sizeof(st), hipMemcpyDeviceToHost, s));
    SX_HIP(hipStreamSynchronize(s));
    rep->final_norm = std::sqrt(std::max(st[0], 0.0));
    SX_REQUIRE(std::isfinite(st[0]) && st[5] != 3.0, SX_ERR_NOCONV, "CG produced a non-finite residual");
    if (st[5] != 0.0) {
      rep->converged         = st[5] == 1.0;  // 2: the preconditioned residual norm went negative
      rep->linear_iterations = int(st[6]);
      return;
    }
  }
  rep->linear_iterations = max_iter;
}

void cg(sx_grad* g, const int32_t* ess_dev, int64_t ness, const double* b, double* x, double rel, double abs_tol,
        int max_iter, Workspace& w, sx_solve_report* rep, const CgHook* hook = nullptr)
{
  cg_core(g, ess_dev, ness, b, x, rel, abs_tol, max_iter, cg_vectors(w, g->form->test->ndof()), rep, hook);
}

// ---------------------------------------------------------------------------
// geometric multigrid
// ---------------------------------------------------------------------------
// y = A_l x: level 0 is J with the essential rows zeroed (x is zero there),
// the vertex levels apply their stencils
void mg_level_apply(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness, int l, const double* x,
                    double* y)
{
  hipStream_t s = g->form->stream;
  if (l == 0) {
    apply_gradient(g, x, y, nullptr, nullptr, sxk::PREC_F64, false, M->lo_active);
    SX_HIP(sxk::set_rows(y, ess, ness, 0.0, s));
  } else {
    SX_HIP(sxm::stencil_apply(M->lv[size_t(l)].S.as<double>(), M->lv[size_t(l)].g, x, y, s));
  }
}

double mg_read_dot(sx_multigrid* M, const double* a, const double* b, int64_t n, hipStream_t s)
{
  SX_HIP(sxk::vec_dot_dev(a, b, n, M->part.as<double>(), M->scal.as<double>(), 0, 0, s));
  double v = 0.0;
  SX_HIP(hipMemcpyAsync(&v, M->scal.as<double>(), sizeof(double), hipMemcpyDeviceToHost, s));
  SX_HIP(hipStreamSynchronize(s));
  return v;
}

// dense inverse of the coarsest operator: symmetrised, constrained rows set to
// identity, inverted by Gauss-Jordan elimination with partial pivoting
void mg_coarse_inverse(sx_multigrid* M, hipStream_t s)
{
  MgLevel&      c  = M->lv.back();
  const int64_t nv = int64_t(c.g.n[0]) * c.g.n[1] * c.g.n[2];
  const int     n  = M->ncoarse;
  std::vector<double>  S(size_t(nv) * 243);
  std::vector<uint8_t> mask(static_cast<size_t>(n));
  SX_HIP(hipMemcpyAsync(S.data(), c.S.ptr, S.size() * 8, hipMemcpyDeviceToHost, s));
  SX_HIP(hipMemcpyAsync(mask.data(), c.mask.ptr, mask.size(), hipMemcpyDeviceToHost, s));
  SX_HIP(hipStreamSynchronize(s));
  std::vector<double> A(size_t(n) * size_t(n), 0.0);
  for (int64_t v = 0; v < nv; v++) {
    const int ix = int(v % c.g.n[0]), iy = int((v / c.g.n[0]) % c.g.n[1]), iz = int(v / (int64_t(c.g.n[0]) * c.g.n[1]));
    for (int o = 0; o < 27; o++) {
      const int jx = ix + o % 3 - 1, jy = iy + (o / 3) % 3 - 1, jz = iz + o / 9 - 1;
      if (jx < 0 || jx >= c.g.n[0] || jy < 0 || jy >= c.g.n[1] || jz < 0 || jz >= c.g.n[2]) continue;
      const int64_t j = (int64_t(jz) * c.g.n[1] + jy) * c.g.n[0] + jx;
      for (int co = 0; co < 3; co++)
        for (int ci = 0; ci < 3; ci++)
          A[size_t(v * 3 + co) * size_t(n) + size_t(j * 3 + ci)] = S[size_t(o * 9 + co * 3 + ci) * size_t(nv) + size_t(v)];
    }
  }
  for (int i = 0; i < n; i++) {
    if (mask[size_t(i)]) A[size_t(i) * n + i] = 1.0;
    for (int j = 0; j < i; j++) {
      const double a = 0.5 * (A[size_t(i) * n + j] + A[size_t(j) * n + i]);
      A[size_t(i) * n + j] = a;
      A[size_t(j) * n + i] = a;
    }
  }
  // X = A^-1 by Gauss-Jordan elimination with partial pivoting (the tangent
  // of a compressed soft material can be indefinite; the inverse of a
  // symmetric matrix stays symmetric to rounding)
  double amax = 0.0;
  for (double a : A) amax = std::max(amax, std::fabs(a));
  std::vector<double> X(static_cast<size_t>(n) * static_cast<size_t>(n), 0.0);
  for (int i = 0; i < n; i++) X[size_t(i) * n + i] = 1.0;
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (std::fabs(A[size_t(r) * n + col]) > std::fabs(A[size_t(piv) * n + col])) piv = r;
    SX_REQUIRE(std::fabs(A[size_t(piv) * n + col]) > 1e-14 * amax && std::isfinite(A[size_t(piv) * n + col]),
               SX_ERR_NOCONV, "multigrid: the coarsest operator is singular");
    if (piv != col)
      for (int c = 0; c < n; c++) {
        std::swap(A[size_t(piv) * n + c], A[size_t(col) * n + c]);
        std::swap(X[size_t(piv) * n + c], X[size_t(col) * n + c]);
      }
    const double inv = 1.0 / A[size_t(col) * n + col];
    for (int c = 0; c < n; c++) {
      A[size_t(col) * n + c] *= inv;
      X[size_t(col) * n + c] *= inv;
    }
    for (int r = 0; r < n; r++) {
      const double f = r == col ? 0.0 : A[size_t(r) * n + col];
      if (f == 0.0) continue;
      for (int c = col; c < n; c++) A[size_t(r) * n + c] -= f * A[size_t(col) * n + c];
      for (int c = 0; c < n; c++) X[size_t(r) * n + c] -= f * X[size_t(col) * n + c];
    }
  }
  M->coarse_inv.upload(X);
}

// With galerkin = 0 the levels above the vertex grid are formed by the
// stencil triple product (sxm::stencil_rap); galerkin = 1 probes every level.
// The vertex-level operator can come element by element from the stored
// tangent (sxm::q1_galerkin) when every part of J that depends on the trial
// field is a domain integral with a GRAD_SYM tangent in FP64 or FP32 and the
// constraints are vertex-compatible (every vertex hat that is nonzero at a
// constrained DOF is constrained itself, so P_0 of the free vertex space
// vanishes on the constraints); otherwise it is probed.
bool mg_elementwise(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness)
{
  sx_form* f = g->form;
  if (M->galerkin != 0 || f->test->nc != 3) return false;
  for (auto& B : f->bintegrals)
    if (depends(*B, g->wrt) && B->has_tangent) return false;
  int count = 0;
  for (auto& I : f->integrals)
    if (depends(*I, g->wrt) && I->has_tangent) {
      if (I->layout != sxk::LAYOUT_GRAD_SYM || (I->tprec != sxk::PREC_F64 && I->tprec != sxk::PREC_F32)) return false;
      count++;
    }
  if (count == 0) return false;
  hipStream_t s = f->stream;
  SX_HIP(hipMemsetAsync(M->flag.ptr, 0, sizeof(int), s));
  SX_HIP(sxm::vertex_compatible(M->G, ess, ness, M->lv[1].mask.as<uint8_t>(), M->flag.as<int>(), s));
  int bad = 0;
  SX_HIP(hipMemcpyAsync(&bad, M->flag.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  SX_HIP(hipStreamSynchronize(s));
  return bad == 0;
}

// Builds the hierarchy for the current linearization of g with the
// essential DOFs ess (device list): masks, the fine Jacobi diagonal, the
// Galerkin operators by probing (27 vertex colours x 3 components per level;
// a colour's vertices are 3 apart, so each vertex sees exactly one of them in
// its 27-point neighbourhood), the Chebyshev intervals from power iteration
// on D^-1 A, and the coarsest inverse.
void mg_setup(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness)
{
  sx_form*    f = g->form;
  hipStream_t s = f->stream;
  SX_REQUIRE(f->test == M->space, SX_ERR_ARG, "the multigrid belongs to another space");
  const int L = int(M->lv.size());
  MgLevel&  l0 = M->lv[0];
  SX_HIP(hipMemsetAsync(l0.mask.ptr, 0, size_t(l0.n), s));
  SX_HIP(sxm::rows_mask(ess, ness, l0.mask.as<uint8_t>(), s));
  SX_HIP(sxm::vertex_mask(M->G, l0.mask.as<uint8_t>(), M->lv[1].mask.as<uint8_t>(), s));
  for (int l = 1; l + 1 < L; l++)
    SX_HIP(sxm::coarsen_mask(M->lv[size_t(l)].mask.as<uint8_t>(), M->lv[size_t(l)].g,
                             M->lv[size_t(l + 1)].mask.as<uint8_t>(), M->lv[size_t(l + 1)].g, s));
  compute_diag(g, ess, ness, M->diag.as<double>());
  SX_HIP(sxm::diag_inverse(M->diag.as<double>(), l0.n, ess, ness, l0.dinv.as<double>(), s));
  if (M->smoother_f32) {
    const size_t bytes = size_t(l0.n) * 4;
    for (DeviceBuffer* b : {&M->fb, &M->fx, &M->fr, &M->fd, &M->ft, &M->fdinv}) b->allocate(bytes);
    SX_HIP(sxm::to_f32(l0.dinv.as<double>(), M->fdinv.as<float>(), nullptr, l0.n, s));
  }
  const bool elementwise = mg_elementwise(M, g, ess, ness);
  // BF16 tangent copies for the level-0 applications (smoothing, power
  // iteration, probing) when asked for and every stored tangent is FP64 or FP32
  M->lo_active = false;
  if (M->smoother_bf16) {
    bool ok = true;
    for (auto& I : f->integrals)
      if (depends(*I, g->wrt) && I->has_tangent)
        ok = ok && (I->tprec == sxk::PREC_F64 || I->tprec == sxk::PREC_F32) &&
             (I->layout == sxk::LAYOUT_GRAD_SYM || I->layout == sxk::LAYOUT_GRAD_FULL) && f->test->nc == 3;
    if (ok) {
      for (auto& I : f->integrals)
        if (depends(*I, g->wrt) && I->has_tangent) {
          const int ds = sxk::tangent_ds(3, I->layout);
          I->d_tangent_lo.allocate(size_t(sxm::bf16_tangent_elems(M->G.p, ds, I->nslots)) * 2);
          SX_HIP(sxm::tangent_to_bf16(M->G.p, ds, I->tprec == sxk::PREC_F32, I->d_tangent.ptr,
                                      I->d_tangent_lo.as<uint16_t>(), I->nslots, s));
        }
      M->lo_active = true;
    }
  }
  for (int l = 0; l + 1 < L; l++) {
    MgLevel& F = M->lv[size_t(l)];
    MgLevel& C = M->lv[size_t(l + 1)];
    SX_HIP(sxk::vec_fill(C.S.as<double>(), 0.0, C.n * 81, s));
    if (l == 0 && elementwise) {
      const sxm::Grid3 gn{{M->G.nn[0], M->G.nn[1], M->G.nn[2]}};
      for (auto& I : f->integrals)
        if (depends(*I, g->wrt) && I->has_tangent)
          SX_HIP(sxm::q1_galerkin(M->G.p, I->tprec == sxk::PREC_F32, I->d_map.as<const int>(), I->d_tangent.ptr,
                                  I->nslots, element_tables(M->G.p + 1), gn, C.g, C.S.as<double>(), s));
      SX_HIP(sxm::stencil_mask(C.S.as<double>(), C.g, C.mask.as<uint8_t>(), s));
      SX_HIP(sxm::stencil_dinv(C.S.as<double>(), C.g, C.mask.as<uint8_t>(), C.dinv.as<double>(), s));
      M->element_setups++;
      continue;
    }
    if (l >= 1 && M->galerkin == 0) {
      // vertex levels: the triple product of the stencils directly
      SX_HIP(sxm::stencil_rap(F.S.as<double>(), F.g, C.S.as<double>(), C.g, s));
      SX_HIP(sxm::stencil_mask(C.S.as<double>(), C.g, C.mask.as<uint8_t>(), s));
      SX_HIP(sxm::stencil_dinv(C.S.as<double>(), C.g, C.mask.as<uint8_t>(), C.dinv.as<double>(), s));
      continue;
    }
    for (int cz = 0; cz < 3; cz++)
      for (int cy = 0; cy < 3; cy++)
        for (int cx = 0; cx < 3; cx++)
          for (int ci = 0; ci < 3; ci++) {
            SX_HIP(sxm::probe_fill(C.x.as<double>(), C.g, cx, cy, cz, ci, C.mask.as<uint8_t>(), s));
            SX_HIP(sxk::vec_fill(F.d.as<double>(), 0.0, F.n, s));
            if (l == 0) {
              SX_HIP(sxm::fe_prolong_add(M->G, C.x.as<double>(), F.d.as<double>(), s));
              SX_HIP(sxk::set_rows(F.d.as<double>(), ess, ness, 0.0, s));
            } else {
              SX_HIP(sxm::grid_prolong_add(C.x.as<double>(), C.g, F.d.as<double>(), F.g, F.mask.as<uint8_t>(), s));
            }
            mg_level_apply(M, g, ess, ness, l, F.d.as<double>(), F.t.as<double>());
            if (l == 0)
              SX_HIP(sxm::fe_restrict(M->G, F.t.as<double>(), C.r.as<double>(), C.mask.as<uint8_t>(), M->rt1.as<double>(),
                                     M->rt2.as<double>(), s));
            else
              SX_HIP(sxm::grid_restrict(F.t.as<double>(), F.g, C.r.as<double>(), C.g, C.mask.as<uint8_t>(), s));
            SX_HIP(sxm::probe_extract(C.r.as<double>(), C.g, cx, cy, cz, ci, C.mask.as<uint8_t>(), C.S.as<double>(),
                                      s));
          }
    SX_HIP(sxm::stencil_dinv(C.S.as<double>(), C.g, C.mask.as<uint8_t>(), C.dinv.as<double>(), s));
  }
  // Chebyshev intervals [upper / range, upper], upper = safety * the power
  // iteration's estimate of the largest eigenvalue of D^-1 A
  for (int l = 0; l + 1 < L; l++) {
    MgLevel& V = M->lv[size_t(l)];
    double*  v = V.d.as<double>();
    SX_HIP(sxm::seed_field(v, V.n, uint64_t(l + 1), V.mask.as<uint8_t>(), s));
    double nrm = std::sqrt(mg_read_dot(M, v, v, V.n, s));
    SX_REQUIRE(nrm > 0.0, SX_ERR_STATE, "multigrid: every DOF of a level is constrained");
    SX_HIP(sxm::scale(v, 1.0 / nrm, V.n, s));
    double lam = 0.0;
    for (int it = 0; it < M->power_iters; it++) {
      mg_level_apply(M, g, ess, ness, l, v, V.t.as<double>());
      SX_HIP(sxm::mul_dinv(V.dinv.as<double>(), V.t.as<double>(), v, V.n, s));
      lam = std::sqrt(mg_read_dot(M, v, v, V.n, s));
      SX_REQUIRE(lam > 0.0 && std::isfinite(lam), SX_ERR_NOCONV, "multigrid: power iteration broke down");
      SX_HIP(sxm::scale(v, 1.0 / lam, V.n, s));
    }
    V.lmax            = lam;
    const double up   = M->safety * lam, lo = up / M->range;
    const double th   = 0.5 * (up + lo), de = 0.5 * (up - lo), sigma = th / de;
    const int    deg  = l == 0 ? M->fine_degree : M->coarse_degree;
    double       rho  = 1.0 / sigma;
    V.theta           = th;
    V.c1.clear();
    V.c2.clear();
    for (int k = 2; k <= deg; k++) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      V.c1.push_back(rn * rho);
      V.c2.push_back(2.0 * rn / de);
      rho = rn;
    }
  }
  mg_coarse_inverse(M, s);
  M->setups++;
}

// Chebyshev smoothing of A_l x = b: from x = 0 (pre; leaves r = b - A x) or
// from the current x (post)
void mg_smooth(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness, int l, const double* b, double* x,
               bool pre)
{
  MgLevel&    V = M->lv[size_t(l)];
  hipStream_t s = g->form->stream;
  double *    r = V.r.as<double>(), *d = V.d.as<double>(), *t = V.t.as<double>();
  if (pre) {
    SX_HIP(sxm::cheb_init(b, V.dinv.as<double>(), V.theta, x, d, r, 0, 1, V.n, s));
  } else {
    mg_level_apply(M, g, ess, ness, l, x, t);
    SX_HIP(sxm::residual(b, t, r, V.n, s));
    SX_HIP(sxm::cheb_init(r, V.dinv.as<double>(), V.theta, x, d, r, 1, 0, V.n, s));
  }
  for (size_t k = 0; k < V.c1.size(); k++) {
    mg_level_apply(M, g, ess, ness, l, d, t);
    SX_HIP(sxm::cheb_step(r, t, d, V.dinv.as<double>(), V.c1[k], V.c2[k], x, V.n, s));
  }
  if (pre) {
    mg_level_apply(M, g, ess, ness, l, d, t);
    SX_HIP(sxm::sub_inplace(r, t, V.n, s));
  }
}

// y = A_0 x on FP32 vectors (x is zero on the essential rows)
void mg_level0_apply_f32(sx_multigrid* M, sx_grad* g, const float* x, float* y)
{
  apply_gradient(g, x, y, nullptr, nullptr, sxk::PREC_F32, false, M->lo_active);
  SX_HIP(sxm::apply_mask(y, M->lv[0].mask.as<uint8_t>(), M->lv[0].n, g->form->stream));
}

// mg_smooth on level 0 with FP32 vectors (b = fb, x = fx)
void mg_smooth0_f32(sx_multigrid* M, sx_grad* g, bool pre)
{
  MgLevel&      V = M->lv[0];
  hipStream_t   s = g->form->stream;
  const int64_t n = V.n;
  float *b = M->fb.as<float>(), *x = M->fx.as<float>(), *r = M->fr.as<float>(), *d = M->fd.as<float>(),
        *t = M->ft.as<float>(), *dinv = M->fdinv.as<float>();
  if (pre) {
    SX_HIP(sxm::cheb_init(b, dinv, V.theta, x, d, r, 0, 1, n, s));
  } else {
    mg_level0_apply_f32(M, g, x, t);
    SX_HIP(sxm::residual(b, t, r, n, s));
    SX_HIP(sxm::cheb_init(r, dinv, V.theta, x, d, r, 1, 0, n, s));
  }
  for (size_t k = 0; k < V.c1.size(); k++) {
    mg_level0_apply_f32(M, g, d, t);
    SX_HIP(sxm::cheb_step(r, t, d, dinv, V.c1[k], V.c2[k], x, n, s));
  }
  if (pre) {
    mg_level0_apply_f32(M, g, d, t);
    SX_HIP(sxm::sub_inplace(r, t, n, s));
  }
}

void mg_vcycle(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness, int l, const double* b, double* x);

// level 0 of the V-cycle on FP32 vectors: fx = V(fb)
void mg_vcycle0_f32(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness)
{
  hipStream_t s = g->form->stream;
  MgLevel&    C = M->lv[1];
  mg_smooth0_f32(M, g, true);
  SX_HIP(sxm::fe_restrict(M->G, M->fr.as<float>(), C.b.as<double>(), C.mask.as<uint8_t>(), M->rt1.as<double>(),
                          M->rt2.as<double>(), s));
  mg_vcycle(M, g, ess, ness, 1, C.b.as<double>(), C.x.as<double>());
  SX_HIP(sxm::fe_prolong_add(M->G, C.x.as<double>(), M->fx.as<float>(), s));
  SX_HIP(sxm::apply_mask(M->fx.as<float>(), M->lv[0].mask.as<uint8_t>(), M->lv[0].n, s));
  mg_smooth0_f32(M, g, false);
}

// x = V-cycle(b) on level l (x is overwritten)
void mg_vcycle(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness, int l, const double* b, double* x)
{
  hipStream_t s = g->form->stream;
  const int   L = int(M->lv.size());
  MgLevel&    V = M->lv[size_t(l)];
  if (l == L - 1) {
    SX_HIP(sxm::dense_mv(M->coarse_inv.as<double>(), M->ncoarse, b, x, s));
    return;
  }
  MgLevel& C = M->lv[size_t(l + 1)];
  mg_smooth(M, g, ess, ness, l, b, x, true);
  if (l == 0)
    SX_HIP(sxm::fe_restrict(M->G, V.r.as<double>(), C.b.as<double>(), C.mask.as<uint8_t>(), M->rt1.as<double>(),
                             M->rt2.as<double>(), s));
  else
    SX_HIP(sxm::grid_restrict(V.r.as<double>(), V.g, C.b.as<double>(), C.g, C.mask.as<uint8_t>(), s));
  mg_vcycle(M, g, ess, ness, l + 1, C.b.as<double>(), C.x.as<double>());
  if (l == 0) {
    SX_HIP(sxm::fe_prolong_add(M->G, C.x.as<double>(), x, s));
    SX_HIP(sxk::set_rows(x, ess, ness, 0.0, s));
  } else {
    SX_HIP(sxm::grid_prolong_add(C.x.as<double>(), C.g, x, V.g, V.mask.as<uint8_t>(), s));
  }
  mg_smooth(M, g, ess, ness, l, b, x, false);
}

// z = M r: the V-cycle on the free DOFs, the identity on the essential ones
void mg_apply(sx_multigrid* M, sx_grad* g, const int32_t* ess, int64_t ness, const double* r, double* z)
{
  hipStream_t s  = g->form->stream;
  MgLevel&    l0 = M->lv[0];
  if (M->smoother_f32 && M->fb.ptr) {
    SX_HIP(sxm::to_f32(r, M->fb.as<float>(), l0.mask.as<uint8_t>(), l0.n, s));
    mg_vcycle0_f32(M, g, ess, ness);
    SX_HIP(sxm::to_f64(M->fx.as<float>(), z, l0.n, s));
    SX_HIP(sxk::copy_rows(z, r, ess, ness, s));
    return;
  }
  SX_HIP(hipMemcpyAsync(l0.b.ptr, r, size_t(l0.n) * 8, hipMemcpyDeviceToDevice, s));
  SX_HIP(sxk::set_rows(l0.b.as<double>(), ess, ness, 0.0, s));
  mg_vcycle(M, g, ess, ness, 0, l0.b.as<double>(), z);
  SX_HIP(sxk::copy_rows(z, r, ess, ness, s));
}

// CG on J_bc x = b preconditioned by the V-cycle (mfem::CGSolver's
// recurrences and stopping rule (r, M r) <= max(rel^2 (r0, M r0), abs^2),
// num/equation_solver.cpp:916-919, with multigrid in place of the Jacobi
// smoother), the convergence test on the device as in cg
// With allow_reuse (the Newton loops, whose essential DOFs stay fixed for the
// whole call) a hierarchy serves up to M->reuse consecutive linearizations.
void mg_pcg(sx_grad* g, sx_multigrid* M, const int32_t* ess, int64_t ness, const double* b, double* x, double rel,
            double abs_tol, int max_iter, Workspace& w, sx_solve_report* rep, bool allow_reuse = false)
{
  sx_form*      f = g->form;
  const int64_t n = f->test->ndof();
  hipStream_t   s = f->stream;
  if (!allow_reuse || M->uses >= M->reuse) {
    mg_setup(M, g, ess, ness);
    M->uses = 0;
  }
  M->uses++;
  w.r.allocate(size_t(n) * 8);
  w.z.allocate(size_t(n) * 8);
  w.d.allocate(size_t(n) * 8);
  w.Ad.allocate(size_t(n) * 8);
  w.part2.allocate(size_t(sxk::reduction_blocks()) * 8);
  double *r = w.r.as<double>(), *z = w.z.as<double>(), *d = w.d.as<double>(), *Ad = w.Ad.as<double>();
  double *part = f->part.as<double>(), *scal = f->scal.as<double>(), *part_rz = w.part2.as<double>();
  double* stop = scal + 8;
  SX_HIP(sxk::vec_fill(x, 0.0, n, s));
  SX_HIP(hipMemcpyAsync(r, b, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  mg_apply(M, g, ess, ness, r, z);
  SX_HIP(sxk::vec_dot_dev(r, z, n, part, scal, 0, 0, s));  // scal[0] = (r, z)
  SX_HIP(hipMemcpyAsync(d, z, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  const double nom0 = read_scalar(f, 0);
  SX_REQUIRE(std::isfinite(nom0) && nom0 >= 0.0, SX_ERR_NOCONV, "multigrid-CG: the preconditioner is not positive");
  rep->initial_norm      = std::sqrt(nom0);
  const double goal      = std::max(nom0 * rel * rel, abs_tol * abs_tol);
  rep->final_norm        = rep->initial_norm;
  rep->converged         = nom0 <= goal;
  rep->linear_iterations = 0;
  if (rep->converged) return;
  SX_HIP(sxk::vec_fill(stop, 0.0, 2, s));
  apply_gradient(g, d, Ad);
  SX_HIP(sxk::copy_rows(Ad, d, ess, ness, s));
  SX_HIP(sxk::dot_partial(d, Ad, n, part, s));
  int it = 1;
  while (it <= max_iter) {
    // a V-cycle costs several J.v, so the host runs at most 4 iterations ahead
    const int chunk = std::min(std::max(it / 8, 1), std::min(4, max_iter - it + 1));
    for (int j = 0; j < chunk; j++, it++) {
      SX_HIP(sxm::pcg_step_xr(x, r, Ad, d, scal, part, sxk::reduction_blocks(), n, stop, s));
      mg_apply(M, g, ess, ness, r, z);
      SX_HIP(sxk::dot_partial(r, z, n, part_rz, s));
      SX_HIP(sxk::cg_step_d(d, z, scal, part_rz, n, goal, it, stop, s));
      apply_gradient(g, d, Ad, nullptr, nullptr, sxk::PREC_F64, true, false, stop);
      SX_HIP(sxk::copy_rows(Ad, d, ess, ness, s));
      SX_HIP(sxk::dot_partial(d, Ad, n, part, s));
    }
    double st[7];  // scal[3..9]
    SX_HIP(hipMemcpyAsync(st, scal + 3, sizeof(st), hipMemcpyDeviceToHost, s));
    SX_HIP(hipStreamSynchronize(s));
    rep->final_norm = std::sqrt(std::max(st[0], 0.0));
    SX_REQUIRE(std::isfinite(st[0]) && st[5] != 3.0, SX_ERR_NOCONV, "multigrid-CG produced a non-finite residual");
    if (st[5] != 0.0) {
      rep->converged         = st[5] == 1.0;  // 2: the V-cycle is not positive definite here
      rep->linear_iterations = int(st[6]);
      return;
    }
  }
  rep->linear_iterations = max_iter;
}

// Columns of a GMRES cycle between two reads of the device stop state: the
// Givens rotations, the residual estimate and the stop test run on the device
// after each column, the host looks every kGmresCheck columns (the J.v of
// columns queued past convergence return at once; their vector work is
// discarded with them).
constexpr int kGmresCheck = 4;

// Runs the Arnoldi columns of one GMRES cycle: column(j) leaves the new
// vector, not yet normalised, in V_{j+1} and the two Gram-Schmidt sums and
// |w|^2 in hb.  Returns the columns done; st receives the device state.
template <class Column>
int gmres_cycle(Column&& column, int m, int max_cols, double* V, int64_t n, double* hb, double* gs, double* st,
                hipStream_t s)
{
  const double* st_dev = gs + (sxk::gmres_state_doubles(m) - 8);
  for (int j = 0; j < m && j < max_cols;) {
    column(j);
    SX_HIP(sxk::gmres_column(hb, j, m, gs, s));
    SX_HIP(sxk::gmres_scale_inv(V + int64_t(j + 1) * n, n, m, gs, s));
    j++;
    if (j % kGmresCheck == 0 || j == m || j == max_cols) {
      SX_HIP(hipMemcpyAsync(st, st_dev, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
      SX_HIP(hipStreamSynchronize(s));
      if (st[3] != 0.0) break;
    }
  }
  return int(st[2]);
}

// Restarted GMRES(m) with left Jacobi preconditioning on J_bc x = b, x0 = 0,
// the mfem::GMRESSolver configured by num/equation_solver.cpp:916-919 (with
// HypreSmoother Jacobi, :1018-1021): Arnoldi on D^-1 J_bc, the residual norm
// ||D^-1 (b - J_bc x)|| goes down to max(rel ||D^-1 b||, abs).  Orthogonalises
// by classical Gram-Schmidt applied twice (two batched dot launches); the
// Givens rotations, the stop test and the triangular solve run on the device
// (gmres_cycle).  Each restart recomputes the true preconditioned residual,
// which also decides convergence.
void gmres(sx_grad* g, const int32_t* ess, int64_t ness, const double* b, double* x, double rel, double abs_tol,
           int max_iter, int m, Workspace& w, sx_solve_report* rep)
{
  sx_form*      f = g->form;
  const int64_t n = f->test->ndof();
  hipStream_t   s = f->stream;
  SX_REQUIRE(m >= 1 && m <= 1000, SX_ERR_ARG, "GMRES restart length must be in 1..1000");
  w.r.allocate(size_t(n) * 8);
  w.t.allocate(size_t(n) * 8);
  w.diag.allocate(size_t(n) * 8);
  w.V.allocate(size_t(m + 1) * size_t(n) * 8);
  const int hn = 2 * (m + 1) + 1;  // h (CGS pass 1), h (pass 2), |w|^2
  w.hb.allocate(size_t(std::max(hn, m + 1)) * 8);
  w.gpart.allocate(size_t(m + 1) * size_t(sxk::multi_dot_blocks()) * 8);
  w.gs.allocate(sxk::gmres_state_doubles(m) * 8);
  double *r = w.r.as<double>(), *t = w.t.as<double>(), *diag = w.diag.as<double>(), *V = w.V.as<double>();
  double *hb = w.hb.as<double>(), *part = w.gpart.as<double>(), *gs = w.gs.as<double>();
  const double* stop = gs + (sxk::gmres_state_doubles(m) - 8) + 3;
  auto          vec  = [&](int j) { return V + int64_t(j) * n; };
  auto          apply_bc = [&](const double* in, double* outv, const double* stp) {
    apply_gradient(g, in, outv, nullptr, nullptr, sxk::PREC_F64, false, false, stp);
    SX_HIP(sxk::copy_rows(outv, in, ess, ness, s));
  };
  auto column = [&](int j) {
    double* wv = vec(j + 1);
    apply_bc(vec(j), t, stop);
    SX_HIP(sxk::jacobi_apply(t, diag, wv, n, s));
    SX_HIP(sxk::multi_dot(V, n, j + 1, wv, n, part, hb, false, s));
    SX_HIP(sxk::multi_axpy(V, n, j + 1, hb, -1.0, wv, n, s));
    SX_HIP(sxk::multi_dot(V, n, j + 1, wv, n, part, hb + (m + 1), false, s));
    SX_HIP(sxk::multi_axpy(V, n, j + 1, hb + (m + 1), -1.0, wv, n, s));
    SX_HIP(sxk::multi_dot(wv, n, 1, wv, n, part, hb + 2 * (m + 1), false, s));
  };
  compute_diag(g, ess, ness, diag);
  SX_HIP(sxk::vec_fill(x, 0.0, n, s));
  double goal  = 0.0;
  int    it    = 0;
  bool   first = true;
  rep->converged = 0;
  for (;;) {
    // z = D^-1 (b - J_bc x)
    apply_bc(x, t, nullptr);
    SX_HIP(hipMemcpyAsync(r, b, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    SX_HIP(sxk::vec_axpy(-1.0, t, r, n, s));
    SX_HIP(sxk::jacobi_apply(r, diag, vec(0), n, s));
    SX_HIP(sxk::multi_dot(vec(0), n, 1, vec(0), n, part, hb, false, s));
    double beta2 = 0.0;
    SX_HIP(hipMemcpyAsync(&beta2, hb, 8, hipMemcpyDeviceToHost, s));
    SX_HIP(hipStreamSynchronize(s));
    const double beta = std::sqrt(std::max(beta2, 0.0));
    SX_REQUIRE(std::isfinite(beta), SX_ERR_NOCONV, "GMRES produced a non-finite residual");
    if (first) {
      rep->initial_norm = beta;
      goal              = std::max(rel * beta, abs_tol);
      first             = false;
    }
    rep->final_norm = beta;
    if (beta <= goal) {
      rep->converged = 1;
      break;
    }
    if (it >= max_iter) break;
    SX_HIP(sxk::scale_copy(1.0 / beta, vec(0), vec(0), n, s));
    SX_HIP(sxk::gmres_cycle_start(gs, m, beta, goal, s));
    double    st[8] = {};
    const int jend  = gmres_cycle(column, m, max_iter - it, V, n, hb, gs, st, s);
    SX_REQUIRE(st[3] != 3.0, SX_ERR_NOCONV, "GMRES produced a non-finite residual");
    it += jend;
    rep->final_norm = st[0];
    // y = H^-1 g (upper triangular), x += V y
    SX_HIP(sxk::gmres_solve_y(gs, m, hb, s));
    SX_HIP(sxk::multi_axpy(V, n, jend, hb, 1.0, x, n, s));
  }
  rep->linear_iterations = it;
}

// ---- learned preconditioner ----
void learned_apply(sx_learned* L, sx_grad* g, const double* diag, const int32_t* ess, int64_t ness, const double* r,
                   double* z)
{
  sx_form*      f  = g->form;
  hipStream_t   s  = f->stream;
  const int64_t n  = f->test->ndof();
  const int64_t nc = L->nv * 3;
  if (L->mg)
    mg_apply(L->mg, g, ess, ness, r, z);  // z = V-cycle(r); its hierarchy is set up by the caller
  else
    SX_HIP(sxk::jacobi_dot(r, diag, z, n, f->part.as<double>(), f->scal.as<double>(), 6, s));  // z = D^-1 r
  if (L->gain != 0.0) {
    SX_HIP(sxu::fe_restrict(L->G, r, L->rc.as<float>(), s));
    SX_HIP(hipMemsetAsync(L->ss.ptr, 0, sizeof(double), s));
    SX_HIP(sxu::sumsq(L->rc.as<float>(), nc, L->ss.as<double>(), s));
    SX_HIP(sxu::normalize(L->rc.as<float>(), L->ss.as<const double>(), nc, nc, L->xin.as<float>(), s));
    SX_HIP(hipStreamSynchronize(s));
    const int rc = sx_unet_forward(L->net, L->xin.as<const float>(), L->y.as<float>());
    if (rc != SX_OK) throw sx_failure(rc, g_last_error);
    SX_HIP(sxu::fe_prolong_add(L->G, L->y.as<const float>(), L->ss.as<const double>(), nc, z, s, L->gain));
  }
  SX_HIP(sxk::set_rows(z, ess, ness, 0.0, s));
}

// the hierarchy of a learned preconditioner's multigrid for this linear
// solve: rebuilt unless a Newton loop lets one serve several linearizations
void learned_mg_setup(sx_learned* L, sx_grad* g, const int32_t* ess, int64_t ness, bool allow_reuse)
{
  if (!L->mg) return;
  sx_multigrid* M = L->mg;
  if (!allow_reuse || M->uses >= M->reuse || M->set_g != g) {
    mg_setup(M, g, ess, ness);
    M->uses = 0;
  }
  M->uses++;
}

// a slot of the pair bank for a new pair: appended while the bank has room,
// then the ring position
int take_pair_slot(sx_learned* L)
{
  int k;
  if (int(L->pr.size()) < L->max_pairs) {
    k = int(L->pr.size());
    L->pr.push_back(std::make_unique<DeviceBuffer>());
    L->pt.push_back(std::make_unique<DeviceBuffer>());
    L->ok.push_back(0);
    L->pr[size_t(k)]->allocate(size_t(L->nv) * 3 * 4);
    L->pt[size_t(k)]->allocate(size_t(L->nv) * 3 * 4);
  } else {
    k       = L->next;
    L->next = (L->next + 1) % L->max_pairs;
  }
  L->ok[size_t(k)] = 0;
  return k;
}

// A random smooth field on the vertex grid: sum over levels m = 2, 3, 5, 9,
// ... (m - 1 doubling up to the grid) of uniform[-1, 1] values on an m^3
// lattice, weighted 1/m and interpolated trilinearly (vertex-aligned).
std::vector<float> multiscale_field(const sxu::FEGrid& G, std::mt19937_64& rng)
{
  const int64_t      nv = int64_t(G.nv[0]) * G.nv[1] * G.nv[2];
  std::vector<float> e(size_t(nv) * 3, 0.0f);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  const int nmax = std::max({G.nv[0], G.nv[1], G.nv[2]});
  for (int m = 2; m <= nmax; m = 2 * m - 1) {
    int md[3];
    for (int d = 0; d < 3; d++) md[d] = std::min(m, G.nv[d]);
    std::vector<double> c(size_t(md[0]) * md[1] * md[2] * 3);
    for (auto& v : c) v = U(rng) / m;
    for (int64_t k = 0; k < G.nv[2]; k++)
      for (int64_t j = 0; j < G.nv[1]; j++)
        for (int64_t i = 0; i < G.nv[0]; i++) {
          const int64_t idx[3] = {i, j, k};
          int           lo[3];
          double        fr[3];
          for (int d = 0; d < 3; d++) {
            const double x = G.nv[d] > 1 ? double(idx[d]) * (md[d] - 1) / double(G.nv[d] - 1) : 0.0;
            lo[d]          = std::min(int(x), std::max(md[d] - 2, 0));
            fr[d]          = md[d] > 1 ? x - lo[d] : 0.0;
          }
          for (int comp = 0; comp < 3; comp++) {
            double acc = 0.0;
            for (int b = 0; b < 8; b++) {
              const int    bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
              const double w  = (bx ? fr[0] : 1 - fr[0]) * (by ? fr[1] : 1 - fr[1]) * (bz ? fr[2] : 1 - fr[2]);
              if (w == 0.0) continue;
              const int64_t ci = (int64_t(lo[2] + bz) * md[1] + (lo[1] + by)) * md[0] + (lo[0] + bx);
              acc += w * c[size_t(ci * 3 + comp)];
            }
            e[size_t(((k * G.nv[1] + j) * G.nv[0] + i) * 3 + comp)] += float(acc);
          }
        }
  }
  return e;
}

// records a pair every `stride` iterations of a Jacobi-CG solve: the input
// P^T r_k and, as the target, the correction the next `window` iterations make,
// avg(x_{k+window}) - avg(x_k), with avg the hat-weighted vertex mean
// (P^T x) / (P^T 1).  A solve that converges sooner closes its open pairs with
// its final iterate; one that stops unconverged leaves them unusable.
struct PairRecorder {
  struct Open {
    int           slot, start;
    DeviceBuffer* x0;  // avg(x_start), FP64
  };
  sx_learned*                                L;
  hipStream_t                                s;
  std::vector<Open>                          open;
  std::vector<std::unique_ptr<DeviceBuffer>> pool;  // FP64 vertex vectors
  std::vector<DeviceBuffer*>                 spare;
  DeviceBuffer                               now;   // avg(x_it), FP64
  PairRecorder(sx_learned* learned, hipStream_t stream) : L(learned), s(stream)
  {
    now.allocate(size_t(L->nv) * 3 * 8);
  }
  DeviceBuffer* take()
  {
    if (spare.empty()) {
      pool.push_back(std::make_unique<DeviceBuffer>());
      pool.back()->allocate(size_t(L->nv) * 3 * 8);
      return pool.back().get();
    }
    DeviceBuffer* b = spare.back();
    spare.pop_back();
    return b;
  }
  void operator()(int it, const double* x, const double* r)
  {
    const bool closes = std::any_of(open.begin(), open.end(), [&](auto& o) { return it - o.start >= L->window; });
    const bool starts = it % L->stride == 0;
    if (!closes && !starts) return;
    SX_HIP(sxu::fe_average(L->G, x, now.as<double>(), s));
    for (auto o = open.begin(); o != open.end();) {
      if (it - o->start >= L->window) {
        close(*o);
        o = open.erase(o);
      } else {
        ++o;
      }
    }
    if (!starts) return;
    const int k = take_pair_slot(L);
    for (auto o = open.begin(); o != open.end();) {
      if (o->slot == k) {
        spare.push_back(o->x0);
        o = open.erase(o);
      } else {
        ++o;
      }
    }
    SX_HIP(sxu::fe_restrict(L->G, r, L->pr[size_t(k)]->as<float>(), s));
    DeviceBuffer* x0 = take();
    SX_HIP(hipMemcpyAsync(x0->ptr, now.ptr, size_t(L->nv) * 3 * 8, hipMemcpyDeviceToDevice, s));
    open.push_back({k, it, x0});
  }
  // pt[slot] = avg(x_now) - avg(x_start), formed in FP64
  void close(const Open& o)
  {
    SX_HIP(sxu::sub_f64(now.as<const double>(), o.x0->as<const double>(), L->pt[size_t(o.slot)]->as<float>(),
                        L->nv * 3, s));
    L->ok[size_t(o.slot)] = 1;
    spare.push_back(o.x0);
  }
  void finish(const double* x, bool converged)
  {
    if (converged && !open.empty()) {
      SX_HIP(sxu::fe_average(L->G, x, now.as<double>(), s));
      for (auto& o : open) close(o);
    }
    open.clear();
    SX_HIP(hipStreamSynchronize(s));
  }
};

// Flexible CG (FCG(1): beta = (z_k, q_{k-1}) / (d_{k-1}, q_{k-1})) with the
// learned preconditioner.  Stops on the Jacobi-preconditioned residual
// measure (r, D^-1 r) <= max(rel^2 (r0, D^-1 r0), abs^2), the rule of the
// Jacobi-CG it replaces; when that measure has not improved for `stall`
// iterations, Jacobi-CG finishes the solve from the current iterate.
void fcg(sx_grad* g, sx_learned* L, const int32_t* ess, int64_t ness, const double* b, double* x, double rel,
         double abs_tol, int max_iter, Workspace& w, sx_solve_report* rep, bool allow_reuse = false)
{
  sx_form*      f = g->form;
  const int64_t n = f->test->ndof();
  hipStream_t   s = f->stream;
  w.r.allocate(size_t(n) * 8);
  w.z.allocate(size_t(n) * 8);
  w.d.allocate(size_t(n) * 8);
  w.Ad.allocate(size_t(n) * 8);
  w.diag.allocate(size_t(n) * 8);
  w.t.allocate(size_t(n) * 8);
  double *r = w.r.as<double>(), *z = w.z.as<double>(), *d = w.d.as<double>(), *q = w.Ad.as<double>(),
         *diag = w.diag.as<double>(), *t = w.t.as<double>();
  double *part = f->part.as<double>(), *scal = f->scal.as<double>();
  auto    dot  = [&](const double* a, const double* c) {
    SX_HIP(sxk::vec_dot_dev(a, c, n, part, scal, 5, 0, s));
    return read_scalar(f, 5);
  };
  auto measure = [&]() {
    SX_HIP(sxk::jacobi_dot(r, diag, t, n, part, scal, 7, s));
    return read_scalar(f, 7);
  };
  auto apply_bc = [&](const double* in, double* outv) {
    apply_gradient(g, in, outv);
    SX_HIP(sxk::copy_rows(outv, in, ess, ness, s));
  };
  compute_diag(g, ess, ness, diag);
  learned_mg_setup(L, g, ess, ness, allow_reuse);
  SX_HIP(sxk::vec_fill(x, 0.0, n, s));
  SX_HIP(hipMemcpyAsync(r, b, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  const double mu0  = measure();
  const double goal = std::max(mu0 * rel * rel, abs_tol * abs_tol);
  rep->initial_norm = std::sqrt(std::max(mu0, 0.0));
  rep->final_norm   = rep->initial_norm;
  rep->converged    = mu0 <= goal;
  rep->linear_iterations = 0;
  if (rep->converged) return;
  learned_apply(L, g, diag, ess, ness, r, z);
  SX_HIP(hipMemcpyAsync(d, z, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
  apply_bc(d, q);
  double dq = dot(d, q), alpha = dot(d, r) / dq;
  double best = mu0, mu = mu0;
  int    since = 0, it = 0;
  bool   stalled = false;
  for (it = 1; it <= max_iter; it++) {
    SX_HIP(sxk::vec_axpy(alpha, d, x, n, s));
    SX_HIP(sxk::vec_axpy(-alpha, q, r, n, s));
    mu              = measure();
    rep->final_norm = std::sqrt(std::max(mu, 0.0));
    SX_REQUIRE(std::isfinite(mu), SX_ERR_NOCONV, "flexible CG produced a non-finite residual");
    if (mu <= goal) {
      rep->converged = 1;
      break;
    }
    if (mu < best) {
      best  = mu;
      since = 0;
    } else if (++since >= L->stall) {
      stalled = true;
      break;
    }
    learned_apply(L, g, diag, ess, ness, r, z);
    const double beta = dot(z, q) / dq;
    SX_HIP(sxk::vec_axpby(1.0, z, -beta, d, n, s));  // d = z - beta d
    apply_bc(d, q);
    dq    = dot(d, q);
    alpha = dot(d, r) / dq;
  }
  rep->linear_iterations = std::min(it, max_iter);
  if (stalled && it < max_iter) {
    // Jacobi-CG from the current iterate: solve A delta = r to the same goal
    L->fallbacks++;
    Workspace       w2;
    DeviceBuffer    rb, delta;
    sx_solve_report lr{};
    rb.allocate(size_t(n) * 8);
    delta.allocate(size_t(n) * 8);
    SX_HIP(hipMemcpyAsync(rb.ptr, r, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    cg(g, ess, ness, rb.as<double>(), delta.as<double>(), std::sqrt(goal / mu), abs_tol, max_iter - it, w2, &lr);
    SX_HIP(sxk::vec_axpy(1.0, delta.as<double>(), x, n, s));
    rep->linear_iterations += lr.linear_iterations;
    rep->converged  = lr.converged;
    rep->final_norm = lr.final_norm;
  }
}

// Flexible GMRES(m) (Saad's FGMRES: right preconditioning with a
// preconditioner that may change every iteration) on J_bc x = b, x0 = 0, with
// the multigrid V-cycle when mg is set, else the learned M(r) (Jacobi alone
// when L is null too).  Minimises ||b - J_bc x||; stops at max(rel ||b||,
// abs), checked on the true residual at each restart.  Stores the m + 1
// Arnoldi vectors and the m preconditioned directions; the least-squares part
// runs on the device (gmres_cycle).
void fgmres(sx_grad* g, sx_learned* L, const int32_t* ess, int64_t ness, const double* b, double* x, double rel,
            double abs_tol, int max_iter, int m, Workspace& w, sx_solve_report* rep, sx_multigrid* mg = nullptr,
            bool allow_reuse = false)
{
  if (mg) {
    if (!allow_reuse || mg->uses >= mg->reuse) {
      mg_setup(mg, g, ess, ness);
      mg->uses = 0;
    }
    mg->uses++;
  }
  sx_form*      f = g->form;
  const int64_t n = f->test->ndof();
  hipStream_t   s = f->stream;
  SX_REQUIRE(m >= 1 && m <= 1000, SX_ERR_ARG, "GMRES restart length must be in 1..1000");
  w.r.allocate(size_t(n) * 8);
  w.t.allocate(size_t(n) * 8);
  w.diag.allocate(size_t(n) * 8);
  w.V.allocate(size_t(m + 1) * size_t(n) * 8);
  w.Zb.allocate(size_t(m) * size_t(n) * 8);
  const int hn = 2 * (m + 1) + 1;
  w.hb.allocate(size_t(std::max(hn, m + 1)) * 8);
  w.gpart.allocate(size_t(m + 1) * size_t(sxk::multi_dot_blocks()) * 8);
  w.gs.allocate(sxk::gmres_state_doubles(m) * 8);
  double *t = w.t.as<double>(), *diag = w.diag.as<double>(), *V = w.V.as<double>();
  double *Z = w.Zb.as<double>(), *hb = w.hb.as<double>(), *part = w.gpart.as<double>(), *gs = w.gs.as<double>();
  const double* stop = gs + (sxk::gmres_state_doubles(m) - 8) + 3;
  auto          vec  = [&](int j) { return V + int64_t(j) * n; };
  auto          zvec = [&](int j) { return Z + int64_t(j) * n; };
  auto          apply_bc = [&](const double* in, double* outv, const double* stp) {
    apply_gradient(g, in, outv, nullptr, nullptr, sxk::PREC_F64, false, false, stp);
    SX_HIP(sxk::copy_rows(outv, in, ess, ness, s));
  };
  auto precondition = [&](const double* in, double* outv) {
    if (mg) {
      mg_apply(mg, g, ess, ness, in, outv);
    } else if (L) {
      learned_apply(L, g, diag, ess, ness, in, outv);
    } else {
      SX_HIP(sxk::jacobi_apply(in, diag, outv, n, s));
    }
  };
  auto column = [&](int j) {
    double* wv = vec(j + 1);
    precondition(vec(j), zvec(j));
    apply_bc(zvec(j), wv, stop);
    SX_HIP(sxk::multi_dot(V, n, j + 1, wv, n, part, hb, false, s));
    SX_HIP(sxk::multi_axpy(V, n, j + 1, hb, -1.0, wv, n, s));
    SX_HIP(sxk::multi_dot(V, n, j + 1, wv, n, part, hb + (m + 1), false, s));
    SX_HIP(sxk::multi_axpy(V, n, j + 1, hb + (m + 1), -1.0, wv, n, s));
    SX_HIP(sxk::multi_dot(wv, n, 1, wv, n, part, hb + 2 * (m + 1), false, s));
  };
  compute_diag(g, ess, ness, diag);
  if (L && !mg) learned_mg_setup(L, g, ess, ness, allow_reuse);
  SX_HIP(sxk::vec_fill(x, 0.0, n, s));
  double goal  = 0.0;
  int    it    = 0;
  bool   first = true;
  rep->converged = 0;
  for (;;) {
    apply_bc(x, t, nullptr);
    SX_HIP(hipMemcpyAsync(vec(0), b, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    SX_HIP(sxk::vec_axpy(-1.0, t, vec(0), n, s));
    SX_HIP(sxk::multi_dot(vec(0), n, 1, vec(0), n, part, hb, false, s));
    double beta2 = 0.0;
    SX_HIP(hipMemcpyAsync(&beta2, hb, 8, hipMemcpyDeviceToHost, s));
    SX_HIP(hipStreamSynchronize(s));
    const double beta = std::sqrt(std::max(beta2, 0.0));
    SX_REQUIRE(std::isfinite(beta), SX_ERR_NOCONV, "flexible GMRES produced a non-finite residual");
    if (first) {
      rep->initial_norm = beta;
      goal              = std::max(rel * beta, abs_tol);
      first             = false;
    }
    rep->final_norm = beta;
    if (beta <= goal) {
      rep->converged = 1;
      break;
    }
    if (it >= max_iter) break;
    SX_HIP(sxk::scale_copy(1.0 / beta, vec(0), vec(0), n, s));
    SX_HIP(sxk::gmres_cycle_start(gs, m, beta, goal, s));
    double    st[8] = {};
    const int jend  = gmres_cycle(column, m, max_iter - it, V, n, hb, gs, st, s);
    SX_REQUIRE(st[3] != 3.0, SX_ERR_NOCONV, "flexible GMRES produced a non-finite residual");
    it += jend;
    rep->final_norm = st[0];
    SX_HIP(sxk::gmres_solve_y(gs, m, hb, s));
    SX_HIP(sxk::multi_axpy(Z, n, jend, hb, 1.0, x, n, s));
  }
  rep->linear_iterations = it;
}

// the linear solve of the Newton loops: Jacobi-CG, Jacobi-CG recording
// training pairs, or flexible CG with the learned preconditioner
void linear_solve(sx_grad* g, const int32_t* ess, int64_t ness, const double* b, double* x, double rel,
                  double abs_tol, int max_iter, Workspace& w, sx_solve_report* rep)
{
  sx_form* f = g->form;
  if (f->learned && f->learn_mode == SX_PRECOND_LEARNED && f->linsolver == SX_LINSOLVER_GMRES) {
    fgmres(g, f->learned, ess, ness, b, x, rel, abs_tol, max_iter, f->gmres_restart, w, rep, nullptr, true);
  } else if (f->learned && f->learn_mode == SX_PRECOND_LEARNED) {
    fcg(g, f->learned, ess, ness, b, x, rel, abs_tol, max_iter, w, rep, true);
  } else if (f->learned && f->learn_mode == SX_PRECOND_RECORD) {
    PairRecorder rec(f->learned, f->stream);
    CgHook       hook = [&rec](int it, const double* xx, const double* rr) { rec(it, xx, rr); };
    cg(g, ess, ness, b, x, rel, abs_tol, max_iter, w, rep, &hook);
    rec.finish(x, rep->converged);
  } else if (f->linsolver == SX_LINSOLVER_GMRES && f->mg) {
    fgmres(g, nullptr, ess, ness, b, x, rel, abs_tol, max_iter, f->gmres_restart, w, rep, f->mg, true);
  } else if (f->linsolver == SX_LINSOLVER_GMRES) {
    gmres(g, ess, ness, b, x, rel, abs_tol, max_iter, f->gmres_restart, w, rep);
  } else if (f->mg) {
    mg_pcg(g, f->mg, ess, ness, b, x, rel, abs_tol, max_iter, w, rep, true);
  } else {
    cg(g, ess, ness, b, x, rel, abs_tol, max_iter, w, rep);
  }
}

double residual_norm(sx_form* f, int wrt, const double* x, double* r, const int32_t* ess_dev, int64_t ness,
                     sx_grad** gout)
{
  std::vector<const double*> in(f->trials.size(), x);
  sx_grad*                   g  = nullptr;
  int                        rc = sx_form_eval(f, 0.0, in.data(), r, gout ? wrt : -1, &g);
  if (rc != SX_OK) throw sx_failure(rc, g_last_error);
  if (gout) *gout = g;
  SX_HIP(sxk::set_rows(r, ess_dev, ness, 0.0, f->stream));
  SX_HIP(sxk::vec_dot_dev(r, r, f->test->ndof(), f->part.as<double>(), f->scal.as<double>(), 5, 0, f->stream));
  return std::sqrt(read_scalar(f, 5));
}

}  // namespace

extern "C" {

int sx_cg_solve(sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x, double rel_tol,
                double abs_tol, int max_iter, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "CG on a gradient before a linearization");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, g->form->test->ndof());
    cg(g, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, w, &rep);
    SX_HIP(hipStreamSynchronize(g->form->stream));
    if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "CG did not converge");
  });
  if (report) *report = rep;
  return rc;
}

// sx_cg_solve with its work vectors, (r, z) partials and essential DOF list
// carved out of the scratch array instead of allocated.  cg_core writes every
// vector before it reads it (x is filled, r copied from b, z and d formed,
// Ad written by the first J.v); the partials are cleared here.
int sx_cg_solve_noalloc(sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                        double rel_tol, double abs_tol, int max_iter, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "CG on a gradient before a linearization");
    SX_REQUIRE(ness >= 0 && (ness == 0 || ess_dofs), SX_ERR_ARG, "null essential DOF list");
    sx_form*      f = g->form;
    const int64_t n = f->test->ndof();
    for (int64_t k = 0; k < ness; k++)
      SX_REQUIRE(ess_dofs[k] >= 0 && ess_dofs[k] < n, SX_ERR_ARG,
                 "essential DOF " + std::to_string(ess_dofs[k]) + " out of range [0, " + std::to_string(n) + ")");
    auto          up    = [](uint64_t bytes) { return (bytes + 255) / 256 * 256; };
    const int64_t nblk  = sxk::reduction_blocks();
    const uint64_t e_b  = up(uint64_t(ness) * 4), p_b = up(uint64_t(nblk) * 8), v_b = up(uint64_t(n) * 8);
    const uint64_t need = e_b + p_b + 5 * v_b;
    SX_REQUIRE(need <= uint64_t(SX_NOALLOC_SCRATCH_BYTES), SX_ERR_ARG,
               "sx_cg_solve_noalloc needs " + std::to_string(need) + " bytes of scratch for " + std::to_string(n) +
                   " DOFs and " + std::to_string(ness) + " essential DOFs; the maximum is " +
                   std::to_string(uint64_t(SX_NOALLOC_SCRATCH_BYTES)));
    unsigned char* base = static_cast<unsigned char*>(noalloc_scratch());
    int32_t*       ess  = reinterpret_cast<int32_t*>(base);
    double*        part = reinterpret_cast<double*>(base + e_b);
    double*        vec  = reinterpret_cast<double*>(base + e_b + p_b);
    const int64_t  vs   = int64_t(v_b / 8);
    hipStream_t    s    = f->stream;
    if (ness > 0) SX_HIP(hipMemcpyAsync(ess, ess_dofs, size_t(ness) * 4, hipMemcpyHostToDevice, s));
    SX_HIP(sxk::vec_fill(part, 0.0, nblk, s));
    const CgVectors v{vec, vec + vs, vec + 2 * vs, vec + 3 * vs, vec + 4 * vs, part};
    cg_core(g, ess, ness, b, x, rel_tol, abs_tol, max_iter, v, &rep, nullptr);
    SX_HIP(hipStreamSynchronize(s));
    if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "CG did not converge");
  });
  if (report) *report = rep;
  return rc;
}

}  // extern "C"

extern "C" int sx_gmres_solve(sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                              double rel_tol, double abs_tol, int max_iter, int restart, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "GMRES on a gradient before a linearization");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, g->form->test->ndof());
    gmres(g, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, restart, w, &rep);
    SX_HIP(hipStreamSynchronize(g->form->stream));
    if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "GMRES did not converge");
  });
  if (report) *report = rep;
  return rc;
}

namespace {

// Newton (num/equation_solver.cpp:81-235): norm goal max(rel ||r0||, abs);
// Jacobi-CG on J_bc c = r, x <- x - c.  Every evaluation is a linearization,
// whose element kernel writes the residual and the tangent in one pass: the
// residual that decides convergence is the right-hand side of the next linear
// solve, so each Newton iteration runs one element pass (the converged check
// pays the tangent once per call).  With f->line_search > 0, the two-way
// backtracking of serac's NewtonSolver (num/equation_solver.cpp:138-185):
// when the full step does not lower the residual norm, halve it (residual
// evaluations only) up to line_search times; then search the opposite
// direction the same way; when both fail, take half the last step of the
// first direction.  The point accepted is then linearized.
void newton_core(sx_form* f, int wrt, const int32_t* ess, int64_t ness, double* x, const sx_solver_options& o,
                 Workspace& w, double* r, double* c, sx_solve_report* rep)
{
  const int64_t n    = f->test->ndof();
  sx_grad*      g    = nullptr;
  DeviceBuffer  x0;  // the iterate before the step (line search only)
  double        norm = residual_norm(f, wrt, x, r, ess, ness, &g);
  rep->initial_norm  = norm;
  const double goal  = std::max(o.rel_tol * norm, o.abs_tol);
  const double norm0 = norm;
  int          it    = 0;
  for (;; it++) {
    if (f->print_level > 0) {
      std::fprintf(stderr, "Newton iteration %3d : ||r|| = %13.6e, ||r||/||r_0|| = %13.6e\n", it, norm,
                   norm0 != 0.0 ? norm / norm0 : norm);
      std::fflush(stderr);
    }
    SX_REQUIRE(std::isfinite(norm), SX_ERR_NOCONV, "Newton residual is not finite");
    if (norm <= goal && it >= f->min_newton) {
      rep->converged = 1;
      break;
    }
    if (it >= o.max_iter) break;
    sx_solve_report lr{};
    linear_solve(g, ess, ness, r, c, o.lin_rel_tol, o.lin_abs_tol, o.lin_max_iter, w, &lr);
    rep->linear_iterations += lr.linear_iterations;
    if (f->print_level > 0) {
      std::fprintf(stderr, "  linear solve: %d iterations, %s, preconditioned norm %.6e -> %.6e\n",
                   lr.linear_iterations, lr.converged ? "converged" : "not converged", lr.initial_norm,
                   lr.final_norm);
      std::fflush(stderr);
    }
    const double prev = norm;
    if (f->line_search > 0) {
      x0.allocate(size_t(n) * 8);
      SX_HIP(hipMemcpyAsync(x0.ptr, x, size_t(n) * 8, hipMemcpyDeviceToDevice, f->stream));
    }
    SX_HIP(sxk::vec_axpy(-1.0, c, x, n, f->stream));
    norm = residual_norm(f, wrt, x, r, ess, ness, &g);  // residual and linearization at the new x
    const int ls_max = f->line_search;
    if (ls_max > 0 && !(norm < prev)) {
      // x = x0 + scale c, residual only (a NaN norm never counts as lower)
      auto step = [&](double scale) {
        SX_HIP(hipMemcpyAsync(x, x0.ptr, size_t(n) * 8, hipMemcpyDeviceToDevice, f->stream));
        SX_HIP(sxk::vec_axpy(scale, c, x, n, f->stream));
        return residual_norm(f, wrt, x, r, ess, ness, nullptr);
      };
      double scale = 1.0;
      int    ls    = 0;
      for (; !(norm < prev) && ls < ls_max; ++ls, ++rep->line_search_steps) {
        scale *= 0.5;
        norm = step(-scale);
      }
      if (ls == ls_max && !(norm < prev)) {
        scale = 1.0;
        norm  = step(scale);
        ls    = 0;
        for (; !(norm < prev) && ls < ls_max; ++ls, ++rep->line_search_steps) {
          scale *= 0.5;
          norm = step(scale);
        }
        if (ls == ls_max && !(norm < prev)) {
          ++rep->line_search_steps;
          scale *= 0.5;
          norm = step(-scale);
        }
      }
      norm = residual_norm(f, wrt, x, r, ess, ness, &g);  // the linearization at the accepted point
    }
  }
  rep->newton_iterations = it;
  rep->final_norm        = norm;
}

}  // namespace

extern "C" {

int sx_newton_solve(sx_form* f, int wrt, const int32_t* ess_dofs, int64_t ness, double* x, double rel_tol,
                    double abs_tol, int max_iter, double lin_rel_tol, double lin_abs_tol, int lin_max_iter,
                    sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(f && x, SX_ERR_ARG, "null argument");
    if (f->mg) f->mg->uses = INT32_MAX;  // a fresh hierarchy for this call's essential DOFs
    if (f->learned && f->learned->mg) f->learned->mg->uses = INT32_MAX;
    SX_REQUIRE(wrt >= 0 && wrt < int(f->trials.size()), SX_ERR_ARG, "wrt out of range");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    const int64_t n = f->test->ndof();
    Workspace     w;
    upload_ess(w.ess, ess_dofs, ness, f->test->ndof());
    DeviceBuffer rbuf, cbuf;
    rbuf.allocate(size_t(n) * 8);
    cbuf.allocate(size_t(n) * 8);
    const sx_solver_options o{rel_tol, abs_tol, max_iter, lin_rel_tol, lin_abs_tol, lin_max_iter};
    newton_core(f, wrt, w.ess.as<const int32_t>(), ness, x, o, w, rbuf.as<double>(), cbuf.as<double>(), &rep);
    SX_HIP(hipStreamSynchronize(f->stream));
    if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "Newton did not converge");
  });
  if (report) *report = rep;
  return rc;
}

int sx_quasistatic_solve(sx_form* f, int wrt, const int32_t* ess_dofs, int64_t ness, const double* ess_final,
                         int nsteps, int warm_start, const sx_solver_options* opts, double* x,
                         sx_solve_report* reports)
{
  return api([&] {
    SX_REQUIRE(f && x && opts && nsteps >= 1, SX_ERR_ARG, "bad argument");
    if (f->mg) f->mg->uses = INT32_MAX;  // a fresh hierarchy for this call's essential DOFs
    if (f->learned && f->learned->mg) f->learned->mg->uses = INT32_MAX;
    SX_REQUIRE(wrt >= 0 && wrt < int(f->trials.size()), SX_ERR_ARG, "wrt out of range");
    SX_REQUIRE(ness == 0 || (ess_dofs && ess_final), SX_ERR_ARG, "null essential DOF data");
    const int64_t n = f->test->ndof();
    hipStream_t   s = f->stream;
    Workspace     w;
    upload_ess(w.ess, ess_dofs, ness, f->test->ndof());
    const int32_t* ess = w.ess.as<const int32_t>();
    DeviceBuffer   rbuf, cbuf, dbc, jdbc, vals;
    rbuf.allocate(size_t(n) * 8);
    cbuf.allocate(size_t(n) * 8);
    dbc.allocate(size_t(n) * 8);
    jdbc.allocate(size_t(n) * 8);
    vals.allocate(size_t(std::max<int64_t>(ness, 1)) * 8);
    double *r = rbuf.as<double>(), *c = cbuf.as<double>();

    // essential values at t = 0 are those x holds
    std::vector<double> start(static_cast<size_t>(ness)), prev(static_cast<size_t>(ness)), target(static_cast<size_t>(ness)),
        inc(static_cast<size_t>(ness));
    {
      std::vector<double> xh(static_cast<size_t>(n));
      SX_HIP(hipMemcpy(xh.data(), x, size_t(n) * 8, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < ness; i++) start[size_t(i)] = xh[size_t(ess_dofs[i])];
    }
    prev = start;
    bool has_qdata = false;
    for (auto& I : f->integrals) has_qdata = has_qdata || I->qdata;
    for (int k = 1; k <= nsteps; k++) {
      const double t = double(k) / double(nsteps);
      for (int64_t i = 0; i < ness; i++) {
        target[size_t(i)] = start[size_t(i)] + t * (ess_final[i] - start[size_t(i)]);
        inc[size_t(i)]    = target[size_t(i)] - prev[size_t(i)];
      }
      sx_solve_report rep{};
      if (warm_start) {
        // du_bc: the boundary increment; rhs = -(R(u) + J du_bc), free rows only
        SX_HIP(hipMemcpyAsync(vals.ptr, inc.data(), size_t(ness) * 8, hipMemcpyHostToDevice, s));
        SX_HIP(sxk::vec_fill(dbc.as<double>(), 0.0, n, s));
        SX_HIP(sxk::set_rows_values(dbc.as<double>(), ess, vals.as<const double>(), ness, s));
        sx_grad* g = nullptr;
        residual_norm(f, wrt, x, r, ess, ness, &g);  // linearize at the converged state
        apply_gradient(g, dbc.as<double>(), jdbc.as<double>());
        SX_HIP(sxk::vec_axpby(-1.0, jdbc.as<double>(), -1.0, r, n, s));  // r = -(R + J du_bc)
        SX_HIP(sxk::set_rows(r, ess, ness, 0.0, s));
        sx_solve_report lr{};
        linear_solve(g, ess, ness, r, c, opts->lin_rel_tol, opts->lin_abs_tol, opts->lin_max_iter, w, &lr);
        rep.linear_iterations += lr.linear_iterations;
        SX_HIP(sxk::vec_axpy(1.0, c, x, n, s));
        SX_HIP(sxk::vec_axpy(1.0, dbc.as<double>(), x, n, s));
      } else {
        SX_HIP(hipMemcpyAsync(vals.ptr, target.data(), size_t(ness) * 8, hipMemcpyHostToDevice, s));
        SX_HIP(sxk::set_rows_values(x, ess, vals.as<const double>(), ness, s));
      }
      const int lin_ws = rep.linear_iterations;
      newton_core(f, wrt, ess, ness, x, *opts, w, r, c, &rep);
      rep.linear_iterations += lin_ws;
      SX_HIP(hipStreamSynchronize(s));
      if (reports) reports[k - 1] = rep;
      if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "Newton did not converge in load step " + std::to_string(k));
      // after equilibrium, evaluate the residual once more with the state
      // buffers committed (phy/solid_mechanics.hpp:1278-1287)
      if (has_qdata) {
        const int saved = f->update_qdata;
        f->update_qdata = 1;
        residual_norm(f, wrt, x, r, ess, ness, nullptr);
        f->update_qdata = saved;
      }
      prev = target;
    }
  });
}

// ---------------------------------------------------------------------------
// geometric multigrid preconditioner
// ---------------------------------------------------------------------------
int sx_mg_create(const sx_space* sp, sx_multigrid** out)
{
  return api([&] {
    SX_REQUIRE(sp && out, SX_ERR_ARG, "null argument");
    const auto& m = sp->mesh->m;
    SX_REQUIRE(m.structured, SX_ERR_UNSUPPORTED, "multigrid needs a structured (cuboid) mesh");
    SX_REQUIRE(sp->native, SX_ERR_UNSUPPORTED, "multigrid needs the library's own (lexicographic) node numbering");
    SX_REQUIRE(sp->nc == 3 && sp->ordering == SX_ORDERING_BYVDIM, SX_ERR_UNSUPPORTED,
               "multigrid acts on 3-component byVDIM spaces");
    auto M   = std::make_unique<sx_multigrid>();
    M->space = sp;
    M->G.p   = sp->order;
    for (int d = 0; d < 3; d++) {
      M->G.nv[d] = m.n[d] + 1;
      M->G.nn[d] = sp->order * m.n[d] + 1;
    }
    const auto xi = sxh::lobatto_points(sp->order + 1);
    for (int k = 0; k <= sp->order; k++) M->G.xi[k] = xi[size_t(k)];
    // level 0: the nodal grid; level 1: the vertex grid; then halve the
    // intervals while every direction has an even count and the level holds
    // more than 200 DOFs
    std::vector<sxm::Grid3> grids(2);
    for (int d = 0; d < 3; d++) grids[1].n[d] = M->G.nv[d];
    for (;;) {
      const sxm::Grid3 c  = grids.back();
      const int64_t    nd = int64_t(c.n[0]) * c.n[1] * c.n[2] * 3;
      bool             ok = nd > 200;
      for (int d = 0; d < 3; d++) ok = ok && (c.n[d] - 1) % 2 == 0 && c.n[d] - 1 >= 2;
      if (!ok) break;
      sxm::Grid3 nc{};
      for (int d = 0; d < 3; d++) nc.n[d] = (c.n[d] - 1) / 2 + 1;
      grids.push_back(nc);
    }
    M->lv      = std::vector<MgLevel>(grids.size());
    M->lv[0].n = sp->ndof();
    for (size_t l = 1; l < M->lv.size(); l++) {
      M->lv[l].g = grids[l];
      M->lv[l].n = int64_t(grids[l].n[0]) * grids[l].n[1] * grids[l].n[2] * 3;
    }
    M->ncoarse = int(M->lv.back().n);
    SX_REQUIRE(M->ncoarse <= 1536, SX_ERR_UNSUPPORTED,
               "multigrid: the coarsest vertex grid is too large for its dense solve (mesh sizes need factors of 2)");
    for (size_t l = 0; l < M->lv.size(); l++) {
      MgLevel& V = M->lv[l];
      V.dinv.allocate(size_t(V.n) * 8);
      V.mask.allocate(size_t(V.n));
      V.r.allocate(size_t(V.n) * 8);
      V.d.allocate(size_t(V.n) * 8);
      V.t.allocate(size_t(V.n) * 8);
      V.b.allocate(size_t(V.n) * 8);
      if (l > 0) {
        V.x.allocate(size_t(V.n) * 8);
        V.S.allocate(size_t(V.n) * 81 * 8);
      }
    }
    M->diag.allocate(size_t(M->lv[0].n) * 8);
    M->rt1.allocate(size_t(sxm::fe_restrict_scratch(M->G, 0)) * 8);
    M->rt2.allocate(size_t(sxm::fe_restrict_scratch(M->G, 1)) * 8);
    M->part.allocate(size_t(sxk::reduction_blocks()) * 8);
    M->scal.allocate(16 * 8);
    M->flag.allocate(sizeof(int));
    *out = M.release();
  });
}

int sx_mg_configure(sx_multigrid* M, int fine_degree, int coarse_degree, double smoothing_range, int power_iterations)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    SX_REQUIRE(fine_degree >= 1 && fine_degree <= 16 && coarse_degree >= 1 && coarse_degree <= 16, SX_ERR_ARG,
               "Chebyshev degrees must be in 1..16");
    SX_REQUIRE(smoothing_range > 1.0, SX_ERR_ARG, "the smoothing range must exceed 1");
    SX_REQUIRE(power_iterations >= 1 && power_iterations <= 1000, SX_ERR_ARG, "power iterations must be in 1..1000");
    M->fine_degree   = fine_degree;
    M->coarse_degree = coarse_degree;
    M->range         = smoothing_range;
    M->power_iters   = power_iterations;
  });
}

int sx_mg_set_galerkin(sx_multigrid* M, int mode)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    SX_REQUIRE(mode == 0 || mode == 1, SX_ERR_ARG, "Galerkin mode is 0 (element-wise where possible) or 1 (probing)");
    M->galerkin = mode;
  });
}

int sx_mg_set_smoother_precision(sx_multigrid* M, int prec)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    SX_REQUIRE(prec == SX_PREC_F64 || prec == SX_PREC_BF16, SX_ERR_ARG,
               "smoother precision is SX_PREC_F64 (the form's tangent) or SX_PREC_BF16");
    M->smoother_bf16 = prec == SX_PREC_BF16;
  });
}

int sx_mg_set_smoother_vectors(sx_multigrid* M, int prec)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    SX_REQUIRE(prec == SX_PREC_F64 || prec == SX_PREC_F32, SX_ERR_ARG,
               "smoother vectors are SX_PREC_F64 or SX_PREC_F32");
    M->smoother_f32 = prec == SX_PREC_F32;
    if (!M->smoother_f32)
      for (DeviceBuffer* b : {&M->fb, &M->fx, &M->fr, &M->fd, &M->ft, &M->fdinv}) b->release();
  });
}

int sx_mg_set_reuse(sx_multigrid* M, int linearizations)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    SX_REQUIRE(linearizations >= 1, SX_ERR_ARG, "a hierarchy serves at least one linearization");
    M->reuse = linearizations;
  });
}

int sx_mg_info(const sx_multigrid* M, int* nlevels, int64_t* setups)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    if (nlevels) *nlevels = int(M->lv.size());
    if (setups) *setups = M->setups;
  });
}

int sx_mg_element_setups(const sx_multigrid* M, int64_t* n)
{
  return api([&] {
    SX_REQUIRE(M && n, SX_ERR_ARG, "null argument");
    *n = M->element_setups;
  });
}

int sx_mg_level(const sx_multigrid* M, int level, int64_t* ndof, int* dims, double* lmax)
{
  return api([&] {
    SX_REQUIRE(M, SX_ERR_ARG, "null multigrid");
    SX_REQUIRE(level >= 0 && level < int(M->lv.size()), SX_ERR_ARG, "level out of range");
    const MgLevel& V = M->lv[size_t(level)];
    if (ndof) *ndof = V.n;
    if (dims)
      for (int d = 0; d < 3; d++) dims[d] = level == 0 ? M->G.nn[d] : V.g.n[d];
    if (lmax) *lmax = V.lmax;
  });
}

int sx_mg_setup(sx_multigrid* M, sx_grad* g, const int32_t* ess_dofs, int64_t ness)
{
  return api([&] {
    SX_REQUIRE(M && g && g->form, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "multigrid setup on a gradient before a linearization");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    for (int64_t k = 0; k < ness; k++)
      SX_REQUIRE(ess_dofs[k] >= 0 && ess_dofs[k] < M->lv[0].n, SX_ERR_ARG, "essential DOF out of range");
    upload_ess(M->ess, ess_dofs, ness, M->lv[0].n);
    M->ness  = ness;
    M->set_g = g;
    mg_setup(M, g, M->ess.as<const int32_t>(), ness);
    SX_HIP(hipStreamSynchronize(g->form->stream));
  });
}

int sx_mg_apply(sx_multigrid* M, const double* r, double* z)
{
  return api([&] {
    SX_REQUIRE(M && r && z, SX_ERR_ARG, "null argument");
    SX_REQUIRE(M->set_g && M->set_g->valid, SX_ERR_STATE, "multigrid applied before sx_mg_setup");
    mg_apply(M, M->set_g, M->ess.as<const int32_t>(), M->ness, r, z);
    SX_HIP(hipStreamSynchronize(M->set_g->form->stream));
  });
}

int sx_mg_level_apply(sx_multigrid* M, int level, const double* x, double* y)
{
  return api([&] {
    SX_REQUIRE(M && x && y, SX_ERR_ARG, "null argument");
    SX_REQUIRE(M->set_g && M->set_g->valid, SX_ERR_STATE, "multigrid applied before sx_mg_setup");
    SX_REQUIRE(level >= 0 && level < int(M->lv.size()), SX_ERR_ARG, "level out of range");
    mg_level_apply(M, M->set_g, M->ess.as<const int32_t>(), M->ness, level, x, y);
    SX_HIP(hipStreamSynchronize(M->set_g->form->stream));
  });
}

int sx_mg_pcg_solve(sx_grad* g, sx_multigrid* M, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                    double rel_tol, double abs_tol, int max_iter, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && M && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "multigrid-CG on a gradient before a linearization");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    for (int64_t k = 0; k < ness; k++)
      SX_REQUIRE(ess_dofs[k] >= 0 && ess_dofs[k] < M->lv[0].n, SX_ERR_ARG, "essential DOF out of range");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, M->lv[0].n);
    mg_pcg(g, M, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, w, &rep);
    SX_HIP(hipStreamSynchronize(g->form->stream));
    if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "multigrid-CG did not converge");
  });
  if (report) *report = rep;
  return rc;
}

int sx_mg_gmres_solve(sx_grad* g, sx_multigrid* M, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                      double rel_tol, double abs_tol, int max_iter, int restart, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && M && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "multigrid-GMRES on a gradient before a linearization");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    SX_REQUIRE(M->space == g->form->test, SX_ERR_ARG, "the multigrid belongs to another space");
    for (int64_t k = 0; k < ness; k++)
      SX_REQUIRE(ess_dofs[k] >= 0 && ess_dofs[k] < M->lv[0].n, SX_ERR_ARG, "essential DOF out of range");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, M->lv[0].n);
    fgmres(g, nullptr, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, restart, w, &rep, M);
    SX_HIP(hipStreamSynchronize(g->form->stream));
    if (!rep.converged) throw sx_failure(SX_ERR_NOCONV, "multigrid-GMRES did not converge");
  });
  if (report) *report = rep;
  return rc;
}

int sx_form_set_shape_displacement(sx_form* f, const double* p)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    SX_REQUIRE(!p || f->test->nc == 3, SX_ERR_UNSUPPORTED,
               "a shape displacement needs a 3-component test space (it is a field of that space)");
    for (const sx_space* tr : f->trials)
      SX_REQUIRE(!p || tr->nnodes == f->test->nnodes, SX_ERR_UNSUPPORTED,
                 "a shape displacement needs trial spaces on the test space's nodes");
    f->shape = p;
    for (auto& g : f->grads) g->valid = false;  // linearizations at another shape are stale
    if (f->shape_grad) f->shape_grad->valid = false;
  });
}

int sx_form_set_multigrid(sx_form* f, sx_multigrid* M)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    SX_REQUIRE(!M || M->space == f->test, SX_ERR_ARG, "the multigrid belongs to another space");
    f->mg = M;
  });
}

int sx_mg_destroy(sx_multigrid* M)
{
  return api([&] { delete M; });
}

// ---------------------------------------------------------------------------
// learned coarse-grid preconditioner
// ---------------------------------------------------------------------------
int sx_learned_create(const sx_space* sp, uint64_t seed, sx_learned** out)
{
  return api([&] {
    SX_REQUIRE(sp && out, SX_ERR_ARG, "null argument");
    const auto& m = sp->mesh->m;
    SX_REQUIRE(m.structured, SX_ERR_UNSUPPORTED, "the learned preconditioner needs a structured (cuboid) mesh");
    SX_REQUIRE(sp->native, SX_ERR_UNSUPPORTED,
               "the learned preconditioner needs the library's own (lexicographic) node numbering");
    SX_REQUIRE(sp->nc == 3 && sp->ordering == SX_ORDERING_BYVDIM, SX_ERR_UNSUPPORTED,
               "the learned preconditioner acts on 3-component byVDIM spaces");
    auto L   = std::make_unique<sx_learned>();
    L->space = sp;
    L->G.p   = sp->order;
    for (int d = 0; d < 3; d++) {
      L->G.nv[d] = m.n[d] + 1;
      L->G.nn[d] = sp->order * m.n[d] + 1;
    }
    const auto xi = sxh::lobatto_points(sp->order + 1);
    for (int k = 0; k <= sp->order; k++) L->G.xi[k] = xi[size_t(k)];
    L->nv = int64_t(L->G.nv[0]) * L->G.nv[1] * L->G.nv[2];
    const int rc = sx_unet_create(L->G.nv[0], L->G.nv[1], L->G.nv[2], seed, &L->net);
    if (rc != SX_OK) throw sx_failure(rc, g_last_error);
    L->rc.allocate(size_t(L->nv) * 3 * 4);
    L->xin.allocate(size_t(L->nv) * 3 * 4);
    L->y.allocate(size_t(L->nv) * 3 * 4);
    L->ss.allocate(sizeof(double));
    *out = L.release();
  });
}

int sx_learned_configure(sx_learned* L, int stride, int window, int max_pairs, int stall_window)
{
  return api([&] {
    SX_REQUIRE(L && stride >= 1 && window >= 1 && max_pairs >= 1 && stall_window >= 1, SX_ERR_ARG, "bad argument");
    SX_REQUIRE(int(L->pr.size()) <= max_pairs, SX_ERR_STATE, "the bank already holds more pairs");
    L->stride    = stride;
    L->window    = window;
    L->max_pairs = max_pairs;
    L->stall     = stall_window;
  });
}

int sx_learned_info(const sx_learned* L, int* npairs, int64_t* fallbacks, sx_unet** net)
{
  return api([&] {
    SX_REQUIRE(L, SX_ERR_ARG, "null argument");
    if (npairs) *npairs = int(std::count(L->ok.begin(), L->ok.end(), char(1)));
    if (fallbacks) *fallbacks = L->fallbacks;
    if (net) *net = L->net;
  });
}

int sx_learned_record_cg(sx_learned* L, sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                         double rel_tol, double abs_tol, int max_iter, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(L && g && g->form && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "solve on a gradient before a linearization");
    SX_REQUIRE(g->form->test == L->space, SX_ERR_ARG, "the preconditioner belongs to another space");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, g->form->test->ndof());
    PairRecorder rec(L, g->form->stream);
    CgHook       hook = [&rec](int it, const double* xx, const double* rr) { rec(it, xx, rr); };
    cg(g, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, w, &rep, &hook);
    rec.finish(x, rep.converged);
    SX_REQUIRE(rep.converged, SX_ERR_NOCONV, "CG did not converge");
  });
  if (report) *report = rep;
  return rc;
}

int sx_learned_get_pair(const sx_learned* L, int k, float* r_host, float* t_host)
{
  return api([&] {
    SX_REQUIRE(L && r_host && t_host && k >= 0, SX_ERR_ARG, "bad argument");
    int j = -1;
    for (size_t i = 0, c = 0; i < L->ok.size(); i++)
      if (L->ok[i] && int(c++) == k) {
        j = int(i);
        break;
      }
    SX_REQUIRE(j >= 0, SX_ERR_ARG, "no such pair");
    const size_t bytes = size_t(L->nv) * 3 * 4;
    SX_HIP(hipMemcpy(r_host, L->pr[size_t(j)]->ptr, bytes, hipMemcpyDeviceToHost));
    SX_HIP(hipMemcpy(t_host, L->pt[size_t(j)]->ptr, bytes, hipMemcpyDeviceToHost));
  });
}

int sx_learned_generate_pairs(sx_learned* L, sx_grad* g, const int32_t* ess_dofs, int64_t ness, int npairs,
                              uint64_t seed)
{
  return api([&] {
    SX_REQUIRE(L && g && g->form && npairs >= 0, SX_ERR_ARG, "bad argument");
    require_trial_grad(g, "pairs from a gradient before a linearization");
    SX_REQUIRE(g->form->test == L->space, SX_ERR_ARG, "the preconditioner belongs to another space");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    sx_form*      f  = g->form;
    hipStream_t   s  = f->stream;
    const int64_t n  = f->test->ndof();
    const int64_t nc = L->nv * 3;
    DeviceBuffer  ess, e, r, ec, one, z1;
    upload_ess(ess, ess_dofs, ness, n);
    e.allocate(size_t(n) * 8);
    r.allocate(size_t(n) * 8);
    ec.allocate(size_t(nc) * 4);
    one.upload(std::vector<double>{double(nc)});  // prolongation scale sqrt(nc / nc) = 1
    if (L->mg) {
      z1.allocate(size_t(n) * 8);
      mg_setup(L->mg, g, ess.as<const int32_t>(), ness);
      L->mg->uses = INT32_MAX;  // the next solve sets its own hierarchy up
    }
    std::mt19937_64 rng(seed);
    for (int p = 0; p < npairs; p++) {
      // e = P e_c (zero on essential DOFs), r = J_bc e: e is the converged
      // correction of residual r
      const std::vector<float> h = multiscale_field(L->G, rng);
      SX_HIP(hipMemcpyAsync(ec.ptr, h.data(), h.size() * 4, hipMemcpyHostToDevice, s));
      SX_HIP(sxk::vec_fill(e.as<double>(), 0.0, n, s));
      SX_HIP(sxu::fe_prolong_add(L->G, ec.as<const float>(), one.as<const double>(), nc, e.as<double>(), s));
      SX_HIP(sxk::set_rows(e.as<double>(), ess.as<const int32_t>(), ness, 0.0, s));
      apply_gradient(g, e.as<const double>(), r.as<double>());
      SX_HIP(sxk::copy_rows(r.as<double>(), e.as<const double>(), ess.as<const int32_t>(), ness, s));
      if (L->mg) {
        // the target is what the V-cycle leaves of the correction: e - MG(r)
        mg_apply(L->mg, g, ess.as<const int32_t>(), ness, r.as<const double>(), z1.as<double>());
        SX_HIP(sxk::vec_axpy(-1.0, z1.as<const double>(), e.as<double>(), n, s));
        SX_HIP(sxk::set_rows(e.as<double>(), ess.as<const int32_t>(), ness, 0.0, s));
      }
      const int k = take_pair_slot(L);
      SX_HIP(sxu::fe_restrict(L->G, r.as<const double>(), L->pr[size_t(k)]->as<float>(), s));
      SX_HIP(sxu::fe_restrict(L->G, e.as<const double>(), L->pt[size_t(k)]->as<float>(), s, true));
      SX_HIP(hipStreamSynchronize(s));  // h is reused by the next pair
      L->ok[size_t(k)] = 1;
    }
  });
}

int sx_learned_train(sx_learned* L, int steps, int batch, double lr, uint64_t seed, double* mean_loss)
{
  return api([&] {
    SX_REQUIRE(L && steps >= 0 && batch >= 1, SX_ERR_ARG, "bad argument");
    std::vector<size_t> usable;
    for (size_t j = 0; j < L->ok.size(); j++)
      if (L->ok[j]) usable.push_back(j);
    SX_REQUIRE(!usable.empty(), SX_ERR_STATE, "no training pairs recorded");
    const int64_t nc = L->nv * 3;
    // the target scale: geometric mean over the bank of rms(target) / rms(input)
    {
      DeviceBuffer sq;
      sq.allocate(2 * usable.size() * sizeof(double));
      SX_HIP(hipMemset(sq.ptr, 0, 2 * usable.size() * sizeof(double)));
      for (size_t i = 0; i < usable.size(); i++) {
        SX_HIP(sxu::sumsq(L->pr[usable[i]]->as<const float>(), nc, sq.as<double>() + 2 * i, nullptr));
        SX_HIP(sxu::sumsq(L->pt[usable[i]]->as<const float>(), nc, sq.as<double>() + 2 * i + 1, nullptr));
      }
      std::vector<double> h(2 * usable.size());
      SX_HIP(hipMemcpy(h.data(), sq.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost));
      double lsum = 0.0;
      int    cnt  = 0;
      for (size_t i = 0; i < usable.size(); i++)
        if (h[2 * i] > 0 && h[2 * i + 1] > 0) {
          lsum += 0.5 * std::log(h[2 * i + 1] / h[2 * i]);
          cnt++;
        }
      L->gain = cnt ? std::exp(lsum / cnt) : 1.0;
    }
    std::mt19937_64 rng(seed);
    DeviceBuffer    tin, tt, ss2;
    tin.allocate(size_t(nc) * 4);
    tt.allocate(size_t(nc) * 4);
    ss2.allocate(sizeof(double));
    double sum = 0.0;
    int    cnt = 0;
    for (int k = 0; k < steps; k++) {
      double step_loss = 0.0;
      for (int b = 0; b < batch; b++) {
        const size_t j = usable[size_t(rng() % usable.size())];
        // input scaled by its RMS s, target by s * gain
        SX_HIP(hipMemset(L->ss.ptr, 0, sizeof(double)));
        SX_HIP(sxu::sumsq(L->pr[j]->as<const float>(), nc, L->ss.as<double>(), nullptr));
        double ssh = 0.0;
        SX_HIP(hipMemcpy(&ssh, L->ss.ptr, sizeof(double), hipMemcpyDeviceToHost));
        const double ssg = ssh * L->gain * L->gain;
        SX_HIP(hipMemcpy(ss2.ptr, &ssg, sizeof(double), hipMemcpyHostToDevice));
        SX_HIP(sxu::normalize(L->pr[j]->as<const float>(), L->ss.as<const double>(), nc, nc, tin.as<float>(), nullptr));
        SX_HIP(sxu::normalize(L->pt[j]->as<const float>(), ss2.as<const double>(), nc, nc, tt.as<float>(), nullptr));
        SX_HIP(hipDeviceSynchronize());
        double loss = 0.0;
        const int rc = sx_unet_backward(L->net, tin.as<const float>(), tt.as<const float>(), &loss);
        if (rc != SX_OK) throw sx_failure(rc, g_last_error);
        step_loss += loss / batch;
      }
      const int rc = sx_unet_adam(L->net, lr);
      if (rc != SX_OK) throw sx_failure(rc, g_last_error);
      if (k >= steps - std::max(1, steps / 10)) {
        sum += step_loss;
        cnt++;
      }
    }
    if (mean_loss) *mean_loss = cnt ? sum / cnt : 0.0;
  });
}

int sx_learned_gain(const sx_learned* L, double* gain)
{
  return api([&] {
    SX_REQUIRE(L && gain, SX_ERR_ARG, "null argument");
    *gain = L->gain;
  });
}

int sx_learned_set_gain(sx_learned* L, double gain)
{
  return api([&] {
    SX_REQUIRE(L && std::isfinite(gain), SX_ERR_ARG, "bad argument");
    L->gain = gain;
  });
}

int sx_learned_set_multigrid(sx_learned* L, sx_multigrid* M)
{
  return api([&] {
    SX_REQUIRE(L, SX_ERR_ARG, "null argument");
    SX_REQUIRE(!M || M->space == L->space, SX_ERR_ARG, "the multigrid belongs to another space");
    L->mg = M;
    if (M) M->uses = INT32_MAX;
  });
}

int sx_learned_apply(sx_learned* L, sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* r, double* z)
{
  return api([&] {
    SX_REQUIRE(L && g && g->form && r && z, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "preconditioner applied before a linearization");
    SX_REQUIRE(g->form->test == L->space, SX_ERR_ARG, "the preconditioner belongs to another space");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, g->form->test->ndof());
    w.diag.allocate(size_t(g->form->test->ndof()) * 8);
    compute_diag(g, w.ess.as<const int32_t>(), ness, w.diag.as<double>());
    learned_mg_setup(L, g, w.ess.as<const int32_t>(), ness, false);
    learned_apply(L, g, w.diag.as<const double>(), w.ess.as<const int32_t>(), ness, r, z);
    SX_HIP(hipStreamSynchronize(g->form->stream));
  });
}

int sx_learned_destroy(sx_learned* L)
{
  delete L;
  return SX_OK;
}

int sx_form_set_preconditioner(sx_form* f, sx_learned* L, int mode)
{
  return api([&] {
    SX_REQUIRE(f, SX_ERR_ARG, "null form");
    SX_REQUIRE(mode == SX_PRECOND_JACOBI || mode == SX_PRECOND_RECORD || mode == SX_PRECOND_LEARNED, SX_ERR_ARG,
               "bad preconditioner mode");
    SX_REQUIRE(mode == SX_PRECOND_JACOBI || L, SX_ERR_ARG, "this mode needs a learned preconditioner");
    SX_REQUIRE(!L || L->space == f->test, SX_ERR_ARG, "the preconditioner belongs to another space");
    f->learned    = mode == SX_PRECOND_JACOBI ? nullptr : L;
    f->learn_mode = mode;
  });
}

int sx_fgmres_solve(sx_grad* g, sx_learned* L, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                    double rel_tol, double abs_tol, int max_iter, int restart, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "solve on a gradient before a linearization");
    SX_REQUIRE(!L || g->form->test == L->space, SX_ERR_ARG, "the preconditioner belongs to another space");
    SX_REQUIRE(ness == 0 || ess_dofs, SX_ERR_ARG, "null essential DOF list");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, g->form->test->ndof());
    fgmres(g, L, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, restart, w, &rep);
    SX_HIP(hipStreamSynchronize(g->form->stream));
    SX_REQUIRE(rep.converged, SX_ERR_NOCONV, "flexible GMRES did not converge");
  });
  if (report) *report = rep;
  return rc;
}

int sx_fcg_solve(sx_grad* g, sx_learned* L, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                 double rel_tol, double abs_tol, int max_iter, sx_solve_report* report)
{
  sx_solve_report rep{};
  int             rc = api([&] {
    SX_REQUIRE(g && g->form && L && b && x, SX_ERR_ARG, "null argument");
    require_trial_grad(g, "solve on a gradient before a linearization");
    SX_REQUIRE(g->form->test == L->space, SX_ERR_ARG, "the preconditioner belongs to another space");
    Workspace w;
    upload_ess(w.ess, ess_dofs, ness, g->form->test->ndof());
    fcg(g, L, w.ess.as<const int32_t>(), ness, b, x, rel_tol, abs_tol, max_iter, w, &rep);
    SX_HIP(hipStreamSynchronize(g->form->stream));
    SX_REQUIRE(rep.converged, SX_ERR_NOCONV, "flexible CG did not converge");
  });
  if (report) *report = rep;
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// quantities of interest: Functional<double(trial)> (fn/functional_qoi.inl:59-477)
// ---------------------------------------------------------------------------
struct sx_qoi {
  struct Term {
    int          id   = 0;
    bool         face = false;  // a boundary integral: slots are faces
    int          dep[2] = {0, 0};  // the trial arguments it reads (DependsOn<i...>)
    double       params[SX_MAX_PARAMS] = {};
    int64_t      nslots = 0;
    DeviceBuffer d_map, d_coords, d_part;
  };
  std::vector<const sx_space*>       trials;
  std::vector<std::unique_ptr<Term>> terms;
  // grads[k]: the gradient in trial k as the residual of derived integrands,
  // a form with test space trials[k] and every trial as an argument
  std::vector<sx_form*> grads;
  ~sx_qoi()
  {
    for (sx_form* g : grads)
      if (g) sx_form_destroy(g);
  }
};

namespace {

void check_trial_pair(const sx_space* a, const sx_space* b)
{
  SX_REQUIRE(a->mesh == b->mesh && a->order == b->order && a->nc == b->nc && a->ordering == b->ordering &&
                 (a == b || (a->native && b->native) || a->emap == b->emap),
             SX_ERR_UNSUPPORTED, "the inner-product QoI needs two trial spaces of the same kind on one mesh");
}

// adds `qoi_id` over `elems` of trial dep[0] (and dep[1])
void qoi_add_domain(sx_qoi* q, int qoi_id, const double* params, int nparams, const int* dep, int ndep,
                    const int* elems, int nelem)
{
  SX_REQUIRE(q, SX_ERR_ARG, "null QoI");
  SX_REQUIRE(nparams >= 0 && nparams <= SX_MAX_PARAMS && (nparams == 0 || params), SX_ERR_ARG,
             "QoI parameters: at most 16 doubles");
  SX_REQUIRE(qoi_id == SX_QOI_MEASURE || qoi_id == SX_QOI_POSITION || qoi_id == SX_QOI_VALUE ||
                 qoi_id == SX_QOI_NEOHOOKEAN_ENERGY || qoi_id == SX_QOI_INNER,
             SX_ERR_UNSUPPORTED, "unknown QoI integrand id " + std::to_string(qoi_id));
  const int want = qoi_id == SX_QOI_INNER ? 2 : 1;
  SX_REQUIRE(dep && ndep == want, SX_ERR_ARG,
             "this QoI integrand depends on " + std::to_string(want) + " trial argument(s) (DependsOn)");
  const int ntr = int(q->trials.size());
  for (int k = 0; k < ndep; k++) SX_REQUIRE(dep[k] >= 0 && dep[k] < ntr, SX_ERR_ARG, "DependsOn index out of range");
  const sx_space* s = q->trials[size_t(dep[0])];
  if (qoi_id == SX_QOI_INNER) check_trial_pair(s, q->trials[size_t(dep[1])]);
  SX_REQUIRE(qoi_id != SX_QOI_NEOHOOKEAN_ENERGY || s->nc == 3, SX_ERR_ARG,
             "the Neo-Hookean strain energy needs a 3-component space");
  if (qoi_id == SX_QOI_POSITION)
    SX_REQUIRE(nparams >= 1 && params[0] >= 0 && params[0] <= 2 && params[0] == std::floor(params[0]), SX_ERR_ARG,
               "position QoI: params {d}, d in 0..2");
  if (qoi_id == SX_QOI_VALUE)
    SX_REQUIRE(nparams >= 1 && params[0] >= 0 && params[0] < s->nc && params[0] == std::floor(params[0]),
               SX_ERR_ARG, "value QoI: params {c}, c a component of the space");
  const auto&      m = s->mesh->m;
  std::vector<int> domain;
  if (elems && nelem > 0) {
    domain.assign(elems, elems + nelem);
    for (int e : domain) SX_REQUIRE(e >= 0 && e < m.ne, SX_ERR_ARG, "element id out of range");
  } else {
    domain.resize(size_t(m.ne));
    for (int64_t e = 0; e < m.ne; e++) domain[size_t(e)] = int(e);
  }
  auto  Tp = std::make_unique<sx_qoi::Term>();
  auto& T  = *Tp;
  T.id     = qoi_id;
  for (int k = 0; k < ndep; k++) T.dep[k] = dep[k];
  for (int k = 0; k < nparams; k++) T.params[k] = params[k];
  T.nslots = int64_t(domain.size());
  std::vector<int>    map(size_t(T.nslots) * size_t(s->npe));
  std::vector<double> xyz(size_t(T.nslots) * 24);
  for (int64_t k = 0; k < T.nslots; k++) {
    const int e = domain[size_t(k)];
    std::copy(s->emap.begin() + long(size_t(e) * size_t(s->npe)), s->emap.begin() + long(size_t(e + 1) * size_t(s->npe)),
              map.begin() + long(size_t(k) * size_t(s->npe)));
    for (int c = 0; c < 8; c++) {
      const int v = m.corner(e, c);
      for (int d = 0; d < 3; d++) xyz[size_t(k) * 24 + size_t(c * 3 + d)] = m.X[size_t(3 * v + d)];
    }
  }
  T.d_map.upload(map);
  T.d_coords.upload(xyz);
  T.d_part.allocate(size_t(std::max<int64_t>(T.nslots, 1)) * 8);
  // the gradients: int u_c dV -> source e_c; the strain energy -> the
  // Neo-Hookean residual (its flux P is d psi / dF); int u_a . u_b dV ->
  // int phi . u_b dV in trial a (a mass residual at u_b) and int phi . u_a dV
  // in trial b; measure and position do not depend on U
  const int* el = elems && nelem > 0 ? elems : nullptr;
  auto       add = [&](int k, int qf, const double* prm, int np, int on) {
    const int d[1] = {on};
    const int rc   = sx_form_add_domain_integral(q->grads[size_t(k)], qf, prm, np, d, 1, el, nelem);
    if (rc != SX_OK) throw sx_failure(rc, g_last_error);
  };
  if (qoi_id == SX_QOI_VALUE) {
    g_internal_qf = true;
    try {
      add(dep[0], sxk::QF_ID_QOI_VALUE, T.params, 1, dep[0]);
    } catch (...) {
      g_internal_qf = false;
      throw;
    }
    g_internal_qf = false;
  } else if (qoi_id == SX_QOI_NEOHOOKEAN_ENERGY) {
    add(dep[0], SX_QF_NEOHOOKEAN, T.params, 3, dep[0]);
  } else if (qoi_id == SX_QOI_INNER) {
    const double mass[2] = {1.0, 0.0};  // {a u, b grad u} with a = 1, b = 0
    add(dep[0], SX_QF_MASS_DIFFUSION, mass, 2, dep[1]);
    add(dep[1], SX_QF_MASS_DIFFUSION, mass, 2, dep[0]);
  }
  q->terms.push_back(std::move(Tp));
}

void qoi_value(sx_qoi* q, const double* const* inputs, double* value)
{
  const sx_space* s0     = q->trials[0];
  hipStream_t     stream = q->grads[0]->stream;
  double          total  = 0.0;
  for (auto& Tp : q->terms) {
    const sx_qoi::Term& T = *Tp;
    const sx_space*     s = q->trials[size_t(T.dep[0])];
    sxk::QoiLaunch      P{};
    P.order    = s->order;
    P.nc       = s->nc;
    P.qoi      = T.id;
    P.ordering = s->ordering;
    P.map      = T.d_map.as<const int>();
    P.coords   = T.d_coords.as<const double>();
    P.nslots   = T.nslots;
    P.nnodes   = s->nnodes;
    P.u        = inputs[T.dep[0]];
    P.u2       = T.id == SX_QOI_INNER ? inputs[T.dep[1]] : nullptr;
    P.tables   = element_tables(s->order + 1);
    for (int k = 0; k < SX_MAX_PARAMS; k++) P.params[k] = T.params[k];
    P.partial = Tp->d_part.as<double>();
    P.stream  = stream;
    SX_HIP(T.face ? sxk::launch_qoi_face(P) : sxk::launch_qoi(P));
    std::vector<double> part(size_t(T.nslots));
    SX_HIP(hipMemcpyAsync(part.data(), T.d_part.ptr, part.size() * 8, hipMemcpyDeviceToHost, stream));
    SX_HIP(hipStreamSynchronize(stream));
    for (double v : part) total += v;  // element order: reproducible
  }
  (void)s0;
  *value = total;
}

void qoi_gradient(sx_qoi* q, double t, const double* const* inputs, int wrt, double* grad)
{
  sx_form* g = q->grads[size_t(wrt)];
  if (g->integrals.empty() && g->bintegrals.empty()) {
    SX_HIP(sxk::vec_fill(grad, 0.0, q->trials[size_t(wrt)]->ndof(), g->stream));
    SX_HIP(hipStreamSynchronize(g->stream));
    return;
  }
  sx_grad*  gr = nullptr;
  const int rc = sx_form_eval(g, t, inputs, grad, -1, &gr);
  if (rc != SX_OK) throw sx_failure(rc, g_last_error);
  SX_HIP(hipStreamSynchronize(g->stream));
}

}  // namespace

extern "C" {

int sx_qoi_create_multi(const sx_space* const* trials, int ntrials, sx_qoi** out)
{
  return api([&] {
    SX_REQUIRE(trials && ntrials >= 1 && out, SX_ERR_ARG, "a QoI needs at least one trial space");
    for (int i = 0; i < ntrials; i++) SX_REQUIRE(trials[i], SX_ERR_ARG, "null trial space");
    for (int i = 1; i < ntrials; i++)
      SX_REQUIRE(trials[i]->mesh == trials[0]->mesh, SX_ERR_UNSUPPORTED, "the trial spaces of a QoI share a mesh");
    auto q = std::make_unique<sx_qoi>();
    q->trials.assign(trials, trials + ntrials);
    q->grads.assign(size_t(ntrials), nullptr);
    for (int k = 0; k < ntrials; k++) {
      const int rc = sx_form_create(trials[k], trials, ntrials, &q->grads[size_t(k)]);
      if (rc != SX_OK) throw sx_failure(rc, g_last_error);
    }
    *out = q.release();
  });
}

int sx_qoi_create(const sx_space* trial, sx_qoi** out)
{
  const sx_space* sp[1] = {trial};
  if (!trial) return api([&] { SX_REQUIRE(false, SX_ERR_ARG, "null argument"); });
  return sx_qoi_create_multi(sp, 1, out);
}

int sx_qoi_add_domain_integral(sx_qoi* q, int qoi_id, const double* params, int nparams, const int* elems,
                               int nelem)
{
  return api([&] {
    const int dep[2] = {0, 0};
    qoi_add_domain(q, qoi_id, params, nparams, dep, qoi_id == SX_QOI_INNER ? 2 : 1, elems, nelem);
  });
}

int sx_qoi_add_domain_integral_dep(sx_qoi* q, int qoi_id, const double* params, int nparams, const int* depends_on,
                                   int ndep, const int* elems, int nelem)
{
  return api([&] { qoi_add_domain(q, qoi_id, params, nparams, depends_on, ndep, elems, nelem); });
}

}  // extern "C"

namespace {

// a boundary term of a QoI on the faces select() picks; the gradient of the
// value QoI is a boundary residual form over the same faces
void qoi_add_boundary(sx_qoi* q, int qoi_id, const double* params, int nparams, const FaceSelector& select)
{
  SX_REQUIRE(q, SX_ERR_ARG, "null QoI");
  SX_REQUIRE(nparams >= 0 && nparams <= SX_MAX_PARAMS && (nparams == 0 || params), SX_ERR_ARG,
             "QoI parameters: at most 16 doubles");
  SX_REQUIRE(qoi_id == SX_QOI_MEASURE || qoi_id == SX_QOI_POSITION || qoi_id == SX_QOI_VALUE, SX_ERR_UNSUPPORTED,
             "boundary QoIs: SX_QOI_MEASURE, SX_QOI_POSITION or SX_QOI_VALUE");
  const sx_space* vs = q->trials[0];
  if (qoi_id == SX_QOI_POSITION)
    SX_REQUIRE(nparams >= 1 && params[0] >= 0 && params[0] <= 2 && params[0] == std::floor(params[0]), SX_ERR_ARG,
               "position QoI: params {d}, d in 0..2");
  if (qoi_id == SX_QOI_VALUE)
    SX_REQUIRE(nparams >= 1 && params[0] >= 0 && params[0] < vs->nc && params[0] == std::floor(params[0]),
               SX_ERR_ARG, "value QoI: params {c}, c a component of the space");
  const auto&                     m     = vs->mesh->m;
  const std::vector<BoundaryFace> faces = select(m);
  const int                       n = vs->order + 1, nf = n * n;
  auto                            Tp = std::make_unique<sx_qoi::Term>();
  auto&                           T  = *Tp;
  T.id                               = qoi_id;
  T.face                             = true;
  for (int k = 0; k < nparams; k++) T.params[k] = params[k];
  T.nslots = int64_t(faces.size());
  std::vector<int>     fmap(faces.size() * size_t(nf));
  std::vector<double>  corners(faces.size() * 12);
  std::vector<int32_t> quads(faces.size() * 4);
  for (size_t i = 0; i < faces.size(); i++) {
    for (int b = 0; b < n; b++)
      for (int a = 0; a < n; a++) {
        int i0, i1, i2;
        face_node(faces[i].lf, a, b, n, i0, i1, i2);
        fmap[i * size_t(nf) + size_t(b * n + a)] =
            vs->emap[size_t(faces[i].e) * size_t(vs->npe) + size_t((i2 * n + i1) * n + i0)];
      }
    for (int k = 0; k < 4; k++) {
      int i0, i1, i2;
      face_node(faces[i].lf, k & 1, k >> 1, 2, i0, i1, i2);
      const int v      = m.corner(faces[i].e, i0 + 2 * i1 + 4 * i2);
      quads[i * 4 + size_t(k)] = v;
      for (int d = 0; d < 3; d++) corners[i * 12 + size_t(k * 3 + d)] = m.X[size_t(3 * v + d)];
    }
  }
  T.d_map.upload(fmap);
  T.d_coords.upload(corners);
  T.d_part.allocate(size_t(std::max<int64_t>(T.nslots, 1)) * 8);
  if (qoi_id == SX_QOI_VALUE) {
    // the gradient int phi_j dA on component c: the traction integrand with
    // A = 0, b = -e_c (residual -(A n + b)), or the flux integrand q0 = 1
    const int dep[1] = {0};
    const int c      = int(params[0]);
    int       rc;
    if (vs->nc == 3) {
      double tp[12] = {0};
      tp[9 + c]     = -1.0;
      rc = sx_form_add_boundary_integral_faces(q->grads[0], SX_BQF_TRACTION, tp, 12, dep, 1, quads.data(),
                                               int64_t(faces.size()));
    } else {
      const double fp[3] = {1.0, 0.0, 0.0};
      rc = sx_form_add_boundary_integral_faces(q->grads[0], SX_BQF_FLUX, fp, 3, dep, 1, quads.data(),
                                               int64_t(faces.size()));
    }
    if (rc != SX_OK) throw sx_failure(rc, g_last_error);
  }
  q->terms.push_back(std::move(Tp));
}

}  // namespace

extern "C" {

int sx_qoi_add_boundary_integral(sx_qoi* q, int qoi_id, const double* params, int nparams, int face_mask,
                                 uint64_t attr_mask)
{
  return api([&] {
    SX_REQUIRE(face_mask >= 0 && face_mask < 64, SX_ERR_ARG, "face_mask selects among 6 planes");
    qoi_add_boundary(q, qoi_id, params, nparams, [&](const sxh::HexMesh& m) {
      return attr_mask ? attribute_faces(m, attr_mask, true) : boundary_faces(m, face_mask);
    });
  });
}

int sx_qoi_add_boundary_integral_pred(sx_qoi* q, int qoi_id, const double* params, int nparams,
                                      sx_face_predicate pred, void* user)
{
  return api([&] {
    SX_REQUIRE(pred, SX_ERR_ARG, "null face predicate");
    qoi_add_boundary(q, qoi_id, params, nparams,
                     [&](const sxh::HexMesh& m) { return predicate_faces(m, pred, user); });
  });
}

int sx_qoi_add_boundary_integral_faces(sx_qoi* q, int qoi_id, const double* params, int nparams,
                                       const int32_t* face_vertices, int64_t nfaces)
{
  return api([&] {
    SX_REQUIRE(nfaces >= 0 && (nfaces == 0 || face_vertices), SX_ERR_ARG, "null face list");
    qoi_add_boundary(q, qoi_id, params, nparams,
                     [&](const sxh::HexMesh& m) { return listed_faces(m, face_vertices, nfaces); });
  });
}

int sx_qoi_eval(sx_qoi* q, double t, const double* u, double* value)
{
  return api([&] {
    SX_REQUIRE(q && u && value, SX_ERR_ARG, "null argument");
    SX_REQUIRE(q->trials.size() == 1, SX_ERR_ARG, "a QoI of several trial arguments: use sx_qoi_eval_multi");
    const double* in[1] = {u};
    (void)t;
    qoi_value(q, in, value);
  });
}

int sx_qoi_eval_multi(sx_qoi* q, double t, const double* const* inputs, double* value)
{
  return api([&] {
    SX_REQUIRE(q && inputs && value, SX_ERR_ARG, "null argument");
    for (size_t i = 0; i < q->trials.size(); i++) SX_REQUIRE(inputs[i], SX_ERR_ARG, "null input vector");
    (void)t;
    qoi_value(q, inputs, value);
  });
}

int sx_qoi_gradient(sx_qoi* q, double t, const double* u, double* grad)
{
  return api([&] {
    SX_REQUIRE(q && u && grad, SX_ERR_ARG, "null argument");
    SX_REQUIRE(q->trials.size() == 1, SX_ERR_ARG, "a QoI of several trial arguments: use sx_qoi_gradient_multi");
    const double* in[1] = {u};
    qoi_gradient(q, t, in, 0, grad);
  });
}

int sx_qoi_gradient_multi(sx_qoi* q, double t, const double* const* inputs, int wrt, double* grad)
{
  return api([&] {
    SX_REQUIRE(q && inputs && grad, SX_ERR_ARG, "null argument");
    SX_REQUIRE(wrt >= 0 && wrt < int(q->trials.size()), SX_ERR_ARG, "wrt out of range");
    for (size_t i = 0; i < q->trials.size(); i++) SX_REQUIRE(inputs[i], SX_ERR_ARG, "null input vector");
    qoi_gradient(q, t, inputs, wrt, grad);
  });
}

int sx_qoi_destroy(sx_qoi* q)
{
  delete q;
  return SX_OK;
}

}  // extern "C"
