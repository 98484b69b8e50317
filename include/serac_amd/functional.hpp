// serac_amd/functional.hpp -- C++ drop-in front end for the matrix-free path.
//
// Reproduces the surface of serac::Functional<test(trials...)>
// (src/serac/numerics/functional/functional.hpp:189-745) over the C-ABI of
// serac_amd.h: H1<p, c>, Dimension<d>, DependsOn<i...>, differentiate_wrt,
// AddDomainIntegral, operator() returning the (reused) output vector or a
// tuple of it and the Gradient, and Gradient::Mult / operator().  Integrands
// are the compiled registry of the library: the reference's functor types
// (MaterialStressFunctor<NeoHookean>, ThermalMaterialIntegrand<...>, ...)
// are mirrored here as plain structs and mapped to qfunction ids by
// qfunction_traits.  Vectors live in device memory (serac_amd::Vector stands
// in for mfem::Vector with device storage); calls are synchronous, like the
// reference's.
//
// Header-only; link against libserac_amd.so.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../serac_amd.h"

namespace serac_amd {

/// Errors from the library, with the status code of serac_amd.h.  The
/// reference aborts through SLIC_ERROR; here the C++ caller gets an exception.
class Error : public std::runtime_error {
 public:
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
  int code;
};

namespace detail {
inline void check(int rc)
{
  if (rc != SX_OK) throw Error(rc, sx_last_error());
}
/// bit a - 1 for each attribute a in 1..64
inline uint64_t attribute_mask(const std::set<int>& attributes)
{
  uint64_t m = 0;
  for (int a : attributes) {
    if (a < 1 || a > 64) throw Error(SX_ERR_ARG, "attribute " + std::to_string(a) + " outside 1..64");
    m |= uint64_t(1) << (a - 1);
  }
  return m;
}
}  // namespace detail

// ---------------------------------------------------------------------------
// device vector
// ---------------------------------------------------------------------------
class Vector {
 public:
  Vector() = default;
  explicit Vector(int64_t n) { SetSize(n); }
  explicit Vector(const std::vector<double>& host) { SetFromHost(host); }
  Vector(const Vector& o) { *this = o; }
  Vector(Vector&& o) noexcept : data_(o.data_), n_(o.n_) { o.data_ = nullptr, o.n_ = 0; }
  Vector& operator=(const Vector& o)
  {
    if (this != &o) {
      SetSize(o.n_);
      if (n_) detail::check(sx_memcpy_dtod(data_, o.data_, size_t(n_) * sizeof(double)));
    }
    return *this;
  }
  Vector& operator=(Vector&& o) noexcept
  {
    std::swap(data_, o.data_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~Vector()
  {
    if (data_) sx_free(data_);
  }

  /// resize to n entries, all zero
  void SetSize(int64_t n)
  {
    if (n != n_) {
      if (data_) sx_free(data_);
      data_ = nullptr;
      n_    = n;
      if (n > 0) {
        void* p = nullptr;
        detail::check(sx_malloc(&p, size_t(n) * sizeof(double)));
        data_ = static_cast<double*>(p);
      }
    }
    if (n_ > 0) detail::check(sx_memset_zero(data_, size_t(n_) * sizeof(double)));
  }
  int64_t       Size() const { return n_; }
  double*       Data() { return data_; }
  const double* Data() const { return data_; }
  const double* Read() const { return data_; }
  double*       ReadWrite() { return data_; }

  void SetFromHost(const std::vector<double>& h)
  {
    if (int64_t(h.size()) != n_) SetSize(int64_t(h.size()));
    if (n_) detail::check(sx_memcpy_htod(data_, h.data(), h.size() * sizeof(double)));
  }
  std::vector<double> ToHost() const
  {
    std::vector<double> h(static_cast<size_t>(n_));
    if (n_) detail::check(sx_memcpy_dtoh(h.data(), data_, h.size() * sizeof(double)));
    return h;
  }
  /// y += a x (this = y)
  void Add(double a, const Vector& x) { detail::check(sx_vec_axpy(a, x.data_, data_, n_)); }
  double operator*(const Vector& o) const
  {
    double r = 0.0;
    detail::check(sx_vec_dot(data_, o.data_, n_, &r));
    return r;
  }

 private:
  double* data_ = nullptr;
  int64_t n_    = 0;
};

// ---------------------------------------------------------------------------
// meshes and spaces (mfem::Mesh / ParFiniteElementSpace stand-ins)
// ---------------------------------------------------------------------------
class Mesh {
 public:
  /// buildCuboidMesh (src/serac/mesh/mesh_utils.cpp:149-153)
  static Mesh Cuboid(int nx, int ny, int nz, double lx = 1.0, double ly = 1.0, double lz = 1.0)
  {
    Mesh m;
    detail::check(sx_mesh_create_cuboid(nx, ny, nz, lx, ly, lz, &m.h_));
    return m;
  }
  /// vertices [nv][3]; hexes [ne][8] in mfem hexahedron vertex order
  Mesh(const std::vector<double>& vertices, const std::vector<int>& hexes)
  {
    detail::check(sx_mesh_create(int(vertices.size() / 3), vertices.data(), int(hexes.size() / 8), hexes.data(), &h_));
  }
  Mesh(Mesh&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
  Mesh& operator=(Mesh&& o) noexcept
  {
    std::swap(h_, o.h_);
    return *this;
  }
  Mesh(const Mesh&)            = delete;
  Mesh& operator=(const Mesh&) = delete;
  ~Mesh()
  {
    if (h_) sx_mesh_destroy(h_);
  }
  /// buildMeshFromFile (src/serac/mesh/mesh_utils.cpp:19-43) for MFEM v1.0 hex meshes
  static Mesh FromFile(const std::string& path)
  {
    Mesh m;
    detail::check(sx_mesh_read_mfem_file(path.c_str(), &m.h_));
    return m;
  }
  void    Perturb(uint64_t seed, double amp) { detail::check(sx_mesh_perturb(h_, seed, amp)); }
  int64_t GetNE() const
  {
    int64_t nv = 0, ne = 0;
    detail::check(sx_mesh_info(h_, &nv, &ne));
    return ne;
  }
  sx_mesh* handle() const { return h_; }

 private:
  Mesh() = default;
  sx_mesh* h_ = nullptr;
};

namespace Ordering {
enum Type { byNODES = SX_ORDERING_BYNODES, byVDIM = SX_ORDERING_BYVDIM };
}

class FiniteElementSpace {
 public:
  FiniteElementSpace(Mesh& mesh, int order, int vdim = 1, Ordering::Type ordering = Ordering::byVDIM)
  {
    detail::check(sx_space_create_h1_hex(mesh.handle(), order, vdim, ordering, &h_));
  }
  /// The space in the caller's node numbering: element_nodes[e * (order+1)^3 + j]
  /// is the node of element e's lexicographic node j (mfem's ElementRestriction
  /// order, fn/element_restriction.cpp:425-467), e.g. mfem's element vdofs of
  /// the first component after its lexicographic permutation.
  FiniteElementSpace(Mesh& mesh, int order, int vdim, Ordering::Type ordering, int64_t nnodes,
                     const std::vector<int32_t>& element_nodes)
  {
    detail::check(
        sx_space_create_h1_hex_map(mesh.handle(), order, vdim, ordering, nnodes, element_nodes.data(), &h_));
  }
  FiniteElementSpace(const FiniteElementSpace&)            = delete;
  FiniteElementSpace& operator=(const FiniteElementSpace&) = delete;
  ~FiniteElementSpace()
  {
    if (h_) sx_space_destroy(h_);
  }
  int64_t GetTrueVSize() const
  {
    int64_t nn = 0, nd = 0, ne = 0;
    int     npe = 0;
    detail::check(sx_space_info(h_, &nn, &nd, &ne, &npe));
    return nd;
  }
  int64_t GetVSize() const { return GetTrueVSize(); }  // one rank: L = T
  int64_t GetNDofs() const
  {
    int64_t nn = 0, nd = 0, ne = 0;
    int     npe = 0;
    detail::check(sx_space_info(h_, &nn, &nd, &ne, &npe));
    return nn;
  }
  std::vector<double> NodeCoordinates() const
  {
    std::vector<double> x(size_t(GetNDofs()) * 3);
    detail::check(sx_space_node_coords(h_, x.data()));
    return x;
  }
  /// essential DOFs on faces of a cuboid mesh (bit 0 x=0, 1 x=max, ...)
  std::vector<int32_t> BoundaryDofs(int face_mask, int comp_mask = 0x7) const
  {
    int64_t n = 0;
    detail::check(sx_space_boundary_dofs(h_, face_mask, comp_mask, nullptr, 0, &n));
    std::vector<int32_t> d(static_cast<size_t>(n));
    detail::check(sx_space_boundary_dofs(h_, face_mask, comp_mask, d.data(), n, &n));
    return d;
  }
  /// essential DOFs on the boundary elements with the given attributes, the
  /// attribute sets of setDisplacementBCs / setTemperatureBCs
  std::vector<int32_t> BoundaryDofs(const std::set<int>& attributes, int comp_mask = 0x7) const
  {
    const uint64_t mask = detail::attribute_mask(attributes);
    int64_t        n    = 0;
    detail::check(sx_space_boundary_dofs_attr(h_, mask, comp_mask, nullptr, 0, &n));
    std::vector<int32_t> d(static_cast<size_t>(n));
    detail::check(sx_space_boundary_dofs_attr(h_, mask, comp_mask, d.data(), n, &n));
    return d;
  }
  sx_space* handle() const { return h_; }

 private:
  sx_space* h_ = nullptr;
};

// ---------------------------------------------------------------------------
// function-space tags and argument wrappers (fn/finite_element.hpp:195-200,
// fn/differentiate_wrt.hpp:12-42)
// ---------------------------------------------------------------------------
template <int p, int c = 1>
struct H1 {
  static constexpr int order      = p;
  static constexpr int components = c;
};
template <int d>
struct Dimension {
  static constexpr int value = d;
};
template <int... i>
struct DependsOn {
};

template <typename T>
struct DifferentiateWRT {
  const T& value;
};
template <typename T>
DifferentiateWRT<T> differentiate_wrt(const T& v)
{
  return DifferentiateWRT<T>{v};
}

/// a subset of the mesh's elements (serac::Domain for volume elements)
struct Domain {
  std::vector<int> elements;  // empty: every element
  static Domain    Elements(std::vector<int> ids) { return Domain{std::move(ids)}; }
};

/// boundary faces (serac::Domain for boundary elements): on planes of the
/// mesh's bounding box (EntireBoundary when face_mask = 0), or the boundary
/// elements with chosen attributes (Domain::ofBoundaryElements by attribute)
struct BoundaryDomain {
  using vec3      = std::array<double, 3>;
  using Predicate = std::function<bool(std::vector<vec3>, int)>;
  int                   face_mask = 0;  // bit 0 x = min, 1 x = max, 2 y = min, 3 y = max, 4 z = min, 5 z = max
  uint64_t              attr_mask = 0;  // bit a - 1 for attribute a; when nonzero it selects the faces
  Predicate             predicate;      // when set: the exterior faces it selects
  std::vector<int32_t>  faces;          // when nonempty: these exterior faces, [n][4] vertex ids
  static BoundaryDomain EntireBoundary() { return BoundaryDomain{0, 0, {}, {}}; }
  static BoundaryDomain Planes(int mask) { return BoundaryDomain{mask, 0, {}, {}}; }
  static BoundaryDomain Attributes(const std::set<int>& attributes)
  {
    return BoundaryDomain{0, detail::attribute_mask(attributes), {}, {}};
  }
  /// Domain::ofBoundaryElements(mesh, predicate) (fn/domain.hpp:113-118): the
  /// predicate sees each exterior face's four corner positions and the
  /// attribute of its boundary element (-1 when the mesh has none there)
  static BoundaryDomain OfBoundaryElements(Predicate pred) { return BoundaryDomain{0, 0, std::move(pred), {}}; }
  /// exterior faces given by their four vertex ids each
  static BoundaryDomain Faces(std::vector<int32_t> quads) { return BoundaryDomain{0, 0, {}, std::move(quads)}; }
};

namespace detail {
inline int call_face_predicate(const double* corners, int attr, void* user)
{
  std::vector<BoundaryDomain::vec3> x(4);
  for (int k = 0; k < 4; k++) x[size_t(k)] = {corners[3 * k], corners[3 * k + 1], corners[3 * k + 2]};
  return (*static_cast<const BoundaryDomain::Predicate*>(user))(x, attr) ? 1 : 0;
}
}  // namespace detail

// ---------------------------------------------------------------------------
// integrands of the compiled registry
// ---------------------------------------------------------------------------
enum class GeometricNonlinearities { On, Off };  // phy/common.hpp:31

namespace heat_transfer {
/// phy/materials/thermal_material.hpp:20-55
struct LinearIsotropicConductor {
  LinearIsotropicConductor(double density = 1.0, double specific_heat_capacity = 1.0, double conductivity = 1.0)
      : density_(density), specific_heat_capacity_(specific_heat_capacity), conductivity_(conductivity)
  {
    if (conductivity_ < 0.0 || density_ < 0.0 || specific_heat_capacity_ < 0.0)
      throw Error(SX_ERR_ARG, "LinearIsotropicConductor: negative property");
  }
  double density_, specific_heat_capacity_, conductivity_;
};
/// phy/heat_transfer.hpp:381-406 (steady state: the heat-capacity term is 0)
template <typename Material>
struct ThermalMaterialIntegrand {
  explicit ThermalMaterialIntegrand(Material m) : material_(m) {}
  Material material_;
};
/// the integrand setSource adds (phy/heat_transfer.hpp:493-503) for a constant source
struct ConstantSource {
  double source;
};
}  // namespace heat_transfer

namespace solid_mechanics {
/// phy/materials/solid_material.hpp:98-127
struct NeoHookean {
  double density;
  double K;
  double G;
};
/// phy/materials/solid_material.hpp:24-50
struct LinearIsotropic {
  double density;
  double K;
  double G;
};
/// phy/materials/solid_material.hpp:132-148
struct LinearHardening {
  double sigma_y;  ///< yield strength
  double Hi;       ///< isotropic hardening modulus
};
/// phy/materials/solid_material.hpp:199-258 (compiled for LinearHardening)
template <typename HardeningType>
struct J2SmallStrain {
  double        E;          ///< Young's modulus
  double        nu;         ///< Poisson's ratio
  HardeningType hardening;  ///< flow stress hardening model
  double        Hk;         ///< kinematic hardening modulus
  double        density;    ///< mass density
  /// the hysteresis variables of one quadrature point (10 doubles, in this order)
  struct State {
    double plastic_strain[3][3];
    double accumulated_plastic_strain;
  };
};
/// phy/solid_mechanics.hpp:808-853 (quasi-static: no acceleration argument)
template <typename Material>
struct MaterialStressFunctor {
  MaterialStressFunctor(Material m, GeometricNonlinearities gn) : material_(m), geom_nonlin_(gn) {}
  Material                material_;
  GeometricNonlinearities geom_nonlin_;
};
}  // namespace solid_mechanics

/// {a u, b grad u}, the integrand of fn/tests/functional_comparisons.cpp:84-203
struct MassDiffusion {
  double a, b;
};

namespace solid_mechanics {
/// the boundary integrand setTraction adds (phy/solid_mechanics.hpp:1023-1037)
/// for traction(X, n) = A n + b
struct LinearTraction {
  double A[3][3];
  double b[3];
};
/// the boundary integrand setPressure adds (phy/solid_mechanics.hpp:1071-1101)
struct Pressure {
  double                  pressure;
  GeometricNonlinearities geom_nonlin;
};
}  // namespace solid_mechanics
namespace heat_transfer {
/// the boundary integrand setFluxBCs adds (phy/heat_transfer.hpp:536-552) for
/// flux = q0 + h (T - T_inf)
struct LinearFlux {
  double q0, h, T_inf;
};
}  // namespace heat_transfer

/// boundary integrand registry (SX_BQF_*)
template <typename Integrand>
struct bqfunction_traits {
  static constexpr bool registered = false;
};
template <>
struct bqfunction_traits<solid_mechanics::LinearTraction> {
  static constexpr bool      registered = true;
  static int                 id(const solid_mechanics::LinearTraction&) { return SX_BQF_TRACTION; }
  static std::vector<double> params(const solid_mechanics::LinearTraction& q)
  {
    std::vector<double> p;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) p.push_back(q.A[i][j]);
    for (int i = 0; i < 3; i++) p.push_back(q.b[i]);
    return p;
  }
};
template <>
struct bqfunction_traits<solid_mechanics::Pressure> {
  static constexpr bool      registered = true;
  static int                 id(const solid_mechanics::Pressure&) { return SX_BQF_PRESSURE; }
  static std::vector<double> params(const solid_mechanics::Pressure& q)
  {
    return {q.pressure, q.geom_nonlin == GeometricNonlinearities::On ? 1.0 : 0.0};
  }
};
template <>
struct bqfunction_traits<heat_transfer::LinearFlux> {
  static constexpr bool      registered = true;
  static int                 id(const heat_transfer::LinearFlux&) { return SX_BQF_FLUX; }
  static std::vector<double> params(const heat_transfer::LinearFlux& q) { return {q.q0, q.h, q.T_inf}; }
};

/// Maps an integrand type to (qfunction id, parameters).  Unregistered types
/// fail at compile time: the C-ABI cannot carry arbitrary lambdas.
template <typename Integrand>
struct qfunction_traits {
  static constexpr bool registered = false;
};

template <>
struct qfunction_traits<heat_transfer::ThermalMaterialIntegrand<heat_transfer::LinearIsotropicConductor>> {
  static constexpr bool registered = true;
  static int            id(const heat_transfer::ThermalMaterialIntegrand<heat_transfer::LinearIsotropicConductor>&)
  {
    return SX_QF_HEAT;
  }
  static std::vector<double> params(const heat_transfer::ThermalMaterialIntegrand<heat_transfer::LinearIsotropicConductor>& q)
  {
    // the third entry (rho c_p) is read only with DependsOn<T, dT/dt>
    return {q.material_.conductivity_, 0.0, q.material_.density_ * q.material_.specific_heat_capacity_};
  }
};
template <>
struct qfunction_traits<heat_transfer::ConstantSource> {
  static constexpr bool      registered = true;
  static int                 id(const heat_transfer::ConstantSource&) { return SX_QF_HEAT; }
  static std::vector<double> params(const heat_transfer::ConstantSource& q) { return {0.0, q.source}; }
};
template <>
struct qfunction_traits<solid_mechanics::MaterialStressFunctor<solid_mechanics::NeoHookean>> {
  static constexpr bool registered = true;
  static int            id(const solid_mechanics::MaterialStressFunctor<solid_mechanics::NeoHookean>& q)
  {
    if (q.geom_nonlin_ != GeometricNonlinearities::On)
      throw Error(SX_ERR_UNSUPPORTED, "NeoHookean is compiled with geometric nonlinearity on");
    return SX_QF_NEOHOOKEAN;
  }
  static std::vector<double> params(const solid_mechanics::MaterialStressFunctor<solid_mechanics::NeoHookean>& q)
  {
    return {q.material_.density, q.material_.K, q.material_.G};
  }
};
template <>
struct qfunction_traits<solid_mechanics::MaterialStressFunctor<solid_mechanics::LinearIsotropic>> {
  static constexpr bool registered = true;
  static int            id(const solid_mechanics::MaterialStressFunctor<solid_mechanics::LinearIsotropic>& q)
  {
    if (q.geom_nonlin_ != GeometricNonlinearities::Off)
      throw Error(SX_ERR_UNSUPPORTED, "LinearIsotropic is compiled with geometric nonlinearity off");
    return SX_QF_LINEAR_ELASTIC;
  }
  static std::vector<double> params(const solid_mechanics::MaterialStressFunctor<solid_mechanics::LinearIsotropic>& q)
  {
    // the density is read only with DependsOn<displacement, acceleration>
    return {q.material_.K - (2.0 / 3.0) * q.material_.G, q.material_.G, q.material_.density};
  }
};
template <>
struct qfunction_traits<MassDiffusion> {
  static constexpr bool      registered = true;
  static int                 id(const MassDiffusion&) { return SX_QF_MASS_DIFFUSION; }
  static std::vector<double> params(const MassDiffusion& q) { return {q.a, q.b}; }
};
template <>
struct qfunction_traits<
    solid_mechanics::MaterialStressFunctor<solid_mechanics::J2SmallStrain<solid_mechanics::LinearHardening>>> {
  using material_type  = solid_mechanics::J2SmallStrain<solid_mechanics::LinearHardening>;
  using integrand_type = solid_mechanics::MaterialStressFunctor<material_type>;
  using state_type     = material_type::State;
  static constexpr bool registered = true;
  static int            id(const integrand_type&) { return SX_QF_J2_LINEAR; }
  static std::vector<double> params(const integrand_type& q)
  {
    const auto& m = q.material_;
    return {m.E,  m.nu, m.hardening.sigma_y, m.hardening.Hi, m.Hk, m.density,
            q.geom_nonlin_ == GeometricNonlinearities::On ? 1.0 : 0.0};
  }
};

/// Assembled sparse matrix (device CSR: int64 row offsets, int32 columns
/// sorted per row, FP64 values), as Gradient::assemble returns it.
class SparseMatrix {
 public:
  explicit SparseMatrix(sx_csr* h) : h_(h)
  {
    int rc = sx_csr_info(h_, &nrows_, &ncols_, &nnz_);
    if (rc != SX_OK) throw Error(rc, sx_last_error());
  }
  SparseMatrix(const SparseMatrix&)            = delete;
  SparseMatrix& operator=(const SparseMatrix&) = delete;
  ~SparseMatrix()
  {
    if (h_) sx_csr_destroy(h_);
  }
  int64_t Height() const { return nrows_; }
  int64_t Width() const { return ncols_; }
  int64_t NumNonZeroElems() const { return nnz_; }
  /// y = A x on the device
  void Mult(const Vector& x, Vector& y) const
  {
    if (y.Size() != nrows_) y.SetSize(nrows_);
    int rc = sx_csr_spmv(h_, x.Data(), y.Data());
    if (rc != SX_OK) throw Error(rc, sx_last_error());
  }
  /// host copies of the CSR arrays
  void ToHost(std::vector<int64_t>& row_ptr, std::vector<int32_t>& col, std::vector<double>& val) const
  {
    row_ptr.resize(size_t(nrows_) + 1);
    col.resize(size_t(nnz_));
    val.resize(size_t(nnz_));
    int rc = sx_csr_download(h_, row_ptr.data(), col.data(), val.data());
    if (rc != SX_OK) throw Error(rc, sx_last_error());
  }
  sx_csr* handle() const { return h_; }

 private:
  sx_csr* h_ = nullptr;
  int64_t nrows_ = 0, ncols_ = 0, nnz_ = 0;
};

/// QuadratureData<State> (fn/quadrature_data.hpp:111-167): one State per
/// quadrature point of every element of a domain, kept on the device.  State
/// must be a struct of doubles.
template <typename State>
class QuadratureData {
 public:
  static_assert(sizeof(State) % sizeof(double) == 0, "State must consist of doubles");
  static constexpr int size = int(sizeof(State) / sizeof(double));
  QuadratureData(int64_t num_elements, int qpts_per_element, const State& init = State{})
      : nelem_(num_elements), nq_(qpts_per_element)
  {
    double v[size];
    std::memcpy(v, &init, sizeof(State));
    int rc = sx_qdata_create(num_elements, qpts_per_element, size, v, &h_);
    if (rc != SX_OK) throw Error(rc, sx_last_error());
  }
  QuadratureData(const QuadratureData&)            = delete;
  QuadratureData& operator=(const QuadratureData&) = delete;
  ~QuadratureData()
  {
    if (h_) sx_qdata_destroy(h_);
  }
  /// host copy, [element][qpoint]
  std::vector<State> Get() const
  {
    std::vector<State> out(size_t(nelem_) * size_t(nq_));
    int                rc = sx_qdata_get(h_, reinterpret_cast<double*>(out.data()));
    if (rc != SX_OK) throw Error(rc, sx_last_error());
    return out;
  }
  void Set(const std::vector<State>& st)
  {
    if (st.size() != size_t(nelem_) * size_t(nq_)) throw Error(SX_ERR_ARG, "QuadratureData::Set: wrong size");
    int rc = sx_qdata_set(h_, reinterpret_cast<const double*>(st.data()));
    if (rc != SX_OK) throw Error(rc, sx_last_error());
  }
  sx_qdata* handle() const { return h_; }

 private:
  sx_qdata* h_ = nullptr;
  int64_t   nelem_;
  int       nq_;
};

namespace detail {
template <typename T>
struct is_dwrt : std::false_type {
};
template <typename T>
struct is_dwrt<DifferentiateWRT<T>> : std::true_type {
};
template <typename T>
const double* data_of(const T& v)
{
  if constexpr (is_dwrt<T>::value)
    return v.value.Data();
  else
    return v.Data();
}
template <typename... T>
constexpr int wrt_index()
{
  int  idx = -1, i = 0;
  bool f[] = {is_dwrt<T>::value..., false};
  for (bool b : f) {
    if (b && idx < 0) idx = i;
    i++;
  }
  return idx;
}
template <typename... T>
constexpr int count_dwrt()
{
  return (0 + ... + int(is_dwrt<T>::value));
}
}  // namespace detail

// ---------------------------------------------------------------------------
// Functional<test(trials...)>
// ---------------------------------------------------------------------------
template <typename T>
class Functional;

template <typename test, typename... trials>
class Functional<test(trials...)> {
  static constexpr int num_trial_spaces = int(sizeof...(trials));

 public:
  /// Functional::Gradient (fn/functional.hpp:548-669): the action of the
  /// stored linearization, its diagonal, and Jacobi-CG on it.
  class Gradient {
   public:
    /// df = J dx  (Gradient::Mult, fn/functional.hpp:556-559)
    void Mult(const Vector& dx, Vector& df) const
    {
      if (df.Size() != test_size_) df.SetSize(test_size_);
      detail::check(sx_grad_apply(h_, dx.Data(), df.Data()));
      detail::check(sx_form_synchronize(form_));
    }
    /// J dx into an internal vector (Gradient::operator(), :562-566)
    Vector& operator()(const Vector& dx)
    {
      Mult(dx, df_);
      return df_;
    }
    /// the diagonal the assembled matrix would have (Jacobi preconditioner)
    Vector AssembleDiagonal() const
    {
      Vector d(test_size_);
      detail::check(sx_grad_diagonal(h_, d.Data()));
      detail::check(sx_form_synchronize(form_));
      return d;
    }
    /// Gradient::assemble (fn/functional.hpp:569-667): the linearization as
    /// a device CSR matrix (the HypreParMatrix of the reference on one rank)
    std::unique_ptr<SparseMatrix> assemble() const
    {
      sx_csr* A = nullptr;
      detail::check(sx_grad_assemble(h_, &A));
      return std::unique_ptr<SparseMatrix>(new SparseMatrix(A));
    }
    int64_t  Height() const { return test_size_; }
    sx_grad* handle() const { return h_; }

   private:
    friend class Functional;
    sx_grad* h_         = nullptr;
    sx_form* form_      = nullptr;
    int64_t  test_size_ = 0;
    Vector   df_;
  };

  /// fn/functional.hpp:219-273
  Functional(const FiniteElementSpace* test_fes, std::array<const FiniteElementSpace*, num_trial_spaces> trial_fes)
  {
    std::array<const sx_space*, num_trial_spaces> t{};
    for (int i = 0; i < num_trial_spaces; i++) t[size_t(i)] = trial_fes[size_t(i)]->handle();
    detail::check(sx_form_create(test_fes->handle(), t.data(), num_trial_spaces, &h_));
    test_      = test_fes->handle();
    test_size_ = test_fes->GetTrueVSize();
    output_.SetSize(test_size_);
  }
  Functional(const Functional&)            = delete;
  Functional& operator=(const Functional&) = delete;
  ~Functional()
  {
    if (h_) sx_form_destroy(h_);
  }

  /// AddDomainIntegral (fn/functional.hpp:286-317), whole mesh or a Domain
  template <int dim, int... args, typename Integrand>
  void AddDomainIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Mesh&)
  {
    AddDomainIntegral(Dimension<dim>{}, DependsOn<args...>{}, integrand, Domain{});
  }
  template <int dim, int... args, typename Integrand>
  void AddDomainIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Domain& domain)
  {
    static_assert(dim == 3, "the compiled kernels integrate over hexahedra (Dimension<3>)");
    static_assert(qfunction_traits<Integrand>::registered,
                  "integrand type is not in the compiled qfunction registry (see qfunction_traits)");
    static_assert(((args >= 0 && args < num_trial_spaces) && ...), "DependsOn index out of range");
    const int                 deps[] = {args...};
    const std::vector<double> p      = qfunction_traits<Integrand>::params(integrand);
    detail::check(sx_form_add_domain_integral(h_, qfunction_traits<Integrand>::id(integrand), p.data(), int(p.size()),
                                              deps, int(sizeof...(args)),
                                              domain.elements.empty() ? nullptr : domain.elements.data(),
                                              int(domain.elements.size())));
  }
  /// AddDomainIntegral with quadrature-point state (the reference's trailing
  /// shared_ptr<QuadratureData<S>> argument); the form keeps a reference
  template <int dim, int... args, typename Integrand, typename State>
  void AddDomainIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Domain& domain,
                         std::shared_ptr<QuadratureData<State>> qdata)
  {
    static_assert(dim == 3, "the compiled kernels integrate over hexahedra (Dimension<3>)");
    static_assert(qfunction_traits<Integrand>::registered,
                  "integrand type is not in the compiled qfunction registry (see qfunction_traits)");
    static_assert(((args >= 0 && args < num_trial_spaces) && ...), "DependsOn index out of range");
    const int                 deps[] = {args...};
    const std::vector<double> p      = qfunction_traits<Integrand>::params(integrand);
    detail::check(sx_form_add_domain_integral_qdata(
        h_, qfunction_traits<Integrand>::id(integrand), p.data(), int(p.size()), deps, int(sizeof...(args)),
        domain.elements.empty() ? nullptr : domain.elements.data(), int(domain.elements.size()),
        qdata ? qdata->handle() : nullptr));
    if (qdata) qdata_keep_.push_back(std::static_pointer_cast<void>(qdata));
  }
  template <int dim, int... args, typename Integrand, typename State>
  void AddDomainIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Mesh&,
                         std::shared_ptr<QuadratureData<State>> qdata)
  {
    AddDomainIntegral(Dimension<dim>{}, DependsOn<args...>{}, integrand, Domain{}, std::move(qdata));
  }
  /// AddBoundaryIntegral (fn/functional.hpp:328-354) over the whole boundary or
  /// the faces on chosen planes
  template <int dim, int... args, typename Integrand>
  void AddBoundaryIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand,
                           const BoundaryDomain& domain)
  {
    static_assert(dim == 2, "boundary integrals on hexahedral meshes are over quadrilateral faces (Dimension<2>)");
    static_assert(bqfunction_traits<Integrand>::registered,
                  "integrand type is not in the compiled boundary qfunction registry (see bqfunction_traits)");
    static_assert(((args >= 0 && args < num_trial_spaces) && ...), "DependsOn index out of range");
    const int                 deps[] = {args...};
    const std::vector<double> p      = bqfunction_traits<Integrand>::params(integrand);
    if (domain.predicate)
      detail::check(sx_form_add_boundary_integral_pred(
          h_, bqfunction_traits<Integrand>::id(integrand), p.data(), int(p.size()), deps, int(sizeof...(args)),
          &detail::call_face_predicate, const_cast<BoundaryDomain::Predicate*>(&domain.predicate)));
    else if (!domain.faces.empty())
      detail::check(sx_form_add_boundary_integral_faces(h_, bqfunction_traits<Integrand>::id(integrand), p.data(),
                                                        int(p.size()), deps, int(sizeof...(args)),
                                                        domain.faces.data(), int64_t(domain.faces.size() / 4)));
    else if (domain.attr_mask)
      detail::check(sx_form_add_boundary_integral_attr(h_, bqfunction_traits<Integrand>::id(integrand), p.data(),
                                                       int(p.size()), deps, int(sizeof...(args)), domain.attr_mask));
    else
      detail::check(sx_form_add_boundary_integral(h_, bqfunction_traits<Integrand>::id(integrand), p.data(),
                                                  int(p.size()), deps, int(sizeof...(args)), domain.face_mask));
  }
  template <int dim, int... args, typename Integrand>
  void AddBoundaryIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Mesh&)
  {
    AddBoundaryIntegral(Dimension<dim>{}, DependsOn<args...>{}, integrand, BoundaryDomain::EntireBoundary());
  }
  /// alias of the reference (fn/functional.hpp:385-393)
  template <int... args, typename Integrand, typename DomainT>
  void AddSurfaceIntegral(DependsOn<args...>, const Integrand& integrand, const DomainT& domain)
  {
    AddBoundaryIntegral(Dimension<2>{}, DependsOn<args...>{}, integrand, domain);
  }

  /// updateQdata (fn/functional.hpp:523)
  void updateQdata(bool update_flag) { detail::check(sx_form_update_qdata(h_, update_flag ? 1 : 0)); }

  /// alias of the reference (fn/functional.hpp:365-373)
  template <int... args, typename Integrand, typename DomainT>
  void AddVolumeIntegral(DependsOn<args...>, const Integrand& integrand, DomainT& domain)
  {
    AddDomainIntegral(Dimension<3>{}, DependsOn<args...>{}, integrand, domain);
  }

  /// scatter mode / XCD block order (serac_amd.h, SX_OPT_*)
  void SetOption(int option, int64_t value) { detail::check(sx_form_set_option(h_, option, value)); }

  /// the residual with the gradient in the shape displacement set on the form
  /// (differentiate_wrt on a ShapeAwareFunctional's shape argument)
  template <typename... T>
  std::tuple<Vector&, Gradient&> EvalShapeGradient(double t, const T&... args)
  {
    static_assert(int(sizeof...(T)) == num_trial_spaces, "wrong number of arguments");
    static_assert(detail::count_dwrt<T...>() == 0, "the shape is the differentiated argument");
    std::array<const double*, sizeof...(T)> in{detail::data_of(args)...};
    sx_grad*                                g = nullptr;
    detail::check(sx_form_eval(h_, t, in.data(), output_.Data(), SX_WRT_SHAPE, &g));
    detail::check(sx_form_synchronize(h_));
    shape_grad_.h_         = g;
    shape_grad_.form_      = h_;
    shape_grad_.test_size_ = test_size_;
    return std::tuple<Vector&, Gradient&>(output_, shape_grad_);
  }

  /// operator()(t, args...) (fn/functional.hpp:501-513).  Returns the output
  /// vector, which the next call overwrites, or a tuple of it and the
  /// Gradient when one argument is wrapped by differentiate_wrt.
  template <typename... T>
  decltype(auto) operator()(double t, const T&... args)
  {
    static_assert(int(sizeof...(T)) == num_trial_spaces, "wrong number of arguments");
    static_assert(detail::count_dwrt<T...>() <= 1, "only one argument may be wrapped by differentiate_wrt");
    constexpr int                              wrt = detail::wrt_index<T...>();
    std::array<const double*, sizeof...(T)>    in{detail::data_of(args)...};
    sx_grad*                                   g = nullptr;
    detail::check(sx_form_eval(h_, t, in.data(), output_.Data(), wrt, &g));
    detail::check(sx_form_synchronize(h_));
    if constexpr (wrt < 0) {
      return static_cast<Vector&>(output_);
    } else {
      grad_.h_         = g;
      grad_.form_      = h_;
      grad_.test_size_ = test_size_;
      return std::tuple<Vector&, Gradient&>(output_, grad_);
    }
  }

  sx_form*  handle() const { return h_; }
  sx_space* test_space() const { return test_; }

 private:
  sx_form*                           h_         = nullptr;
  sx_space*                          test_      = nullptr;
  int64_t                            test_size_ = 0;
  Vector                             output_;
  Gradient                           grad_;
  Gradient                           shape_grad_;
  std::vector<std::shared_ptr<void>> qdata_keep_;  // the form borrows the state buffers
};

// ---------------------------------------------------------------------------
// ShapeAwareFunctional<shape, test(trials...)> (fn/shape_aware_functional.hpp:292-577)
// ---------------------------------------------------------------------------
/// The domain integrals evaluated on the configuration X + p of a shape
/// displacement p: with J = I + dp/dX, trial gradients become du/dX J^-1,
/// fluxes f J^-T det J and sources s det J (ShapeCorrection,
/// fn/shape_aware_functional.hpp:37-125; boundary integrals see the faces of
/// X + p).  p lives in the test space (a 3-component H1 space of the mesh);
/// differentiate_wrt applies to a trial argument or to p (domain integrals).
template <typename shape, typename T>
class ShapeAwareFunctional;

template <typename shape, typename test, typename... trials>
class ShapeAwareFunctional<shape, test(trials...)> {
 public:
  using Gradient = typename Functional<test(trials...)>::Gradient;

  ShapeAwareFunctional(const FiniteElementSpace* shape_fes, const FiniteElementSpace* test_fes,
                       std::array<const FiniteElementSpace*, sizeof...(trials)> trial_fes)
      : f_(test_fes, trial_fes)
  {
    if (shape_fes->handle() != test_fes->handle())
      throw Error(SX_ERR_UNSUPPORTED, "the shape displacement space must be the test space");
  }
  template <int dim, int... args, typename Integrand>
  void AddDomainIntegral(Dimension<dim> d, DependsOn<args...> dep, const Integrand& integrand, const Mesh& mesh)
  {
    f_.AddDomainIntegral(d, dep, integrand, mesh);
  }
  template <typename... A>
  void AddBoundaryIntegral(A&&... a)
  {
    f_.AddBoundaryIntegral(std::forward<A>(a)...);
  }
  /// operator()(t, shape_displacement, args...) (fn/shape_aware_functional.hpp:531-560)
  template <typename... T>
  decltype(auto) operator()(double t, const Vector& p, const T&... args)
  {
    detail::check(sx_form_set_shape_displacement(f_.handle(), p.Data()));
    return f_(t, args...);
  }
  /// operator()(t, differentiate_wrt(p), args...): the residual and the
  /// Gradient in the shape displacement, whose action is dR/dp dp
  template <typename... T>
  std::tuple<Vector&, Gradient&> operator()(double t, const DifferentiateWRT<Vector>& p, const T&... args)
  {
    detail::check(sx_form_set_shape_displacement(f_.handle(), p.value.Data()));
    return f_.EvalShapeGradient(t, args...);
  }
  sx_form* handle() const { return f_.handle(); }

 private:
  Functional<test(trials...)> f_;
};

// ---------------------------------------------------------------------------
// quantities of interest: Functional<double(trial)> (fn/functional_qoi.inl:59-477)
// ---------------------------------------------------------------------------
namespace qoi {
struct Measure {  ///< int 1 dV (TrivialIntegrator of fn/tests/functional_qoi.cpp)
};
struct Position {  ///< int X_d dV, the first moment (ZeroIndexIntegrator for d = 0)
  int direction;
};
struct Value {  ///< int u_c dV (GetZeroIntegrator for c = 0)
  int component;
};
}  // namespace qoi
namespace solid_mechanics {
/// int psi(I + grad u) dV, the strain energy of a hyperelastic material
template <typename Material>
struct StrainEnergy {
  Material material;
};
}  // namespace solid_mechanics

/// maps QoI integrand types to the compiled registry (SX_QOI_*)
template <typename Q>
struct qoi_traits {
  static constexpr bool registered = false;
};
template <>
struct qoi_traits<qoi::Measure> {
  static constexpr bool      registered = true;
  static int                 id(const qoi::Measure&) { return SX_QOI_MEASURE; }
  static std::vector<double> params(const qoi::Measure&) { return {}; }
};
template <>
struct qoi_traits<qoi::Position> {
  static constexpr bool      registered = true;
  static int                 id(const qoi::Position&) { return SX_QOI_POSITION; }
  static std::vector<double> params(const qoi::Position& q) { return {double(q.direction)}; }
};
template <>
struct qoi_traits<qoi::Value> {
  static constexpr bool      registered = true;
  static int                 id(const qoi::Value&) { return SX_QOI_VALUE; }
  static std::vector<double> params(const qoi::Value& q) { return {double(q.component)}; }
};
template <>
struct qoi_traits<solid_mechanics::StrainEnergy<solid_mechanics::NeoHookean>> {
  static constexpr bool registered = true;
  static int id(const solid_mechanics::StrainEnergy<solid_mechanics::NeoHookean>&) { return SX_QOI_NEOHOOKEAN_ENERGY; }
  static std::vector<double> params(const solid_mechanics::StrainEnergy<solid_mechanics::NeoHookean>& q)
  {
    return {q.material.density, q.material.K, q.material.G};
  }
};

template <typename trial>
class Functional<double(trial)> {
 public:
  /// Functional<double(...)>::Gradient (fn/functional_qoi.inl:407-477)
  class Gradient {
   public:
    /// the directional derivative dq/dU . dx
    double operator()(const Vector& dx) const
    {
      double v = 0.0;
      detail::check(sx_vec_dot(g_.Data(), dx.Data(), g_.Size(), &v));
      return v;
    }
    /// dq/dU (the HypreParVector of assemble(), fn/functional_qoi.inl:422)
    Vector assemble() const { return g_; }

   private:
    friend class Functional;
    Vector g_;
  };

  explicit Functional(std::array<const FiniteElementSpace*, 1> trial_space)
  {
    detail::check(sx_qoi_create(trial_space[0]->handle(), &h_));
  }
  Functional(const Functional&)            = delete;
  Functional& operator=(const Functional&) = delete;
  ~Functional()
  {
    if (h_) sx_qoi_destroy(h_);
  }

  /// AddDomainIntegral (fn/functional_qoi.inl:160-190)
  template <int dim, int... args, typename Integrand>
  void AddDomainIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Domain& domain)
  {
    static_assert(dim == 3, "this build integrates over hexahedra (Dimension<3>)");
    static_assert(qoi_traits<Integrand>::registered,
                  "QoI integrand type is not in the compiled registry (see qoi_traits)");
    const std::vector<double> p = qoi_traits<Integrand>::params(integrand);
    detail::check(sx_qoi_add_domain_integral(h_, qoi_traits<Integrand>::id(integrand), p.data(), int(p.size()),
                                             domain.elements.empty() ? nullptr : domain.elements.data(),
                                             int(domain.elements.size())));
  }
  template <int dim, int... args, typename Integrand>
  void AddDomainIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Mesh&)
  {
    AddDomainIntegral(Dimension<dim>{}, DependsOn<args...>{}, integrand, Domain{});
  }

  /// AddBoundaryIntegral (fn/functional_qoi.inl:204-230): measure, position or value over boundary faces
  template <int dim, int... args, typename Integrand>
  void AddBoundaryIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand,
                           const BoundaryDomain& domain)
  {
    static_assert(dim == 2, "boundary integrals on hexahedral meshes are over quadrilateral faces (Dimension<2>)");
    static_assert(qoi_traits<Integrand>::registered,
                  "QoI integrand type is not in the compiled registry (see qoi_traits)");
    const std::vector<double> p = qoi_traits<Integrand>::params(integrand);
    detail::check(sx_qoi_add_boundary_integral(h_, qoi_traits<Integrand>::id(integrand), p.data(), int(p.size()),
                                               domain.face_mask, domain.attr_mask));
  }
  template <int dim, int... args, typename Integrand>
  void AddBoundaryIntegral(Dimension<dim>, DependsOn<args...>, const Integrand& integrand, const Mesh&)
  {
    AddBoundaryIntegral(Dimension<dim>{}, DependsOn<args...>{}, integrand, BoundaryDomain::EntireBoundary());
  }

  /// operator()(t, U) -> q(U) (fn/functional_qoi.inl:380-400)
  double operator()(double t, const Vector& U)
  {
    double v = 0.0;
    detail::check(sx_qoi_eval(h_, t, U.Data(), &v));
    return v;
  }
  /// operator()(t, differentiate_wrt(U)) -> {q(U), Gradient&} (fn/functional_qoi.inl:323-378)
  std::tuple<double, Gradient&> operator()(double t, DifferentiateWRT<Vector> U)
  {
    const double v = (*this)(t, U.value);
    if (grad_.g_.Size() != U.value.Size()) grad_.g_.SetSize(U.value.Size());
    detail::check(sx_qoi_gradient(h_, t, U.value.Data(), grad_.g_.Data()));
    return {v, grad_};
  }

 private:
  sx_qoi*  h_ = nullptr;
  Gradient grad_;
};

// ---------------------------------------------------------------------------
// EquationSolver: Newton + Jacobi-CG / Jacobi-GMRES / multigrid-CG with the
// reference's option names (num/solver_config.hpp:342-400,
// num/equation_solver.cpp:81-235).  Preconditioner::HypreAMG selects the
// geometric multigrid V-cycle (3-component spaces on cuboid meshes), the
// stand-in for BoomerAMG (num/equation_solver.cpp:1014-1021) with CG.
// ---------------------------------------------------------------------------
// Newton: plain Newton; NewtonLineSearch: with the two-way backtracking line
// search (max_line_search_iterations cut-backs); the other solvers of
// num/solver_config.hpp:142-155 (L-BFGS, trust region, KINSOL, PETSc) are not built
enum class NonlinearSolver { Newton, NewtonLineSearch };
enum class LinearSolver { CG, GMRES };
enum class Preconditioner { HypreJacobi, HypreAMG };

struct LinearSolverOptions {
  LinearSolver   linear_solver  = LinearSolver::CG;
  Preconditioner preconditioner = Preconditioner::HypreJacobi;
  double         relative_tol   = 1.0e-8;
  double         absolute_tol   = 1.0e-12;
  int            max_iterations = 300;
  int            gmres_restart  = 50;  // mfem::GMRESSolver's default Krylov dimension
};
// the field order of num/solver_config.hpp:373-400, so aggregate
// initialisers written for the reference fill the same fields
struct NonlinearSolverOptions {
  NonlinearSolver nonlin_solver              = NonlinearSolver::Newton;
  double          relative_tol               = 1.0e-8;
  double          absolute_tol               = 1.0e-12;
  int             min_iterations             = 0;
  int             max_iterations             = 20;
  int             max_line_search_iterations = 0;
  int             print_level                = 0;
};

class EquationSolver {
 public:
  EquationSolver(NonlinearSolverOptions nl = {}, LinearSolverOptions lin = {}) : nl_(nl), lin_(lin) {}
  /// Solve residual(x) = 0 with essential rows zeroed; x holds the essential
  /// values on entry.  Returns the report; throws Error on non-convergence.
  template <typename Form>
  sx_solve_report solve(Form& residual, Vector& x, const std::vector<int32_t>& essential_dofs, int wrt = 0)
  {
    sx_solve_report rep{};
    detail::check(sx_form_set_option(residual.handle(), SX_OPT_LINEAR_SOLVER,
                                     lin_.linear_solver == LinearSolver::GMRES ? SX_LINSOLVER_GMRES : SX_LINSOLVER_CG));
    detail::check(sx_form_set_option(residual.handle(), SX_OPT_GMRES_RESTART, lin_.gmres_restart));
    detail::check(sx_form_set_option(
        residual.handle(), SX_OPT_LINE_SEARCH,
        nl_.nonlin_solver == NonlinearSolver::NewtonLineSearch ? nl_.max_line_search_iterations : 0));
    detail::check(sx_form_set_option(residual.handle(), SX_OPT_MIN_NEWTON_ITERATIONS, nl_.min_iterations));
    detail::check(sx_form_set_option(residual.handle(), SX_OPT_PRINT_LEVEL, nl_.print_level));
    const bool amg = lin_.preconditioner == Preconditioner::HypreAMG;
    if (amg) {
      if (!mg_ || mg_space_ != residual.test_space()) {
        sx_multigrid* m = nullptr;
        detail::check(sx_mg_create(residual.test_space(), &m));
        mg_.reset(m);
        mg_space_ = residual.test_space();
      }
      detail::check(sx_form_set_multigrid(residual.handle(), mg_.get()));
    }
    int rc = sx_newton_solve(residual.handle(), wrt, essential_dofs.data(), int64_t(essential_dofs.size()), x.Data(),
                             nl_.relative_tol, nl_.absolute_tol, nl_.max_iterations, lin_.relative_tol,
                             lin_.absolute_tol, lin_.max_iterations, &rep);
    if (amg) detail::check(sx_form_set_multigrid(residual.handle(), nullptr));
    last_ = rep;
    detail::check(rc);
    return rep;
  }
  const sx_solve_report& report() const { return last_; }

 private:
  struct MgDelete {
    void operator()(sx_multigrid* m) const { sx_mg_destroy(m); }
  };
  NonlinearSolverOptions                      nl_;
  LinearSolverOptions                         lin_;
  sx_solve_report                             last_{};
  std::unique_ptr<sx_multigrid, MgDelete>     mg_;
  sx_space*                                   mg_space_ = nullptr;
};

}  // namespace serac_amd
