// oracle_capi.cpp -- extern "C" wrapper of the CPU oracle for ctypes.
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker, never by the product library.
#include <chrono>

#include "serac_oracle_form.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;
template <typename F>
int guard(F&& f)
{
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
struct OForm {
  Form form;
  OForm(const Mesh& m, const Space& s) : form(m, s) {}
};
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

double orc_rng_u01(uint64_t seed, uint64_t stream, uint64_t idx) { return rng_u01(seed, stream, idx); }

int orc_gl_rule(int q, double* x, double* w)
{
  return guard([&] { gl_rule(q, x, w); });
}
int orc_gll_nodes(int n, double* x)
{
  return guard([&] { gll_nodes(n, x); });
}
int orc_gll_basis(int n, double x, double* B, double* G)
{
  return guard([&] {
    gll_basis(n, x, B);
    gll_dbasis(n, x, G);
  });
}

// ---- meshes
void* orc_mesh_cuboid(int nx, int ny, int nz, double lx, double ly, double lz)
{
  return new Mesh(cuboid(nx, ny, nz, lx, ly, lz));
}
void* orc_mesh_create(int nv, const double* X, int ne, const int* hex)
{
  auto* m = new Mesh();
  m->nv   = nv;
  m->ne   = ne;
  m->X.assign(X, X + size_t(nv) * 3);
  m->hex.assign(hex, hex + size_t(ne) * 8);
  return m;
}
int orc_mesh_perturb(void* m, uint64_t seed, double amp)
{
  return guard([&] { perturb(*static_cast<Mesh*>(m), seed, amp); });
}
void orc_mesh_info(void* m, int64_t* out)
{
  auto* M = static_cast<Mesh*>(m);
  out[0]  = M->nv;
  out[1]  = M->ne;
}
void orc_mesh_vertices(void* m, double* X)
{
  auto* M = static_cast<Mesh*>(m);
  std::memcpy(X, M->X.data(), M->X.size() * sizeof(double));
}
void orc_mesh_free(void* m) { delete static_cast<Mesh*>(m); }

// ---- spaces
void* orc_space(void* m, int p, int nc, int ordering)
{
  try {
    return new Space(make_space(*static_cast<Mesh*>(m), p, nc, ordering));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_space_info(void* s, int64_t* out)
{
  auto* S = static_cast<Space*>(s);
  out[0]  = S->nnodes;
  out[1]  = S->ndof();
  out[2]  = S->ne;
  out[3]  = S->npe;
}
void orc_space_emap(void* s, int* out)
{
  auto* S = static_cast<Space*>(s);
  std::memcpy(out, S->emap.data(), S->emap.size() * sizeof(int));
}
void orc_space_node_coords(void* m, void* s, double* out)
{
  auto v = node_coords(*static_cast<Mesh*>(m), *static_cast<Space*>(s));
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}
void orc_space_free(void* s) { delete static_cast<Space*>(s); }
void orc_gather(void* s, const double* L, double* E) { gather(*static_cast<Space*>(s), L, E); }
void orc_scatter_add(void* s, const double* E, double* L) { scatter_add(*static_cast<Space*>(s), E, L); }

// ---- element kernels (single element, n nodes and q points per direction)
int orc_interpolate(int n, int q, int nc, const double* X, double* value, double* grad, int brute)
{
  return guard([&] {
    Basis1D b(n, q);
    if (brute)
      interpolate_brute(b, nc, X, value, grad);
    else
      interpolate(b, nc, X, value, grad);
  });
}
int orc_integrate(int n, int q, int nc, const double* source, const double* flux, double* r, int brute)
{
  return guard([&] {
    Basis1D b(n, q);
    if (brute)
      integrate_brute(b, nc, source, flux, r);
    else
      integrate(b, nc, source, flux, r);
  });
}
int orc_geometric_factors(void* m, int Q, double* Xq, double* Jq)
{
  return guard([&] {
    std::vector<double> X, J;
    geometric_factors(*static_cast<Mesh*>(m), Q, X, J);
    std::memcpy(Xq, X.data(), X.size() * sizeof(double));
    std::memcpy(Jq, J.data(), J.size() * sizeof(double));
  });
}
// physical-space qfunction (no geometry): value[nc], grad[nc*3] -> source, flux
int orc_qf_eval(int kind, const double* params, int nc, const double* value, const double* grad, double* source,
                double* flux)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    qf_eval(P, nc, value, grad, source, flux);
  });
}

// as orc_qf_eval for a qfunction with quadrature-point state (in/out)
int orc_qf_eval_state(int kind, const double* params, int nc, const double* value, const double* grad,
                      double* source, double* flux, double* state)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    qf_eval(P, nc, value, grad, source, flux, state);
  });
}
// forward-mode tangent at one point: dflux[a][b] = d flux_a / d grad_b
// ((3 nc) x (3 nc), row-major); the state is read, not updated
int orc_qf_tangent(int kind, const double* params, int nc, const double* value, const double* grad,
                   const double* state, double* dflux)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    dual<9> v[3], g[9], s[3], f[9];
    for (int i = 0; i < nc; i++) v[i] = dual<9>(value[i]);
    for (int k = 0; k < 3 * nc; k++) {
      g[k]             = dual<9>(grad[k]);
      g[k].gradient[k] = 1.0;
    }
    double st[16] = {0};
    const int ss  = qf_state_size(kind);
    if (ss) std::copy_n(state, ss, st);
    qf_eval(P, nc, v, g, s, f, ss ? st : nullptr);
    for (int a = 0; a < 3 * nc; a++)
      for (int b = 0; b < 3 * nc; b++) dflux[a * 3 * nc + b] = f[a].gradient[b];
  });
}

// ---- forms
void* orc_form(void* m, void* s)
{
  return new OForm(*static_cast<Mesh*>(m), *static_cast<Space*>(s));
}
void orc_form_free(void* f) { delete static_cast<OForm*>(f); }
int  orc_form_add_integral(void* f, int kind, const double* params, const int* elems, int nelem)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    std::vector<int> el;
    if (elems && nelem > 0) el.assign(elems, elems + nelem);
    static_cast<OForm*>(f)->form.add_integral(P, el);
  });
}
// AddDomainIntegral with QuadratureData: every point's state starts at state_init
int orc_form_add_integral_state(void* f, int kind, const double* params, const int* elems, int nelem,
                                const double* state_init)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    std::vector<int> el;
    if (elems && nelem > 0) el.assign(elems, elems + nelem);
    static_cast<OForm*>(f)->form.add_integral(P, el, state_init);
  });
}
// AddDomainIntegral of two trial arguments (DependsOn<0, 1>): the integrand
// also reads the second argument's value (qf_eval2); state_init as above, or
// null for qfunctions without state
int orc_form_add_integral2(void* f, int kind, const double* params, const int* elems, int nelem,
                           const double* state_init)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    if (kind != QF_NEOHOOKEAN && kind != QF_J2_LINEAR && kind != QF_LINEAR_ELASTIC && kind != QF_HEAT)
      throw std::runtime_error("this qfunction has no two-argument form");
    std::vector<int> el;
    if (elems && nelem > 0) el.assign(elems, elems + nelem);
    auto& form = static_cast<OForm*>(f)->form;
    form.add_integral(P, el, state_init);
    form.integrals.back().two = true;
  });
}
// residual / linearization in the first argument, with U1 the second argument
int orc_form_eval2(void* f, const double* U, const double* U1, double* R, int linearize)
{
  return guard([&] {
    auto& form = static_cast<OForm*>(f)->form;
    form.U1    = U1;
    form.eval(U, R, linearize != 0);
    form.U1 = nullptr;
  });
}
// AddBoundaryIntegral on the bounding-box planes of face_mask (0: whole boundary)
int orc_form_add_boundary_integral(void* f, int kind, const double* params, int face_mask)
{
  return guard([&] {
    QFParams P;
    P.kind = kind;
    for (int i = 0; i < 16; i++) P.p[i] = params[i];
    static_cast<OForm*>(f)->form.add_boundary_integral(P, face_mask);
  });
}
int orc_form_nfaces(void* f, int bi) { return int(static_cast<OForm*>(f)->form.bintegrals[size_t(bi)].faces.size()); }
// face k of boundary integral bi: out = {element, local face}
void orc_form_face(void* f, int bi, int k, int* out)
{
  const auto& F = static_cast<OForm*>(f)->form.bintegrals[size_t(bi)].faces[size_t(k)];
  out[0]        = F[0];
  out[1]        = F[1];
}
// boundary face matrix over its element's DOFs (after a linearization)
int orc_form_face_matrix(void* f, int bi, int k, double* K)
{
  return guard([&] { static_cast<OForm*>(f)->form.face_matrix(bi, size_t(k), K); });
}
void orc_form_update_qdata(void* f, int flag) { static_cast<OForm*>(f)->form.update_qdata = flag != 0; }
int  orc_form_ssize(void* f, int integral) { return static_cast<OForm*>(f)->form.integrals[size_t(integral)].ssize; }
// state of an integral, [elem-in-domain][q][ssize]
int orc_form_get_qdata(void* f, int integral, double* out)
{
  return guard([&] {
    auto& I = static_cast<OForm*>(f)->form.integrals[size_t(integral)];
    std::memcpy(out, I.state.data(), I.state.size() * sizeof(double));
  });
}
int orc_form_set_qdata(void* f, int integral, const double* in)
{
  return guard([&] {
    auto& I = static_cast<OForm*>(f)->form.integrals[size_t(integral)];
    std::memcpy(I.state.data(), in, I.state.size() * sizeof(double));
  });
}
int orc_form_eval(void* f, const double* U, double* R, int linearize)
{
  return guard([&] { static_cast<OForm*>(f)->form.eval(U, R, linearize != 0); });
}
// residual / linearization over nthreads host threads (node-band scatter)
int orc_form_eval_threads(void* f, const double* U, double* R, int linearize, int nthreads)
{
  return guard([&] { static_cast<OForm*>(f)->form.eval_threads(U, R, linearize != 0, nthreads); });
}
// Functional<double(trials...)> value of a compiled QoI kind over elems
// (all elements when nelem = 0); U1 only for the inner product (kind 5)
int orc_form_qoi(void* f, int kind, const double* params, const int* elems, int nelem, const double* U0,
                 const double* U1, double* value)
{
  return guard([&] {
    std::vector<int> dom(elems, elems + (elems ? nelem : 0));
    *value = static_cast<OForm*>(f)->form.qoi(kind, params, dom, U0, U1);
  });
}
int orc_form_jvp(void* f, const double* dU, double* dR)
{
  return guard([&] { static_cast<OForm*>(f)->form.jvp(dU, dR); });
}
int orc_form_diagonal(void* f, double* diag)
{
  return guard([&] { static_cast<OForm*>(f)->form.diagonal(diag); });
}
int orc_form_dsize(void* f, int integral) { return static_cast<OForm*>(f)->form.integrals[size_t(integral)].dsize; }
// stored derivative of (integral, slot in domain, q): dsize doubles
int orc_form_derivative(void* f, int integral, int64_t slot, int q, double* out)
{
  return guard([&] {
    auto& F = static_cast<OForm*>(f)->form;
    auto& I = F.integrals[size_t(integral)];
    std::memcpy(out, &I.D[(size_t(slot) * F.nq() + q) * I.dsize], sizeof(double) * I.dsize);
  });
}
int orc_form_element_matrix(void* f, int integral, int64_t slot, double* K)
{
  return guard([&] { static_cast<OForm*>(f)->form.element_matrix(integral, size_t(slot), K); });
}
int orc_form_geometry(void* f, double* Xq, double* Jq)
{
  return guard([&] {
    auto& F = static_cast<OForm*>(f)->form;
    std::memcpy(Xq, F.Xq.data(), F.Xq.size() * sizeof(double));
    std::memcpy(Jq, F.Jq.data(), F.Jq.size() * sizeof(double));
  });
}

// Jacobi-PCG on the BC-eliminated stored linearization; report[4] =
// {iterations, initial_norm, final_norm, converged}
int orc_cg(void* f, const int* ess, int ness, const double* b, double* x, double rel, double abs_tol, int maxit,
           double* report)
{
  return guard([&] {
    auto&               F = static_cast<OForm*>(f)->form;
    size_t              n = size_t(F.space.ndof());
    std::vector<char>   is_ess(n, 0);
    for (int i = 0; i < ness; i++) is_ess[size_t(ess[i])] = 1;
    std::vector<double> bb(b, b + n), xx(n);
    auto                rep = pcg_jacobi(F, is_ess, bb, xx, rel, abs_tol, maxit);
    std::memcpy(x, xx.data(), n * sizeof(double));
    report[0] = rep.iterations;
    report[1] = rep.initial_norm;
    report[2] = rep.final_norm;
    report[3] = rep.converged;
  });
}

// Newton + Jacobi-CG; x in/out holds the essential values. report[5] =
// {newton its, linear its, initial norm, final norm, converged}
int orc_newton_ls(void* f, const int* ess, int ness, double* x, double rel, double abs_tol, int maxit,
                  double lin_rel, double lin_abs, int lin_max, int max_ls, double* report)
{
  return guard([&] {
    auto&               F = static_cast<OForm*>(f)->form;
    size_t              n = size_t(F.space.ndof());
    std::vector<char>   is_ess(n, 0);
    for (int i = 0; i < ness; i++) is_ess[size_t(ess[i])] = 1;
    std::vector<double> xx(x, x + n);
    auto                rep = newton(F, is_ess, xx, rel, abs_tol, maxit, lin_rel, lin_abs, lin_max, max_ls);
    std::memcpy(x, xx.data(), n * sizeof(double));
    report[0] = rep.newton_iterations;
    report[1] = rep.total_linear_iterations;
    report[2] = rep.initial_norm;
    report[3] = rep.final_norm;
    report[4] = rep.converged;
    report[5] = rep.line_search_steps;
  });
}
int orc_newton(void* f, const int* ess, int ness, double* x, double rel, double abs_tol, int maxit, double lin_rel,
               double lin_abs, int lin_max, double* report)
{
  return guard([&] {
    auto&               F = static_cast<OForm*>(f)->form;
    size_t              n = size_t(F.space.ndof());
    std::vector<char>   is_ess(n, 0);
    for (int i = 0; i < ness; i++) is_ess[size_t(ess[i])] = 1;
    std::vector<double> xx(x, x + n);
    auto                rep = newton(F, is_ess, xx, rel, abs_tol, maxit, lin_rel, lin_abs, lin_max);
    std::memcpy(x, xx.data(), n * sizeof(double));
    report[0] = rep.newton_iterations;
    report[1] = rep.total_linear_iterations;
    report[2] = rep.initial_norm;
    report[3] = rep.final_norm;
    report[4] = rep.converged;
  });
}

// Wall time (seconds) of `reps` oracle J.v applications (the CPU baseline).
double orc_time_jvp(void* f, const double* dU, double* dR, int reps)
{
  auto& F  = static_cast<OForm*>(f)->form;
  auto  t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps; i++) F.jvp(dU, dR);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// the same with the elements split over nthreads host threads (OpenMP)
double orc_time_jvp_threads(void* f, const double* dU, double* dR, int reps, int nthreads)
{
  auto& F = static_cast<OForm*>(f)->form;
  F.prepare_threads(nthreads);  // the per-thread bands are allocated outside the timed region
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < reps; i++) F.jvp_threads(dU, dR, nthreads);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
