/*
 * serac_amd.h -- C-ABI of the MI355X (gfx950) matrix-free residual / J.v path.
 *
 * Every entry point returns SX_OK (0) or an error code; sx_last_error() gives a
 * thread-local message.  Vectors passed to compute entry points are DEVICE
 * pointers (use sx_malloc / sx_memcpy_* when the caller has no HIP runtime of
 * its own).  Work is enqueued on the form's HIP stream and is asynchronous
 * unless stated otherwise; sx_form_synchronize() waits for it.
 *
 * Reference interfaces each entry point replaces are cited as
 * path:line relative to the LLNL/serac checkout, with
 *   fn/ = src/serac/numerics/functional/, num/ = src/serac/numerics/,
 *   phy/ = src/serac/physics/.
 */
#ifndef SERAC_AMD_H
#define SERAC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sx_mesh  sx_mesh;
typedef struct sx_space sx_space;
typedef struct sx_form  sx_form;
typedef struct sx_grad  sx_grad;
typedef struct sx_qdata sx_qdata;
typedef struct sx_csr   sx_csr;

/* status codes */
enum {
  SX_OK              = 0,
  SX_ERR_ARG         = 1, /* invalid argument (the reference aborts via SLIC_ERROR) */
  SX_ERR_HIP         = 2, /* HIP runtime error (message carries hipGetErrorString) */
  SX_ERR_UNSUPPORTED = 3, /* combination not compiled (order, qfunction, precision) */
  SX_ERR_STATE       = 4, /* call out of order (e.g. J.v before a linearization) */
  SX_ERR_NOCONV      = 5  /* solver did not converge (results still written) */
};

/* L-vector orderings, fn/element_restriction.cpp:425-432 */
enum { SX_ORDERING_BYNODES = 0, SX_ORDERING_BYVDIM = 1 };

/* compiled quadrature-point functions (the integrand registry) */
enum {
  SX_QF_HEAT           = 1, /* params {k, s, rho c_p}: source -s, flux k grad T (rho c_p read
                               with two trial arguments only: source rho c_p dT/dt - s)
                               phy/heat_transfer.hpp:381-406,493-503 +
                               phy/materials/thermal_material.hpp:50-55 */
  SX_QF_MASS_DIFFUSION = 2, /* params {a, b}: {a u, b grad u}
                               fn/tests/functional_comparisons.cpp:32-33 */
  SX_QF_NEOHOOKEAN     = 3, /* params {density, K, G}: MaterialStressFunctor<NeoHookean>,
                               geometric nonlinearity on, quasi-static
                               phy/solid_mechanics.hpp:834-852,
                               phy/materials/solid_material.hpp:112-122 */
  SX_QF_LINEAR_ELASTIC = 4, /* params {lambda, mu, density}: sigma = lambda tr(eps) I + 2 mu eps
                               (density read with two trial arguments only) */
  SX_QF_J2_LINEAR      = 5  /* params {E, nu, sigma_y, Hi, Hk, density, geom_nonlin (0/1)}:
                               MaterialStressFunctor<J2SmallStrain<LinearHardening>>,
                               quasi-static; state of 10 doubles per quadrature point
                               {plastic strain [3][3] row-major, accumulated plastic strain}
                               phy/materials/solid_material.hpp:132-148,199-258,
                               phy/solid_mechanics.hpp:834-852 */
};

/* compiled boundary (face) qfunctions; the source they return is multiplied
 * by |dX/dxi1 x dX/dxi2| and integrated against the test function values
 * (fn/boundary_integral_kernels.hpp:106-205) */
enum {
  SX_BQF_TRACTION = 101, /* params {A (3x3 row-major), b (3)}: setTraction with
                            traction(X, n) = A n + b, source = -(A n_hat + b)
                            phy/solid_mechanics.hpp:1023-1037 */
  SX_BQF_PRESSURE = 102, /* params {p, geom_nonlin (0/1)}: setPressure, source =
                            p (dx/dxi1 x dx/dxi2) / |dX/dxi1 x dX/dxi2| with x = X + u
                            when geom_nonlin, else x = X  phy/solid_mechanics.hpp:1071-1101 */
  SX_BQF_FLUX     = 103  /* params {q0, h, T_inf}, scalar: setFluxBCs with
                            flux = q0 + h (T - T_inf) (constant flux or Robin)
                            phy/heat_transfer.hpp:536-552 */
};

#define SX_MAX_PARAMS 16

/* ---- runtime helpers (no reference counterpart: device memory plumbing) ---- */
const char* sx_last_error(void);
int         sx_init(int device);                 /* select the device; idempotent */
int         sx_device_name(char* buf, size_t n); /* e.g. "AMD Instinct MI355X gfx950" */
int         sx_malloc(void** ptr, size_t bytes);
int         sx_free(void* ptr);
int         sx_memcpy_htod(void* dst, const void* src, size_t bytes);
int         sx_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int         sx_memcpy_dtod(void* dst, const void* src, size_t bytes);
int         sx_memset_zero(void* ptr, size_t bytes);
int         sx_device_synchronize(void);

/* ---- meshes ----
 * Replaces buildCuboidMesh -> mfem::Mesh::MakeCartesian3D
 * (src/serac/mesh/mesh_utils.cpp:149-153): N^3 hexes on [0,lx]x[0,ly]x[0,lz]
 * with trilinear geometry, lexicographic vertex and element ids. */
int sx_mesh_create_cuboid(int nx, int ny, int nz, double lx, double ly, double lz, sx_mesh** out);
/* General hex mesh: coords [nverts][3]; hex_vertices [nelem][8] in mfem's
 * hexahedron vertex order (the order of the .mesh files in data/meshes). */
int sx_mesh_create(int nverts, const double* coords, int nelem, const int* hex_vertices, sx_mesh** out);
/* Moves interior vertices of a cuboid mesh by h_d * U[-amp, amp] per
 * coordinate (counter-based RNG keyed by seed, vertex id and coordinate). */
int sx_mesh_perturb(sx_mesh* mesh, uint64_t seed, double amp);
/* Replaces buildMeshFromFile -> mfem::Mesh(std::istream&)
 * (src/serac/mesh/mesh_utils.cpp:19-43) for MFEM mesh v1.0 text (the format of
 * the .mesh files under data/meshes): 3D hexahedra (geometry 5), quadrilateral boundary
 * elements (geometry 3), straight (vertex) geometry.  Element and boundary
 * attributes are kept.  Curved meshes (`nodes`) and other element types give
 * SX_ERR_UNSUPPORTED; malformed text gives SX_ERR_ARG.  Cuboid meshes carry
 * mfem's Cartesian boundary attributes: 1 z=0, 2 y=0, 3 x=max, 4 y=max,
 * 5 x=0, 6 z=max. */
int sx_mesh_read_mfem(const char* text, int64_t nbytes, sx_mesh** out);
int sx_mesh_read_mfem_file(const char* path, sx_mesh** out);
int sx_mesh_attributes(const sx_mesh* mesh, int32_t* attrs_out);                    /* host [nelem] */
int sx_mesh_boundary_info(const sx_mesh* mesh, int64_t* nbdr);
int sx_mesh_boundary(const sx_mesh* mesh, int32_t* quads_out, int32_t* attrs_out); /* host [nbdr][4], [nbdr] */
int sx_mesh_info(const sx_mesh* mesh, int64_t* nverts, int64_t* nelem);
int sx_mesh_vertices(const sx_mesh* mesh, double* coords_out); /* host [nverts][3] */
int sx_mesh_destroy(sx_mesh* mesh);

/* ---- H1 spaces on hexes ----
 * Replaces mfem::ParFiniteElementSpace + serac::ElementRestriction
 * (fn/element_restriction.cpp:216-277,425-467): per-element node lists in
 * lexicographic order, E layout [element][component][node]. order 1..4. */
int sx_space_create_h1_hex(sx_mesh* mesh, int order, int components, int ordering, sx_space** out);
/* The same space with the caller's node numbering: elem_nodes[e][j] (host,
 * [nelem][(order+1)^3]) is the node of element e's lexicographic node j, the
 * order mfem's ElementRestriction builds from the FE space's element vdofs with
 * its lexicographic permutation (fn/element_restriction.cpp:216-277 GetElementVDofs
 * + lexicographic_permutations :18-59, :425-467), so a serac binding passes
 * mfem's own numbering and the L-vectors are serac's T-vectors on one rank.
 * The map must relabel the mesh's conforming order-p nodes one to one
 * (checked exactly; SX_ERR_ARG names the first element and node that do not).
 * Everything defined on spaces works on it, except multigrid and the learned
 * preconditioner, which need the library's own lexicographic grid numbering
 * (SX_ERR_UNSUPPORTED). */
int sx_space_create_h1_hex_map(sx_mesh* mesh, int order, int components, int ordering, int64_t nnodes,
                               const int32_t* elem_nodes, sx_space** out);
int sx_space_info(const sx_space* s, int64_t* nnodes, int64_t* ndofs, int64_t* nelem, int* nodes_per_elem);
int sx_space_element_nodes(const sx_space* s, int32_t* map_out); /* host [nelem][nodes_per_elem] */
int sx_space_node_coords(const sx_space* s, double* xyz_out);    /* host [nnodes][3] */
/* Essential DOFs on the boundary faces selected by face_mask
 * (bit 0 x=0, 1 x=max, 2 y=0, 3 y=max, 4 z=0, 5 z=max) of a cuboid mesh,
 * for the components selected by comp_mask. Writes up to cap sorted DOF ids. */
int sx_space_boundary_dofs(const sx_space* s, int face_mask, int comp_mask, int32_t* dofs_out, int64_t cap,
                           int64_t* count);
/* Essential DOFs on the boundary elements whose attribute a has bit a-1 set
 * in attr_mask (the essential boundary attribute sets of serac's
 * setDisplacementBCs / setTemperatureBCs, phy/solid_mechanics.hpp), for the
 * components in comp_mask; works on any hex mesh with boundary elements. */
int sx_space_boundary_dofs_attr(const sx_space* s, uint64_t attr_mask, int comp_mask, int32_t* dofs_out, int64_t cap,
                                int64_t* count);
/* ElementRestriction::Gather / ScatterAdd (fn/element_restriction.cpp:443-467),
 * device pointers, bit-exact with the reference's serial loops. */
int sx_space_gather(const sx_space* s, const double* L, double* E);
int sx_space_scatter_add(const sx_space* s, const double* E, double* L);
int sx_space_destroy(sx_space* s);

/* ---- quadrature-point state: QuadratureData<State> (fn/quadrature_data.hpp:111-167) ---- */
/* nelem x nq records of `size` doubles, each set to init[0..size) */
int sx_qdata_create(int64_t nelem, int nq, int size, const double* init, sx_qdata** out);
/* out[3] = {nelem, nq, size} */
int sx_qdata_info(const sx_qdata* q, int64_t* out);
/* host copies in the reference's order [elem][q][size]; synchronize the device */
int sx_qdata_get(const sx_qdata* q, double* host);
int sx_qdata_set(sx_qdata* q, const double* host);
int sx_qdata_destroy(sx_qdata* q);

/* ---- forms: Functional<test(trials...)> (fn/functional.hpp:189-745) ---- */
/* constructor, fn/functional.hpp:219-273 (single rank: P = identity) */
int sx_form_create(const sx_space* test, const sx_space* const* trials, int ntrials, sx_form** out);
/* AddDomainIntegral, fn/functional.hpp:286-317 (+ fn/integral.hpp:184-260).
 * elems == NULL / nelem == 0 means the whole mesh (EntireDomain).
 * depends_on holds one trial index (DependsOn<i>), or two (DependsOn<a, b>)
 * for the integrands that read a rate as their second argument, as serac's
 * physics modules add them: the solid materials (SX_QF_NEOHOOKEAN,
 * SX_QF_LINEAR_ELASTIC, SX_QF_J2_LINEAR) with the acceleration, source
 * density * a (MaterialStressFunctor, phy/solid_mechanics.hpp:834-852), and
 * SX_QF_HEAT with the temperature rate, source rho c_p * dT/dt - s
 * (ThermalMaterialIntegrand, phy/heat_transfer.hpp:380-406, 432-436).  The
 * density is params[0] (Neo-Hookean), params[5] (J2) or params[2] (linear
 * elastic); rho c_p is params[2] (heat).  Gradients in either argument, J.v,
 * diagonals and assembly follow; the E-vector entry points refuse such an
 * integral (SX_ERR_UNSUPPORTED).  The index k of the E-vector entry points
 * counts the integrals as added. */
int sx_form_add_domain_integral(sx_form* f, int qfunc_id, const double* params, int nparams, const int* depends_on,
                                int ndep, const int* elems, int nelem);
/* AddBoundaryIntegral, fn/functional.hpp:328-354 (fn/integral.hpp:275-347).
 * Integrates over the faces that belong to one element, selected by
 * face_mask among the planes of the mesh's bounding box (bit 0 x = min,
 * 1 x = max, 2 y = min, 3 y = max, 4 z = min, 5 z = max); face_mask = 0 is the
 * whole boundary (EntireBoundary).  qfunc_id is one of SX_BQF_*. */
int sx_form_add_boundary_integral(sx_form* f, int qfunc_id, const double* params, int nparams,
                                  const int* depends_on, int ndep, int face_mask);
/* The same on the boundary elements whose attribute a has bit a-1 set in
 * attr_mask (Domain::ofBoundaryElements by attribute).  Each selected boundary
 * element must be an exterior face of the mesh. */
int sx_form_add_boundary_integral_attr(sx_form* f, int qfunc_id, const double* params, int nparams,
                                       const int* depends_on, int ndep, uint64_t attr_mask);
/* The same on the exterior faces a caller predicate selects:
 * Domain::ofBoundaryElements(mesh, predicate) (fn/domain.hpp:113-118,
 * fn/domain.cpp:312-386).  The library calls pred(corners, attr, user) on the
 * host once per exterior face (a face that belongs to one element), in
 * (element, local face) order: corners [4][3] are the face's vertex
 * coordinates, attr the attribute of the mesh's boundary element on that face
 * (-1 when it has none); a nonzero return selects the face. */
typedef int (*sx_face_predicate)(const double* corners, int attr, void* user);
int sx_form_add_boundary_integral_pred(sx_form* f, int qfunc_id, const double* params, int nparams,
                                       const int* depends_on, int ndep, sx_face_predicate pred, void* user);
/* The same on an explicit list of exterior faces, each given by its four
 * vertex ids (in any order): face_vertices [nfaces][4], host.  A face that is
 * not an exterior face of the mesh is refused (SX_ERR_ARG). */
int sx_form_add_boundary_integral_faces(sx_form* f, int qfunc_id, const double* params, int nparams,
                                        const int* depends_on, int ndep, const int32_t* face_vertices, int64_t nfaces);
/* AddDomainIntegral with a QuadratureData argument (fn/functional.hpp:286-317,
 * fn/quadrature_data.hpp:111-167).  qdata must hold one record per element of
 * the domain (in domain order) and quadrature point; it is borrowed and must
 * outlive the form.  Required for qfunctions with state, refused otherwise. */
int sx_form_add_domain_integral_qdata(sx_form* f, int qfunc_id, const double* params, int nparams,
                                      const int* depends_on, int ndep, const int* elems, int nelem, sx_qdata* qdata);
/* updateQdata(bool), fn/functional.hpp:523: while set, every evaluation writes
 * the qfunction's updated state back to its quadrature data. */
int sx_form_update_qdata(sx_form* f, int flag);
/* operator()(t, U...) / operator()(t, differentiate_wrt(U)), fn/functional.hpp:446-513.
 * inputs: ntrials device pointers (T = L on one rank). wrt = -1: residual only;
 * wrt = i: also stores the quadrature-point linearization w.r.t. trial i and
 * returns its gradient operator in *grad_out (owned by the form).
 * wrt = SX_WRT_SHAPE: differentiate_wrt on the shape argument of a
 * ShapeAwareFunctional (fn/shape_aware_functional.hpp:292-577): returns the
 * gradient in the shape displacement p (sx_form_set_shape_displacement; zero
 * when none) at these inputs, whose sx_grad_apply(g, dp, df) gives
 * df = dR/dp dp, recomputed per application by forward-mode duals through
 * the geometry (no stored tangent).  Domain integrals of 3-component forms;
 * diagonal, assembly and the solvers refuse it (SX_ERR_UNSUPPORTED). */
#define SX_WRT_SHAPE (-2)
int sx_form_eval(sx_form* f, double t, const double* const* inputs, double* out, int wrt, sx_grad** grad_out);
/* Form options (no reference counterpart: the reference runs serially).
 *   SX_OPT_SCATTER: SX_SCATTER_COLORED (default) launches one kernel per
 *     element colour and adds into the output with plain read-modify-writes:
 *     results are bitwise reproducible.  SX_SCATTER_ATOMIC launches all
 *     elements at once and adds with FP64 atomics: fewer L-vector passes,
 *     last-bit rounding depends on the order the hardware applies them.
 *     Must be set before the first AddDomainIntegral.
 *   SX_OPT_XCD_REMAP: 1 gives each of the 8 XCDs a contiguous run of element
 *     blocks (L2 locality for shared nodes); 0 (default) keeps the hardware's
 *     round-robin order, so the XCDs work on neighbouring elements at the same
 *     time (faster for every J.v variant measured with the address-order
 *     scatter). */
enum {
  SX_OPT_SCATTER           = 1,
  SX_OPT_XCD_REMAP         = 2,
  SX_OPT_TANGENT_PRECISION = 3,
  SX_OPT_JVP_ARITHMETIC    = 4,
  SX_OPT_LINEAR_SOLVER     = 5, /* SX_LINSOLVER_CG (default) or _GMRES: the linear solves of the Newton loops */
  SX_OPT_GMRES_RESTART     = 6, /* GMRES restart length (default 50, mfem's GMRESSolver default) */
  SX_OPT_SCATTER_ORDER     = 7, /* 1: the element kernels add each element's values in address
                                   order through LDS (whole rows per wave instruction);
                                   2 (default): the same, after adding the values of nodes shared
                                   with the previous element of the workgroup in LDS (one add per
                                   such node);
                                   0: one component per instruction from the node's own lane */
  SX_OPT_LINE_SEARCH       = 8, /* max_line_search_iterations of NonlinearSolverOptions
                                   (num/solver_config.hpp:389-391, num/equation_solver.cpp:21-185):
                                   0 (default) plain Newton; k > 0 the two-way backtracking line
                                   search of serac's NewtonSolver in the Newton loops */
  SX_OPT_MIN_NEWTON_ITERATIONS = 9, /* min_iterations (num/solver_config.hpp:384-385): the Newton
                                       loops do not stop before this many iterations (default 0) */
  SX_OPT_PRINT_LEVEL       = 10  /* print_level (num/solver_config.hpp:393-394): > 0 makes the Newton
                                    loops write each iteration's residual norm and linear solve to
                                    stderr (default 0: silent) */
};
/* LinearSolver::CG / LinearSolver::GMRES of num/equation_solver.cpp:916-919 */
enum { SX_LINSOLVER_CG = 0, SX_LINSOLVER_GMRES = 1 };
enum { SX_SCATTER_COLORED = 0, SX_SCATTER_ATOMIC = 1 };
/* Precisions.  SX_OPT_TANGENT_PRECISION selects how the next linearization
 * stores the quadrature-point tangent (computed in FP64, rounded once to
 * FP32 or BF16); J.v then streams that format.  SX_OPT_JVP_ARITHMETIC
 * (SX_PREC_F64 default, or SX_PREC_F32) sets the element arithmetic of J.v on
 * FP64 vectors (sx_grad_apply and the solvers): with SX_PREC_F32 the element
 * work runs in FP32 while the vectors, the assembly into them and every
 * Krylov dot product and norm stay FP64 (the precision-split operator of the
 * mixed-precision Newton). */
enum { SX_PREC_F64 = 0, SX_PREC_F32 = 1, SX_PREC_BF16 = 2 };
int sx_form_set_option(sx_form* f, int option, int64_t value);
int sx_form_synchronize(sx_form* f);
int sx_form_stream(sx_form* f, void** hip_stream); /* the form's hipStream_t */
int sx_form_destroy(sx_form* f);

/* ---- gradient operator: Functional::Gradient (fn/functional.hpp:548-669) ---- */
/* Gradient::Mult / operator() (fn/functional.hpp:556-566 -> ActionOfGradient :406-432) */
int sx_grad_apply(sx_grad* g, const double* dx, double* df);
/* The Integral seam (fn/integral.hpp:127-142): domain integral k (in the order
 * added) on E-vectors in the reference layout [element][component][node]
 * (fn/element_restriction.cpp:448, lexicographic nodes) over the integral's
 * domain in the order its elements were given (the whole mesh: element order),
 * without gather or scatter: sx_form_eval_e is eval_func (linearize != 0:
 * the evaluation with AD, storing the integral's tangent for J.v),
 * sx_grad_apply_e is jvp.  Device pointers of nelem * ncomp * nodes_per_elem
 * doubles (sx_form_integral_info).  Not for shape-aware forms. */
int sx_form_integral_info(const sx_form* f, int k, int64_t* nelem, int* nodes_per_elem, int* ncomp);
int sx_form_eval_e(sx_form* f, int k, double t, const double* u_E, double* r_E, int linearize);
int sx_grad_apply_e(sx_grad* g, int k, const double* du_E, double* dr_E);
/* The same action with FP32 vectors and FP32 arithmetic (the precision-split
 * operator of the mixed-precision solvers; no reference counterpart). */
int sx_grad_apply_f32(sx_grad* g, const float* dx, float* df);
/* diagonal of the assembled gradient (the values Gradient::assemble,
 * fn/functional.hpp:569-667, would put on the diagonal), for Jacobi */
int sx_grad_diagonal(sx_grad* g, double* diag);
/* bytes of stored linearization (tangent) held by this gradient */
int sx_grad_tangent_bytes(const sx_grad* g, int64_t* bytes);
/* Gradient::assemble (fn/functional.hpp:569-667 with element_gradient_kernel,
 * fn/domain_integral_kernels.hpp:327-361): the linearization as a sparse
 * matrix, rows = test DOFs, columns = trial DOFs, in CSR form on the device
 * (int64 row offsets, int32 columns sorted within each row, FP64 values).
 * The pattern couples every pair of DOFs that share an element.  With
 * coloured scatter the values are bitwise reproducible.  The caller owns *out
 * (sx_csr_destroy).  Synchronous. */
int sx_grad_assemble(sx_grad* g, sx_csr** out);
int sx_csr_info(const sx_csr* A, int64_t* nrows, int64_t* ncols, int64_t* nnz);
/* device pointers into the matrix (valid until sx_csr_destroy) */
int sx_csr_arrays(const sx_csr* A, const int64_t** row_ptr, const int32_t** col, const double** val);
/* host copies: row_ptr[nrows + 1], col[nnz], val[nnz]; any may be NULL */
int sx_csr_download(const sx_csr* A, int64_t* row_ptr, int32_t* col, double* val);
/* y = A x (device vectors; one wavefront per row, reproducible). Synchronous. */
int sx_csr_spmv(const sx_csr* A, const double* x, double* y);
int sx_csr_destroy(sx_csr* A);
/* Runs nrep applications back to back on the form's stream, bracketed by HIP
 * events; writes the elapsed milliseconds of the whole batch and of the
 * element kernels alone (sum over launches). vec_precision: SX_PREC_F64
 * (double vectors) or SX_PREC_F32 (float vectors). Synchronous. */
int sx_grad_apply_timed(sx_grad* g, int vec_precision, const void* dx, void* df, int nrep, double* ms_total,
                        double* ms_kernels, int* launches_per_apply);

/* ---- solvers (num/equation_solver.cpp:22-240, 893-950, 1018-1021) ---- */
typedef struct sx_solve_report {
  int    newton_iterations;
  int    linear_iterations; /* summed over Newton iterations */
  double initial_norm;
  double final_norm;
  int    converged;
  int    line_search_steps; /* residual evaluations of the line search (SX_OPT_LINE_SEARCH) */
} sx_solve_report;

/* Jacobi-preconditioned CG on the gradient with essential rows/columns
 * eliminated (phy/solid_mechanics.hpp:1133-1139). Solves J_bc x = b. Stopping
 * rule of mfem::CGSolver: (r, z) <= max(rel^2 (r0, z0), abs^2); a negative
 * (r, z) stops unconverged (SX_ERR_NOCONV), as mfem's does. */
int sx_cg_solve(sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x, double rel_tol,
                double abs_tol, int max_iter, sx_solve_report* report);
/* sx_cg_solve without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The five work vectors, the (r, z) partials and the essential DOF list live
 * in a library-owned device array of SX_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends: 40 ndof + 8 * 1024 + 4 ness bytes (each part
 * rounded up to 256) must fit, else SX_ERR_ARG.  The form must have been
 * linearized (its tangent is the form's own, not scratch).  Every call on one
 * GPU shares that array: calls that overlap on one GPU race.  Same results,
 * iteration counts and errors as sx_cg_solve. */
#define SX_NOALLOC_SCRATCH_BYTES (128u << 20)
int sx_cg_solve_noalloc(sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                        double rel_tol, double abs_tol, int max_iter, sx_solve_report* report);
/* Restarted GMRES(restart) with left Jacobi preconditioning on J_bc x = b
 * (x0 = 0), mfem::GMRESSolver + HypreSmoother Jacobi as built by
 * buildLinearSolverAndPreconditioner (num/equation_solver.cpp:916-919,
 * 1018-1021): stops when ||D^-1 (b - J_bc x)|| <= max(rel ||D^-1 b||, abs).
 * For non-symmetric tangents (J2 under geometric nonlinearity). */
int sx_gmres_solve(sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x, double rel_tol,
                   double abs_tol, int max_iter, int restart, sx_solve_report* report);
/* Newton (num/equation_solver.cpp:81-235, no line search) with Jacobi-CG
 * linear solves on R(x) = 0 with essential rows zeroed. x (device) holds the
 * essential values on entry and the solution on exit. Form trial wrt. */
int sx_newton_solve(sx_form* f, int wrt, const int32_t* ess_dofs, int64_t ness, double* x, double rel_tol,
                    double abs_tol, int max_iter, double lin_rel_tol, double lin_abs_tol, int lin_max_iter,
                    sx_solve_report* report);

/* Quasi-static load stepping (phy/solid_mechanics.hpp:1620-1630 quasiStaticSolve,
 * :1727-1770 warmStartDisplacement): the essential values ramp linearly from
 * their values in x (t = 0) to ess_final (host array, t = 1) over nsteps; each
 * step optionally warm-starts with one linear solve on the previous
 * linearization, du = du_bc + J_bc^{-1} (-R(u) - J du_bc), then runs Newton +
 * Jacobi-CG.  After each converged step, integrals with quadrature data get
 * one more residual evaluation with the state committed (as the reference's
 * time step does, phy/solid_mechanics.hpp:1278-1287).  x (device) is updated in
 * place; reports[nsteps] (host) receive per-step Newton / CG iteration counts
 * and residual norms. */
typedef struct sx_solver_options {
  double rel_tol;     /* Newton, relative to the step's initial residual */
  double abs_tol;     /* Newton, absolute */
  int    max_iter;    /* Newton */
  double lin_rel_tol; /* Jacobi-CG */
  double lin_abs_tol;
  int    lin_max_iter;
} sx_solver_options;
int sx_quasistatic_solve(sx_form* f, int wrt, const int32_t* ess_dofs, int64_t ness, const double* ess_final,
                         int nsteps, int warm_start, const sx_solver_options* opts, double* x,
                         sx_solve_report* reports);

/* ---- learned coarse-grid preconditioner: the U-Net (no reference counterpart) ----
 * A 4-level 3D convolutional U-Net (16 -> 32 -> 64 -> 64 channels, 3x3x3
 * kernels, BF16 weights and activations, FP32 accumulation) on a vertex grid
 * of nx x ny x nz nodes (each 2^3 m + 1), mapping 3 channels to 3 channels,
 * channels-last [z][y][x][3].  Parameters: per layer [27][cin][cout] weights
 * then [cout] biases, FP32 masters. */
typedef struct sx_unet sx_unet;
int sx_unet_create(int nx, int ny, int nz, uint64_t seed, sx_unet** out);
int sx_unet_info(const sx_unet* net, int64_t* nparam, int64_t* nvox);
int sx_unet_get_params(const sx_unet* net, float* host);
int sx_unet_set_params(sx_unet* net, const float* host);
int sx_unet_get_grads(const sx_unet* net, float* host);
/* y = N(x); device FP32 [nvox][3] */
int sx_unet_forward(sx_unet* net, const float* x, float* y);
/* forward + backward of loss = |N(x) - t|^2 / |t|^2; gradients accumulate */
int sx_unet_backward(sx_unet* net, const float* x, const float* t, double* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) with the accumulated gradients, which it clears */
int sx_unet_adam(sx_unet* net, double lr);
/* activation of the hidden layers: SX_UNET_RELU (default) or SX_UNET_IDENTITY
 * (a linear network: the map it learns, residual -> correction, is linear);
 * re-initialises the parameters for it (variance gain/fan_in, gain 2 for
 * ReLU, 1 for identity with the output layer scaled by 1e-2) and clears Adam */
enum { SX_UNET_RELU = 0, SX_UNET_IDENTITY = 1 };
/* forward convolutions with BF16 input and output on the matrix cores
 * (implicit GEMM, mfma_f32_16x16x32_bf16; default on) or on the vector units */
int sx_unet_set_mfma(sx_unet* net, int on);
int sx_unet_set_activation(sx_unet* net, int activation, uint64_t seed);
int sx_unet_destroy(sx_unet* net);

/* ---- learned coarse-grid preconditioner (no reference counterpart) ----
 * For a 3-component byVDIM space on a structured (cuboid) mesh of order p:
 *   M(r) = D^-1 r + gain s P N(P^T r / s),   s = rms(P^T r),
 * D the Jacobi diagonal, P the interpolation from the mesh's vertex (Q1) grid
 * to the order-p nodes and N the U-Net above on the vertex grid (each
 * dimension 8 m + 1 vertices).  Used inside flexible CG; Jacobi-CG finishes a
 * solve whose (r, D^-1 r) stalls for `stall_window` iterations.  N trains on
 * pairs recorded from Jacobi-CG solves every `stride` iterations: the input
 * P^T r_k and the target avg(x_{k+window}) - avg(x_k), the correction of the
 * next `window` iterations as a hat-weighted vertex mean (P^T x) / (P^T 1). */
typedef struct sx_learned   sx_learned;
typedef struct sx_multigrid sx_multigrid;
enum { SX_PRECOND_JACOBI = 0, SX_PRECOND_RECORD = 1, SX_PRECOND_LEARNED = 2 };
int sx_learned_create(const sx_space* s, uint64_t seed, sx_learned** out);
int sx_learned_configure(sx_learned* L, int stride, int window, int max_pairs, int stall_window);
/* pairs in the bank, fallbacks to Jacobi-CG so far, the network (owned by L) */
int sx_learned_info(const sx_learned* L, int* npairs, int64_t* fallbacks, sx_unet** net);
/* Jacobi-CG on J_bc x = b (as sx_cg_solve) recording pairs into L */
int sx_learned_record_cg(sx_learned* L, sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                         double rel_tol, double abs_tol, int max_iter, sx_solve_report* report);
/* copy usable pair k (0 <= k < npairs) to host FP32 arrays [nvertex][3]:
 * the restricted residual P^T r_k and the vertex-mean correction */
int sx_learned_get_pair(const sx_learned* L, int k, float* r_host, float* t_host);
/* `npairs` pairs from the linearization g itself: e = P e_c for random smooth
 * vertex fields e_c (a multiscale sum of trilinearly interpolated uniform
 * lattices; zero on the essential DOFs) with r = J_bc e.  e is then the
 * converged correction of residual r, so the pair is (P^T r, avg(e)). */
int sx_learned_generate_pairs(sx_learned* L, sx_grad* g, const int32_t* ess_dofs, int64_t ness, int npairs,
                              uint64_t seed);
/* `steps` Adam steps, each on `batch` random pairs (gradients summed); mean
 * loss of the last tenth.  Sets the target scale `gain` first, the geometric
 * mean over the bank of rms(target) / rms(input): N learns t / (s gain) from
 * P^T r / s, and M(r) = D^-1 r + gain s P N(P^T r / s). */
int sx_learned_train(sx_learned* L, int steps, int batch, double lr, uint64_t seed, double* mean_loss);
int sx_learned_gain(const sx_learned* L, double* gain);
/* overrides the target scale; gain 0 switches the network's correction off,
 * leaving M(r) = B(r) (the ablation of the same solver without the network) */
int sx_learned_set_gain(sx_learned* L, double gain);
/* Attaches a multigrid of the same space (borrowed; NULL detaches): M(r)
 * becomes the V-cycle plus the network's correction, MG(r) + gain s P N(P^T r
 * / s), in place of D^-1 r + ....  Pairs generated afterwards target what the
 * V-cycle leaves of the converged correction, avg(e - MG(J_bc e)); the
 * solves set the hierarchy up per linearization (per Newton call with
 * sx_mg_set_reuse).  No reference counterpart: a learned correction of the
 * multigrid that stands in for Preconditioner::HypreAMG
 * (num/equation_solver.cpp:1014-1021). */
int sx_learned_set_multigrid(sx_learned* L, sx_multigrid* M);
/* z = M(r) for the current linearization (device vectors) */
int sx_learned_apply(sx_learned* L, sx_grad* g, const int32_t* ess_dofs, int64_t ness, const double* r, double* z);
int sx_learned_destroy(sx_learned* L);
/* the linear solves of sx_newton_solve / sx_quasistatic_solve on this form:
 * SX_PRECOND_JACOBI (Jacobi-CG), SX_PRECOND_RECORD (Jacobi-CG recording
 * pairs into L), SX_PRECOND_LEARNED (flexible CG with L).  L is borrowed. */
int sx_form_set_preconditioner(sx_form* f, sx_learned* L, int mode);
/* flexible GMRES(restart) (right preconditioning) on J_bc x = b (x0 = 0) with
 * the learned preconditioner, or with Jacobi when L is NULL; minimises
 * ||b - J_bc x|| down to max(rel ||b||, abs).  The Newton loops use it when
 * the form has SX_PRECOND_LEARNED and SX_OPT_LINEAR_SOLVER = GMRES. */
int sx_fgmres_solve(sx_grad* g, sx_learned* L, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                    double rel_tol, double abs_tol, int max_iter, int restart, sx_solve_report* report);
/* flexible CG with the learned preconditioner on J_bc x = b (x0 = 0) */
int sx_fcg_solve(sx_grad* g, sx_learned* L, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                 double rel_tol, double abs_tol, int max_iter, sx_solve_report* report);

/* ShapeAwareFunctional (fn/shape_aware_functional.hpp:292-577): the domain
 * integrals of f are evaluated on the configuration X + p, p a shape
 * displacement in f's test space (a 3-component device L-vector, borrowed; it
 * must stay valid while f is evaluated; NULL restores X).  At each quadrature
 * point dx/dxi = dX/dxi + dp/dxi, which is the reference's ShapeCorrection
 * (trial gradients du/dX (I + dp/dX)^-1, fluxes f (I + dp/dX)^-T det, sources
 * times det); boundary integrals of 3-component fields see the faces of
 * X + p (position, tangents and area weight, compute_boundary_area_correction
 * of fn/shape_aware_functional.hpp:156-180).  Derivatives are with respect to
 * the trial arguments.  Setting p invalidates earlier linearizations. */
int sx_form_set_shape_displacement(sx_form* f, const double* p);

/* ---- geometric multigrid preconditioner ----
 * Stands where the reference's linear solver options offer hypre's AMG
 * (num/equation_solver.cpp:1022-1051, Preconditioner::HypreAMG): for a
 * 3-component byVDIM space on a structured (cuboid) mesh of order p, a V-cycle
 * over the order-p nodal grid (level 0, the form's matrix-free J_bc) and the
 * mesh's vertex grids (level 1; each further level halves the intervals while
 * every direction has an even count and the level holds more than 200 DOFs).
 * Coarse operators are Galerkin P^T A P (27-point 3x3-block stencils; level 1
 * from the stored element tangents where possible, see sx_mg_set_galerkin,
 * otherwise and above by probing: 81 applications of the finer operator per
 * level), smoothing is
 * Chebyshev-Jacobi of degree `fine_degree` on level 0 and `coarse_degree`
 * above, on [upper / smoothing_range, upper] with upper = 1.15 x a
 * `power_iterations`-step power-iteration estimate of the largest eigenvalue
 * of D^-1 A, and the coarsest level (at most 1536 DOFs) is solved with a
 * dense inverse.  Essential DOFs are eliminated on every level (a vertex
 * component is constrained when its own node's is) and M is the identity on
 * them.  Defaults: degrees 2 and 2, range 12, 12 power iterations. */
int sx_mg_create(const sx_space* s, sx_multigrid** out);
int sx_mg_configure(sx_multigrid* M, int fine_degree, int coarse_degree, double smoothing_range, int power_iterations);
/* inside one sx_newton_solve / sx_quasistatic_solve call, one hierarchy
 * serves up to `linearizations` consecutive linear solves (the smoother's
 * level-0 operator is always the current J unless it uses BF16 copies; the
 * coarse levels, diagonals, eigenvalue bounds and BF16 copies lag).  Default 1: a setup per linearization, as the
 * reference's preconditioners are set up on each SetOperator. */
int sx_mg_set_reuse(sx_multigrid* M, int linearizations);
/* how the Galerkin operators are formed: 0 (default) level 1 element by
 * element from the stored tangent when every trial-dependent part of J is a
 * domain integral with a GRAD_SYM tangent in FP64 or FP32 and the essential
 * DOFs are vertex-compatible (else by probing), and the levels above by the
 * stencil triple product; 1 every level by probing */
int sx_mg_set_galerkin(sx_multigrid* M, int mode);
/* SX_PREC_BF16: the level-0 applications of the V-cycle (smoothing, power
 * iteration, probing) use BF16 copies of the stored tangents, made at each
 * setup (FP64 or FP32 GRAD_SYM / GRAD_FULL tangents only; otherwise the form's
 * own tangent is used); the Krylov solver's J.v keeps the form's tangent.
 * SX_PREC_F64 (default): the form's tangent. */
int sx_mg_set_smoother_precision(sx_multigrid* M, int prec);
/* SX_PREC_F32: the level-0 part of the V-cycle (Chebyshev vectors, the
 * level-0 operator's J.v, the order-p transfers) runs on FP32 vectors
 * (allocated at the next setup); SX_PREC_F64 (default): FP64 vectors */
int sx_mg_set_smoother_vectors(sx_multigrid* M, int prec);
/* number of levels and of setups so far */
int sx_mg_info(const sx_multigrid* M, int* nlevels, int64_t* setups);
/* setups whose level-1 operator was formed element by element */
int sx_mg_element_setups(const sx_multigrid* M, int64_t* n);
/* level l: DOFs, grid points per direction (dims[3]) and the eigenvalue estimate of the last setup */
int sx_mg_level(const sx_multigrid* M, int level, int64_t* ndof, int* dims, double* lmax);
/* builds the hierarchy for the current linearization of g (host DOF list) */
int sx_mg_setup(sx_multigrid* M, sx_grad* g, const int32_t* ess_dofs, int64_t ness);
/* z = M r with the hierarchy of the last sx_mg_setup (device vectors) */
int sx_mg_apply(sx_multigrid* M, const double* r, double* z);
/* y = A_l x on level l of the last setup: level 0 is J with the essential
 * rows zeroed (x must be zero there), level l >= 1 the Galerkin stencil */
int sx_mg_level_apply(sx_multigrid* M, int level, const double* x, double* y);
/* CG on J_bc x = b (x0 = 0) preconditioned by M: the recurrences and the
 * stopping rule (r, M r) <= max(rel^2 (r0, M r0), abs^2) of mfem::CGSolver;
 * sets the hierarchy up first */
int sx_mg_pcg_solve(sx_grad* g, sx_multigrid* M, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                    double rel_tol, double abs_tol, int max_iter, sx_solve_report* report);
/* flexible GMRES(restart) on J_bc x = b (x0 = 0) right-preconditioned by M
 * (for nonsymmetric tangents); minimises ||b - J_bc x|| down to
 * max(rel ||b||, abs); sets the hierarchy up first */
int sx_mg_gmres_solve(sx_grad* g, sx_multigrid* M, const int32_t* ess_dofs, int64_t ness, const double* b, double* x,
                      double rel_tol, double abs_tol, int max_iter, int restart, sx_solve_report* report);
/* the linear solves of sx_newton_solve / sx_quasistatic_solve on this form
 * use M (set up once per linearization): multigrid-CG, or multigrid-FGMRES
 * with SX_OPT_LINEAR_SOLVER = GMRES; NULL returns to Jacobi.  M is borrowed:
 * detach it before sx_mg_destroy. */
int sx_form_set_multigrid(sx_form* f, sx_multigrid* M);
int sx_mg_destroy(sx_multigrid* M);

/* ---- quantities of interest: Functional<double(trial)> (fn/functional_qoi.inl:59-477) ----
 * q(U) = sum over domain integrals of  int q(X, u, grad u) dV  on one trial
 * space, with the compiled integrands below; the value is summed element by
 * element in domain order (reproducible).  sx_qoi_gradient writes dq/dU (the
 * HypreParVector of Gradient::assemble, fn/functional_qoi.inl:422); the
 * directional derivative Gradient::operator()(dU) is its dot product with dU. */
typedef struct sx_qoi sx_qoi;
enum {
  SX_QOI_MEASURE           = 1, /* int 1 dV (TrivialIntegrator of fn/tests/functional_qoi.cpp) */
  SX_QOI_POSITION          = 2, /* int X_d dV, params {d}: the first moment (ZeroIndexIntegrator) */
  SX_QOI_VALUE             = 3, /* int u_c dV, params {c} (GetZeroIntegrator) */
  SX_QOI_NEOHOOKEAN_ENERGY = 4, /* int psi(I + grad u) dV, params {density, K, G}: the strain energy whose
                                   first Piola stress is the flux of SX_QF_NEOHOOKEAN */
  SX_QOI_INNER             = 5  /* int u_a . u_b dV of two trial arguments DependsOn<a, b> (a = b: int |u|^2) */
};
int sx_qoi_create(const sx_space* trial, sx_qoi** out);
/* Functional<double(trials...)> with several trial arguments
 * (fn/functional_qoi.inl:307-472); all on one mesh.  The single-trial entry
 * points below act on trial 0 and need ntrials = 1. */
int sx_qoi_create_multi(const sx_space* const* trials, int ntrials, sx_qoi** out);
/* elems == NULL / nelem == 0: the whole mesh */
int sx_qoi_add_domain_integral(sx_qoi* q, int qoi_id, const double* params, int nparams, const int* elems, int nelem);
/* AddDomainIntegral(Dimension<3>, DependsOn<i...>, ...) of a QoI: the trial
 * arguments the integrand reads (one; two for SX_QOI_INNER, whose spaces must
 * have the same order, components, ordering and numbering) */
int sx_qoi_add_domain_integral_dep(sx_qoi* q, int qoi_id, const double* params, int nparams, const int* depends_on,
                                   int ndep, const int* elems, int nelem);
/* AddBoundaryIntegral (fn/functional_qoi.inl:204-230) of SX_QOI_MEASURE,
 * _POSITION or _VALUE over the boundary faces on the bounding-box planes of
 * face_mask (0: the whole boundary) or, when attr_mask != 0, over the
 * boundary elements with those attributes; dA = |dX/dxi1 x dX/dxi2| */
int sx_qoi_add_boundary_integral(sx_qoi* q, int qoi_id, const double* params, int nparams, int face_mask,
                                 uint64_t attr_mask);
/* The same over the exterior faces pred selects (as
 * sx_form_add_boundary_integral_pred: Domain::ofBoundaryElements, fn/domain.cpp:312-386)
 * or over an explicit list of exterior faces ([nfaces][4] vertex ids, host). */
int sx_qoi_add_boundary_integral_pred(sx_qoi* q, int qoi_id, const double* params, int nparams,
                                      sx_face_predicate pred, void* user);
int sx_qoi_add_boundary_integral_faces(sx_qoi* q, int qoi_id, const double* params, int nparams,
                                       const int32_t* face_vertices, int64_t nfaces);
/* Functional<double(trial)>::operator()(t, U); u is a device L-vector */
int sx_qoi_eval(sx_qoi* q, double t, const double* u, double* value);
/* dq/dU at U into grad (device L-vector) */
int sx_qoi_gradient(sx_qoi* q, double t, const double* u, double* grad);
/* operator()(t, U...) and operator()(t, ..., differentiate_wrt(U_wrt), ...)'s
 * gradient for a QoI of several trial arguments: inputs[ntrials] device
 * L-vectors; grad is an L-vector of trial wrt */
int sx_qoi_eval_multi(sx_qoi* q, double t, const double* const* inputs, double* value);
int sx_qoi_gradient_multi(sx_qoi* q, double t, const double* const* inputs, int wrt, double* grad);
int sx_qoi_destroy(sx_qoi* q);

/* ---- vector kernels used by the solvers, exposed for callers ---- */
int sx_vec_dot(const double* a, const double* b, int64_t n, double* result); /* synchronous, deterministic */
int sx_vec_axpy(double alpha, const double* x, double* y, int64_t n);        /* y += alpha x */

#ifdef __cplusplus
}
#endif

#endif /* SERAC_AMD_H */
