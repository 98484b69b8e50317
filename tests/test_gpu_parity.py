"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on the
same seeded inputs.  Integer maps and E-vector gather/scatter are bit-exact;
FP64 residuals, J.v and diagonals agree to 1e-12 relative (max-norm)."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu

RTOL = 1e-12

QF = {
    "heat": (sx.SX_QF_HEAT, O.QF_HEAT, 1, (1.0, 1.0)),
    "mass_diffusion": (sx.SX_QF_MASS_DIFFUSION, O.QF_MASS_DIFFUSION, 1, (1.7, 2.1)),
    "mass_diffusion_vec": (sx.SX_QF_MASS_DIFFUSION, O.QF_MASS_DIFFUSION, 3, (1.7, 2.1)),
    "neohookean": (sx.SX_QF_NEOHOOKEAN, O.QF_NEOHOOKEAN, 3, (1.0, 10.0, 0.25)),
    "linear_elastic": (sx.SX_QF_LINEAR_ELASTIC, O.QF_LINEAR_ELASTIC, 3, (1.2, 0.8)),
}


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


class Pair:
    """The same problem set up on the GPU library and on the oracle."""

    def __init__(self, n, p, nc, perturb=0.2, seed=42, ordering=1, mesh_arrays=None):
        if mesh_arrays is None:
            self.gm = sx.Mesh.cuboid(n, n, n)
            self.om = O.Mesh.cuboid(n, n, n)
            if perturb > 0:
                self.gm.perturb(seed, perturb)
                self.om.perturb(seed, perturb)
        else:
            X, hexes = mesh_arrays
            self.gm = sx.Mesh.from_arrays(X, hexes)
            self.om = O.Mesh.from_arrays(X, hexes)
        self.gs = sx.Space(self.gm, p, nc, ordering)
        self.os = O.Space(self.om, p, nc, ordering)
        self.nc, self.p = nc, p

    def forms(self, qf_name, elems=None):
        gid, oid, nc, params = QF[qf_name]
        assert nc == self.nc
        gf = sx.Functional(self.gs, [self.gs])
        gf.add_domain_integral(gid, params, elems=elems)
        of = O.Form(self.os)
        of.add_domain_integral(oid, params, elems)
        return gf, of


def field(space, scale, seed, stream):
    return W.nodal_field(space.nnodes, space.nc, seed, stream, -scale, scale, space.ordering)


def test_mesh_perturbation_bit_exact(gpu_lib):
    gm, om = sx.Mesh.cuboid(5, 4, 3), O.Mesh.cuboid(5, 4, 3)
    gm.perturb(7, 0.2)
    om.perturb(7, 0.2)
    assert np.array_equal(gm.vertices(), om.vertices())


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("nc,ordering", [(1, 1), (3, 1), (3, 0)])
def test_gather_scatter_bit_exact(gpu_lib, p, nc, ordering):
    pr = Pair(3, p, nc, ordering=ordering)
    assert np.array_equal(pr.gs.element_nodes(), pr.os.emap())
    L = field(pr.gs, 1.0, 3, 2000)
    E_gpu = pr.gs.gather(sx.DeviceVector.from_numpy(L)).numpy()
    E_ref = pr.os.gather(L)
    assert np.array_equal(E_gpu, E_ref)
    L0 = field(pr.gs, 1.0, 4, 2000)
    Lg = sx.DeviceVector.from_numpy(L0)
    pr.gs.scatter_add(sx.DeviceVector.from_numpy(E_ref), Lg)
    assert np.array_equal(Lg.numpy(), pr.os.scatter_add(E_ref, L0))


def test_unstructured_numbering_bit_exact(gpu_lib, patch_mesh):
    X, hexes = patch_mesh
    for p in (1, 2, 3, 4):
        pr = Pair(0, p, 3, mesh_arrays=(X, hexes))
        assert np.array_equal(pr.gs.element_nodes(), pr.os.emap())
        assert np.allclose(pr.gs.node_coords(), pr.os.node_coords(), rtol=0, atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("qf", ["heat", "mass_diffusion", "mass_diffusion_vec", "neohookean", "linear_elastic"])
def test_residual_jvp_diagonal_parity(gpu_lib, p, qf):
    nc = QF[qf][2]
    n = {1: 4, 2: 3, 3: 3, 4: 2}[p]
    pr = Pair(n, p, nc)
    gf, of = pr.forms(qf)
    xyz = pr.gs.node_coords()
    if qf == "neohookean":
        U = W.neohookean_displacement(xyz) * 10.0
    else:
        U = field(pr.gs, 1.0, 11, 2000)
    dU = field(pr.gs, 1.0, 12, 3000)
    Ug, dUg = sx.DeviceVector.from_numpy(U), sx.DeviceVector.from_numpy(dU)

    R_ref = of.residual(U)
    assert relerr(gf(0.0, Ug).numpy(), R_ref) < RTOL

    R_lin = of.linearize(U)
    out, grad = gf(0.0, Ug, wrt=0)
    assert relerr(out.numpy(), R_lin) < RTOL
    assert relerr(grad(dUg).numpy(), of.jvp(dU)) < RTOL
    assert relerr(grad.diagonal().numpy(), of.diagonal()) < RTOL


GOLDEN = {
    # name: (mesh, order, components, qfunction)
    "heat_p2_cube3": ("cube3", 2, 1, sx.SX_QF_HEAT, (1.0, 1.0)),
    "neohookean_p3_cube2": ("cube2", 3, 3, sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25)),
    "neohookean_p2_patch": ("patch", 2, 3, sx.SX_QF_NEOHOOKEAN, (1.0, 1.0, 1.0)),
    "massdiff_p4_cube2": ("cube2", 4, 1, sx.SX_QF_MASS_DIFFUSION, (1.7, 2.1)),
}


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_vectors(gpu_lib, name, patch_mesh):
    import os

    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz"))
    kind, p, nc, qf, params = GOLDEN[name]
    if kind == "patch":
        mesh = sx.Mesh.from_arrays(*patch_mesh)
    else:
        n = int(kind[-1])
        mesh = sx.Mesh.cuboid(n, n, n)
        mesh.perturb(42, 0.2)
    space = sx.Space(mesh, p, nc)
    assert np.array_equal(space.element_nodes(), gold[f"{name}/emap"])
    f = sx.Functional(space, [space])
    f.add_domain_integral(qf, params)
    U = sx.DeviceVector.from_numpy(gold[f"{name}/U"])
    assert relerr(f(0.0, U).numpy(), gold[f"{name}/R"]) < RTOL
    _, grad = f(0.0, U, wrt=0)
    assert relerr(grad(sx.DeviceVector.from_numpy(gold[f"{name}/dU"])).numpy(), gold[f"{name}/JdU"]) < RTOL
    assert relerr(grad.diagonal().numpy(), gold[f"{name}/diag"]) < RTOL


def test_byNODES_ordering_parity(gpu_lib):
    pr = Pair(3, 2, 3, ordering=0)
    gf, of = pr.forms("neohookean")
    U = field(pr.gs, 0.05, 5, 2000)
    dU = field(pr.gs, 1.0, 6, 3000)
    of.linearize(U)
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    assert relerr(out.numpy(), of.residual(U)) < RTOL
    assert relerr(grad(sx.DeviceVector.from_numpy(dU)).numpy(), of.jvp(dU)) < RTOL


def test_domain_subset_parity(gpu_lib):
    pr = Pair(4, 2, 1)
    elems = np.array([0, 3, 5, 17, 18, 40, 63], dtype=np.int32)
    gf, of = pr.forms("heat", elems=elems)
    U = field(pr.gs, 1.0, 5, 2000)
    dU = field(pr.gs, 1.0, 6, 3000)
    of.linearize(U)
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    assert relerr(out.numpy(), of.residual(U)) < RTOL
    assert relerr(grad(sx.DeviceVector.from_numpy(dU)).numpy(), of.jvp(dU)) < RTOL


def test_unstructured_patch_mesh_parity(gpu_lib, patch_mesh):
    X, hexes = patch_mesh
    pr = Pair(0, 3, 3, mesh_arrays=(X, hexes))
    gf, of = pr.forms("neohookean")
    U = field(pr.gs, 0.05, 5, 2000)
    dU = field(pr.gs, 1.0, 6, 3000)
    of.linearize(U)
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    assert relerr(out.numpy(), of.residual(U)) < RTOL
    assert relerr(grad(sx.DeviceVector.from_numpy(dU)).numpy(), of.jvp(dU)) < RTOL


@pytest.mark.parametrize("qf,p", [("heat", 2), ("neohookean", 3), ("mass_diffusion_vec", 1),
                                  ("neohookean", 4), ("neohookean", 2)])
@pytest.mark.parametrize("remap", [0, 1])
@pytest.mark.parametrize("order", [0, 1, 2])  # SX_OPT_SCATTER_ORDER: per lane, address order, merged faces
@pytest.mark.parametrize("ordering", [1, 0])
def test_atomic_scatter_parity(gpu_lib, qf, p, remap, order, ordering):
    nc = QF[qf][2]
    pr = Pair(5, p, nc, ordering=ordering)
    gid, oid, _, params = QF[qf]
    gf = sx.Functional(pr.gs, [pr.gs])
    gf.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_ATOMIC)
    gf.set_option(sx.capi.SX_OPT_XCD_REMAP, remap)
    gf.set_option(sx.capi.SX_OPT_SCATTER_ORDER, order)
    gf.add_domain_integral(gid, params)
    of = O.Form(pr.os)
    of.add_domain_integral(oid, params)
    U = field(pr.gs, 0.005, 5, 2000)  # small enough that no element inverts
    dU = field(pr.gs, 1.0, 6, 3000)
    R_ref = of.linearize(U)
    assert np.isfinite(R_ref).all()
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    assert relerr(out.numpy(), R_ref) < RTOL
    assert relerr(grad(sx.DeviceVector.from_numpy(dU)).numpy(), of.jvp(dU)) < RTOL
    assert relerr(grad.diagonal().numpy(), of.diagonal()) < RTOL


def rel2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# (tangent precision, vector precision): tolerance on ||Jv - Jv_ref|| / ||Jv_ref||
PRECISIONS = {
    ("f64", "f64"): 1e-13,
    ("f32", "f64"): 1e-6,
    ("f32", "f32"): 1e-5,
    ("bf16", "f32"): 1.5e-2,
}
_PREC = {"f64": 0, "f32": 1, "bf16": 2}


@pytest.mark.parametrize("tp,vp", sorted(PRECISIONS))
@pytest.mark.parametrize("qf,p", [("neohookean", 3), ("heat", 2), ("mass_diffusion_vec", 4), ("neohookean", 4)])
def test_precision_split_jvp(gpu_lib, tp, vp, qf, p):
    nc = QF[qf][2]
    pr = Pair(3 if p < 4 else 2, p, nc)
    gid, oid, _, params = QF[qf]
    gf = sx.Functional(pr.gs, [pr.gs])
    gf.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, _PREC[tp])
    gf.add_domain_integral(gid, params)
    of = O.Form(pr.os)
    of.add_domain_integral(oid, params)
    U = field(pr.gs, 0.005, 5, 2000)
    dU = field(pr.gs, 1.0, 6, 3000)
    R_ref = of.linearize(U)
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    assert relerr(out.numpy(), R_ref) < RTOL  # the residual stays FP64
    dtype = np.float32 if vp == "f32" else np.float64
    jv = grad(sx.DeviceVector.from_numpy(dU, dtype=dtype)).numpy().astype(np.float64)
    assert rel2(jv, of.jvp(dU)) < PRECISIONS[(tp, vp)]
    diag_tol = {"f64": 1e-12, "f32": 1e-6, "bf16": 1e-2}[tp]
    assert rel2(grad.diagonal().numpy(), of.diagonal()) < diag_tol


def test_scatter_mode_must_precede_integrals(gpu_lib):
    pr = Pair(2, 1, 1)
    gf = sx.Functional(pr.gs, [pr.gs])
    gf.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    with pytest.raises(sx.SeracAmdError) as e:
        gf.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_ATOMIC)
    assert e.value.code == sx.capi.SX_ERR_STATE


def test_run_twice_bitwise_identical(gpu_lib):
    pr = Pair(4, 3, 3)
    gf, _ = pr.forms("neohookean")
    U = sx.DeviceVector.from_numpy(W.neohookean_displacement(pr.gs.node_coords()))
    dU = sx.DeviceVector.from_numpy(field(pr.gs, 1.0, 6, 3000))
    r1 = gf(0.0, U).numpy()
    r2 = gf(0.0, U).numpy()
    assert np.array_equal(r1, r2)
    _, grad = gf(0.0, U, wrt=0)
    a, b = grad(dU).numpy(), grad(dU).numpy()
    assert np.array_equal(a, b)


def test_jacobi_cg_heat_matches_oracle(gpu_lib):
    pr = Pair(6, 2, 1, perturb=0.0)
    gf, of = pr.forms("heat")
    xyz = pr.gs.node_coords()
    T = W.heat_temperature(xyz)
    ess = pr.gs.boundary_dofs(0b000001, 0x1)  # x = 0 face
    assert np.array_equal(ess, np.sort(np.nonzero(np.abs(xyz[:, 0]) < 1e-12)[0]).astype(np.int32))
    R = of.linearize(T)
    R[ess] = 0.0
    x_ref, rep_ref = of.cg(ess, R, rel=1e-10)
    Tg = sx.DeviceVector.from_numpy(T)
    out, grad = gf(0.0, Tg, wrt=0)
    b = out.numpy()
    b[ess] = 0.0
    x = sx.DeviceVector(pr.gs.ndof)
    rep = grad.cg(ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-10)
    assert rep["converged"] and rep_ref["converged"]
    assert abs(rep["linear_iterations"] - rep_ref["iterations"]) <= 2
    assert relerr(x.numpy(), x_ref) < 1e-8


@pytest.mark.parametrize("warm_start", [0, 1])
def test_quasistatic_load_steps_match_oracle(gpu_lib, warm_start):
    pr = Pair(3, 2, 3, perturb=0.2)
    gf, of = pr.forms("neohookean")
    ess0 = pr.gs.boundary_dofs(0b000001, 0x7)
    ess1 = pr.gs.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    final = np.where(np.isin(ess, ess1), -0.02, 0.0)
    nsteps = 4
    xg = sx.DeviceVector(pr.gs.ndof)
    reps = sx.quasistatic_solve(gf, xg, ess, final, nsteps, warm_start=bool(warm_start), rel_tol=1e-10,
                                lin_rel_tol=1e-10)
    assert all(r["converged"] for r in reps)
    x = np.zeros(pr.gs.ndof)
    for k in range(1, nsteps + 1):
        x[ess] = final * k / nsteps
        x, rep = of.newton(ess, x, rel=1e-10, abs_tol=0.0, lin_rel=1e-10, lin_abs=0.0)
        assert rep["converged"]
    assert relerr(xg.numpy(), x) < 1e-8
    R = of.residual(xg.numpy())
    R[ess] = 0.0
    assert np.linalg.norm(R) < 1e-8 * np.linalg.norm(of.residual(np.zeros(pr.gs.ndof)) + 1e-30) + 1e-12


def test_newton_neohookean_matches_oracle(gpu_lib):
    pr = Pair(3, 2, 3, perturb=0.2)
    gf, of = pr.forms("neohookean")
    xyz = pr.gs.node_coords()
    # clamp x = 0, compress x = 1 by 1% in x (y, z free on that face)
    ess0 = pr.gs.boundary_dofs(0b000001, 0x7)
    ess1 = pr.gs.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    x0 = np.zeros(pr.gs.ndof)
    x0[ess1] = -0.01
    x_ref, rep_ref = of.newton(ess, x0, rel=1e-10, abs_tol=1e-14, lin_rel=1e-10, lin_abs=0.0)
    xg = sx.DeviceVector.from_numpy(x0)
    rep = gf.newton(xg, ess, rel_tol=1e-10, abs_tol=1e-14, lin_rel_tol=1e-10, lin_abs_tol=0.0)
    assert rep["converged"] and rep_ref["converged"]
    assert rep["newton_iterations"] == rep_ref["newton_iterations"]
    assert relerr(xg.numpy(), x_ref) < 1e-9
    # the converged state satisfies the oracle's residual
    R = of.residual(xg.numpy())
    R[ess] = 0.0
    assert np.linalg.norm(R) <= 1e-9 * rep_ref["initial_norm"]
