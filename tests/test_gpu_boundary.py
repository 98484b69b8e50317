"""GPU parity for boundary integrals (AddBoundaryIntegral) through the C-ABI:
residual, linearization, J.v (FP64 and FP32), Jacobi diagonal and CSR
assembly against the CPU oracle, the Neo-Hookean traction patch test, and a
follower-pressure Newton solve."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_assemble import lidx
from test_gpu_parity import relerr
from test_oracle_physics import A_PATCH, B_PATCH

pytestmark = pytest.mark.gpu

CASES = {
    "traction": (sx.SX_BQF_TRACTION, O.BQF_TRACTION, 3, list(np.arange(9.0) / 7.0) + [0.3, -1.1, 2.0]),
    "pressure": (sx.SX_BQF_PRESSURE, O.BQF_PRESSURE, 3, [2.5, 1.0]),
    "pressure_small": (sx.SX_BQF_PRESSURE, O.BQF_PRESSURE, 3, [2.5, 0.0]),
    "robin": (sx.SX_BQF_FLUX, O.BQF_FLUX, 1, [0.7, 1.3, 0.2]),
}
VOLUME = {3: (sx.SX_QF_NEOHOOKEAN, O.QF_NEOHOOKEAN, (1.0, 10.0, 0.25)), 1: (sx.SX_QF_HEAT, O.QF_HEAT, (1.0, 1.0))}


def setup(n, p, case, ordering=1, face_mask=0, volume=True, perturb=0.2):
    gid, oid, nc, params = CASES[case]
    gm, om = sx.Mesh.cuboid(n, n, n), O.Mesh.cuboid(n, n, n)
    gm.perturb(11, perturb)
    om.perturb(11, perturb)
    gs, os_ = sx.Space(gm, p, nc, ordering), O.Space(om, p, nc, ordering)
    gf, of = sx.Functional(gs, [gs]), O.Form(os_)
    if volume:
        vg, vo, vp = VOLUME[nc]
        gf.add_domain_integral(vg, vp)
        of.add_domain_integral(vo, vp)
    gf.add_boundary_integral(gid, params, face_mask=face_mask)
    of.add_boundary_integral(oid, params, face_mask)
    return gs, os_, gf, of, (gm, om)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("p,ordering", [(1, 1), (2, 0), (3, 1)])
def test_boundary_parity(gpu_lib, case, p, ordering):
    nc = CASES[case][2]
    gs, os_, gf, of, _ = setup(2, p, case, ordering)
    # small enough that no element inverts at p = 3 (the volume term is Neo-Hookean)
    U = W.nodal_field(gs.nnodes, nc, 3, 2000, -0.005, 0.005, ordering)
    Ug = sx.DeviceVector.from_numpy(U)
    R_ref = of.residual(U)
    assert np.isfinite(R_ref).all()
    assert relerr(gf(0.0, Ug).numpy(), R_ref) < 1e-12
    out, grad = gf(0.0, Ug, wrt=0)
    assert relerr(out.numpy(), of.linearize(U)) < 1e-12
    dU = W.nodal_field(gs.nnodes, nc, 5, 3000, -1.0, 1.0, ordering)
    ref = of.jvp(dU)
    assert relerr(grad(sx.DeviceVector.from_numpy(dU)).numpy(), ref) < 1e-12
    jv32 = grad(sx.DeviceVector.from_numpy(dU, dtype=np.float32)).numpy()
    assert np.linalg.norm(jv32 - ref) / np.linalg.norm(ref) < 1e-5
    assert relerr(grad.diagonal().numpy(), of.diagonal()) < 1e-12
    # CSR: volume element matrices + boundary face matrices of the oracle
    D = grad.assemble().to_scipy().toarray()
    K = np.zeros_like(D)
    emap = os_.emap()
    for e in range(os_.ne):
        d = np.array([lidx(os_, emap[e, j], c) for c in range(nc) for j in range(os_.npe)])
        K[np.ix_(d, d)] += of.element_matrix(0, e)
    for k, (e, lf) in enumerate(of.faces(0)):
        d = np.array([lidx(os_, emap[e, j], c) for c in range(nc) for j in range(os_.npe)])
        K[np.ix_(d, d)] += of.face_matrix(0, k)
    assert np.abs(D - K).max() <= 1e-12 * np.abs(K).max()


@pytest.mark.parametrize("mask", [0b000010, 0b101010, 0b010101])
def test_boundary_face_selection(gpu_lib, mask):
    gs, os_, gf, of, _ = setup(3, 2, "pressure", face_mask=mask, volume=False)
    U = W.nodal_field(gs.nnodes, 3, 3, 2000, -0.05, 0.05, 1)
    assert relerr(gf(0.0, sx.DeviceVector.from_numpy(U)).numpy(), of.residual(U)) < 1e-12


def test_neohookean_traction_patch_gpu(gpu_lib, patch_mesh):
    _, P = O.qf_eval(O.QF_NEOHOOKEAN, [1.0, 1.0, 1.0], 3, np.zeros(3), A_PATCH.ravel())
    trac = list(np.asarray(P).ravel()) + [0.0, 0.0, 0.0]
    meshes = []
    gm = sx.Mesh.cuboid(3, 3, 3)
    gm.perturb(7, 0.2)
    meshes.append(gm)
    X, hexes = patch_mesh
    meshes.append(sx.Mesh.from_arrays(X, hexes))
    for gm in meshes:
        for p in (1, 2, 3):
            gs = sx.Space(gm, p, 3, 1)
            gf = sx.Functional(gs, [gs])
            gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 1.0, 1.0))
            gf.add_boundary_integral(sx.SX_BQF_TRACTION, trac, face_mask=0b101010)
            x = gs.node_coords()
            U = (x @ A_PATCH.T + B_PATCH).ravel()
            R = gf(0.0, sx.DeviceVector.from_numpy(U)).numpy().reshape(-1, 3)
            free = np.all(x > 1e-12, axis=1)
            assert np.abs(R[free]).max() < 1e-13


def test_follower_pressure_newton_matches_oracle(gpu_lib):
    # a block clamped at z = 0 under follower pressure on z = 1
    CASES["pressure_newton"] = (sx.SX_BQF_PRESSURE, O.BQF_PRESSURE, 3, [0.02, 1.0])
    gs, os_, gf, of, _ = setup(2, 2, "pressure_newton", face_mask=0b100000)
    ess = gs.boundary_dofs(0b010000, 0x7)
    x0 = np.zeros(gs.ndof)
    xg = sx.DeviceVector.from_numpy(x0)
    rep = gf.newton(xg, ess, rel_tol=1e-10, abs_tol=1e-14, lin_rel_tol=1e-12, lin_abs_tol=0.0)
    x_ref, rep_ref = of.newton(ess, x0, rel=1e-10, abs_tol=1e-14, lin_rel=1e-12, lin_abs=0.0)
    assert rep["converged"] and rep_ref["converged"]
    assert rep["newton_iterations"] == rep_ref["newton_iterations"] >= 2
    assert relerr(xg.numpy(), x_ref) < 1e-9
    assert np.abs(x_ref).max() > 1e-3  # the block deformed
