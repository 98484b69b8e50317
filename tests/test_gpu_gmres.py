"""GMRES(m) with left Jacobi preconditioning (mfem::GMRESSolver as
num/equation_solver.cpp:916-919 builds it) on the GPU: against a direct sparse
solve of the assembled, BC-eliminated matrix, against a numpy restatement of
the same Krylov method on that matrix (iteration counts), on the non-symmetric
J2 tangent under geometric nonlinearity, and inside the Newton loop."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_j2 import J2_PARAMS, J2Pair, relerr

pytestmark = pytest.mark.gpu


def eliminated(A, ess):
    """J_bc: rows and columns of ess replaced by the identity."""
    keep = np.ones(A.shape[0])
    keep[ess] = 0.0
    K = sp.diags(keep)
    return (K @ A @ K + sp.diags(1.0 - keep)).tocsr()


def gmres_reference(A, b, rel, m, maxit):
    """Left-Jacobi GMRES(m) from x0 = 0 on a scipy matrix: iterations used."""
    d = A.diagonal()
    x = np.zeros_like(b)
    it, goal = 0, None
    while True:
        z = (b - A @ x) / d
        beta = np.linalg.norm(z)
        goal = rel * beta if goal is None else goal
        if beta <= goal or it >= maxit:
            return x, it
        V = [z / beta]
        H = np.zeros((m + 1, m))
        for j in range(m):
            w = (A @ V[j]) / d
            for _ in range(2):
                h = np.array([v @ w for v in V])
                w = w - sum(hi * v for hi, v in zip(h, V))
                H[: j + 1, j] += h
            H[j + 1, j] = np.linalg.norm(w)
            V.append(w / H[j + 1, j])
            it += 1
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y, *_ = np.linalg.lstsq(H[: j + 2, : j + 1], e1, rcond=None)
            if np.linalg.norm(e1 - H[: j + 2, : j + 1] @ y) <= goal or it >= maxit:
                break
        x = x + np.array(V[: len(y)]).T @ y


@pytest.mark.parametrize("restart", [50, 8])
def test_gmres_heat_matches_direct_solve_and_reference_iterations(gpu_lib, restart):
    m = sx.Mesh.cuboid(6, 6, 6)
    m.perturb(42, 0.2)
    s = sx.Space(m, 2, 1, 1)
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    T = W.nodal_field(s.nnodes, 1, 3, 2000, -1.0, 1.0, 1)
    _, grad = f(0.0, sx.DeviceVector.from_numpy(T), wrt=0)
    ess = s.boundary_dofs(0b000001, 0x1)
    b = W.nodal_field(s.nnodes, 1, 4, 2001, -1.0, 1.0, 1)
    b[ess] = 0.0
    x = sx.DeviceVector(s.ndof)
    rep = grad.gmres(ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-10, restart=restart)
    A = eliminated(grad.assemble().to_scipy(), ess)
    x_direct = spla.spsolve(A.tocsc(), b)
    assert rep["converged"]
    assert relerr(x.numpy(), x_direct) < 1e-8
    # the true preconditioned residual meets the goal
    d = A.diagonal()
    assert np.linalg.norm((b - A @ x.numpy()) / d) <= 1e-10 * np.linalg.norm(b / d) * 1.01
    _, it_ref = gmres_reference(A, b, 1e-10, restart, 10000)
    assert abs(rep["linear_iterations"] - it_ref) <= 2, (rep["linear_iterations"], it_ref)


def test_fgmres_jacobi_heat_matches_direct_solve(gpu_lib):
    # flexible GMRES (right preconditioning) with plain Jacobi as its preconditioner
    m = sx.Mesh.cuboid(6, 6, 6)
    m.perturb(42, 0.2)
    s = sx.Space(m, 2, 1, 1)
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    _, grad = f(0.0, sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 1, 3, 2000, -1.0, 1.0, 1)), wrt=0)
    ess = s.boundary_dofs(0b000001, 0x1)
    b = W.nodal_field(s.nnodes, 1, 4, 2001, -1.0, 1.0, 1)
    b[ess] = 0.0
    A = eliminated(grad.assemble().to_scipy(), ess)
    for restart in (50, 8):
        x = sx.DeviceVector(s.ndof)
        rep = grad.fgmres(None, ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-12, restart=restart)
        assert rep["converged"]
        assert np.linalg.norm(b - A @ x.numpy()) <= 1.01e-12 * np.linalg.norm(b)
        assert relerr(x.numpy(), spla.spsolve(A.tocsc(), b)) < 1e-9


def test_gmres_solves_the_nonsymmetric_j2_tangent(gpu_lib):
    pr = J2Pair(3, 2, J2_PARAMS["c4"], state_seed=4)
    U = pr.shear_field(0.02, 1e-3, 15)
    _, grad = pr.gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    A0 = grad.assemble().to_scipy()
    asym = abs(A0 - A0.T).max() / abs(A0).max()
    assert asym > 1e-6  # geometric nonlinearity: the J2 tangent is not symmetric
    ess = pr.gs.boundary_dofs(0b010000, 0x7)
    b = W.nodal_field(pr.gs.nnodes, 3, 5, 2002, -1.0, 1.0, 1)
    b[ess] = 0.0
    x = sx.DeviceVector(pr.gs.ndof)
    rep = grad.gmres(ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-12, max_iter=20000)
    A = eliminated(A0, ess)
    assert rep["converged"]
    assert relerr(x.numpy(), spla.spsolve(A.tocsc(), b)) < 1e-8


def test_newton_with_gmres_matches_cg_newton(gpu_lib):
    out = []
    for solver in (sx.capi.SX_LINSOLVER_CG, sx.capi.SX_LINSOLVER_GMRES):
        pr = J2Pair(2, 2, J2_PARAMS["c4"], perturb=0.1)
        pr.gf.set_option(sx.capi.SX_OPT_LINEAR_SOLVER, solver)
        bot = pr.gs.boundary_dofs(0b010000, 0x7)
        top = pr.gs.boundary_dofs(0b100000, 0x7)
        ess = np.union1d(bot, top).astype(np.int32)
        target = np.where(np.isin(ess, pr.gs.boundary_dofs(0b100000, 0x1)), 0.006, 0.0)
        xg = sx.DeviceVector(pr.gs.ndof)
        reps = sx.quasistatic_solve(pr.gf, xg, ess, target, 2, warm_start=True, rel_tol=1e-10, lin_rel_tol=1e-12,
                                    lin_max_iter=20000)
        assert all(r["converged"] for r in reps)
        out.append((xg.numpy(), pr.qd.numpy()))
    assert relerr(out[1][0], out[0][0]) < 1e-8
    assert np.abs(out[1][1] - out[0][1]).max() <= 1e-8 * np.abs(out[0][1]).max()


def test_gmres_options_are_checked(gpu_lib):
    s = sx.Space(sx.Mesh.cuboid(2, 2, 2), 1, 1, 1)
    f = sx.Functional(s, [s])
    with pytest.raises(sx.SeracAmdError):
        f.set_option(sx.capi.SX_OPT_LINEAR_SOLVER, 7)
    with pytest.raises(sx.SeracAmdError):
        f.set_option(sx.capi.SX_OPT_GMRES_RESTART, 0)
