"""Solver guards on the GPU: Jacobi-CG stops unconverged, at the oracle's
iteration and iterate, when the Jacobi preconditioner is not positive
definite (mfem::CGSolver's rule), and every solver entry point rejects
essential DOFs outside [0, ndof) with SX_ERR_ARG before anything reaches the
device."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import capi
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(a):
    gm, om = sx.Mesh.cuboid(3, 3, 3), O.Mesh.cuboid(3, 3, 3)
    gm.perturb(42, 0.2)
    om.perturb(42, 0.2)
    gs, os_ = sx.Space(gm, 2, 1), O.Space(om, 2, 1)
    gf = sx.Functional(gs, [gs])
    gf.add_domain_integral(sx.SX_QF_MASS_DIFFUSION, (a, 1.0))
    of = O.Form(os_)
    of.add_domain_integral(O.QF_MASS_DIFFUSION, (a, 1.0))
    of.linearize(np.zeros(os_.ndof))
    _, grad = gf(0.0, sx.DeviceVector(gs.ndof), wrt=0)
    return gs, gf, grad, of


@pytest.mark.parametrize("a", [-300.0, -700.0])
def test_cg_indefinite_jacobi_stops_like_oracle(gpu_lib, a):
    gs, gf, grad, of = _pair(a)
    b = W.direction(gs.nnodes, 1)
    none = np.array([], dtype=np.int32)
    x_ref, rep_ref = of.cg(none, b, rel=1e-10, maxit=500)
    x = sx.DeviceVector.from_numpy(np.full(gs.ndof, 7.0))
    rep = grad.cg(none, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-10, max_iter=500, raise_on_noconv=False)
    assert not rep_ref["converged"] and not rep["converged"]
    assert rep["linear_iterations"] == rep_ref["iterations"]
    xg = x.numpy()
    if a == -700.0:
        assert not xg.any()  # x = 0, as mfem returns it
    else:
        assert np.abs(xg - x_ref).max() <= 1e-10 * np.abs(x_ref).max()
    with pytest.raises(sx.SeracAmdError) as e:
        grad.cg(none, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-10, max_iter=500)
    assert e.value.code == capi.SX_ERR_NOCONV


@pytest.mark.parametrize("bad", [-1, "ndof"])
def test_solvers_reject_out_of_range_essential_dofs(gpu_lib, bad):
    gs, gf, grad, _ = _pair(1.7)
    n = gs.ndof
    ess = np.array([0, n if bad == "ndof" else bad], dtype=np.int32)
    b = sx.DeviceVector.from_numpy(W.direction(gs.nnodes, 1))
    x = sx.DeviceVector(n)
    calls = [
        lambda: grad.cg(ess, b, x),
        lambda: grad.gmres(ess, b, x),
        lambda: gf.newton(x, ess),
        lambda: sx.quasistatic_solve(gf, x, ess, np.zeros(len(ess)), 2),
    ]
    for call in calls:
        with pytest.raises(sx.SeracAmdError) as e:
            call()
        assert e.value.code == capi.SX_ERR_ARG
        assert "out of range" in str(e.value)
    # a valid list still solves
    rep = grad.cg(np.array([0], dtype=np.int32), b, x)
    assert rep["converged"]
