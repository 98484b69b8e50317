"""Newton with the two-way backtracking line search of serac's NewtonSolver
(num/equation_solver.cpp:21-185, max_line_search_iterations) on the GPU
against the oracle's restatement of it.  In these large-deformation cases
plain Newton steps invert elements (a NaN residual, no convergence, in both
implementations); with the line search both converge, taking the same Newton
iterations and line-search steps, to the same solution."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import capi

from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

NH = (1.0, 10.0, 0.25)
# (mesh intervals, order, compression, shear of the x = 1 face, start from the linear field)
CASES = [(2, 1, 0.1, 0.5, False), (3, 1, 0.2, 0.0, False), (3, 1, 0.3, 0.6, True)]


def _start(X, ess_nodes, strain, shear, linear):
    x0 = np.zeros(X.shape[0] * 3)
    x1n = ess_nodes
    x0[x1n * 3] = -strain
    x0[x1n * 3 + 1] = shear
    if linear:
        x0[0::3] = -strain * X[:, 0]
        x0[1::3] = shear * X[:, 0]
    return x0


@pytest.mark.parametrize("case", CASES)
def test_line_search_newton_matches_oracle(gpu_lib, case):
    n, p, strain, shear, linear = case
    gm, om = sx.Mesh.cuboid(n, n, n), O.Mesh.cuboid(n, n, n)
    gm.perturb(42, 0.2)
    om.perturb(42, 0.2)
    gs, os_ = sx.Space(gm, p, 3, 1), O.Space(om, p, 3, 1)
    X = gs.node_coords()
    x0n = np.where(np.abs(X[:, 0]) < 1e-12)[0]
    x1n = np.where(np.abs(X[:, 0] - 1.0) < 1e-12)[0]
    ess = np.union1d((x0n[:, None] * 3 + np.arange(3)).ravel(), (x1n[:, None] * 3 + np.arange(3)).ravel())
    ess = ess.astype(np.int32)
    x0 = _start(X, x1n, strain, shear, linear)
    of = O.Form(os_)
    of.add_domain_integral(O.QF_NEOHOOKEAN, NH)
    x_ref, rep_ref = of.newton(ess, x0, rel=1e-8, abs_tol=1e-12, maxit=40, lin_rel=1e-10, lin_max=20000,
                               max_line_search=8)
    _, rep_plain = of.newton(ess, x0, rel=1e-8, abs_tol=1e-12, maxit=40, lin_rel=1e-10, lin_max=20000)
    assert rep_ref["converged"] and rep_ref["line_search_steps"] > 0 and not rep_plain["converged"]
    gf = sx.Functional(gs, [gs])
    gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH)
    # plain Newton fails on the GPU too
    x = sx.DeviceVector.from_numpy(x0)
    rep = gf.newton(x, ess, rel_tol=1e-8, abs_tol=1e-12, max_iter=40, lin_rel_tol=1e-10, lin_max_iter=20000,
                    raise_on_noconv=False)
    assert not rep["converged"] and rep["line_search_steps"] == 0
    gf.set_option(capi.SX_OPT_LINE_SEARCH, 8)
    x = sx.DeviceVector.from_numpy(x0)
    rep = gf.newton(x, ess, rel_tol=1e-8, abs_tol=1e-12, max_iter=40, lin_rel_tol=1e-10, lin_max_iter=20000)
    print(f"{case}: GPU Newton {rep['newton_iterations']} its, {rep['line_search_steps']} line-search steps; "
          f"oracle {rep_ref['newton_iterations']}, {rep_ref['line_search_steps']}")
    assert rep["converged"]
    assert rep["newton_iterations"] == rep_ref["newton_iterations"]
    assert rep["line_search_steps"] == rep_ref["line_search_steps"]
    assert relerr(x.numpy(), x_ref) < 1e-8


def test_line_search_option_bounds(gpu_lib):
    m = sx.Mesh.cuboid(2, 2, 2)
    s = sx.Space(m, 1, 3, 1)
    f = sx.Functional(s, [s])
    for bad in (-1, 65):
        with pytest.raises(sx.SeracAmdError) as e:
            f.set_option(capi.SX_OPT_LINE_SEARCH, bad)
        assert e.value.code == capi.SX_ERR_ARG
