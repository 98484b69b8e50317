"""The oracle's Newton line search (the two-way backtracking of serac's
NewtonSolver, num/equation_solver.cpp:138-185), pinned by what it guarantees
on large-deformation Neo-Hookean steps where plain Newton inverts elements:

- with no line-search iterations it is plain Newton, bit for bit;
- every accepted step lowers the residual norm (the residual norms after
  1, 2, ... Newton iterations, from runs capped there, decrease strictly),
  which a step taken in the wrong direction or left unscaled breaks;
- it reaches a point whose residual meets the Newton goal where plain Newton
  produces a NaN residual."""
import numpy as np
import pytest

import oracle as O

NH = (1.0, 10.0, 0.25)


def problem(n, p, strain, shear, linear):
    m = O.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = O.Space(m, p, 3)
    f = O.Form(s)
    f.add_domain_integral(O.QF_NEOHOOKEAN, NH)
    X = s.node_coords()
    x0n = np.where(np.abs(X[:, 0]) < 1e-12)[0]
    x1n = np.where(np.abs(X[:, 0] - 1.0) < 1e-12)[0]
    ess = np.union1d((x0n[:, None] * 3 + np.arange(3)).ravel(), (x1n[:, None] * 3 + np.arange(3)).ravel())
    x0 = np.zeros(s.nnodes * 3)
    x0[x1n * 3], x0[x1n * 3 + 1] = -strain, shear
    if linear:
        x0[0::3], x0[1::3] = -strain * X[:, 0], shear * X[:, 0]
    return f, ess.astype(np.int32), x0


@pytest.mark.parametrize("case", [(2, 1, 0.1, 0.5, False), (3, 1, 0.2, 0.0, False), (2, 1, 0.3, 0.3, False)])
def test_line_search_decreases_and_converges(case):
    f, ess, x0 = problem(*case)
    kw = dict(rel=1e-8, abs_tol=1e-12, lin_rel=1e-10, lin_max=20000)
    x_plain, rep_plain = f.newton(ess, x0, maxit=40, **kw)
    x_zero, rep_zero = f.newton(ess, x0, maxit=40, max_line_search=0, **kw)
    assert np.array_equal(x_plain, x_zero, equal_nan=True) and rep_zero["line_search_steps"] == 0
    assert not rep_plain["converged"]
    _, rep = f.newton(ess, x0, maxit=40, max_line_search=8, **kw)
    assert rep["converged"] and rep["line_search_steps"] > 0
    assert rep["final_norm"] <= 1e-8 * rep["initial_norm"]
    norms = [rep["initial_norm"]]
    for k in range(1, rep["newton_iterations"] + 1):
        _, rk = f.newton(ess, x0, maxit=k, max_line_search=8, **kw)
        norms.append(rk["final_norm"])
    assert all(b < a for a, b in zip(norms, norms[1:])), norms
