"""The oracle's Jacobi-CG stops the way mfem's CGSolver does when the Jacobi
preconditioner is not positive definite: (r0, D^-1 r0) < 0 returns x = 0
unconverged, and a negative (r_k, D^-1 r_k) later stops unconverged at that
iteration.  Checked against a dense numpy PCG on the matrix the oracle's own
J.v builds column by column (mass + diffusion with a negative mass weight,
so the diagonal has mixed or negative signs)."""
import numpy as np
import pytest

import oracle as O
from serac_amd import workloads as W


def _problem(a):
    m = O.Mesh.cuboid(3, 3, 3)
    m.perturb(42, 0.2)
    s = O.Space(m, 2, 1)
    f = O.Form(s)
    f.add_domain_integral(O.QF_MASS_DIFFUSION, (a, 1.0))
    f.linearize(np.zeros(s.ndof))
    A = np.stack([f.jvp(e) for e in np.eye(s.ndof)], axis=1)
    return s, f, A


def dense_pcg(A, b, rel, maxit):
    """Preconditioned CG with D = diag(A), stopping rules of mfem::CGSolver."""
    d = np.diag(A)
    x = np.zeros_like(b)
    r = b.copy()
    z = r / d
    p = z.copy()
    nom = r @ z
    if nom < 0:
        return x, 0, False
    goal = nom * rel * rel
    for it in range(1, maxit + 1):
        Ap = A @ p
        alpha = nom / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        z = r / d
        betanom = r @ z
        if betanom < 0:
            return x, it, False
        if betanom <= goal:
            return x, it, True
        p = z + betanom / nom * p
        nom = betanom
    return x, maxit, False


def test_diagonal_signs_of_the_cases():
    for a, want in ((-300.0, "mixed"), (-700.0, "negative")):
        _, f, A = _problem(a)
        d = f.diagonal()
        assert np.allclose(d, np.diag(A), rtol=1e-12, atol=1e-15)
        kind = "negative" if (d < 0).all() else "mixed" if (d < 0).any() else "positive"
        assert kind == want


@pytest.mark.parametrize("a", [-300.0, -700.0])
def test_indefinite_jacobi_stops_unconverged(a):
    s, f, A = _problem(a)
    b = W.direction(s.nnodes, 1)
    x, rep = f.cg(np.array([], dtype=np.int32), b, rel=1e-10, maxit=500)
    xr, itr, conv = dense_pcg(A, b, 1e-10, 500)
    assert not rep["converged"] and not conv
    assert rep["iterations"] == itr
    if a == -700.0:
        assert itr == 0 and not x.any()
    else:
        assert itr >= 1
        assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()


def test_positive_definite_case_still_converges():
    s, f, A = _problem(1.7)
    b = W.direction(s.nnodes, 1)
    x, rep = f.cg(np.array([], dtype=np.int32), b, rel=1e-10, maxit=500)
    assert rep["converged"]
    assert np.abs(A @ x - b).max() <= 1e-8 * np.abs(b).max()
