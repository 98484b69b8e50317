"""Pins the oracle's sum-factorized interpolate / integrate
(fn/detail/hexahedron_H1.inl:169-280) and geometric factors
(fn/geometric_factors.cpp:17-58) without re-stating them: full-basis
evaluation, adjointness, polynomial reproduction and exact geometry."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n", [2, 3, 4, 5])
@pytest.mark.parametrize("nc", [1, 3])
def test_sum_factorization_equals_full_basis(n, nc):
    rng = np.random.default_rng(n * 10 + nc)
    X = rng.uniform(-1, 1, (nc, n ** 3))
    v1, g1 = O.interpolate(n, n, nc, X)
    v2, g2 = O.interpolate(n, n, nc, X, brute=True)
    assert np.allclose(v1, v2, rtol=0, atol=1e-13)
    assert np.allclose(g1, g2, rtol=0, atol=1e-12)
    s = rng.uniform(-1, 1, (nc, n ** 3))
    f = rng.uniform(-1, 1, (nc, 3, n ** 3))
    assert np.allclose(O.integrate(n, n, nc, s, f), O.integrate(n, n, nc, s, f, brute=True), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_integrate_is_the_weighted_adjoint_of_interpolate(n):
    # sum_q w_q (s_q u_q + f_q . grad u_q) == sum_j u_j r_j with r = integrate(s, f)
    rng = np.random.default_rng(3)
    q = n
    u = rng.uniform(-1, 1, (1, n ** 3))
    s = rng.uniform(-1, 1, (1, q ** 3))
    f = rng.uniform(-1, 1, (1, 3, q ** 3))
    val, grd = O.interpolate(n, q, 1, u)
    _, w = O.gl_rule(q)
    W = np.einsum("i,j,k->kji", w, w, w).ravel()  # q = (qz*Q + qy)*Q + qx
    lhs = np.sum(W * (s[0] * val[0] + np.sum(f[0] * grd[0], axis=0)))
    rhs = np.dot(u[0], O.integrate(n, q, 1, s, f)[0])
    assert abs(lhs - rhs) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_interpolation_reproduces_polynomials(p):
    n = p + 1
    nodes = O.gll_nodes(n)
    xg, _ = O.gl_rule(n)
    c = [0.3, -1.1, 0.7]

    def f(x, y, z):  # degree <= p in each variable: inside Q_p
        return (c[0] + x ** p) * (c[1] + y ** p) * (c[2] + z ** p) + x ** (p - 1) * y * z

    def grad(x, y, z, h=1e-6):
        return np.array([(f(x + h, y, z) - f(x - h, y, z)) / (2 * h), (f(x, y + h, z) - f(x, y - h, z)) / (2 * h),
                         (f(x, y, z + h) - f(x, y, z - h)) / (2 * h)])

    U = np.array([[f(nodes[i], nodes[j], nodes[k]) for k in range(n) for j in range(n) for i in range(n)]])
    val, grd = O.interpolate(n, n, 1, U)
    for q in range(n ** 3):
        qx, qy, qz = q % n, (q // n) % n, q // (n * n)
        pt = (xg[qx], xg[qy], xg[qz])
        assert abs(val[0, q] - f(*pt)) < 1e-13
        assert np.allclose(grd[0, :, q], grad(*pt), atol=1e-7)


def test_affine_mesh_has_constant_jacobian():
    A = np.array([[1.2, 0.1, -0.2], [0.05, 0.9, 0.3], [0.0, -0.1, 1.1]])
    b = np.array([0.3, -0.2, 1.0])
    m0 = O.Mesh.cuboid(2, 3, 2, 1.0, 1.5, 0.5)
    X = m0.vertices() @ A.T + b
    hexes = _hexes_of(m0)
    m = O.Mesh.from_arrays(X, hexes)
    for Q in (2, 3, 4, 5):
        Xq, Jq = m.geometric_factors(Q)
        # Jq[e][j][i][q] = d x_i / d xi_j; on the cuboid d X / d xi = diag(h)
        h = np.array([1.0 / 2, 1.5 / 3, 0.5 / 2])
        expect = (A * h[None, :]).T  # [j][i]
        assert np.allclose(Jq, expect[None, :, :, None], atol=1e-14)


def test_volume_of_perturbed_unit_cube():
    m = O.Mesh.cuboid(5, 4, 3)
    m.perturb(42, 0.2)
    for Q in (2, 3, 4, 5):
        _, Jq = m.geometric_factors(Q)
        _, w = O.gl_rule(Q)
        W = np.einsum("i,j,k->kji", w, w, w).ravel()
        det = np.linalg.det(np.transpose(Jq, (0, 3, 1, 2)))
        assert np.all(det > 0)
        assert abs(np.sum(det * W[None, :]) - 1.0) < 1e-13


def test_perturbation_moves_only_interior_vertices():
    m0 = O.Mesh.cuboid(4, 4, 4)
    m = O.Mesh.cuboid(4, 4, 4)
    m.perturb(9, 0.2)
    X0, X = m0.vertices(), m.vertices()
    bnd = np.any((X0 < 1e-12) | (X0 > 1 - 1e-12), axis=1)
    assert np.array_equal(X0[bnd], X[bnd])
    d = np.abs(X - X0)[~bnd]
    assert d.max() <= 0.2 * 0.25 + 1e-15 and d.max() > 0.1 * 0.25


def _hexes_of(mesh):
    """Element -> vertex lists of a cuboid mesh, in mfem hexahedron order."""
    # reconstruct from a space of order 1: lexicographic corners -> mfem slots
    s = O.Space(mesh, 1, 1)
    em = s.emap()  # [e][lex corner] node ids = vertex ids for the structured numbering
    lex2v = [0, 1, 3, 2, 4, 5, 7, 6]
    hexes = np.zeros_like(em)
    for c in range(8):
        hexes[:, lex2v[c]] = em[:, c]
    return hexes
