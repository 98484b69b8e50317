"""The oracle's QoI values (Functional<double(trials...)>, the restatement the
GPU QoI kernels are checked against), pinned by closed forms on the perturbed
unit cube, whose trilinear cells still tile [0,1]^3 exactly:

- measure 1 and first moments 1/2 for p = 1..4;
- int u_c for u linear in x (in the space for every p >= 1): a0 + a . (1/2);
- int u . v for u, v linear in x at p >= 2, where the quadrature is exact:
  a0 b0 + (a0 sum b + b0 sum a)/2 + sum_i a_i b_i / 3 + sum_{i != j} a_i b_j / 4;
- the Neo-Hookean strain energy of u = A x: psi(I + A) for a general A."""
import numpy as np
import pytest

import oracle as O

NH = (1.0, 10.0, 0.25)


def cube(p, nc, n=3):
    m = O.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = O.Space(m, p, nc)
    return s, O.Form(s)


def linear_field(s, coeffs):
    """byVDIM field, component c = coeffs[c][0] + coeffs[c][1:] . x"""
    xyz = s.node_coords()
    U = np.zeros((s.nnodes, s.nc))
    for c in range(s.nc):
        U[:, c] = coeffs[c][0] + xyz @ np.asarray(coeffs[c][1:])
    return U.ravel()


def psi(A, K, G):
    F = np.eye(3) + A
    J = np.linalg.det(F)
    lam = K - 2.0 * G / 3.0
    return 0.5 * G * (np.trace(F @ F.T) - 3.0) - G * np.log(J) + 0.5 * lam * np.log(J) ** 2


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_measure_and_moments(p):
    s, f = cube(p, 1)
    assert abs(f.qoi(O.QOI_MEASURE) - 1.0) < 1e-13
    for d in range(3):
        assert abs(f.qoi(O.QOI_POSITION, (d,)) - 0.5) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 4])
def test_value_of_a_linear_field(p):
    s, f = cube(p, 3)
    co = [(0.3, 2.0, -1.5, 0.7), (-1.1, 0.2, 0.4, -0.9), (0.5, -0.3, 1.2, 0.1)]
    U = linear_field(s, co)
    for c in range(3):
        assert abs(f.qoi(O.QOI_VALUE, (c,), U=U) - (co[c][0] + 0.5 * sum(co[c][1:]))) < 1e-13


@pytest.mark.parametrize("p", [2, 3, 4])
def test_inner_product_of_linear_fields(p):
    s, f = cube(p, 3)
    ca = [(0.3, 2.0, -1.5, 0.7), (-1.1, 0.2, 0.4, -0.9), (0.5, -0.3, 1.2, 0.1)]
    cb = [(1.0, -0.5, 0.25, 2.0), (0.2, 1.1, -0.7, 0.3), (-0.4, 0.6, 0.9, -1.3)]
    U, V = linear_field(s, ca), linear_field(s, cb)
    want = 0.0
    for a, b in zip(ca, cb):
        a0, av, b0, bv = a[0], np.array(a[1:]), b[0], np.array(b[1:])
        cross = np.outer(av, bv)
        want += a0 * b0 + 0.5 * (a0 * bv.sum() + b0 * av.sum()) + np.trace(cross) / 3 + (cross.sum() - np.trace(cross)) / 4
    assert abs(f.qoi(O.QOI_INNER, U=U, U1=V) - want) < 1e-13 * max(1.0, abs(want))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_neohookean_energy_of_an_affine_deformation(p):
    s, f = cube(p, 3)
    A = np.array([[0.02, 0.03, -0.01], [-0.015, 0.01, 0.02], [0.005, -0.025, -0.02]])
    U = (s.node_coords() @ A.T).ravel()
    assert abs(f.qoi(O.QOI_NEOHOOKEAN_ENERGY, NH, U=U) - psi(A, NH[1], NH[2])) < 1e-14
    # and a different deformation gives a different energy
    assert abs(f.qoi(O.QOI_NEOHOOKEAN_ENERGY, NH, U=2 * U) - psi(A, NH[1], NH[2])) > 1e-6


def test_domain_subset_is_the_sum_of_its_elements():
    s, f = cube(2, 1)
    a = f.qoi(O.QOI_MEASURE, elems=[0, 5, 9])
    b = sum(f.qoi(O.QOI_MEASURE, elems=[e]) for e in (0, 5, 9))
    assert abs(a - b) < 1e-15 and 0.0 < a < 1.0
