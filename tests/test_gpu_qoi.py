"""Quantities of interest, Functional<double(trial)> (fn/functional_qoi.inl), on
the GPU: the measure and first moments of the perturbed unit cube and of the
reference's beam-hex.mesh (closed forms: the perturbed cells still tile the
cube, and trilinear geometry is integrated exactly), the integral of fields
that are linear in x, the Neo-Hookean strain energy of affine deformations,
and the gradients against the Neo-Hookean residual and central differences
(fn/tests/functional_qoi.cpp: measure, moment, check_gradient)."""
import os

import numpy as np
import pytest

import serac_amd as sx
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NH = (1.0, 10.0, 0.25)


def cube(n=4, perturb=0.2, p=2, nc=3):
    m = sx.Mesh.cuboid(n, n, n)
    if perturb:
        m.perturb(42, perturb)
    return m, sx.Space(m, p, nc, 1)


def psi(A, K, G):
    """Neo-Hookean strain energy density at F = I + A."""
    F = np.eye(3) + A
    J = np.linalg.det(F)
    lam = K - 2.0 * G / 3.0
    return 0.5 * G * (np.trace(F @ F.T) - 3.0) - G * np.log(J) + 0.5 * lam * np.log(J) ** 2


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("nc", [1, 3])
def test_measure_and_moments_of_the_perturbed_cube(gpu_lib, p, nc):
    m, s = cube(3, 0.2, p, nc)
    U = sx.DeviceVector(s.ndof)
    q = sx.QoI(s)
    q.add_domain_integral(sx.capi.SX_QOI_MEASURE)
    assert abs(q(0.0, U) - 1.0) < 1e-13
    for d in range(3):
        qm = sx.QoI(s)
        qm.add_domain_integral(sx.capi.SX_QOI_POSITION, (d,))
        assert abs(qm(0.0, U) - 0.5) < 1e-13
        assert np.abs(qm.gradient(0.0, U).numpy()).max() == 0.0


def test_measure_and_moment_of_beam_hex(gpu_lib):
    m = sx.Mesh.read_mfem(os.path.join(ROOT, "tests", "golden", "meshes", "beam-hex.mesh"))
    s = sx.Space(m, 2, 1, 1)
    U = sx.DeviceVector(s.ndof)
    q = sx.QoI(s)
    q.add_domain_integral(sx.capi.SX_QOI_MEASURE)
    q.add_domain_integral(sx.capi.SX_QOI_POSITION, (0,))
    assert abs(q(0.0, U) - (8.0 + 32.0)) < 1e-12
    half = sx.QoI(s)
    half.add_domain_integral(sx.capi.SX_QOI_MEASURE, elems=[0, 1, 2])
    assert abs(half(0.0, U) - 3.0) < 1e-13


@pytest.mark.parametrize("p", [1, 3])
def test_integral_of_a_linear_field_and_its_gradient(gpu_lib, p):
    m, s = cube(4, 0.2, p, 3)
    xyz = s.node_coords()
    Uh = np.zeros(s.ndof)
    Uh[1::3] = 0.3 + 2.0 * xyz[:, 0] - 1.5 * xyz[:, 2]  # component 1, linear in x
    U = sx.DeviceVector.from_numpy(Uh)
    q = sx.QoI(s)
    q.add_domain_integral(sx.capi.SX_QOI_VALUE, (1,))
    assert abs(q(0.0, U) - (0.3 + 1.0 - 0.75)) < 1e-13
    g = q.gradient(0.0, U).numpy()
    # the gradient is int phi_j dV on component 1: it sums to the volume and
    # the QoI is linear, so q(U + V) - q(U) = g . V
    assert abs(g[1::3].sum() - 1.0) < 1e-13
    assert np.abs(g[0::3]).max() == 0.0 and np.abs(g[2::3]).max() == 0.0
    V = W.nodal_field(s.nnodes, 3, 5, 2000, -1.0, 1.0, 1)
    dq = q(0.0, sx.DeviceVector.from_numpy(Uh + V)) - q(0.0, U)
    assert abs(dq - g @ V) < 1e-12 * max(1.0, abs(dq))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_neohookean_energy_of_affine_deformations(gpu_lib, p):
    m, s = cube(3, 0.2, p, 3)
    xyz = s.node_coords()
    A = np.array([[0.02, 0.01, -0.005], [0.0, -0.01, 0.004], [0.003, 0.002, 0.015]])
    Uh = (xyz @ A.T).ravel()
    q = sx.QoI(s)
    q.add_domain_integral(sx.capi.SX_QOI_NEOHOOKEAN_ENERGY, NH)
    val = q(0.0, sx.DeviceVector.from_numpy(Uh))
    ref = psi(A, NH[1], NH[2])
    assert abs(val - ref) <= 1e-11 * abs(ref)
    assert abs(q(0.0, sx.DeviceVector(s.ndof))) < 1e-15  # no energy at rest


def test_energy_gradient_is_the_neohookean_residual_and_matches_central_differences(gpu_lib):
    m, s = cube(3, 0.2, 2, 3)
    Uh = W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    U = sx.DeviceVector.from_numpy(Uh)
    q = sx.QoI(s)
    q.add_domain_integral(sx.capi.SX_QOI_NEOHOOKEAN_ENERGY, NH)
    g = q.gradient(0.0, U).numpy()
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH)
    R = f(0.0, U).numpy()
    assert np.abs(g - R).max() <= 1e-14 * np.abs(R).max()
    V = W.nodal_field(s.nnodes, 3, 4, 2001, -1.0, 1.0, 1)

    def fd_error(eps):
        fd = q(0.0, sx.DeviceVector.from_numpy(Uh + eps * V)) - q(0.0, sx.DeviceVector.from_numpy(Uh - eps * V))
        return abs(fd / (2 * eps) - g @ V) / abs(g @ V)

    # central differences converge at second order (check_gradient.hpp:47-89)
    e1, e2 = fd_error(1e-3), fd_error(5e-4)
    assert 3.5 < e1 / e2 < 4.5, (e1, e2)
    assert fd_error(1e-5) < 1e-6


def test_qoi_arguments_are_checked(gpu_lib):
    m, s1 = cube(2, 0.0, 1, 1)
    q = sx.QoI(s1)
    with pytest.raises(sx.SeracAmdError):
        q.add_domain_integral(sx.capi.SX_QOI_NEOHOOKEAN_ENERGY, NH)
    with pytest.raises(sx.SeracAmdError):
        q.add_domain_integral(sx.capi.SX_QOI_POSITION, (3,))
    with pytest.raises(sx.SeracAmdError):
        q.add_domain_integral(sx.capi.SX_QOI_VALUE, (1,))
    with pytest.raises(sx.SeracAmdError):
        q.add_domain_integral(99)
    # the internal integrand of the value gradient is not a public qfunction
    f = sx.Functional(s1, [s1])
    with pytest.raises(sx.SeracAmdError):
        f.add_domain_integral(6, (0.0,))


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("nc", [1, 3])
def test_boundary_measures_moments_and_values(gpu_lib, p, nc):
    # the reference's SumOfMeasures: int 1 dV + int 1 dA.  The perturbation
    # moves interior vertices only, so the boundary is the flat unit-cube
    # surface (area 6; int x dA = 0 + 1 + 4 * 0.5)
    m, s = cube(3, 0.2, p, nc)
    xyz = s.node_coords()
    Uh = np.zeros(s.ndof)
    c = nc - 1
    Uh[c::nc] = 0.4 + 1.2 * xyz[:, 0]
    U = sx.DeviceVector.from_numpy(Uh)
    q = sx.QoI(s)
    q.add_domain_integral(sx.capi.SX_QOI_MEASURE)
    q.add_boundary_integral(sx.capi.SX_QOI_MEASURE)
    assert abs(q(0.0, U) - 7.0) < 1e-13
    qx = sx.QoI(s)
    qx.add_boundary_integral(sx.capi.SX_QOI_POSITION, (0,))
    assert abs(qx(0.0, U) - 3.0) < 1e-13
    qa = sx.QoI(s)
    qa.add_boundary_integral(sx.capi.SX_QOI_MEASURE, attrs=[3, 5])  # the x = max and x = 0 faces
    assert abs(qa(0.0, U) - 2.0) < 1e-13
    qv = sx.QoI(s)
    qv.add_boundary_integral(sx.capi.SX_QOI_VALUE, (c,))
    assert abs(qv(0.0, U) - (0.4 * 6 + 1.2 * 3)) < 1e-12
    g = qv.gradient(0.0, U).numpy()
    assert abs(g[c::nc].sum() - 6.0) < 1e-12
    V = W.nodal_field(s.nnodes, nc, 5, 2000, -1.0, 1.0, 1)
    dq = qv(0.0, sx.DeviceVector.from_numpy(Uh + V)) - qv(0.0, U)
    assert abs(dq - g @ V) < 1e-12 * max(1.0, abs(dq))


def test_beam_hex_boundary_by_attribute(gpu_lib):
    m = sx.Mesh.read_mfem(os.path.join(ROOT, "tests", "golden", "meshes", "beam-hex.mesh"))
    s = sx.Space(m, 2, 3, 1)
    U = sx.DeviceVector(s.ndof)
    for attrs, area in (([1], 1.0), ([2], 1.0), ([3], 32.0), ([1, 2, 3], 34.0)):
        q = sx.QoI(s)
        q.add_boundary_integral(sx.capi.SX_QOI_MEASURE, attrs=attrs)
        assert abs(q(0.0, U) - area) < 1e-12, attrs
