"""Oracle pinning for boundary integrals (AddBoundaryIntegral).

- the reference's own Neo-Hookean patch test (phy/tests/solid_statics_patch.cpp:
  30-100): the affine solution with traction P(A) . N on part of the boundary
  leaves a zero residual at every DOF that is not Dirichlet, to 1e-13, on the
  perturbed cube and on the reference's patch3D_hexes mesh;
- closed forms: total traction -(A n + b) |face|, pressure p |face| n_out per
  planar face and zero over the closed surface (face orientation), total flux
  (q0 + h (T - T_inf)) |boundary|;
- follower-pressure and Robin tangents against central differences.
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_mesh_fixture
from serac_amd import workloads as W
from test_oracle_physics import A_PATCH, B_PATCH

FACES = {0: (0, -1.0), 1: (0, 1.0), 2: (1, -1.0), 3: (1, 1.0), 4: (2, -1.0), 5: (2, 1.0)}


def cube_space(n, p, nc, perturb=0.2, seed=7):
    m = O.Mesh.cuboid(n, n, n)
    if perturb:
        m.perturb(seed, perturb)
    return m, O.Space(m, p, nc)


@pytest.mark.parametrize("bit", range(6))
def test_traction_total_per_face(bit):
    m, s = cube_space(2, 2, 3)
    A = np.arange(9.0).reshape(3, 3) / 7.0
    b = np.array([0.3, -1.1, 2.0])
    f = O.Form(s)
    f.add_boundary_integral(O.BQF_TRACTION, list(A.ravel()) + list(b), 1 << bit)
    R = f.residual(np.zeros(s.ndof)).reshape(-1, 3)
    d, sgn = FACES[bit]
    n = np.zeros(3)
    n[d] = sgn
    assert np.abs(R.sum(axis=0) - (-(A @ n + b))).max() < 1e-14
    assert len(f.faces()) == 4


@pytest.mark.parametrize("p", [1, 2, 3])
def test_pressure_outward_and_closed_surface(p):
    m, s = cube_space(3, p, 3)
    for bit in range(6):
        f = O.Form(s)
        f.add_boundary_integral(O.BQF_PRESSURE, [2.5, 1.0], 1 << bit)
        R = f.residual(np.zeros(s.ndof)).reshape(-1, 3)
        d, sgn = FACES[bit]
        n = np.zeros(3)
        n[d] = sgn
        assert np.abs(R.sum(axis=0) - 2.5 * n).max() < 1e-13
    f = O.Form(s)
    f.add_boundary_integral(O.BQF_PRESSURE, [2.5, 1.0], 0)
    assert len(f.faces()) == 6 * 9
    R = f.residual(np.zeros(s.ndof)).reshape(-1, 3)
    assert np.abs(R.sum(axis=0)).max() < 1e-13


def test_flux_total():
    m, s = cube_space(2, 2, 1)
    f = O.Form(s)
    q0, h, tinf, c = 0.7, 1.3, 0.2, 0.9
    f.add_boundary_integral(O.BQF_FLUX, [q0, h, tinf], 0)
    R = f.residual(np.full(s.ndof, c))
    assert abs(R.sum() - 6 * (q0 + h * (c - tinf))) < 1e-13


def patch_residual(m, s):
    """NH patch: traction P(A) . N on the x, y, z = max faces; Dirichlet on the min faces."""
    _, P = O.qf_eval(O.QF_NEOHOOKEAN, [1.0, 1.0, 1.0], 3, np.zeros(3), A_PATCH.ravel())
    f = O.Form(s)
    f.add_domain_integral(O.QF_NEOHOOKEAN, [1.0, 1.0, 1.0])
    # traction(X, N) = P N; the residual carries -traction
    f.add_boundary_integral(O.BQF_TRACTION, list(np.asarray(P).ravel()) + [0, 0, 0], 0b101010)
    x = s.node_coords()
    U = (x @ A_PATCH.T + B_PATCH).ravel()
    R = f.residual(U).reshape(-1, 3)
    free = np.all(x > 1e-12, axis=1)  # nodes off the Dirichlet (min) faces
    return R, free


@pytest.mark.parametrize("p", [1, 2, 3])
def test_neohookean_patch_with_traction_perturbed_cube(p):
    m, s = cube_space(3, p, 3)
    R, free = patch_residual(m, s)
    assert free.sum() > 0
    assert np.abs(R[free]).max() < 1e-13
    assert np.abs(R[~free]).max() > 1e-3


@pytest.mark.parametrize("p", [1, 2])
def test_neohookean_patch_with_traction_reference_mesh(p):
    X, hexes = load_mesh_fixture("patch3D_hexes")
    m = O.Mesh.from_arrays(X, hexes)
    s = O.Space(m, p, 3)
    R, free = patch_residual(m, s)
    assert np.abs(R[free]).max() < 1e-13


def fd_check(f, s, U, seed):
    dU = W.nodal_field(s.nnodes, s.nc, seed, 3000, -1.0, 1.0, 1)
    f.linearize(U)
    jv = f.jvp(dU)
    h = 1e-6
    fd = (f.residual(U + h * dU) - f.residual(U - h * dU)) / (2 * h)
    return np.abs(jv - fd).max() / np.abs(jv).max()


def test_follower_pressure_tangent():
    m, s = cube_space(2, 2, 3)
    f = O.Form(s)
    f.add_boundary_integral(O.BQF_PRESSURE, [3.0, 1.0], 0)
    U = W.nodal_field(s.nnodes, 3, 4, 2000, -0.05, 0.05, 1)
    assert fd_check(f, s, U, 5) < 1e-7
    # without geometric nonlinearity the load does not depend on u
    g = O.Form(s)
    g.add_boundary_integral(O.BQF_PRESSURE, [3.0, 0.0], 0)
    g.linearize(U)
    assert np.abs(g.jvp(np.ones(s.ndof))).max() == 0.0


def test_robin_tangent_is_boundary_mass():
    m, s = cube_space(2, 2, 1)
    f = O.Form(s)
    h = 1.7
    f.add_boundary_integral(O.BQF_FLUX, [0.0, h, 0.0], 0)
    U = W.nodal_field(s.nnodes, 1, 4, 2000, -1, 1, 1)
    assert fd_check(f, s, U, 6) < 1e-8
    f.linearize(U)
    # J 1 = h * (boundary "lumped" mass) sums to h |boundary|
    assert abs(f.jvp(np.ones(s.ndof)).sum() - 6 * h) < 1e-13
