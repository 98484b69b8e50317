"""ShapeAwareFunctional (fn/shape_aware_functional.hpp:292-577) on the GPU:
the domain integrals evaluated on X + p.  For a shape displacement p that is
the trilinear (Q1) interpolant of vertex displacements d, X + p is exactly the
mesh whose vertices moved by d, so the shape-aware residual and J.v on X must
equal the plain ones on the moved mesh (nodes matched by position).  Also:
p = 0 reproduces the plain form bit for bit, the shape-aware J.v against
central differences of the shape-aware residual, and the argument checks."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import cKDTree

import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_learned import interp_1d

pytestmark = pytest.mark.gpu
NH = (1.0, 10.0, 0.25)


def mfem_hexes(n):
    """Connectivity of the n^3 cuboid in mfem's hex vertex order."""
    v = lambda i, j, k: (k * (n + 1) + j) * (n + 1) + i  # noqa: E731
    hexes = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                hexes.append([v(i, j, k), v(i + 1, j, k), v(i + 1, j + 1, k), v(i, j + 1, k),
                              v(i, j, k + 1), v(i + 1, j, k + 1), v(i + 1, j + 1, k + 1), v(i, j + 1, k + 1)])
    return np.array(hexes, dtype=np.int32)


def form(space):
    f = sx.Functional(space, [space])
    f.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_COLORED)
    f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH)
    return f


@pytest.mark.parametrize("p", [1, 2, 3])
def test_q1_shape_displacement_equals_the_moved_mesh(gpu_lib, p):
    n = 3
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    Xv = m.vertices()
    d = W.nodal_field(Xv.shape[0], 3, 9, 4000, -0.04, 0.04, 1).reshape(-1, 3)
    P1 = sp.csr_matrix(interp_1d(n, p))
    pn = (sp.kron(P1, sp.kron(P1, P1)) @ d)  # the Q1 interpolant at the order-p nodes
    m2 = sx.Mesh.from_arrays(Xv + d, mfem_hexes(n))
    s2 = sx.Space(m2, p, 3, 1)
    # node k of s sits at node perm[k] of s2
    dist, perm = cKDTree(s2.node_coords()).query(s.node_coords() + pn)
    assert dist.max() < 1e-12 and len(set(perm.tolist())) == s.nnodes
    dof2 = (3 * perm[:, None] + np.arange(3)).ravel()  # byVDIM
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    U2 = np.zeros_like(U)
    U2[dof2] = U
    V = W.nodal_field(s.nnodes, 3, 5, 2001, -1.0, 1.0, 1)
    V2 = np.zeros_like(V)
    V2[dof2] = V
    f, f2 = form(s), form(s2)
    pv = sx.DeviceVector.from_numpy(pn.ravel())
    f.set_shape_displacement(pv)
    R, g = f(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    R2, g2 = f2(0.0, sx.DeviceVector.from_numpy(U2), wrt=0)
    R, R2 = R.numpy(), R2.numpy()[dof2]
    assert np.abs(R - R2).max() <= 1e-11 * np.abs(R2).max()
    Jv = g(sx.DeviceVector.from_numpy(V)).numpy()
    Jv2 = g2(sx.DeviceVector.from_numpy(V2)).numpy()[dof2]
    assert np.abs(Jv - Jv2).max() <= 1e-11 * np.abs(Jv2).max()
    # and the shape changes the answer
    f.set_shape_displacement(None)
    R0 = f(0.0, sx.DeviceVector.from_numpy(U)).numpy()
    assert np.abs(R0 - R).max() > 1e-3 * np.abs(R).max()


def test_zero_shape_is_the_plain_form_and_jv_matches_differences(gpu_lib):
    n, p = 3, 3
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    f = form(s)
    U = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1))
    R0 = f(0.0, U).numpy()
    zero = sx.DeviceVector(s.ndof)
    f.set_shape_displacement(zero)
    assert np.array_equal(f(0.0, U).numpy(), R0)
    # a smooth order-p shape field (not Q1): J.v against central differences
    xyz = s.node_coords()
    pn = np.stack([0.03 * np.sin(2 * xyz[:, 1]), 0.02 * xyz[:, 0] * xyz[:, 2], -0.025 * xyz[:, 0] ** 2], axis=1)
    pv = sx.DeviceVector.from_numpy(pn.ravel())
    f.set_shape_displacement(pv)
    Uh = U.numpy()
    V = W.nodal_field(s.nnodes, 3, 4, 2001, -1.0, 1.0, 1)
    _, g = f(0.0, U, wrt=0)
    Jv = g(sx.DeviceVector.from_numpy(V)).numpy()

    def fd(eps):
        rp = f(0.0, sx.DeviceVector.from_numpy(Uh + eps * V)).numpy()
        rm = f(0.0, sx.DeviceVector.from_numpy(Uh - eps * V)).numpy()
        return np.abs((rp - rm) / (2 * eps) - Jv).max() / np.abs(Jv).max()

    e1, e2 = fd(1e-3), fd(5e-4)
    assert 3.5 < e1 / e2 < 4.5, (e1, e2)
    assert fd(1e-5) < 1e-6


@pytest.mark.parametrize("bqf", ["traction", "pressure"])
def test_boundary_integrals_see_the_moved_faces(gpu_lib, bqf):
    # traction and follower pressure on the z = max face: the vertices there
    # keep z (so the face selection by planes is the same on the moved mesh)
    # and move in x and y
    n, p = 3, 2
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    Xv = m.vertices()
    d = W.nodal_field(Xv.shape[0], 3, 11, 4000, -0.04, 0.04, 1).reshape(-1, 3)
    d[np.abs(Xv[:, 2] - 1.0) < 1e-12, 2] = 0.0
    P1 = sp.csr_matrix(interp_1d(n, p))
    pn = sp.kron(P1, sp.kron(P1, P1)) @ d
    m2 = sx.Mesh.from_arrays(Xv + d, mfem_hexes(n))
    s2 = sx.Space(m2, p, 3, 1)
    dist, perm = cKDTree(s2.node_coords()).query(s.node_coords() + pn)
    assert dist.max() < 1e-12
    dof2 = (3 * perm[:, None] + np.arange(3)).ravel()
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    U2 = np.zeros_like(U)
    U2[dof2] = U
    V = W.nodal_field(s.nnodes, 3, 5, 2001, -1.0, 1.0, 1)
    V2 = np.zeros_like(V)
    V2[dof2] = V
    fs = []
    for sp_ in (s, s2):
        f = form(sp_)
        if bqf == "traction":
            f.add_boundary_integral(sx.SX_BQF_TRACTION, list(np.arange(9.0) / 7.0) + [0.3, -1.1, 2.0], face_mask=1 << 5)
        else:
            f.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.7, 1.0), face_mask=1 << 5)
        fs.append(f)
    pv = sx.DeviceVector.from_numpy(pn.ravel())
    fs[0].set_shape_displacement(pv)
    R, g = fs[0](0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    R2, g2 = fs[1](0.0, sx.DeviceVector.from_numpy(U2), wrt=0)
    R, R2 = R.numpy(), R2.numpy()[dof2]
    assert np.abs(R - R2).max() <= 1e-11 * np.abs(R2).max()
    Jv = g(sx.DeviceVector.from_numpy(V)).numpy()
    Jv2 = g2(sx.DeviceVector.from_numpy(V2)).numpy()[dof2]
    assert np.abs(Jv - Jv2).max() <= 1e-11 * np.abs(Jv2).max()


def test_shape_displacement_arguments_are_checked(gpu_lib):
    m = sx.Mesh.cuboid(2, 2, 2)
    s1 = sx.Space(m, 2, 1, 1)
    heat = sx.Functional(s1, [s1])
    heat.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0, 1.0, 0.0))
    with pytest.raises(sx.SeracAmdError):
        heat.set_shape_displacement(sx.DeviceVector(s1.ndof))  # one component
    s3 = sx.Space(m, 2, 3, 1)
    f = form(s3)
    f.set_shape_displacement(sx.DeviceVector(s3.ndof))
    R0 = f(0.0, sx.DeviceVector(s3.ndof)).numpy()
    f.set_shape_displacement(None)
    assert np.array_equal(f(0.0, sx.DeviceVector(s3.ndof)).numpy(), R0)
