"""differentiate_wrt on the shape argument of ShapeAwareFunctional
(fn/shape_aware_functional.hpp:292-577) on the GPU: dR/dp dp, recomputed by
forward-mode duals through the geometry.

- against central differences of the shape-aware residual (second-order
  convergence of the difference, and agreement at small steps), for
  Neo-Hookean, linear elasticity and J2 plasticity with state, p = 1..3;
- the moved-mesh identity: for Q1 shape fields p = P d and dp = P dd
  (trilinear interpolants of vertex displacements), dR/dp dp is the
  derivative of the plain residual on the mesh whose vertices moved by
  d + eps dd, taken by central differences on those meshes;
- linearity in dp, the state is not committed, and the error paths."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import cKDTree

import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_learned import interp_1d
from test_gpu_shape_aware import mfem_hexes

pytestmark = pytest.mark.gpu

QF = {
    "neohookean": (sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25)),
    "linear_elastic": (sx.SX_QF_LINEAR_ELASTIC, (1.2, 0.8)),
    "j2": (sx.SX_QF_J2_LINEAR, (200.0, 0.3, 0.5, 2.0, 0.0, 1.0, 1.0)),
}


def make(space, name, qd=None):
    qf, params = QF[name]
    f = sx.Functional(space, [space])
    f.add_domain_integral(qf, params, qdata=qd)
    return f


def smooth_shape(xyz, a=1.0):
    return (a * np.stack([0.03 * np.sin(2 * xyz[:, 1]), 0.02 * xyz[:, 0] * xyz[:, 2],
                          -0.025 * xyz[:, 0] ** 2], axis=1)).ravel()


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("name", ["neohookean", "linear_elastic", "j2"])
def test_shape_derivative_against_central_differences(gpu_lib, p, name):
    m = sx.Mesh.cuboid(3, 3, 3)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3)
    qd = sx.QuadratureData(s.ne, (p + 1) ** 3, np.zeros(10)) if name == "j2" else None
    f = make(s, name, qd)
    xyz = s.node_coords()
    if name == "j2":
        # simple shear beyond yield: the return map is active at every point
        Uh = np.zeros(s.ndof)
        Uh[0::3] = 0.02 * xyz[:, 1]
        Uh += W.nodal_field(s.nnodes, 3, 3, 2000, -1e-4, 1e-4, 1)
    else:
        Uh = W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    U = sx.DeviceVector.from_numpy(Uh)
    ph = smooth_shape(xyz)
    dp = W.nodal_field(s.nnodes, 3, 4, 2001, -0.01, 0.01, 1)
    pv = sx.DeviceVector.from_numpy(ph)
    f.set_shape_displacement(pv)
    _, g = f(0.0, U, wrt="shape")
    Jp = g(sx.DeviceVector.from_numpy(dp)).numpy()
    assert np.isfinite(Jp).all() and np.abs(Jp).max() > 0

    def R_at(shape):
        v = sx.DeviceVector.from_numpy(shape)
        f.set_shape_displacement(v)
        r = f(0.0, U).numpy().copy()
        f.set_shape_displacement(pv)
        return r

    def err(eps):
        fd = (R_at(ph + eps * dp) - R_at(ph - eps * dp)) / (2 * eps)
        return np.abs(fd - Jp).max() / np.abs(Jp).max()

    e1, e2 = err(0.2), err(0.1)
    if name != "j2":  # the return map's kinks make the large steps noisy
        assert 3.5 < e1 / e2 < 4.5, (e1, e2)
    assert err(1e-4) < 1e-6
    if qd is not None:
        assert not qd.numpy().any()  # nothing committed
    # linear in dp (setting a shape displacement invalidated the gradient: linearize again)
    _, g = f(0.0, U, wrt="shape")
    assert np.array_equal(g(sx.DeviceVector.from_numpy(dp)).numpy(), Jp)
    dq = W.nodal_field(s.nnodes, 3, 5, 2001, -0.01, 0.01, 1)
    Jq = g(sx.DeviceVector.from_numpy(dq)).numpy()
    Jc = g(sx.DeviceVector.from_numpy(2.0 * dp - 3.0 * dq)).numpy()
    assert np.abs(Jc - (2.0 * Jp - 3.0 * Jq)).max() <= 1e-12 * (np.abs(Jp).max() + np.abs(Jq).max())


def test_shape_derivative_is_the_moved_mesh_derivative(gpu_lib):
    n, p = 3, 2
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3)
    Xv = m.vertices()
    d = W.nodal_field(Xv.shape[0], 3, 9, 4000, -0.03, 0.03, 1).reshape(-1, 3)
    dd = W.nodal_field(Xv.shape[0], 3, 10, 4000, -0.03, 0.03, 1).reshape(-1, 3)
    P1 = sp.csr_matrix(interp_1d(n, p))
    P3 = sp.kron(P1, sp.kron(P1, P1))
    pn, dpn = P3 @ d, P3 @ dd
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    f = make(s, "neohookean")
    f.set_shape_displacement(sx.DeviceVector.from_numpy(pn.ravel()))
    _, g = f(0.0, sx.DeviceVector.from_numpy(U), wrt="shape")
    Jp = g(sx.DeviceVector.from_numpy(dpn.ravel())).numpy()

    def moved_residual(eps):
        m2 = sx.Mesh.from_arrays(Xv + d + eps * dd, mfem_hexes(n))
        s2 = sx.Space(m2, p, 3)
        _, perm = cKDTree(s2.node_coords()).query(s.node_coords() + pn + eps * dpn)
        dof2 = (3 * perm[:, None] + np.arange(3)).ravel()
        U2 = np.zeros_like(U)
        U2[dof2] = U
        return make(s2, "neohookean")(0.0, sx.DeviceVector.from_numpy(U2)).numpy()[dof2].copy()

    eps = 1e-5
    fd = (moved_residual(eps) - moved_residual(-eps)) / (2 * eps)
    assert np.abs(fd - Jp).max() <= 1e-6 * np.abs(Jp).max()


def test_shape_gradient_without_shape_and_errors(gpu_lib):
    m = sx.Mesh.cuboid(2, 2, 2)
    m.perturb(42, 0.2)
    s = sx.Space(m, 2, 3)
    f = make(s, "neohookean")
    U = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 3, 2000, -0.01, 0.01, 1))
    # no shape displacement set: the derivative at p = 0 equals the one at an explicit zero field
    _, g0 = f(0.0, U, wrt="shape")
    dp = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 4, 2001, -0.01, 0.01, 1))
    a = g0(dp).numpy()
    zero = sx.DeviceVector(s.ndof)
    f.set_shape_displacement(zero)
    _, g1 = f(0.0, U, wrt="shape")
    assert np.array_equal(g1(dp).numpy(), a)
    for call in (g1.diagonal, lambda: g1.cg(np.array([], dtype=np.int32), dp, sx.DeviceVector(s.ndof))):
        with pytest.raises(sx.SeracAmdError, match="shape gradient"):
            call()
    # setting another shape invalidates the shape gradient
    f.set_shape_displacement(None)
    with pytest.raises(sx.SeracAmdError):
        g1(dp)
    # scalar forms and forms with boundary integrals are refused
    s1 = sx.Space(m, 2, 1)
    heat = sx.Functional(s1, [s1])
    heat.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    with pytest.raises(sx.SeracAmdError, match="3-component"):
        heat(0.0, sx.DeviceVector(s1.ndof), wrt="shape")
    fb = make(s, "neohookean")
    fb.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.7, 1.0), face_mask=1 << 5)
    with pytest.raises(sx.SeracAmdError, match="boundary"):
        fb(0.0, U, wrt="shape")
