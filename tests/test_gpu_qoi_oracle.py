"""QoI values on the GPU against the oracle's restatement
(Functional<double(trials...)>, fn/functional_qoi.inl) on perturbed meshes and
random fields, and QoIs of two trial arguments (DependsOn<0, 1>): the inner
product int u0 . u1 dV, whose gradients are mass residuals of the other
argument, checked against the oracle's mass integrand and central differences."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu
NH = (1.0, 10.0, 0.25)


def pair(p, nc, n=3):
    gm, om = sx.Mesh.cuboid(n, n, n), O.Mesh.cuboid(n, n, n)
    gm.perturb(42, 0.2)
    om.perturb(42, 0.2)
    return sx.Space(gm, p, nc), O.Form(O.Space(om, p, nc))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_qoi_values_against_oracle(gpu_lib, p):
    gs, of = pair(p, 3)
    U = W.nodal_field(gs.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    V = W.nodal_field(gs.nnodes, 3, 4, 2000, -1.0, 1.0, 1)
    Ud, Vd = sx.DeviceVector.from_numpy(U), sx.DeviceVector.from_numpy(V)
    elems = [0, 3, 7, 8, 20, 26]
    cases = [(sx.capi.SX_QOI_MEASURE, (), O.QOI_MEASURE), (sx.capi.SX_QOI_POSITION, (2,), O.QOI_POSITION),
             (sx.capi.SX_QOI_VALUE, (1,), O.QOI_VALUE),
             (sx.capi.SX_QOI_NEOHOOKEAN_ENERGY, NH, O.QOI_NEOHOOKEAN_ENERGY)]
    for gid, prm, oid in cases:
        for el in (None, elems):
            q = sx.QoI(gs)
            q.add_domain_integral(gid, prm, elems=el)
            assert rel(q(0.0, Ud), of.qoi(oid, prm, elems=el, U=U)) < 1e-12
    q2 = sx.QoI([gs, gs])
    q2.add_domain_integral(sx.capi.SX_QOI_INNER, depends_on=(0, 1))
    assert rel(q2(0.0, Ud, Vd), of.qoi(O.QOI_INNER, U=U, U1=V)) < 1e-12


@pytest.mark.parametrize("nc", [1, 3])
def test_two_argument_qoi_gradients(gpu_lib, nc):
    gs, of = pair(2, nc)
    U = W.nodal_field(gs.nnodes, nc, 3, 2000, -0.01, 0.01, 1)
    V = W.nodal_field(gs.nnodes, nc, 4, 2000, -1.0, 1.0, 1)
    Ud, Vd = sx.DeviceVector.from_numpy(U), sx.DeviceVector.from_numpy(V)
    q = sx.QoI([gs, gs])
    q.add_domain_integral(sx.capi.SX_QOI_INNER, depends_on=(0, 1))
    if nc == 3:
        q.add_domain_integral(sx.capi.SX_QOI_NEOHOOKEAN_ENERGY, NH, depends_on=(0,))
    val = q(0.0, Ud, Vd)
    ref = of.qoi(O.QOI_INNER, U=U, U1=V) + (of.qoi(O.QOI_NEOHOOKEAN_ENERGY, NH, U=U) if nc == 3 else 0.0)
    assert rel(val, ref) < 1e-12
    g0 = q.gradient(0.0, Ud, Vd, wrt=0).numpy()
    g1 = q.gradient(0.0, Ud, Vd, wrt=1).numpy()
    # the oracle's mass residual int phi . w dV at w = V (for wrt 0) and w = U (for wrt 1)
    mass = O.Form(of.space)
    mass.add_domain_integral(O.QF_MASS_DIFFUSION, (1.0, 0.0))
    want0 = mass.residual(V)
    if nc == 3:
        nh = O.Form(of.space)
        nh.add_domain_integral(O.QF_NEOHOOKEAN, NH)
        want0 = want0 + nh.residual(U)
    want1 = mass.residual(U)
    assert np.abs(g0 - want0).max() <= 1e-12 * np.abs(want0).max()
    assert np.abs(g1 - want1).max() <= 1e-12 * np.abs(want1).max()
    # central differences of the value
    for wrt, g in ((0, g0), (1, g1)):
        D = W.nodal_field(gs.nnodes, nc, 7 + wrt, 2001, -1.0, 1.0, 1)
        eps = 1e-6
        args_p = [U + eps * D, V] if wrt == 0 else [U, V + eps * D]
        args_m = [U - eps * D, V] if wrt == 0 else [U, V - eps * D]
        fd = (q(0.0, *[sx.DeviceVector.from_numpy(a) for a in args_p]) -
              q(0.0, *[sx.DeviceVector.from_numpy(a) for a in args_m])) / (2 * eps)
        assert abs(fd - g @ D) <= 1e-7 * max(abs(g @ D), 1e-12)


def test_inner_product_of_one_argument_with_itself(gpu_lib):
    gs, of = pair(3, 3)
    U = W.nodal_field(gs.nnodes, 3, 3, 2000, -1.0, 1.0, 1)
    Ud = sx.DeviceVector.from_numpy(U)
    q = sx.QoI(gs)
    q.add_domain_integral(sx.capi.SX_QOI_INNER, depends_on=(0, 0))
    assert rel(q(0.0, Ud), of.qoi(O.QOI_INNER, U=U, U1=U)) < 1e-12
    mass = O.Form(of.space)
    mass.add_domain_integral(O.QF_MASS_DIFFUSION, (1.0, 0.0))
    want = 2.0 * mass.residual(U)
    g = q.gradient(0.0, Ud).numpy()
    assert np.abs(g - want).max() <= 1e-12 * np.abs(want).max()


def test_multi_argument_errors(gpu_lib):
    gs, _ = pair(2, 3)
    s1 = sx.Space(gs.mesh, 2, 1)
    q = sx.QoI([gs, s1])
    with pytest.raises(sx.SeracAmdError, match="same kind"):
        q.add_domain_integral(sx.capi.SX_QOI_INNER, depends_on=(0, 1))
    with pytest.raises(sx.SeracAmdError, match="DependsOn"):
        q.add_domain_integral(sx.capi.SX_QOI_VALUE, (0,), depends_on=(0, 1))
    with pytest.raises(sx.SeracAmdError, match="out of range"):
        q.add_domain_integral(sx.capi.SX_QOI_VALUE, (0,), depends_on=(2,))
    q.add_domain_integral(sx.capi.SX_QOI_VALUE, (0,), depends_on=(1,))
    a, b = sx.DeviceVector(gs.ndof), sx.DeviceVector.from_numpy(np.ones(s1.ndof))
    assert abs(q(0.0, a, b) - 1.0) < 1e-13
    with pytest.raises(sx.SeracAmdError, match="several trial"):
        capi_q = sx.QoI([gs, s1])
        capi_q.add_domain_integral(sx.capi.SX_QOI_MEASURE, depends_on=(0,))
        sx.capi.call("sx_qoi_eval", capi_q.h, 0.0, a.ptr, __import__("ctypes").byref(__import__("ctypes").c_double()))
