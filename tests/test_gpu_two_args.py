"""Domain integrals of two trial arguments (DependsOn<0, 1>) on the GPU
against the oracle, whose integrand reads both arguments in one call (the
solid materials with their acceleration, phy/solid_mechanics.hpp:834-852; the
heat integrand with the temperature rate, phy/heat_transfer.hpp:380-406):

- the residual R(U, A);
- differentiate_wrt(U): the linearized residual, J.v and the diagonal;
- differentiate_wrt(A): J.v (the oracle's R(U, dA) - R(U, 0), exact since A
  enters linearly) and the diagonal (the oracle's mass integrand);
- the E-vector entry points refuse such an integral and count the integrals
  as added; argument errors."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import capi
from serac_amd import workloads as W

from test_gpu_parity import relerr

pytestmark = pytest.mark.gpu

CASES = {
    "neohookean": (sx.SX_QF_NEOHOOKEAN, O.QF_NEOHOOKEAN, (1.3, 10.0, 0.25), 3, 1.3, 2),
    "linear_elastic": (sx.SX_QF_LINEAR_ELASTIC, O.QF_LINEAR_ELASTIC, (1.2, 0.8, 0.7), 3, 0.7, 2),
    "j2": (sx.SX_QF_J2_LINEAR, O.QF_J2_LINEAR, (200.0, 0.3, 0.5, 2.0, 0.0, 1.9, 1.0), 3, 1.9, 2),
    "heat": (sx.SX_QF_HEAT, O.QF_HEAT, (1.1, 0.4, 2.5), 1, 2.5, 3),
}


def _setup(case, ordering=1):
    gid, oid, params, nc, coef, p = CASES[case]
    gm, om = sx.Mesh.cuboid(3, 3, 3), O.Mesh.cuboid(3, 3, 3)
    gm.perturb(42, 0.2)
    om.perturb(42, 0.2)
    gs, os_ = sx.Space(gm, p, nc, ordering), O.Space(om, p, nc, ordering)
    gf, of = sx.Functional(gs, [gs, gs]), O.Form(os_)
    qd = None
    if oid == O.QF_J2_LINEAR:
        qd = sx.QuadratureData(gs.ne, (p + 1) ** 3, np.zeros(10))
        gf.add_domain_integral(gid, params, depends_on=(0, 1), qdata=qd)
        of.add_domain_integral2(oid, params, state_init=np.zeros(10))
    else:
        gf.add_domain_integral(gid, params, depends_on=(0, 1))
        of.add_domain_integral2(oid, params)
    mass = O.Form(os_)
    mass.add_domain_integral(O.QF_MASS_DIFFUSION, (coef, 0.0))
    return gs, gf, of, mass, qd


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("ordering", [1, 0])
def test_two_argument_integral_against_oracle(gpu_lib, case, ordering):
    gs, gf, of, mass, _ = _setup(case, ordering)
    nc = gs.nc
    amp = 1.0 if case == "heat" else 0.01
    U = W.nodal_field(gs.nnodes, nc, 3, 2000, -amp, amp, ordering)
    A = W.nodal_field(gs.nnodes, nc, 4, 2001, -1.0, 1.0, ordering)
    dU = W.nodal_field(gs.nnodes, nc, 5, 2002, -1.0, 1.0, ordering)
    dA = W.nodal_field(gs.nnodes, nc, 6, 2003, -1.0, 1.0, ordering)
    Ud, Ad = sx.DeviceVector.from_numpy(U), sx.DeviceVector.from_numpy(A)
    R_ref = of.residual2(U, A)
    assert relerr(gf(0.0, Ud, Ad).numpy(), R_ref) < 1e-12
    # differentiate_wrt(U)
    out, g0 = gf(0.0, Ud, Ad, wrt=0)
    assert relerr(out.numpy(), of.linearize2(U, A)) < 1e-12
    assert relerr(g0(sx.DeviceVector.from_numpy(dU)).numpy(), of.jvp(dU)) < 1e-12
    assert relerr(g0.diagonal().numpy(), of.diagonal()) < 1e-12
    # differentiate_wrt(A): the mass term
    out1, g1 = gf(0.0, Ud, Ad, wrt=1)
    assert relerr(out1.numpy(), R_ref) < 1e-12
    jv1_ref = of.residual2(U, dA) - of.residual2(U, np.zeros_like(dA))
    assert relerr(g1(sx.DeviceVector.from_numpy(dA)).numpy(), jv1_ref) < 1e-11
    mass.linearize(np.zeros_like(U))
    assert relerr(g1.diagonal().numpy(), mass.diagonal()) < 1e-12


def test_two_argument_errors_and_evector_indexing(gpu_lib):
    gs, gf, _, _, _ = _setup("neohookean")
    # a one-argument integral after the two-argument one is E-vector integral 1
    gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.3, 10.0, 0.25), depends_on=(0,))
    assert gf.integral_info(0) == gf.integral_info(1)
    with pytest.raises(sx.SeracAmdError) as e:
        uE = sx.DeviceVector(gs.ne * gs.npe * 3)
        capi.call("sx_form_eval_e", gf.h, 0, 0.0, uE.ptr, uE.ptr, 0)
    assert e.value.code == capi.SX_ERR_UNSUPPORTED
    with pytest.raises(sx.SeracAmdError) as e:
        gf.add_domain_integral(sx.SX_QF_MASS_DIFFUSION, (1.0, 1.0), depends_on=(0, 1))
    assert e.value.code == capi.SX_ERR_UNSUPPORTED
    with pytest.raises(sx.SeracAmdError) as e:
        gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.3, 10.0, 0.25), depends_on=(0, 0))
    assert e.value.code == capi.SX_ERR_ARG
    with pytest.raises(sx.SeracAmdError) as e:
        gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.3, 10.0, 0.25), depends_on=(0, 2))
    assert e.value.code == capi.SX_ERR_ARG
