"""The Integral seam on E-vectors (fn/integral.hpp:127-142): domain integrals
evaluated on [element][component][node] vectors with no gather or scatter.
scatter_add(eval_e(gather(u))) is the residual; the E-vector J.v of each
element equals the oracle's element matrix times that element's values; the
order of a domain subset's elements is the order they were given in."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [("heat", sx.SX_QF_HEAT, O.QF_HEAT, 1, (1.0, 1.0), 3),
         ("neohookean", sx.SX_QF_NEOHOOKEAN, O.QF_NEOHOOKEAN, 3, (1.0, 10.0, 0.25), 2),
         ("mass_diffusion", sx.SX_QF_MASS_DIFFUSION, O.QF_MASS_DIFFUSION, 3, (1.7, 2.1), 1)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_evector_seam(gpu_lib, case):
    _, gid, oid, nc, params, p = case
    gm, om = sx.Mesh.cuboid(3, 2, 3), O.Mesh.cuboid(3, 2, 3)
    gm.perturb(42, 0.2)
    om.perturb(42, 0.2)
    gs, os_ = sx.Space(gm, p, nc), O.Space(om, p, nc)
    elems = [5, 0, 17, 9]  # a domain subset, in a given order
    for el in (None, elems):
        gf = sx.Functional(gs, [gs])
        gf.add_domain_integral(gid, params, elems=el)
        of = O.Form(os_)
        of.add_domain_integral(oid, params, el)
        U = W.nodal_field(gs.nnodes, nc, 3, 2000, -0.01, 0.01, 1)
        dU = W.nodal_field(gs.nnodes, nc, 4, 3000, -1.0, 1.0, 1)
        R_ref = of.linearize(U)
        # E-vectors over the integral's domain, in its order
        EU, EdU = os_.gather(U).reshape(gs.ne, -1), os_.gather(dU).reshape(gs.ne, -1)
        dom = list(range(gs.ne)) if el is None else el
        uE = sx.DeviceVector.from_numpy(EU[dom].ravel())
        duE = sx.DeviceVector.from_numpy(EdU[dom].ravel())
        assert gf.integral_info(0) == (len(dom), gs.npe, nc)
        rE = gf.eval_e(uE, linearize=True).numpy().reshape(len(dom), -1)
        # scatter of the E residual is the residual
        full = np.zeros((gs.ne, nc * gs.npe))
        full[dom] = rE
        R = os_.scatter_add(full.ravel())
        assert np.abs(R - R_ref).max() <= 1e-12 * np.abs(R_ref).max()
        # E-vector J.v = element matrix x element values
        _, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        drE = grad.apply_e(duE).numpy().reshape(len(dom), -1)
        for k, e in enumerate(dom):
            Ke = of.element_matrix(0, k)
            want = Ke @ EdU[e]
            assert np.abs(drE[k] - want).max() <= 1e-12 * max(np.abs(want).max(), 1e-300)
        # and scattering it gives the L-vector J.v
        full[dom] = drE
        assert np.abs(os_.scatter_add(full.ravel()) - of.jvp(dU)).max() <= 1e-12 * np.abs(of.jvp(dU)).max()


def test_evector_errors(gpu_lib):
    gm = sx.Mesh.cuboid(2, 2, 2)
    gs = sx.Space(gm, 2, 3)
    gf = sx.Functional(gs, [gs])
    gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    ne, npe, nc = gf.integral_info(0)
    uE = sx.DeviceVector(ne * npe * nc)
    with pytest.raises(sx.SeracAmdError, match="out of range"):
        gf.eval_e(uE, k=1)
    _, g = gf(0.0, sx.DeviceVector(gs.ndof), wrt=0)
    g.apply_e(uE)  # linearized through the L-vector call: fine
    gf.set_shape_displacement(sx.DeviceVector(gs.ndof))
    with pytest.raises(sx.SeracAmdError, match="shape-aware"):
        gf.eval_e(uE)
