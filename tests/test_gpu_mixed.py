"""Mixed-precision operator and iterative-refinement Newton: FP32 element
arithmetic on FP64 vectors (SX_OPT_JVP_ARITHMETIC) with FP32 / BF16 stored
tangents, FP64 residuals, dot products and norms.  The converged solution must
be the FP64 one: Newton on the FP64 residual corrects the FP32 operator."""
import numpy as np
import pytest

import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_parity import Pair, relerr

pytestmark = pytest.mark.gpu


def mixed(form, tprec):
    form.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, tprec)
    form.set_option(sx.capi.SX_OPT_JVP_ARITHMETIC, sx.capi.SX_PREC_F32)


@pytest.mark.parametrize("tprec,tol", [(sx.capi.SX_PREC_F64, 1e-6), (sx.capi.SX_PREC_F32, 1e-6),
                                       (sx.capi.SX_PREC_BF16, 1.5e-2)])
@pytest.mark.parametrize("scatter", [0, 1])
def test_mixed_operator_against_oracle(gpu_lib, tprec, tol, scatter):
    pr = Pair(3, 3, 3, perturb=0.2)
    gf = sx.Functional(pr.gs, [pr.gs])
    gf.set_option(sx.capi.SX_OPT_SCATTER, scatter)
    gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    mixed(gf, tprec)
    import oracle as O

    of = O.Form(pr.os)
    of.add_domain_integral(O.QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    U = W.nodal_field(pr.gs.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    assert relerr(out.numpy(), of.linearize(U)) < 1e-12  # the residual stays FP64
    dU = W.nodal_field(pr.gs.nnodes, 3, 5, 3000, -1.0, 1.0, 1)
    jv = grad(sx.DeviceVector.from_numpy(dU))
    assert jv.dtype == np.float64
    ref = of.jvp(dU)
    assert np.linalg.norm(jv.numpy() - ref) / np.linalg.norm(ref) < tol


# The BF16 tangent (about 1e-3 relative per entry) is an operator option only:
# rounding the 45 entries independently perturbs the near-rigid modes of this
# ill-conditioned elasticity operator enough that Newton on it does not
# converge, so the mixed-precision Newton uses the FP32 tangent.
# p = 4 (config 3's order) runs the production path: serac's quasiStaticSolve
# warm start from the stress-free state (an SPD tangent), then Newton, with the
# multigrid preconditioner the mixed path is meant for.  Started from the
# loaded face alone, the first tangent is indefinite (the compressive stress
# exceeds the shear modulus, K = 10, G = 0.25), and a preconditioned CG on it
# with the FP32-arithmetic operator (about 2e-7 relative) converges or breaks
# down depending on the last bits of its inputs: Jacobi-CG stalled near 5e-2
# with one of two diagonals equal to 1e-16, and multigrid-CG took 165 (FP64)
# against 316-328 (mixed, varying between runs) iterations for that solve.
@pytest.mark.parametrize("tprec", [sx.capi.SX_PREC_F32])
@pytest.mark.parametrize("p", [2, 4])
def test_mixed_newton_reaches_fp64_solution(gpu_lib, tprec, p):
    pr = Pair(3, p, 3, perturb=0.2)
    ess0 = pr.gs.boundary_dofs(0b000001, 0x7)
    ess1 = pr.gs.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    x0 = np.zeros(pr.gs.ndof)
    # the x = max face moves by 2 % (p = 2, no warm start) or 0.4 % (p = 4: the
    # first Lobatto node layer sits at 0.17 h)
    x0[ess1] = -0.02 if p == 2 else -0.004
    sols = []
    for mode in ("f64", "mixed"):
        gf = sx.Functional(pr.gs, [pr.gs])
        gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        if mode == "mixed":
            mixed(gf, tprec)
        if p == 4:
            mg = sx.Multigrid(pr.gs)
            gf.set_multigrid(mg)
            x = sx.DeviceVector(pr.gs.ndof)
            reps = sx.quasistatic_solve(gf, x, ess, x0[ess], 1, warm_start=True, rel_tol=1e-10, lin_rel_tol=1e-6)
            rep = reps[-1]
        else:
            x = sx.DeviceVector.from_numpy(x0)
            rep = gf.newton(x, ess, rel_tol=1e-10, abs_tol=0.0, max_iter=30, lin_rel_tol=1e-6, lin_abs_tol=0.0,
                            lin_max_iter=20000)
        assert rep["converged"], (mode, rep)
        sols.append(x.numpy())
        # the FP64 residual (essential rows zeroed) meets the tolerance
        R = gf(0.0, x).numpy()
        R[ess] = 0.0
        assert np.linalg.norm(R) <= 1e-10 * rep["initial_norm"] * 1.0001
        gf.set_multigrid(None)
    assert relerr(sols[1], sols[0]) < 1e-8


def test_mixed_quasistatic_matches_fp64(gpu_lib):
    pr = Pair(3, 2, 3, perturb=0.2)
    ess0 = pr.gs.boundary_dofs(0b000001, 0x7)
    ess1 = pr.gs.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    final = np.where(np.isin(ess, ess1), -0.03, 0.0)
    sols = []
    for mode in ("f64", "mixed"):
        gf = sx.Functional(pr.gs, [pr.gs])
        gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        if mode == "mixed":
            mixed(gf, sx.capi.SX_PREC_F32)
        x = sx.DeviceVector(pr.gs.ndof)
        reps = sx.quasistatic_solve(gf, x, ess, final, 3, warm_start=True, rel_tol=1e-10, lin_rel_tol=1e-6)
        assert all(r["converged"] for r in reps)
        sols.append(x.numpy())
    assert relerr(sols[1], sols[0]) < 1e-8


def test_jvp_arithmetic_option_validation(gpu_lib):
    pr = Pair(2, 1, 1)
    gf, _ = pr.forms("heat")
    with pytest.raises(sx.SeracAmdError):
        gf.set_option(sx.capi.SX_OPT_JVP_ARITHMETIC, sx.capi.SX_PREC_BF16)
