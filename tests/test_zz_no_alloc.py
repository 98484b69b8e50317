"""The allocation-free Jacobi-CG (sx_cg_solve_noalloc) against sx_cg_solve.

Both run the same CG on the same linearized gradient; with the coloured
(reproducible) scatter their iterates are identical, so the solutions must
agree bit for bit and take the same iteration count.  Two calls in a row, the
second on a smaller space, show that nothing the first call left in the shared
scratch array leaks into the second; a space too large for the scratch array
is refused with SX_ERR_ARG before anything runs.
"""
import ctypes

import numpy as np
import pytest

import serac_amd as sx
from serac_amd import capi
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu


def _neohookean_gradient(n, p):
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3)
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    U = W.neohookean_displacement(s.node_coords())
    _, grad = f(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    ess = s.boundary_dofs(0b000001, 0x7)  # x = 0 face clamped
    return s, f, grad, np.ascontiguousarray(ess, dtype=np.int32)


def _noalloc(grad, ess, b, x, rel_tol, max_iter):
    rep = capi.SolveReport()
    iptr = ess.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    capi.call("sx_cg_solve_noalloc", grad.h, iptr, len(ess), b.ptr, x.ptr, rel_tol, 0.0, max_iter, ctypes.byref(rep))
    return rep.as_dict()


@pytest.mark.parametrize("sizes", [((6, 2), (3, 2)), ((4, 3), (2, 1))])
def test_noalloc_cg_equals_cg_on_two_calls(gpu_lib, sizes):
    for n, p in sizes:  # the second call is on a smaller space
        s, f, grad, ess = _neohookean_gradient(n, p)
        b_h = W.direction(s.nnodes, 3)
        b_h[ess] = 0.0
        b = sx.DeviceVector.from_numpy(b_h)
        x_ref = sx.DeviceVector.from_numpy(np.full(s.ndof, 3.0))
        x_na = sx.DeviceVector.from_numpy(np.full(s.ndof, -5.0))
        rep_ref = grad.cg(ess, b, x_ref, rel_tol=1e-8, max_iter=40000)
        rep_na = _noalloc(grad, ess, b, x_na, 1e-8, 40000)
        assert rep_na["converged"] and rep_ref["converged"]
        assert rep_na["linear_iterations"] == rep_ref["linear_iterations"] > 0
        assert rep_na["final_norm"] == rep_ref["final_norm"]
        assert np.array_equal(x_na.numpy(), x_ref.numpy()), (n, p)


def test_noalloc_cg_refuses_a_space_beyond_the_scratch(gpu_lib):
    # 1-component p=4 on 40^3: 161^3 = 4.17M DOFs, 167 MB of work vectors
    m = sx.Mesh.cuboid(40, 40, 40)
    s = sx.Space(m, 4, 1)
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    _, grad = f(0.0, sx.DeviceVector(s.ndof), wrt=0)
    b = sx.DeviceVector(s.ndof)
    x = sx.DeviceVector(s.ndof)
    none = np.zeros(0, dtype=np.int32)
    rep = capi.SolveReport()
    rc = capi.call_rc("sx_cg_solve_noalloc", grad.h, none.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 0, b.ptr,
                      x.ptr, 1e-10, 0.0, 10, ctypes.byref(rep))
    assert rc == capi.SX_ERR_ARG
    assert b"bytes of scratch" in capi.load().sx_last_error()
    # out-of-range essential DOFs are refused as sx_cg_solve refuses them
    bad = np.array([0, s.ndof], dtype=np.int32)
    rc = capi.call_rc("sx_cg_solve_noalloc", grad.h, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 2, b.ptr,
                      x.ptr, 1e-10, 0.0, 10, ctypes.byref(rep))
    assert rc == capi.SX_ERR_ARG
