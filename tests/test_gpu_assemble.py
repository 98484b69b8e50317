"""Gradient::assemble parity: the device CSR matrix against the oracle's element
matrices (element_gradient_kernel) assembled densely, its sparsity pattern
against the DOF pairs that share an element (bit-exact), and CSR SpMV against
the matrix-free J.v."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_parity import QF, Pair, relerr

pytestmark = pytest.mark.gpu


def lidx(sp, node, c):
    return node * sp.nc + c if sp.ordering == 1 else c * sp.nnodes + node


def oracle_dense(pr, of, elems=None):
    sp = pr.os
    emap = sp.emap()
    n = sp.ndof
    K = np.zeros((n, n))
    dom = range(sp.ne) if elems is None else elems
    for slot, e in enumerate(dom):
        Ke = of.element_matrix(0, slot)
        d = np.array([lidx(sp, emap[e, j], c) for c in range(sp.nc) for j in range(sp.npe)])
        K[np.ix_(d, d)] += Ke
    return K


def expected_pattern(pr):
    sp = pr.os
    emap = sp.emap()
    n = sp.ndof
    P = np.zeros((n, n), dtype=bool)
    for e in range(sp.ne):
        d = np.array([lidx(sp, emap[e, j], c) for c in range(sp.nc) for j in range(sp.npe)])
        P[np.ix_(d, d)] = True
    return P


@pytest.mark.parametrize("name,p,ordering", [
    ("heat", 2, 1), ("mass_diffusion", 2, 1), ("mass_diffusion_vec", 1, 0), ("neohookean", 1, 1),
    ("neohookean", 2, 0), ("neohookean", 3, 1), ("linear_elastic", 2, 1),
])
@pytest.mark.parametrize("scatter", [0, 1])
def test_assemble_matches_oracle(gpu_lib, name, p, ordering, scatter):
    gid, oid, nc, params = QF[name]
    pr = Pair(2, p, nc, ordering=ordering)
    gf = sx.Functional(pr.gs, [pr.gs])
    gf.set_option(sx.capi.SX_OPT_SCATTER, scatter)
    gf.add_domain_integral(gid, params)
    of = O.Form(pr.os)
    of.add_domain_integral(oid, params)
    U = W.nodal_field(pr.gs.nnodes, nc, 3, 2000, -0.01, 0.01, ordering)
    _, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    of.linearize(U)
    A = grad.assemble()
    rp, col, val = A.arrays()
    n = pr.gs.ndof
    assert A.shape == (n, n)
    # pattern: exactly the DOF pairs sharing an element, sorted within rows
    P = np.zeros((n, n), dtype=bool)
    rows = np.repeat(np.arange(n), np.diff(rp))
    P[rows, col] = True
    assert np.array_equal(P, expected_pattern(pr))
    assert all(np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0) for i in range(n))
    # values
    K = oracle_dense(pr, of)
    D = A.to_scipy().toarray()
    assert np.abs(D - K).max() <= 1e-12 * np.abs(K).max()
    # SpMV equals the matrix-free action
    dU = W.nodal_field(pr.gs.nnodes, nc, 5, 3000, -1.0, 1.0, ordering)
    y = sx.DeviceVector(n)
    A.mult(sx.DeviceVector.from_numpy(dU), y)
    jv = grad(sx.DeviceVector.from_numpy(dU)).numpy()
    assert relerr(y.numpy(), jv) < 1e-12
    # the diagonal entries are the Jacobi diagonal
    assert relerr(np.diag(D), grad.diagonal().numpy()) < 1e-12
    if name in ("neohookean", "linear_elastic", "heat"):
        assert np.abs(D - D.T).max() <= 1e-12 * np.abs(D).max()


def test_assemble_colored_is_reproducible(gpu_lib):
    pr = Pair(3, 2, 3)
    gf, of = pr.forms("neohookean")
    U = W.nodal_field(pr.gs.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    _, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    v1 = grad.assemble().arrays()[2]
    v2 = grad.assemble().arrays()[2]
    assert np.array_equal(v1, v2)


def test_assemble_j2_nonsymmetric(gpu_lib):
    from test_gpu_j2 import J2Pair, J2_PARAMS

    pr = J2Pair(2, 2, J2_PARAMS["c4"], state_seed=3)
    U = pr.shear_field(0.02, 1e-3, 12)
    _, grad = pr.gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    pr.of.linearize(U)
    D = grad.assemble().to_scipy().toarray()
    sp = pr.os
    emap = sp.emap()
    K = np.zeros_like(D)
    for e in range(sp.ne):
        Ke = pr.of.element_matrix(0, e)
        d = np.array([lidx(sp, emap[e, j], c) for c in range(3) for j in range(sp.npe)])
        K[np.ix_(d, d)] += Ke
    assert np.abs(D - K).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(D - D.T).max() > 1e-6 * np.abs(D).max()


def test_assemble_null_gradient_is_an_error(gpu_lib):
    pr = Pair(2, 1, 1)
    gf2 = sx.Functional(pr.gs, [pr.gs])
    gf2.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    with pytest.raises(sx.SeracAmdError):
        sx.Gradient(gf2, sx.capi._vp()).assemble()
