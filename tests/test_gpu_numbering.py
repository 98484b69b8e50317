"""Spaces in the caller's node numbering (sx_space_create_h1_hex_map): a
serac binding passes mfem's element -> node map (ElementRestriction's
lexicographic order, fn/element_restriction.cpp:425-467) and its T-vectors.
With coloured scatter every node sums its element contributions in the same
order whatever the numbering, so residual, J.v and diagonal on a randomly
permuted numbering equal the permuted originals bit for bit."""
import numpy as np
import pytest

import serac_amd as sx
from serac_amd import capi
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu


def _perm_space(mesh, p, nc, ordering, seed):
    base = sx.Space(mesh, p, nc, ordering)
    own = base.element_nodes()
    perm = np.random.default_rng(seed).permutation(base.nnodes).astype(np.int32)  # own node -> caller node
    alt = sx.Space(mesh, p, nc, ordering, element_nodes=perm[own])
    assert alt.nnodes == base.nnodes and np.array_equal(alt.element_nodes(), perm[own])
    return base, alt, perm


def _dof_perm(perm, nc, ordering):
    """index array q with v_alt[q] = v_base"""
    nn = len(perm)
    node = np.arange(nn)
    if ordering == sx.SX_ORDERING_BYVDIM:
        return (perm[:, None] * nc + np.arange(nc)[None, :]).ravel()
    return (np.arange(nc)[:, None] * nn + perm[None, :]).ravel() if nc > 1 else perm[node]


CASES = [
    ("heat", sx.SX_QF_HEAT, 1, (1.0, 1.0)),
    ("neohookean", sx.SX_QF_NEOHOOKEAN, 3, (1.0, 10.0, 0.25)),
    ("mass_diffusion", sx.SX_QF_MASS_DIFFUSION, 3, (1.7, 2.1)),
]


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ordering", [sx.SX_ORDERING_BYVDIM, sx.SX_ORDERING_BYNODES])
def test_permuted_numbering_bitwise(gpu_lib, p, case, ordering):
    name, qf, nc, params = case
    mesh = sx.Mesh.cuboid(3, 2, 3)
    mesh.perturb(42, 0.2)
    base, alt, perm = _perm_space(mesh, p, nc, ordering, seed=p)
    q = _dof_perm(perm, nc, ordering)
    scale = 0.01 if nc == 3 else 1.0
    U = W.nodal_field(base.nnodes, nc, 3, 2000, -scale, scale, ordering)
    dU = W.nodal_field(base.nnodes, nc, 4, 3000, -1.0, 1.0, ordering)
    U2, dU2 = np.empty_like(U), np.empty_like(dU)
    U2[q], dU2[q] = U, dU
    outs = []
    for sp, u, du in ((base, U, dU), (alt, U2, dU2)):
        f = sx.Functional(sp, [sp])
        f.add_domain_integral(qf, params)
        r0 = f(0.0, sx.DeviceVector.from_numpy(u)).numpy()
        r, g = f(0.0, sx.DeviceVector.from_numpy(u), wrt=0)
        outs.append((r0, r.numpy(), g(sx.DeviceVector.from_numpy(du)).numpy(), g.diagonal().numpy()))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(b[q], a)


def test_permuted_numbering_unstructured_mesh(gpu_lib, patch_mesh):
    X, hexes = patch_mesh
    mesh = sx.Mesh.from_arrays(X, hexes)
    base, alt, perm = _perm_space(mesh, 2, 3, sx.SX_ORDERING_BYVDIM, seed=7)
    q = _dof_perm(perm, 3, sx.SX_ORDERING_BYVDIM)
    U = W.nodal_field(base.nnodes, 3, 3, 2000, -0.01, 0.01, 1)
    U2 = np.empty_like(U)
    U2[q] = U
    rs = []
    for sp, u in ((base, U), (alt, U2)):
        f = sx.Functional(sp, [sp])
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        rs.append(f(0.0, sx.DeviceVector.from_numpy(u)).numpy())
    assert np.array_equal(rs[1][q], rs[0])
    # node coordinates follow the caller's numbering
    assert np.array_equal(alt.node_coords()[perm], base.node_coords())


def test_boundary_dofs_and_solve_in_caller_numbering(gpu_lib):
    mesh = sx.Mesh.cuboid(3, 3, 3)
    mesh.perturb(42, 0.2)
    base, alt, perm = _perm_space(mesh, 2, 3, sx.SX_ORDERING_BYVDIM, seed=11)
    q = _dof_perm(perm, 3, sx.SX_ORDERING_BYVDIM)
    for face in (0b000001, 0b000010, 0b110000):
        assert np.array_equal(np.sort(q[base.boundary_dofs(face, 0x7)]), alt.boundary_dofs(face, 0x7))
    assert np.array_equal(np.sort(q[base.boundary_dofs_attr([1, 2, 3], 0x5)]), alt.boundary_dofs_attr([1, 2, 3], 0x5))
    # the load-stepped Newton solve gives the permuted solution
    sols = []
    for sp in (base, alt):
        f = sx.Functional(sp, [sp])
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        ess0, ess1 = sp.boundary_dofs(0b000001, 0x7), sp.boundary_dofs(0b000010, 0x1)
        ess = np.union1d(ess0, ess1).astype(np.int32)
        x = sx.DeviceVector(sp.ndof)
        reps = sx.quasistatic_solve(f, x, ess, np.where(np.isin(ess, ess1), -0.02, 0.0), 2, rel_tol=1e-10,
                                    lin_rel_tol=1e-10)
        assert all(r["converged"] for r in reps)
        sols.append(x.numpy())
    assert np.abs(sols[1][q] - sols[0]).max() <= 1e-9 * np.abs(sols[0]).max()


def test_map_errors(gpu_lib):
    mesh = sx.Mesh.cuboid(2, 2, 2)
    base = sx.Space(mesh, 2, 3)
    own = base.element_nodes()
    bad = own.copy()
    bad[0, -1], bad[0, -2] = bad[0, -2], bad[0, -1]  # a shared node renamed in one element only
    with pytest.raises(sx.SeracAmdError, match="consistently"):
        sx.Space(mesh, 2, 3, element_nodes=bad)
    bad = own.copy()
    bad[1, 3] = -1
    with pytest.raises(sx.SeracAmdError, match="out of range"):
        sx.Space(mesh, 2, 3, element_nodes=bad)
    import ctypes

    h = ctypes.c_void_p()
    rc = capi.call_rc("sx_space_create_h1_hex_map", mesh.h, 2, 3, 1, base.nnodes + 1,
                      own.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(h))
    assert rc == capi.SX_ERR_ARG and b"conforming space" in capi.load().sx_last_error()
    # structured-grid features refuse a foreign numbering, and test / trial must agree
    perm = np.arange(base.nnodes, dtype=np.int32)[::-1].copy()
    alt = sx.Space(mesh, 2, 3, element_nodes=perm[own])
    with pytest.raises(sx.SeracAmdError, match="numbering"):
        sx.Multigrid(alt)
    f = sx.Functional(alt, [base])
    with pytest.raises(sx.SeracAmdError, match="number their nodes alike"):
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    # the identity map is the library's own numbering: multigrid accepts it
    same = sx.Space(mesh, 2, 3, element_nodes=own)
    sx.Multigrid(same)
