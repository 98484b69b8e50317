"""The geometric multigrid preconditioner (the GPU stand-in for the reference's
Preconditioner::HypreAMG, num/equation_solver.cpp:1014-1021) against a numpy
restatement built from the assembled Jacobian: the Galerkin level operators
P^T A P that probing finds, the whole V-cycle z = M r (Chebyshev-Jacobi
smoothing with the GPU's eigenvalue estimates, masked transfers, the dense
coarsest solve), its symmetry and positivity, and multigrid-CG and the Newton
loops against direct solves and Jacobi-CG."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_gmres import eliminated
from test_gpu_learned import interp_1d, setup

pytestmark = pytest.mark.gpu

SAFETY = 1.15


def trilinear_1d(nc):
    """Vertex level l + 1 -> l on one direction: (2 (nc - 1) + 1) x nc."""
    P = np.zeros((2 * (nc - 1) + 1, nc))
    for i in range(nc):
        P[2 * i, i] = 1.0
        if 2 * i - 1 >= 0:
            P[2 * i - 1, i] = 0.5
        if 2 * i + 1 < P.shape[0]:
            P[2 * i + 1, i] = 0.5
    return P


def kron3(Px, Py=None, Pz=None):
    """The 3-component tensor-product transfer (x fastest), DOF = node * 3 + c."""
    Py = Px if Py is None else Py
    Pz = Px if Pz is None else Pz
    P = sp.kron(sp.csr_matrix(Pz), sp.kron(sp.csr_matrix(Py), sp.csr_matrix(Px))).tocsr()
    return sp.kron(P, sp.identity(3)).tocsr()


def coinciding(fine_dims, idx):
    """Flat indices of the fine points idx[d] (per direction) in a [z][y][x] grid of fine_dims (x, y, z)."""
    ix, iy, iz = idx
    return ((iz[:, None, None] * fine_dims[1] + iy[None, :, None]) * fine_dims[0] + ix[None, None, :]).ravel()


def hierarchy(n, p, grad, ess, info):
    """The numpy restatement: level operators, masks and transfers; n is the
    interval count per direction (an int for a cube, or (nx, ny, nz))."""
    ns = (n, n, n) if np.isscalar(n) else tuple(n)
    nn = [p * k + 1 for k in ns]
    keep0 = np.ones(3 * int(np.prod(nn)))
    keep0[ess] = 0.0
    Z0 = sp.diags(keep0)
    A0 = (Z0 @ grad.assemble().to_scipy() @ Z0).tocsr()
    levels = [{"A": A0, "keep": keep0}]
    # level 1: the vertex grid; a vertex component is constrained when the
    # node at the vertex is
    nv = [k + 1 for k in ns]
    node = coinciding(nn, [p * np.arange(k) for k in nv])
    keep1 = keep0.reshape(-1, 3)[node].ravel()
    Ps = [(Z0 @ kron3(*[interp_1d(k, p) for k in ns]) @ sp.diags(keep1)).tocsr()]
    keeps = [keep1]
    dims = [nv]
    for lvl in info["levels"][2:]:
        nc = list(lvl["dims"])
        nf = dims[-1]
        vert = coinciding(nf, [2 * np.arange(k) for k in nc])
        kc = keeps[-1].reshape(-1, 3)[vert].ravel()
        Ps.append((sp.diags(keeps[-1]) @ kron3(*[trilinear_1d(k) for k in nc]) @ sp.diags(kc)).tocsr())
        keeps.append(kc)
        dims.append(nc)
    for P, k in zip(Ps, keeps):
        A = (P.T @ levels[-1]["A"] @ P).tocsr()
        levels.append({"A": A, "keep": k, "P": P})
    return levels


def chebyshev_coeffs(lmax, rng, deg):
    up = SAFETY * lmax
    lo = up / rng
    th, de = 0.5 * (up + lo), 0.5 * (up - lo)
    sigma, rho, c = th / de, de / th, []
    for _ in range(2, deg + 1):
        rn = 1.0 / (2.0 * sigma - rho)
        c.append((rn * rho, 2.0 * rn / de))
        rho = rn
    return th, c


def vcycle_numpy(levels, info, fine_deg, coarse_deg, rng, r, ess):
    L = len(levels)
    for lvl, gi in zip(levels, info["levels"]):
        d = lvl["A"].diagonal()
        lvl["dinv"] = np.where(lvl["keep"] > 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
        if lvl is not levels[-1]:
            lvl["theta"], lvl["c"] = chebyshev_coeffs(gi["lmax"], rng, fine_deg if lvl is levels[0] else coarse_deg)
    Ac = levels[-1]["A"].toarray()
    Ac = 0.5 * (Ac + Ac.T) + np.diag(1.0 - levels[-1]["keep"])
    Ainv = np.linalg.inv(Ac)

    def smooth(l, b, x, pre):
        V = levels[l]
        A, dinv = V["A"], V["dinv"]
        if pre:
            d = dinv * b / V["theta"]
            x, rr = d.copy(), b.copy()
        else:
            rr = b - A @ x
            d = dinv * rr / V["theta"]
            x = x + d
        for c1, c2 in V["c"]:
            rr = rr - A @ d
            d = c1 * d + c2 * dinv * rr
            x = x + d
        if pre:
            rr = rr - A @ d
        return x, rr

    def vcycle(l, b):
        if l == L - 1:
            return Ainv @ b
        x, rr = smooth(l, b, None, True)
        P = levels[l + 1]["P"]
        xc = vcycle(l + 1, P.T @ rr)
        x = x + P @ xc
        x, _ = smooth(l, b, x, False)
        return x

    b = r.copy()
    b[ess] = 0.0
    z = vcycle(0, b)
    z[ess] = r[ess]
    return z


def linearized(n, p, amp=0.0002):
    m, s, f, ess = setup(n, p)
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -amp, amp, 1)
    _, grad = f(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    return m, s, f, ess, grad


@pytest.mark.parametrize("n,p", [(4, 2), (8, 1), (4, 3)])
def test_levels_are_galerkin_products_and_the_vcycle_matches_numpy(gpu_lib, n, p):
    m, s, f, ess, grad = linearized(n, p)
    mg = sx.Multigrid(s, fine_degree=3, coarse_degree=2, smoothing_range=10.0)
    mg.setup(grad, ess)
    info = mg.info()
    assert info["levels"][0]["dims"] == (p * n + 1,) * 3
    assert info["levels"][1]["dims"] == (n + 1,) * 3
    assert info["levels"][-1]["ndof"] <= 200 or (info["levels"][-1]["dims"][0] - 1) % 2 == 1
    levels = hierarchy(n, p, grad, ess, info)
    assert [lv["A"].shape[0] for lv in levels] == [lv["ndof"] for lv in info["levels"]]
    rng = np.random.default_rng(5)
    for lvl, lv in enumerate(levels):
        x = rng.standard_normal(lv["A"].shape[0]) * lv["keep"]
        y = sx.DeviceVector(len(x))
        mg.level_apply(lvl, sx.DeviceVector.from_numpy(x), y)
        ref = lv["A"] @ x
        assert np.abs(y.numpy() - ref).max() <= 1e-12 * np.abs(ref).max(), lvl
    # the power-iteration estimates bracket the largest eigenvalue of D^-1 A
    for lv, gi in zip(levels[:-1], info["levels"][:-1]):
        A = lv["A"].toarray()
        keep = lv["keep"] > 0
        Af = A[np.ix_(keep, keep)]
        d = np.diag(Af)
        lam = np.linalg.eigvalsh(Af / np.sqrt(np.outer(d, d))).max()
        assert 0.6 * lam <= gi["lmax"] <= 1.25 * lam, (gi["lmax"], lam)
    r = W.nodal_field(s.nnodes, 3, 7, 3000, -1.0, 1.0, 1)
    z = sx.DeviceVector(s.ndof)
    mg.apply(sx.DeviceVector.from_numpy(r), z)
    zref = vcycle_numpy(levels, info, 3, 2, 10.0, r, ess)
    assert np.abs(z.numpy() - zref).max() <= 1e-10 * np.abs(zref).max()


@pytest.mark.parametrize("p,tprec", [(1, "f64"), (2, "f64"), (3, "f64"), (4, "f64"), (3, "f32"), (2, "f32")])
def test_elementwise_vertex_operator_equals_the_probed_one(gpu_lib, p, tprec):
    n = 4
    m, s, f, ess = setup(n, p)
    if tprec == "f32":
        f.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, sx.capi.SX_PREC_F32)
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.0002, 0.0002, 1)
    _, grad = f(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    mgs = []
    for mode in (0, 1):
        mg = sx.Multigrid(s)
        mg.set_galerkin(mode)
        mg.setup(grad, ess)
        mgs.append(mg)
    assert mgs[0].info()["element_setups"] == 1 and mgs[1].info()["element_setups"] == 0
    nv1 = mgs[0].info()["levels"][1]["ndof"]
    x = sx.DeviceVector.from_numpy(np.random.default_rng(2).standard_normal(nv1))
    ys = []
    for mg in mgs:
        y = sx.DeviceVector(nv1)
        mg.level_apply(1, x, y)
        ys.append(y.numpy())
    assert np.abs(ys[0] - ys[1]).max() <= 1e-11 * np.abs(ys[1]).max()


def test_constraints_off_the_vertices_fall_back_to_probing(gpu_lib):
    m, s, f, ess = setup(4, 2)
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.0002, 0.0002, 1)
    _, grad = f(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    # one more constrained DOF at an edge-midpoint node inside the mesh
    nn = 2 * 4 + 1
    node = (4 * nn + 4) * nn + 3
    ess2 = np.union1d(ess, [3 * node]).astype(np.int32)
    mg = sx.Multigrid(s)
    mg.setup(grad, ess2)
    assert mg.info()["element_setups"] == 0
    mg.setup(grad, ess)
    assert mg.info()["element_setups"] == 1


def test_non_cubic_mesh_levels_and_vcycle_match_numpy(gpu_lib):
    # different interval counts per direction (8 x 4 x 6, p = 2): vertex grids
    # 9x5x7 and 5x3x4 (dense); catches a swapped direction in any transfer
    ns, p = (8, 4, 6), 2
    m = sx.Mesh.cuboid(*ns)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    ess = s.boundary_dofs(0b000001, 0x7)
    U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.0002, 0.0002, 1)
    _, grad = f(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
    for galerkin in (0, 1):
        mg = sx.Multigrid(s, fine_degree=3, coarse_degree=2, smoothing_range=10.0)
        mg.set_galerkin(galerkin)
        mg.setup(grad, ess)
        info = mg.info()
        assert [lv["dims"] for lv in info["levels"]] == [(17, 9, 13), (9, 5, 7), (5, 3, 4)]
        assert info["element_setups"] == (1 if galerkin == 0 else 0)
        levels = hierarchy(ns, p, grad, ess, info)
        rng = np.random.default_rng(5)
        for lvl, lv in enumerate(levels):
            x = rng.standard_normal(lv["A"].shape[0]) * lv["keep"]
            y = sx.DeviceVector(len(x))
            mg.level_apply(lvl, sx.DeviceVector.from_numpy(x), y)
            ref = lv["A"] @ x
            assert np.abs(y.numpy() - ref).max() <= 1e-12 * np.abs(ref).max(), (galerkin, lvl)
        r = W.nodal_field(s.nnodes, 3, 7, 3000, -1.0, 1.0, 1)
        z = sx.DeviceVector(s.ndof)
        mg.apply(sx.DeviceVector.from_numpy(r), z)
        zref = vcycle_numpy(levels, info, 3, 2, 10.0, r, ess)
        assert np.abs(z.numpy() - zref).max() <= 1e-10 * np.abs(zref).max()


def test_bf16_smoother_level0_is_the_bf16_linearization(gpu_lib):
    # the BF16 copy of an FP64 tangent rounds the same doubles as a BF16
    # linearization, so the level-0 operator equals that form's J.v bit for bit
    n, p = 4, 3
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    ess = s.boundary_dofs(0b000001, 0x7)
    U = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 3, 2000, -0.0002, 0.0002, 1))
    grads = []
    for tprec in (sx.capi.SX_PREC_F64, sx.capi.SX_PREC_BF16):
        f = sx.Functional(s, [s])
        f.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_COLORED)
        f.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, tprec)
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        _, g = f(0.0, U, wrt=0)
        grads.append((f, g))
    mg = sx.Multigrid(s)
    mg.set_smoother_precision(sx.capi.SX_PREC_BF16)
    mg.setup(grads[0][1], ess)
    x = np.random.default_rng(4).standard_normal(s.ndof)
    x[ess] = 0.0
    y = sx.DeviceVector(s.ndof)
    mg.level_apply(0, sx.DeviceVector.from_numpy(x), y)
    yref = grads[1][1](sx.DeviceVector.from_numpy(x)).numpy()
    yref[ess] = 0.0
    assert np.array_equal(y.numpy(), yref)
    # and differs from the FP64 operator at BF16 rounding
    y64 = grads[0][1](sx.DeviceVector.from_numpy(x)).numpy()
    y64[ess] = 0.0
    err = np.abs(y.numpy() - y64).max() / np.abs(y64).max()
    assert 1e-6 < err < 1e-2


def test_bf16_smoother_vcycle_is_symmetric_and_newton_matches(gpu_lib):
    m, s, f, ess, grad = linearized(8, 2)
    mg = sx.Multigrid(s)
    mg.set_smoother_precision(sx.capi.SX_PREC_BF16)
    mg.setup(grad, ess)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(s.ndof), rng.standard_normal(s.ndof)
    Mu, Mv = sx.DeviceVector(s.ndof), sx.DeviceVector(s.ndof)
    mg.apply(sx.DeviceVector.from_numpy(u), Mu)
    mg.apply(sx.DeviceVector.from_numpy(v), Mv)
    Mu, Mv = Mu.numpy(), Mv.numpy()
    assert abs(v @ Mu - u @ Mv) <= 1e-10 * abs(v @ Mu)
    assert u @ Mu > 0 and v @ Mv > 0
    # Newton on the 5 % compression path in mixed precision with the BF16 smoother
    ess0 = s.boundary_dofs(0b000001, 0x7)
    ess1 = s.boundary_dofs(0b000010, 0x1)
    essc = np.union1d(ess0, ess1).astype(np.int32)
    final = np.where(np.isin(essc, ess1), -0.05, 0.0)
    xs = []
    for use_mg in (False, True):
        fm = sx.Functional(s, [s])
        fm.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        mgm = None
        if use_mg:
            fm.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, sx.capi.SX_PREC_F32)
            fm.set_option(sx.capi.SX_OPT_JVP_ARITHMETIC, sx.capi.SX_PREC_F32)
            mgm = sx.Multigrid(s)
            mgm.set_smoother_precision(sx.capi.SX_PREC_BF16)
        fm.set_multigrid(mgm)
        x = sx.DeviceVector(s.ndof)
        reps = sx.quasistatic_solve(fm, x, essc, final, 2, rel_tol=1e-10, lin_rel_tol=1e-8, lin_max_iter=50000)
        assert all(r["converged"] for r in reps)
        xs.append(x.numpy())
    assert np.abs(xs[1] - xs[0]).max() <= 1e-7 * np.abs(xs[0]).max()


@pytest.mark.parametrize("bf16", [False, True])
def test_fp32_smoother_vectors_match_the_fp64_vcycle(gpu_lib, bf16):
    m, s, f, ess, grad = linearized(8, 3)
    zs = []
    r = W.nodal_field(s.nnodes, 3, 7, 3000, -1.0, 1.0, 1)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(s.ndof), rng.standard_normal(s.ndof)
    for prec in (sx.capi.SX_PREC_F64, sx.capi.SX_PREC_F32):
        mg = sx.Multigrid(s)
        if bf16:
            mg.set_smoother_precision(sx.capi.SX_PREC_BF16)
        mg.set_smoother_vectors(prec)
        mg.setup(grad, ess)
        z = sx.DeviceVector(s.ndof)
        mg.apply(sx.DeviceVector.from_numpy(r), z)
        zs.append(z.numpy())
        if prec == sx.capi.SX_PREC_F32:
            Mu, Mv = sx.DeviceVector(s.ndof), sx.DeviceVector(s.ndof)
            mg.apply(sx.DeviceVector.from_numpy(u), Mu)
            mg.apply(sx.DeviceVector.from_numpy(v), Mv)
            Mu, Mv = Mu.numpy(), Mv.numpy()
            assert abs(v @ Mu - u @ Mv) <= 1e-5 * abs(v @ Mu)
            assert u @ Mu > 0 and v @ Mv > 0
    assert np.abs(zs[1] - zs[0]).max() <= 1e-4 * np.abs(zs[0]).max()
    assert np.array_equal(zs[1][ess], r[ess])


def test_vcycle_is_symmetric_and_positive(gpu_lib):
    m, s, f, ess, grad = linearized(8, 2)
    mg = sx.Multigrid(s)
    mg.setup(grad, ess)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(s.ndof), rng.standard_normal(s.ndof)
    Mu, Mv = sx.DeviceVector(s.ndof), sx.DeviceVector(s.ndof)
    mg.apply(sx.DeviceVector.from_numpy(u), Mu)
    mg.apply(sx.DeviceVector.from_numpy(v), Mv)
    Mu, Mv = Mu.numpy(), Mv.numpy()
    assert abs(v @ Mu - u @ Mv) <= 1e-11 * abs(v @ Mu)
    assert u @ Mu > 0 and v @ Mv > 0


@pytest.mark.parametrize("p", [1, 2, 3])
def test_multigrid_cg_matches_a_direct_solve_in_far_fewer_iterations(gpu_lib, p):
    n = 8
    m, s, f, ess, grad = linearized(n, p)
    b = W.nodal_field(s.nnodes, 3, 8, 3000, -1.0, 1.0, 1)
    b[ess] = 0.0
    A = eliminated(grad.assemble().to_scipy(), ess)
    x_ref = spla.spsolve(A.tocsc(), b)
    mg = sx.Multigrid(s)
    x = sx.DeviceVector(s.ndof)
    rep = grad.mg_cg(mg, ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-10, max_iter=2000)
    assert rep["converged"]
    assert np.abs(x.numpy() - x_ref).max() <= 1e-7 * np.abs(x_ref).max()
    xj = sx.DeviceVector(s.ndof)
    repj = grad.cg(ess, sx.DeviceVector.from_numpy(b), xj, rel_tol=1e-10, max_iter=50000)
    assert rep["linear_iterations"] * 5 < repj["linear_iterations"], (rep, repj)
    assert mg.info()["setups"] == 1


@pytest.mark.parametrize("arith", ["f64", "mixed_f32"])
def test_newton_with_multigrid_matches_jacobi_newton(gpu_lib, arith):
    n, p = 8, 2
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    ess0 = s.boundary_dofs(0b000001, 0x7)
    ess1 = s.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    final = np.where(np.isin(ess, ess1), -0.05, 0.0)

    def solve(mg):
        f = sx.Functional(s, [s])
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        if arith == "mixed_f32" and mg is not None:
            f.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, sx.capi.SX_PREC_F32)
            f.set_option(sx.capi.SX_OPT_JVP_ARITHMETIC, sx.capi.SX_PREC_F32)
        f.set_multigrid(mg)
        x = sx.DeviceVector(s.ndof)
        reps = sx.quasistatic_solve(f, x, ess, final, 2, rel_tol=1e-10, lin_rel_tol=1e-8, lin_max_iter=50000)
        return x.numpy(), reps

    x_j, reps_j = solve(None)
    mg = sx.Multigrid(s)
    x_m, reps_m = solve(mg)
    assert np.abs(x_m - x_j).max() <= 1e-7 * np.abs(x_j).max()
    assert all(r["converged"] for r in reps_m)
    assert sum(r["linear_iterations"] for r in reps_m) * 5 < sum(r["linear_iterations"] for r in reps_j)
    # one setup per linearization: the warm-start solve of each load step and each Newton step
    assert mg.info()["setups"] == sum(r["newton_iterations"] + 1 for r in reps_m)


def test_a_hierarchy_reused_across_a_load_step_reaches_the_same_solution(gpu_lib):
    n, p = 8, 2
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = sx.Space(m, p, 3, 1)
    ess0 = s.boundary_dofs(0b000001, 0x7)
    ess1 = s.boundary_dofs(0b000010, 0x1)
    ess = np.union1d(ess0, ess1).astype(np.int32)
    final = np.where(np.isin(ess, ess1), -0.05, 0.0)
    xs, setups = [], []
    for reuse in (1, 3):
        f = sx.Functional(s, [s])
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        mg = sx.Multigrid(s)
        mg.set_reuse(reuse)
        f.set_multigrid(mg)
        x = sx.DeviceVector(s.ndof)
        reps = sx.quasistatic_solve(f, x, ess, final, 2, rel_tol=1e-10, lin_rel_tol=1e-8)
        xs.append(x.numpy())
        nsolves = sum(r["newton_iterations"] + 1 for r in reps)
        setups.append((mg.info()["setups"], nsolves))
    assert setups[0][0] == setups[0][1]
    assert setups[1][0] == -(-setups[1][1] // 3)  # one hierarchy per three linear solves
    assert np.abs(xs[1] - xs[0]).max() <= 1e-7 * np.abs(xs[0]).max()
    with pytest.raises(sx.SeracAmdError):
        mg.set_reuse(0)


def test_j2_plastic_path_with_multigrid_matches_jacobi(gpu_lib):
    # J2 plasticity (small strain, kinematic + isotropic hardening): the 81-entry
    # tangent takes the probed vertex level; cyclic shear of the top face with
    # the state committed after every step, against the same path with Jacobi-CG
    from test_gpu_j2 import J2_PARAMS

    n, p = 4, 2
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.1)
    s = sx.Space(m, p, 3, 1)
    bot = s.boundary_dofs(0b010000, 0x7)
    top = s.boundary_dofs(0b100000, 0x7)
    ess = np.union1d(bot, top).astype(np.int32)
    top_x = np.isin(ess, s.boundary_dofs(0b100000, 0x1))
    runs = []
    for use_mg in (False, True):
        qd = sx.QuadratureData(s.ne, (p + 1) ** 3, np.zeros(10))
        f = sx.Functional(s, [s])
        f.add_domain_integral(sx.SX_QF_J2_LINEAR, J2_PARAMS["kinematic_small_strain"], qdata=qd)
        mg = sx.Multigrid(s) if use_mg else None
        f.set_multigrid(mg)
        x = sx.DeviceVector(s.ndof)
        its = 0
        for g in (0.004, 0.008, 0.0):
            reps = sx.quasistatic_solve(f, x, ess, np.where(top_x, g, 0.0), 1, rel_tol=1e-10, lin_rel_tol=1e-10,
                                        lin_max_iter=20000)
            assert reps[0]["converged"]
            its += reps[0]["linear_iterations"]
        runs.append((x.numpy(), qd.numpy(), its, mg.info() if mg else None))
        f.set_multigrid(None)
    (xj, sj, itj, _), (xm, sm, itm, info) = runs
    assert np.abs(xm - xj).max() <= 1e-7 * np.abs(xj).max()
    assert np.abs(sm - sj).max() <= 1e-7 * np.abs(sj).max()
    assert sj[..., 9].max() > 0  # the path yielded
    assert info["element_setups"] == 0 and info["setups"] > 0
    assert itm < itj


def test_multigrid_gmres_meets_the_true_residual_goal(gpu_lib):
    m, s, f, ess, grad = linearized(8, 2)
    b = W.nodal_field(s.nnodes, 3, 8, 3000, -1.0, 1.0, 1)
    b[ess] = 0.0
    A = eliminated(grad.assemble().to_scipy(), ess)
    mg = sx.Multigrid(s)
    x = sx.DeviceVector(s.ndof)
    rep = grad.mg_gmres(mg, ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-9, max_iter=500, restart=50)
    assert rep["converged"] and rep["linear_iterations"] < 100
    assert np.linalg.norm(b - A @ x.numpy()) <= 1.01e-9 * np.linalg.norm(b)


def test_j2_geometrically_nonlinear_path_with_multigrid_gmres(gpu_lib):
    # the J2 tangent under geometric nonlinearity is not symmetric: Newton with
    # multigrid-FGMRES against Newton with Jacobi-CG on the same cyclic path
    from test_gpu_j2 import J2_PARAMS

    n, p = 4, 2
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.1)
    s = sx.Space(m, p, 3, 1)
    bot = s.boundary_dofs(0b010000, 0x7)
    top = s.boundary_dofs(0b100000, 0x7)
    ess = np.union1d(bot, top).astype(np.int32)
    top_x = np.isin(ess, s.boundary_dofs(0b100000, 0x1))
    runs = []
    for use_mg in (False, True):
        qd = sx.QuadratureData(s.ne, (p + 1) ** 3, np.zeros(10))
        f = sx.Functional(s, [s])
        f.add_domain_integral(sx.SX_QF_J2_LINEAR, J2_PARAMS["c4"], qdata=qd)
        mg = None
        if use_mg:
            mg = sx.Multigrid(s)
            f.set_multigrid(mg)
            f.set_option(sx.capi.SX_OPT_LINEAR_SOLVER, sx.capi.SX_LINSOLVER_GMRES)
        x = sx.DeviceVector(s.ndof)
        for g in (0.004, 0.008, 0.0):
            reps = sx.quasistatic_solve(f, x, ess, np.where(top_x, g, 0.0), 1, rel_tol=1e-10, lin_rel_tol=1e-10,
                                        lin_max_iter=20000)
            assert reps[0]["converged"]
        runs.append((x.numpy(), qd.numpy(), mg.info() if mg else None))
        f.set_multigrid(None)
    (xj, sj, _), (xm, sm, info) = runs
    assert np.abs(xm - xj).max() <= 1e-7 * np.abs(xj).max()
    assert np.abs(sm - sj).max() <= 1e-7 * np.abs(sj).max()
    assert info["setups"] > 0


def test_multigrid_arguments_are_checked(gpu_lib, patch_mesh):
    s1 = sx.Space(sx.Mesh.cuboid(4, 4, 4), 2, 1, 1)
    with pytest.raises(sx.SeracAmdError):
        sx.Multigrid(s1)  # one component
    with pytest.raises(sx.SeracAmdError):
        X, hexes = patch_mesh
        sx.Multigrid(sx.Space(sx.Mesh.from_arrays(X, hexes), 2, 3, 1))  # not a cuboid
    with pytest.raises(sx.SeracAmdError):
        sx.Multigrid(sx.Space(sx.Mesh.cuboid(17, 17, 17), 1, 3, 1))  # 18^3 x 3 coarsest DOFs
    s = sx.Space(sx.Mesh.cuboid(4, 4, 4), 2, 3, 1)
    mg = sx.Multigrid(s)
    with pytest.raises(sx.SeracAmdError):
        mg.configure(fine_degree=0)
    with pytest.raises(sx.SeracAmdError):
        mg.set_galerkin(2)
    with pytest.raises(sx.SeracAmdError):
        mg.set_smoother_precision(sx.capi.SX_PREC_F32)
    with pytest.raises(sx.SeracAmdError):
        mg.set_smoother_vectors(sx.capi.SX_PREC_BF16)
    with pytest.raises(sx.SeracAmdError):
        mg.apply(sx.DeviceVector(s.ndof), sx.DeviceVector(s.ndof))  # before a setup
    f = sx.Functional(s1, [s1])
    with pytest.raises(sx.SeracAmdError):
        f.set_multigrid(mg)  # another space
