"""Configs 1, 3, 4 and 5 at their own sizes on the GPU (config 2's are in
test_gpu_fullsize.py):

- c1, heat p=2 on 16^3 (35,937 DOFs): residual, J.v and Jacobi-CG to 1e-10
  straight against the oracle on the whole mesh;
- c3, Neo-Hookean p=4 on 48^3 perturbed hexes (21.6M DOFs): the FP64 operator
  is linear, symmetric and the central difference of the residual, oracle
  parity on sampled elements; the FP32 tangent with FP32 element arithmetic
  stays within 1e-6 of the FP64 J.v;
- c4, J2 plasticity p=2 on 96^3 (21.6M DOFs, 23.9M points x 10 state
  doubles, 81-entry tangent): linear, not symmetric (geometric
  nonlinearity), the central difference of the residual, and sampled-element
  parity with the state committed by updateQdata;
- c5, Neo-Hookean p=2 on 128^3 (50.9M DOFs): linear, symmetric, the central
  difference of the residual, oracle parity on sampled elements.

Sampled elements always include the last one, whose nodes carry the largest
ids, so the full-size index ranges are exercised."""
import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

pytestmark = pytest.mark.gpu

NH = (1.0, 10.0, 0.25)
J2 = (200.0, 0.3, 0.5, 2.0, 0.0, 1.0, 1.0)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def sample(ne, k, seed):
    e = np.sort(np.random.default_rng(seed).choice(ne, k, replace=False)).astype(np.int32)
    e[-1] = ne - 1
    e[0] = 0
    return np.unique(e)


def linear_and_symmetric(grad, nnodes, symmetric, tol=1e-12):
    v = W.direction(nnodes, 3, seed=1)
    w = W.direction(nnodes, 3, seed=2)
    Jv = grad(sx.DeviceVector.from_numpy(v)).numpy()
    Jw = grad(sx.DeviceVector.from_numpy(w)).numpy()
    Jc = grad(sx.DeviceVector.from_numpy(2.0 * v - 3.0 * w)).numpy()
    scale = np.abs(Jv).max() + np.abs(Jw).max()
    assert np.abs(Jc - (2.0 * Jv - 3.0 * Jw)).max() < tol * scale
    a, b = float(w @ Jv), float(v @ Jw)
    bound = np.sqrt(float(Jv @ Jv) * float(w @ w))
    if symmetric:
        assert abs(a - b) < 1e-11 * bound
    else:
        assert abs(a - b) > 1e-6 * bound
    return v, Jv


def central_difference(form, U, v, Jv, eps=1e-6, tol=1e-6):
    Rp = form(0.0, sx.DeviceVector.from_numpy(U + eps * v)).numpy().copy()
    Rm = form(0.0, sx.DeviceVector.from_numpy(U - eps * v)).numpy().copy()
    fd = (Rp - Rm) / (2 * eps)
    assert np.linalg.norm(fd - Jv) < tol * np.linalg.norm(Jv)


# ---------------------------------------------------------------- c1
def test_c1_heat_16_whole_mesh_against_oracle(gpu_lib):
    n, p = 16, 2
    gm, om = sx.Mesh.cuboid(n, n, n), O.Mesh.cuboid(n, n, n)
    gs, os_ = sx.Space(gm, p, 1), O.Space(om, p, 1)
    assert gs.ndof == 35937
    gf = sx.Functional(gs, [gs])
    gf.add_domain_integral(sx.SX_QF_HEAT, (1.0, 1.0))
    of = O.Form(os_)
    of.add_domain_integral(O.QF_HEAT, (1.0, 1.0))
    T = W.heat_temperature(gs.node_coords())
    dT = W.direction(gs.nnodes, 1)
    out, grad = gf(0.0, sx.DeviceVector.from_numpy(T), wrt=0)
    R_ref = of.linearize(T)
    assert rel(out.numpy(), R_ref) < 1e-12
    assert rel(grad(sx.DeviceVector.from_numpy(dT)).numpy(), of.jvp(dT)) < 1e-12
    assert rel(grad.diagonal().numpy(), of.diagonal()) < 1e-12
    # Jacobi-CG to 1e-10 on the Newton step of the linear problem, T = 0 on x = 0
    ess = gs.boundary_dofs(0b000001, 0x1)
    b = -R_ref
    b[ess] = 0.0
    x_ref, rep_ref = of.cg(ess, b, rel=1e-10)
    x = sx.DeviceVector(gs.ndof)
    rep = grad.cg(ess, sx.DeviceVector.from_numpy(b), x, rel_tol=1e-10)
    assert rep["converged"] and rep_ref["converged"]
    assert abs(rep["linear_iterations"] - rep_ref["iterations"]) <= 2
    assert rel(x.numpy(), x_ref) < 1e-8
    # T + x solves the heat problem: the oracle residual vanishes off the Dirichlet face
    R = of.residual(T + x.numpy())
    R[ess] = 0.0
    assert np.linalg.norm(R) <= 1e-8 * np.linalg.norm(R_ref)


# ---------------------------------------------------------------- c3
class TestC3:
    n, p = 48, 4

    @pytest.fixture(scope="class")
    def c3(self, gpu_lib):
        mesh = sx.Mesh.cuboid(self.n, self.n, self.n)
        mesh.perturb(42, 0.2)
        space = sx.Space(mesh, self.p, 3)
        assert space.ndof == 21567171
        form = sx.Functional(space, [space])
        form.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_ATOMIC)
        form.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH)
        U = W.neohookean_displacement(space.node_coords(), noise=W.noise_for_mesh(self.n * self.p / 3))
        _, grad = form(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        return space, form, grad, U

    def test_linear_symmetric_and_fd(self, c3):
        space, form, grad, U = c3
        v, Jv = linear_and_symmetric(grad, space.nnodes, symmetric=True)
        central_difference(form, U, v, Jv)

    def test_fp32_operator_error_at_full_size(self, c3):
        space, form, grad, U = c3
        f32 = sx.Functional(space, [space])
        f32.set_option(sx.capi.SX_OPT_TANGENT_PRECISION, sx.capi.SX_PREC_F32)
        f32.set_option(sx.capi.SX_OPT_JVP_ARITHMETIC, sx.capi.SX_PREC_F32)
        f32.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH)
        out32, g32 = f32(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        out64 = form(0.0, sx.DeviceVector.from_numpy(U)).numpy()
        assert rel(out32.numpy(), out64) < 1e-12  # the residual stays FP64
        v = W.direction(space.nnodes, 3, seed=5)
        a = g32(sx.DeviceVector.from_numpy(v)).numpy()
        b = grad(sx.DeviceVector.from_numpy(v)).numpy()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6

    def test_sampled_elements_match_oracle(self, c3):
        space, _, _, U = c3
        elems = sample(space.ne, 48, 0)
        gf = sx.Functional(space, [space])
        gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH, elems=elems)
        om = O.Mesh.cuboid(self.n, self.n, self.n)
        om.perturb(42, 0.2)
        of = O.Form(O.Space(om, self.p, 3))
        of.add_domain_integral(O.QF_NEOHOOKEAN, NH, elems)
        v = W.direction(space.nnodes, 3, seed=4)
        out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        assert rel(out.numpy(), of.linearize(U)) < 1e-12
        assert rel(grad(sx.DeviceVector.from_numpy(v)).numpy(), of.jvp(v)) < 1e-12
        assert rel(grad.diagonal().numpy(), of.diagonal()) < 1e-12


# ---------------------------------------------------------------- c4
def shear(xyz, gamma, noise, seed, nnodes):
    U = np.zeros(3 * nnodes)
    U[0::3] = gamma * xyz[:, 1] + 0.3 * gamma * xyz[:, 2] ** 2
    U[1::3] = 0.5 * gamma * xyz[:, 0] * xyz[:, 2]
    U[2::3] = -0.2 * gamma * np.sin(2 * xyz[:, 0])
    return U + W.nodal_field(nnodes, 3, seed, 2000, -noise, noise, 1)


class TestC4:
    n, p = 96, 2

    @pytest.fixture(scope="class")
    def c4(self, gpu_lib):
        mesh = sx.Mesh.cuboid(self.n, self.n, self.n)
        space = sx.Space(mesh, self.p, 3)
        assert space.ndof == 21567171
        nq = (self.p + 1) ** 3
        qd = sx.QuadratureData(space.ne, nq, np.zeros(10))
        form = sx.Functional(space, [space])
        form.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_ATOMIC)
        form.add_domain_integral(sx.SX_QF_J2_LINEAR, J2, qdata=qd)
        U = shear(space.node_coords(), 0.03, 1e-5, 11, space.nnodes)
        _, grad = form(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        return space, form, grad, U, qd

    def test_linear_nonsymmetric_and_fd(self, c4):
        space, form, grad, U, qd = c4
        v, Jv = linear_and_symmetric(grad, space.nnodes, symmetric=False)
        # the return map is smooth inside the plastic regime: every point yields
        # under 3% shear, and +-1e-7 moves no point across the yield surface
        central_difference(form, U, v, Jv, eps=1e-7, tol=1e-5)
        assert np.abs(qd.numpy()[:4]).max() == 0.0  # nothing committed

    def test_sampled_elements_and_commit_match_oracle(self, c4):
        space, _, _, U, _ = c4
        elems = sample(space.ne, 64, 1)
        nq = (self.p + 1) ** 3
        qd = sx.QuadratureData(len(elems), nq, np.zeros(10))
        gf = sx.Functional(space, [space])
        gf.add_domain_integral(sx.SX_QF_J2_LINEAR, J2, elems=elems, qdata=qd)
        om = O.Mesh.cuboid(self.n, self.n, self.n)
        of = O.Form(O.Space(om, self.p, 3))
        of.add_domain_integral_state(O.QF_J2_LINEAR, J2, np.zeros(10), elems)
        v = W.direction(space.nnodes, 3, seed=6)
        out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        assert rel(out.numpy(), of.linearize(U)) < 1e-12
        assert rel(grad(sx.DeviceVector.from_numpy(v)).numpy(), of.jvp(v)) < 1e-12
        # commit the state (updateQdata) on both sides
        gf.update_qdata(True)
        of.update_qdata(True)
        gf(0.0, sx.DeviceVector.from_numpy(U))
        of.residual(U)
        st, st_ref = qd.numpy(), of.qdata(nelem=len(elems))
        assert np.abs(st_ref[..., 9]).min() > 0.0  # every sampled point yielded
        assert rel(st, st_ref) < 1e-12


# ---------------------------------------------------------------- c5
class TestC5:
    n, p = 128, 2

    @pytest.fixture(scope="class")
    def c5(self, gpu_lib):
        mesh = sx.Mesh.cuboid(self.n, self.n, self.n)
        space = sx.Space(mesh, self.p, 3)
        assert space.ndof == 50923779
        form = sx.Functional(space, [space])
        form.set_option(sx.capi.SX_OPT_SCATTER, sx.capi.SX_SCATTER_ATOMIC)
        form.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH)
        U = W.neohookean_displacement(space.node_coords(), noise=W.noise_for_mesh(self.n * self.p / 3))
        _, grad = form(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        return space, form, grad, U

    def test_linear_symmetric_and_fd(self, c5):
        space, form, grad, U = c5
        v, Jv = linear_and_symmetric(grad, space.nnodes, symmetric=True)
        central_difference(form, U, v, Jv)
        d = grad.diagonal().numpy()
        assert d.min() > 0.0 and np.isfinite(d).all()

    def test_sampled_elements_match_oracle(self, c5):
        space, _, _, U = c5
        elems = sample(space.ne, 64, 2)
        gf = sx.Functional(space, [space])
        gf.add_domain_integral(sx.SX_QF_NEOHOOKEAN, NH, elems=elems)
        of = O.Form(O.Space(O.Mesh.cuboid(self.n, self.n, self.n), self.p, 3))
        of.add_domain_integral(O.QF_NEOHOOKEAN, NH, elems)
        v = W.direction(space.nnodes, 3, seed=7)
        out, grad = gf(0.0, sx.DeviceVector.from_numpy(U), wrt=0)
        assert rel(out.numpy(), of.linearize(U)) < 1e-12
        assert rel(grad(sx.DeviceVector.from_numpy(v)).numpy(), of.jvp(v)) < 1e-12
