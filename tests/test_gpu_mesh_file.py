"""Meshes read from MFEM files and boundary selection by attribute, on the GPU:
attribute-selected DOFs and boundary integrals against the bounding-box plane
selection, and the reference's J2 beam (physics/tests/solid.cpp: beam-hex.mesh,
support attribute 1, tip attribute 2) against the oracle."""
import os

import numpy as np
import pytest

import oracle as O
import serac_amd as sx
from serac_amd import workloads as W

from golden.make_mesh_fixtures import parse_mfem_mesh
from test_gpu_j2 import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHES = os.path.join(ROOT, "tests", "golden", "meshes")
# mfem Cartesian boundary attribute -> face_mask bit of the same plane
ATTR_PLANE = {1: 4, 2: 2, 3: 1, 4: 3, 5: 0, 6: 5}


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("nc,ordering", [(1, 1), (3, 1), (3, 0)])
def test_boundary_dofs_by_attribute_match_planes(gpu_lib, p, nc, ordering):
    m = sx.Mesh.cuboid(3, 4, 2)
    m.perturb(5, 0.2)
    s = sx.Space(m, p, nc, ordering)
    comp = 0x1 if nc == 1 else 0x5
    for attr, bit in ATTR_PLANE.items():
        assert np.array_equal(s.boundary_dofs_attr([attr], comp), s.boundary_dofs(1 << bit, comp))
    assert np.array_equal(s.boundary_dofs_attr([1, 3, 5], comp),
                          s.boundary_dofs((1 << 4) | (1 << 1) | (1 << 0), comp))
    assert np.array_equal(s.boundary_dofs_attr(range(1, 7), comp), s.boundary_dofs(0x3f, comp))
    assert len(s.boundary_dofs_attr([7], comp)) == 0


def test_boundary_integrals_by_attribute_match_planes(gpu_lib):
    m = sx.Mesh.cuboid(3, 3, 3)
    m.perturb(9, 0.2)
    s = sx.Space(m, 2, 3, 1)
    U = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 4, 2000, -0.01, 0.01, 1))
    dU = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 5, 2001, -1.0, 1.0, 1))
    traction = (0.1, 0.0, 0.0, 0.0, 0.2, 0.0, 0.0, 0.0, 0.3, 0.05, -0.02, 0.01)
    out = []
    for sel in ({"attrs": [3]}, {"face_mask": 1 << 1}):
        f = sx.Functional(s, [s])
        f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
        f.add_boundary_integral(sx.SX_BQF_TRACTION, traction, **sel)
        f.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.3, 1.0), **({"attrs": range(1, 7)} if "attrs" in sel
                                                                   else {"face_mask": 0}))
        r, g = f(0.0, U, wrt=0)
        out.append((r.numpy(), g(dU).numpy(), g.diagonal().numpy()))
    for a, b in zip(out[0], out[1]):
        assert relerr(a, b) < 1e-13


def test_boundary_integral_needs_exterior_faces(gpu_lib):
    # a boundary element on the interface between the two halves of twohex
    text = open(os.path.join(MESHES, "twohex.mesh")).read()
    verts, hexes, attrs, quads, battrs = parse_mfem_mesh(None, text)
    shared = sorted(set(hexes[0]) & set(hexes[1]))
    assert len(shared) == 4
    i = text.index("boundary")
    n = int(text[i:].split("\n")[1])
    text = text.replace(f"boundary\n{n}\n", f"boundary\n{n + 1}\n7 3 {shared[0]} {shared[1]} {shared[3]} {shared[2]}\n")
    m = sx.Mesh.read_mfem(text=text)
    s = sx.Space(m, 1, 3, 1)
    # DOF selection accepts the interior face, a boundary integral does not
    assert len(s.boundary_dofs_attr([7])) == 12
    f = sx.Functional(s, [s])
    with pytest.raises(sx.SeracAmdError, match="interior face"):
        f.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.3, 1.0), attrs=[7])


def test_j2_beam_from_mesh_file_matches_oracle(gpu_lib):
    # physics/tests/solid.cpp: p = 2, J2SmallStrain<LinearHardening> with
    # E = 10000, nu = 0.25, sigma_y = 50, Hi = 50, Hk = 5, density 1, small
    # strain; support (attribute 1) fixed, tip (attribute 2) moved by
    # (0, 0, t (t - 1)) over 10 steps to t = 1
    params = (10000.0, 0.25, 50.0, 50.0, 5.0, 1.0, 0.0)
    path = os.path.join(MESHES, "beam-hex.mesh")
    gm = sx.Mesh.read_mfem(path)
    verts, hexes = parse_mfem_mesh(path)[:2]
    om = O.Mesh.from_arrays(np.array(verts), np.array(hexes, dtype=np.int32))
    gs, os_ = sx.Space(gm, 2, 3, 1), O.Space(om, 2, 3, 1)
    support, tip = gs.boundary_dofs_attr([1]), gs.boundary_dofs_attr([2])
    xyz = os_.node_coords()

    def plane_dofs(x):
        nodes = np.nonzero(np.abs(xyz[:, 0] - x) < 1e-12)[0]
        return np.sort(np.concatenate([3 * nodes + c for c in range(3)])).astype(np.int32)

    assert np.array_equal(support, plane_dofs(0.0)) and np.array_equal(tip, plane_dofs(8.0))
    ess = np.union1d(support, tip).astype(np.int32)
    tip_z = np.isin(ess, tip[2::3])
    nq = 27
    qd = sx.QuadratureData(gs.ne, nq, np.zeros(10))
    gf = sx.Functional(gs, [gs])
    gf.add_domain_integral(sx.SX_QF_J2_LINEAR, params, qdata=qd)
    of = O.Form(os_)
    of.add_domain_integral_state(O.QF_J2_LINEAR, params, np.zeros(10), None)
    xg = sx.DeviceVector(gs.ndof)
    x = np.zeros(gs.ndof)
    for k in range(1, 11):
        t = 0.1 * k
        target = np.where(tip_z, t * (t - 1.0), 0.0)
        # elastic steps: the warm start already solves the step, and the residual
        # left (~1e-10 against forces ~1e2) is round-off, so the goal is absolute
        reps = sx.quasistatic_solve(gf, xg, ess, target, 1, warm_start=True, rel_tol=1e-10, abs_tol=1e-8,
                                    lin_rel_tol=1e-12, lin_max_iter=20000)
        assert reps[0]["converged"]
        du = np.zeros_like(x)
        du[ess] = target - x[ess]
        R = of.linearize(x)
        rhs = -(R + of.jvp(du))
        rhs[ess] = 0.0
        d, crep = of.cg(ess, rhs, rel=1e-12, maxit=20000)
        assert crep["converged"]
        x = x + d + du
        x, rep = of.newton(ess, x, rel=1e-10, abs_tol=1e-8, lin_rel=1e-12, lin_abs=0.0, lin_max=20000)
        assert rep["converged"]
        of.update_qdata(True)
        of.residual(x)
        of.update_qdata(False)
        assert relerr(xg.numpy(), x) < 1e-7, k
    st, st_ref = qd.numpy(), of.qdata()
    assert np.abs(st - st_ref).max() <= 1e-7 * np.abs(st_ref).max()
    # the beam yielded and keeps a permanent plastic strain after unloading
    assert st_ref[..., 9].max() > 1e-4
