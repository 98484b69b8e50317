"""Boundary integrals over caller-selected faces on the GPU:
Domain::ofBoundaryElements(mesh, predicate) (fn/domain.cpp:312-386) through a
C predicate on each exterior face's corners and attribute, and explicit lists
of exterior faces, against the plane and attribute selections (the same faces
give the same residual, J.v and diagonal: bit for bit when the faces come in
the same order, to rounding otherwise), for residual forms and QoIs."""
import os

import numpy as np
import pytest

import serac_amd as sx
from serac_amd import workloads as W

from test_gpu_j2 import relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRACTION = (0.1, 0.0, 0.0, 0.0, 0.2, 0.0, 0.0, 0.0, 0.3, 0.05, -0.02, 0.01)


def _eval(s, U, dU, **sel):
    f = sx.Functional(s, [s])
    f.add_domain_integral(sx.SX_QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    f.add_boundary_integral(sx.SX_BQF_TRACTION, TRACTION, **sel)
    f.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.3, 1.0), **sel)
    r, g = f(0.0, U, wrt=0)
    return r.numpy(), g(dU).numpy(), g.diagonal().numpy()


def _cube(n=3, p=2):
    m = sx.Mesh.cuboid(n, n, n)
    m.perturb(9, 0.2)
    s = sx.Space(m, p, 3, 1)
    U = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 4, 2000, -0.01, 0.01, 1))
    dU = sx.DeviceVector.from_numpy(W.nodal_field(s.nnodes, 3, 5, 2001, -1.0, 1.0, 1))
    return m, s, U, dU


def test_predicate_on_corners_equals_planes_bitwise(gpu_lib):
    m, s, U, dU = _cube()
    calls = []

    def x_max_or_z_min(corners, attr):
        calls.append(attr)
        return bool(np.all(np.abs(corners[:, 0] - 1.0) < 1e-12) or np.all(np.abs(corners[:, 2]) < 1e-12))

    a = _eval(s, U, dU, predicate=x_max_or_z_min)
    b = _eval(s, U, dU, face_mask=(1 << 1) | (1 << 4))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # called once per exterior face and integral (6 sides of 3 x 3 faces, two integrals), with the
    # cuboid's boundary attributes
    assert len(calls) == 2 * 54 and set(calls) == set(range(1, 7))


def test_predicate_on_attributes_equals_attribute_selection(gpu_lib):
    m, s, U, dU = _cube()
    a = _eval(s, U, dU, predicate=lambda corners, attr: attr in (3, 5))
    b = _eval(s, U, dU, attrs=[3, 5])
    for x, y in zip(a, b):
        assert relerr(x, y) < 1e-13


def test_face_list_in_any_order(gpu_lib):
    m, s, U, dU = _cube()
    quads, attrs = m.boundary()
    sel = quads[np.isin(attrs, [1, 4])]
    rng = np.random.default_rng(0)
    shuffled = np.array([rng.permutation(q) for q in sel[rng.permutation(len(sel))]], dtype=np.int32)
    a = _eval(s, U, dU, faces=shuffled)
    b = _eval(s, U, dU, attrs=[1, 4])
    for x, y in zip(a, b):
        assert relerr(x, y) < 1e-13


def test_face_selection_errors(gpu_lib):
    m, s, U, dU = _cube(n=2, p=1)
    f = sx.Functional(s, [s])
    X = m.vertices()
    # the interface between elements 0 and 1 (the x = 1/2 plane, y, z < 1/2) is interior
    inner = np.where((np.abs(X[:, 0] - 0.5) < 0.3) & (X[:, 1] < 0.7) & (X[:, 2] < 0.7))[0]
    assert len(inner) == 4
    with pytest.raises(sx.SeracAmdError, match="not an exterior face"):
        f.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.3, 1.0), faces=inner.reshape(1, 4))
    with pytest.raises(sx.SeracAmdError, match="not an exterior face"):
        f.add_boundary_integral(sx.SX_BQF_PRESSURE, (0.3, 1.0), faces=[[0, 1, 2, 3]])


def test_qoi_over_selected_faces(gpu_lib):
    m, s, U, dU = _cube()
    quads, attrs = m.boundary()
    values = []
    for sel in ({"predicate": lambda c, a: bool(np.all(np.abs(c[:, 1] - 1.0) < 1e-12))},
                {"faces": quads[attrs == 4]}, {"face_mask": 1 << 3}):
        for qid, prm in ((sx.capi.SX_QOI_MEASURE, ()), (sx.capi.SX_QOI_POSITION, (0,)),
                         (sx.capi.SX_QOI_VALUE, (2,))):
            q = sx.QoI(s)
            q.add_boundary_integral(qid, prm, **sel)
            values.append((q(0.0, U), q.gradient(0.0, U).numpy()))
    k = len(values) // 3
    for i in range(k):
        for j in (1, 2):
            v0, g0 = values[i]
            v1, g1 = values[j * k + i]
            assert abs(v0 - v1) <= 1e-13 * max(abs(v1), 1.0)
            assert np.abs(g0 - g1).max() <= 1e-13 * max(np.abs(g1).max(), 1.0)
    assert abs(values[0][0] - 1.0) < 1e-13  # the area of the y = 1 face
    assert abs(values[1][0] - 0.5) < 1e-13  # its x moment
