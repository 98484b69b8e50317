"""The MFEM mesh reader (buildMeshFromFile, src/serac/mesh/mesh_utils.cpp:19-43)
against an independent Python parser of the same files (the reference's
data/meshes, copied as fixtures), its refusals, and the cuboid mesh's mfem
Cartesian boundary attributes.  Mesh set-up is host code: no GPU needed."""
import os

import numpy as np
import pytest

import serac_amd as sx
from serac_amd import capi

from golden.make_mesh_fixtures import parse_mfem_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHES = os.path.join(ROOT, "tests", "golden", "meshes")


@pytest.mark.parametrize("name", ["onehex", "twohex", "patch3D_hexes", "beam-hex", "elongatedhex"])
def test_reader_matches_independent_parser(name):
    path = os.path.join(MESHES, name + ".mesh")
    verts, hexes, attrs, quads, battrs = parse_mfem_mesh(path)
    m = sx.Mesh.read_mfem(path)
    assert (m.nv, m.ne) == (len(verts), len(hexes))
    assert np.array_equal(m.vertices(), np.array(verts))
    assert np.array_equal(m.attributes(), np.array(attrs))
    q, a = m.boundary()
    assert np.array_equal(q, np.array(quads, dtype=np.int32).reshape(-1, 4))
    assert np.array_equal(a, np.array(battrs))
    # the text entry point reads the same mesh
    m2 = sx.Mesh.read_mfem(text=open(path).read())
    assert np.array_equal(m2.vertices(), m.vertices())


def test_patch_mesh_matches_the_oracle_fixture(patch_mesh):
    X, hexes = patch_mesh
    m = sx.Mesh.read_mfem(os.path.join(MESHES, "patch3D_hexes.mesh"))
    assert np.array_equal(m.vertices(), X)


@pytest.mark.parametrize("name,what", [("toroid-hex", b"curved"), ("patch3D_tets_and_hexes", b"hexahedra")])
def test_unsupported_meshes_are_refused(name, what):
    rc = capi.call_rc("sx_mesh_read_mfem_file", os.path.join(MESHES, name + ".mesh").encode(),
                      capi.ctypes.byref(capi.ctypes.c_void_p()))
    assert rc == capi.SX_ERR_UNSUPPORTED
    assert what in capi.load().sx_last_error()


ONEHEX = open(os.path.join(MESHES, "onehex.mesh")).read()


@pytest.mark.parametrize("text,what", [
    (ONEHEX.replace("MFEM mesh v1.0", "MFEM mesh v9"), b"header"),
    (ONEHEX.replace("vertices", "vertexes"), b"unknown section"),
    (ONEHEX[: ONEHEX.index("vertices") + len("vertices\n8\n3\n0 0 0\n")], b"truncated"),
    (ONEHEX.replace("1 5 0 1 3 2 4 5 7 6", "1 5 0 1 3 2 4 5 7 60"), b"out of range"),
    (ONEHEX.replace("1 5 0 1 3 2 4 5 7 6", "1 5 0 1 3 2 4 5 7"), b"expected 10"),
])
def test_malformed_text_is_an_argument_error(text, what):
    assert "1 5 0 1 3 2 4 5 7 6" in ONEHEX
    raw = text.encode()
    rc = capi.call_rc("sx_mesh_read_mfem", raw, len(raw), capi.ctypes.byref(capi.ctypes.c_void_p()))
    assert rc == capi.SX_ERR_ARG
    assert what in capi.load().sx_last_error()


def test_missing_file_is_an_argument_error():
    rc = capi.call_rc("sx_mesh_read_mfem_file", b"/nonexistent/x.mesh", capi.ctypes.byref(capi.ctypes.c_void_p()))
    assert rc == capi.SX_ERR_ARG
    assert b"cannot open" in capi.load().sx_last_error()


def test_cuboid_boundary_attributes_follow_mfem_cartesian_convention():
    nx, ny, nz = 3, 4, 5
    m = sx.Mesh.cuboid(nx, ny, nz, 1.0, 2.0, 3.0)
    X = m.vertices()
    q, a = m.boundary()
    assert len(q) == 2 * (nx * ny + ny * nz + nx * nz)
    # attribute -> (coordinate, plane value)
    planes = {1: (2, 0.0), 2: (1, 0.0), 3: (0, 1.0), 4: (1, 2.0), 5: (0, 0.0), 6: (2, 3.0)}
    counts = {1: nx * ny, 6: nx * ny, 2: nx * nz, 4: nx * nz, 3: ny * nz, 5: ny * nz}
    for attr, (d, val) in planes.items():
        sel = q[a == attr]
        assert len(sel) == counts[attr]
        assert np.all(np.abs(X[sel, d] - val) < 1e-14)
        # counter-clockwise seen from outside: the quad normal points out of the box
        P = X[sel]
        n = np.cross(P[:, 1] - P[:, 0], P[:, 3] - P[:, 0])
        outward = np.zeros(3)
        outward[d] = 1.0 if val > 0 else -1.0
        assert np.all(n @ outward > 0)
    assert np.all(m.attributes() == 1)
