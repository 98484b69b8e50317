"""Mesh fixtures from the reference's data files.

Reads data/meshes/*.mesh (MFEM mesh v1.0 text) from the reference checkout --
data, not code.  It writes two things next to this script:
- JSON with the vertices and hex connectivity (mfem vertex order) of the
  meshes the oracle tests use;
- verbatim copies of the mesh files the reader tests use, under meshes/.
Run it in the container that has the reference mounted:
    python tests/golden/make_mesh_fixtures.py
parse_mfem_mesh is also the tests' independent parser: a line-based reading of
the same format, written separately from the library's C++ reader.
"""
import json
import os
import shutil
import sys

REF = os.environ.get("SERAC_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
JSON_MESHES = ("patch3D_hexes", "onehex", "twohex")
COPIED_MESHES = ("patch3D_hexes", "onehex", "twohex", "beam-hex", "elongatedhex", "toroid-hex",
                 "patch3D_tets_and_hexes")


def parse_mfem_mesh(path, text=None):
    """(vertices, hexes, element attributes, boundary quads, boundary attributes) of a hex mesh file."""
    src = open(path).read() if text is None else text
    lines = [l.split("#")[0].strip() for l in src.split("\n")]
    lines = [l for l in lines if l]
    i = lines.index("elements")
    ne = int(lines[i + 1])
    hexes, attrs = [], []
    for k in range(ne):
        f = [int(v) for v in lines[i + 2 + k].split()]
        assert f[1] == 5, "only hexahedra (geometry 5)"
        attrs.append(f[0])
        hexes.append(f[2:10])
    quads, battrs = [], []
    if "boundary" in lines:
        b = lines.index("boundary")
        for k in range(int(lines[b + 1])):
            f = [int(v) for v in lines[b + 2 + k].split()]
            assert f[1] == 3, "quadrilateral boundary elements (geometry 3)"
            battrs.append(f[0])
            quads.append(f[2:6])
    j = lines.index("vertices")
    nv, dim = int(lines[j + 1]), int(lines[j + 2])
    assert dim == 3
    verts = [[float(v) for v in lines[j + 3 + k].split()] for k in range(nv)]
    return verts, hexes, attrs, quads, battrs


def main():
    os.makedirs(os.path.join(HERE, "meshes"), exist_ok=True)
    for name in COPIED_MESHES:
        src = os.path.join(REF, "data", "meshes", name + ".mesh")
        if not os.path.exists(src):
            print("missing", src, file=sys.stderr)
            continue
        shutil.copyfile(src, os.path.join(HERE, "meshes", name + ".mesh"))
        if name in JSON_MESHES:
            verts, hexes = parse_mfem_mesh(src)[:2]
            out = os.path.join(HERE, name + ".json")
            with open(out, "w") as fh:
                json.dump({"source": "data/meshes/" + name + ".mesh", "vertices": verts, "hexes": hexes}, fh,
                          indent=1)
            print("wrote", out, len(verts), "vertices", len(hexes), "hexes")


if __name__ == "__main__":
    main()
