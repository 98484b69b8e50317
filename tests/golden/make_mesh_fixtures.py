"""Convert the reference's hex mesh data files into JSON fixtures.

Reads data/meshes/{patch3D_hexes,onehex,twohex}.mesh (MFEM mesh v1.0 text) from
the reference checkout -- data, not code -- and writes vertices plus hex
connectivity (mfem vertex order) next to this script.  Run in the container
that has the reference mounted:  python tests/golden/make_mesh_fixtures.py
"""
import json
import os
import sys

REF = os.environ.get("SERAC_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_mfem_mesh(path):
    toks = open(path).read().split("\n")
    lines = [l.strip() for l in toks if l.strip() and not l.strip().startswith("#")]
    i = lines.index("elements")
    ne = int(lines[i + 1])
    hexes = []
    for k in range(ne):
        f = [int(v) for v in lines[i + 2 + k].split()]
        assert f[1] == 5, "only hexahedra (geometry 5)"
        hexes.append(f[2:10])
    j = lines.index("vertices")
    nv, dim = int(lines[j + 1]), int(lines[j + 2])
    assert dim == 3
    verts = [[float(v) for v in lines[j + 3 + k].split()] for k in range(nv)]
    return verts, hexes


def main():
    for name in ("patch3D_hexes", "onehex", "twohex"):
        src = os.path.join(REF, "data", "meshes", name + ".mesh")
        if not os.path.exists(src):
            print("missing", src, file=sys.stderr)
            continue
        verts, hexes = parse_mfem_mesh(src)
        out = os.path.join(HERE, name + ".json")
        with open(out, "w") as fh:
            json.dump({"source": "data/meshes/" + name + ".mesh", "vertices": verts, "hexes": hexes}, fh, indent=1)
        print("wrote", out, len(verts), "vertices", len(hexes), "hexes")


if __name__ == "__main__":
    main()
