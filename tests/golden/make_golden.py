"""Writes tests/golden/oracle_vectors.npz: small seeded problems with their
oracle residual, linearized J.v and diagonal, used as fixed regression
vectors by the CPU tests (oracle) and the GPU tests (HIP path).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from conftest import load_mesh_fixture  # noqa: E402
from serac_amd import workloads as W  # noqa: E402

CASES = {
    # name: (mesh, order, components, qfunction, params, field amplitude)
    "heat_p2_cube3": ("cube3", 2, 1, O.QF_HEAT, (1.0, 1.0), 1.0),
    "neohookean_p3_cube2": ("cube2", 3, 3, O.QF_NEOHOOKEAN, (1.0, 10.0, 0.25), 0.02),
    "neohookean_p2_patch": ("patch", 2, 3, O.QF_NEOHOOKEAN, (1.0, 1.0, 1.0), 0.05),
    "massdiff_p4_cube2": ("cube2", 4, 1, O.QF_MASS_DIFFUSION, (1.7, 2.1), 1.0),
}


def make_mesh(kind):
    if kind == "patch":
        X, hexes = load_mesh_fixture("patch3D_hexes")
        return O.Mesh.from_arrays(X, hexes), (X, hexes)
    n = int(kind[-1])
    m = O.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    return m, None


def compute(name):
    kind, p, nc, qf, params, amp = CASES[name]
    m, arrays = make_mesh(kind)
    s = O.Space(m, p, nc)
    f = O.Form(s)
    f.add_domain_integral(qf, params)
    U = W.nodal_field(s.nnodes, nc, 11, 2000, -amp, amp)
    dU = W.nodal_field(s.nnodes, nc, 12, 3000, -1.0, 1.0)
    R = f.residual(U)
    f.linearize(U)
    return dict(U=U, dU=dU, R=R, JdU=f.jvp(dU), diag=f.diagonal(), emap=s.emap())


def main():
    out = {}
    for name in CASES:
        for k, v in compute(name).items():
            out[f"{name}/{k}"] = v
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
