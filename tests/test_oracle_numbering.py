"""Pins the oracle's H1 node numbering and ElementRestriction
(fn/element_restriction.cpp:216-277, 425-467): conformity (shared nodes sit at
one physical point whichever element computes them), node counts, gather /
scatter adjointness and multiplicities, and invariance of the physics under
element re-orientation."""
import numpy as np
import pytest

import oracle as O
from conftest import load_mesh_fixture

LEX2V = [0, 1, 3, 2, 4, 5, 7, 6]


def element_node_positions(mesh_X, hexes, p):
    """Physical position of each element's lexicographic nodes from that element's map."""
    g = O.gll_nodes(p + 1)
    n = p + 1
    out = np.zeros((len(hexes), n ** 3, 3))
    for e, hx in enumerate(hexes):
        corners = mesh_X[np.array(hx)[LEX2V]]
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    w = np.array([(g[i] if a else 1 - g[i]) * (g[j] if b else 1 - g[j]) * (g[k] if c else 1 - g[k])
                                  for c in (0, 1) for b in (0, 1) for a in (0, 1)])
                    out[e, i + n * (j + n * k)] = w @ corners
    return out


def assert_conforming(space, X, hexes):
    em = space.emap()
    pos = element_node_positions(X, hexes, space.p)
    first = np.full((space.nnodes, 3), np.nan)
    for e in range(len(hexes)):
        for j, node in enumerate(em[e]):
            if np.isnan(first[node, 0]):
                first[node] = pos[e, j]
            else:
                assert np.allclose(first[node], pos[e, j], atol=1e-13), (e, j, node)
    assert not np.isnan(first).any()
    # distinct ids are distinct points
    u = np.unique(np.round(first, 9), axis=0)
    assert len(u) == space.nnodes


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_reference_patch_mesh_numbering_conforms(p):
    X, hexes = load_mesh_fixture("patch3D_hexes")
    m = O.Mesh.from_arrays(X, hexes)
    s = O.Space(m, p, 1)
    assert_conforming(s, X, hexes)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_structured_and_unstructured_numberings_agree_geometrically(p):
    m = O.Mesh.cuboid(3, 2, 2)
    m.perturb(1, 0.2)
    X = m.vertices()
    s1 = O.Space(m, p, 1)
    em = s1.emap()
    hexes = np.zeros((m.ne, 8), dtype=np.int32)
    s_lin = O.Space(m, 1, 1).emap()
    for c in range(8):
        hexes[:, LEX2V[c]] = s_lin[:, c]
    mu = O.Mesh.from_arrays(X, hexes)
    s2 = O.Space(mu, p, 1)
    assert s1.nnodes == s2.nnodes == (3 * p + 1) * (2 * p + 1) * (2 * p + 1)
    assert_conforming(s1, X, hexes)
    assert_conforming(s2, X, hexes)
    # the two numberings differ by a permutation of nodes
    x1, x2 = s1.node_coords(), s2.node_coords()
    perm = np.zeros(s1.nnodes, dtype=np.int64)
    perm[em.ravel()] = s2.emap().ravel()
    assert np.allclose(x1, x2[perm], atol=1e-14)


def rotated_hexes(hexes, seed):
    """Re-orient every element by a random orientation-preserving symmetry of the cube."""
    rng = np.random.default_rng(seed)
    # lexicographic corner coordinates
    cor = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)])
    rots = []
    for perm in ([0, 1, 2], [1, 2, 0], [2, 0, 1], [1, 0, 2], [0, 2, 1], [2, 1, 0]):
        for flips in range(8):
            f = np.array([(flips >> d) & 1 for d in range(3)])
            P = np.zeros((3, 3))
            for i, pi in enumerate(perm):
                P[i, pi] = -1 if f[i] else 1
            if np.linalg.det(P) > 0:
                rots.append((perm, f))
    out = np.array(hexes).copy()
    for e in range(len(hexes)):
        perm, f = rots[rng.integers(len(rots))]
        new_lex = np.zeros(8, dtype=np.int64)
        for c in range(8):
            src = cor[c][perm] ^ f
            new_lex[c] = hexes[e][LEX2V[src[0] + 2 * src[1] + 4 * src[2]]]
        for c in range(8):
            out[e][LEX2V[c]] = new_lex[c]
    return out


@pytest.mark.parametrize("p", [2, 3, 4])
def test_residual_invariant_under_element_reorientation(p):
    X, hexes = load_mesh_fixture("patch3D_hexes")
    h2 = rotated_hexes(hexes, seed=p)
    s1 = O.Space(O.Mesh.from_arrays(X, hexes), p, 3)
    s2 = O.Space(O.Mesh.from_arrays(X, h2), p, 3)
    assert_conforming(s2, X, h2)
    x1, x2 = s1.node_coords(), s2.node_coords()
    u = lambda x: np.stack([np.sin(x[:, 0]) * 0.02, x[:, 1] * x[:, 2] * 0.01, np.cos(x[:, 2]) * 0.01], axis=1).ravel()
    R1, R2 = [], []
    for s, x, R in ((s1, x1, R1), (s2, x2, R2)):
        f = O.Form(s)
        f.add_domain_integral(O.QF_NEOHOOKEAN, [1.0, 10.0, 0.25])
        R.append(f.residual(u(x)).reshape(-1, 3))
    # match nodes by position
    key1 = {tuple(np.round(v, 9)): i for i, v in enumerate(x1)}
    perm = np.array([key1[tuple(np.round(v, 9))] for v in x2])
    assert np.allclose(R2[0], R1[0][perm], rtol=0, atol=1e-13)


@pytest.mark.parametrize("nc,ordering", [(1, 1), (3, 1), (3, 0)])
def test_gather_scatter_adjoint_and_multiplicity(nc, ordering):
    m = O.Mesh.cuboid(3, 2, 2)
    s = O.Space(m, 2, nc, ordering)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, s.ndof)
    y = rng.uniform(-1, 1, s.ne * nc * s.npe)
    assert abs(np.dot(s.gather(x), y) - np.dot(x, s.scatter_add(y))) < 1e-12
    mult = s.scatter_add(np.ones(s.ne * nc * s.npe))
    # structured grid: a node's multiplicity is the product over directions of
    # (2 if it sits on an interior element boundary in that direction else 1)
    em = s.emap()
    count = np.bincount(em.ravel(), minlength=s.nnodes)
    for c in range(nc):
        idx = np.arange(s.nnodes) * nc + c if ordering == 1 else c * s.nnodes + np.arange(s.nnodes)
        assert np.array_equal(mult[idx], count.astype(float))
    assert set(np.unique(count)) <= {1, 2, 4, 8}
