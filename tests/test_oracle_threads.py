"""The oracle's multi-threaded J.v and linearization (the all-cores CPU
baseline) equal its serial ones: per-thread node bands and a final sum only
reorder additions."""
import numpy as np
import pytest

import oracle as O
from serac_amd import workloads as W


def _form(ordering=1, qf=O.QF_NEOHOOKEAN, nc=3, dims=(4, 3, 5)):
    m = O.Mesh.cuboid(*dims)
    m.perturb(42, 0.2)
    s = O.Space(m, 2, nc, ordering)
    f = O.Form(s)
    f.add_domain_integral(qf, (1.0, 10.0, 0.25) if nc == 3 else (1.0, 1.0))
    return m, s, f


@pytest.mark.parametrize("nthreads", [1, 3, 8])
@pytest.mark.parametrize("ordering", [0, 1])
def test_threaded_jvp_matches_serial(nthreads, ordering):
    m, s, f = _form(ordering)
    f.linearize(W.nodal_field(s.nnodes, 3, 3, 2000, -0.003, 0.003, 1))
    for seed in (1, 2):  # the bands are reused: a second call must start from zero
        dU = W.nodal_field(s.nnodes, 3, seed, 3000, -1.0, 1.0, 1)
        ref = f.jvp(dU)
        out = f.jvp_threads(dU, nthreads)
        assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_threaded_jvp_more_threads_than_elements():
    m, s, f = _form(dims=(2, 1, 1))
    f.linearize(W.nodal_field(s.nnodes, 3, 3, 2000, -0.003, 0.003, 1))
    dU = W.direction(s.nnodes, 3)
    ref = f.jvp(dU)
    assert np.abs(f.jvp_threads(dU, 5) - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("nthreads", [2, 7])
def test_threaded_linearize_matches_serial(nthreads):
    U = None
    outs = []
    for threaded in (False, True):
        m, s, f = _form()
        U = W.nodal_field(s.nnodes, 3, 3, 2000, -0.003, 0.003, 1)
        r = f.linearize_threads(U, nthreads) if threaded else f.linearize(U)
        dU = W.direction(s.nnodes, 3)
        outs.append((r, f.jvp(dU)))
    (r0, j0), (r1, j1) = outs
    assert np.abs(r1 - r0).max() <= 1e-13 * np.abs(r0).max()
    assert np.array_equal(j1, j0)  # the stored tangent is the same per element


def test_threaded_heat():
    m, s, f = _form(qf=O.QF_HEAT, nc=1)
    U = W.nodal_field(s.nnodes, 1, 3, 2000, -1.0, 1.0, 1)
    r = f.linearize(U)
    dU = W.direction(s.nnodes, 1)
    ref = f.jvp(dU)
    assert np.abs(f.jvp_threads(dU, 4) - ref).max() <= 1e-13 * np.abs(ref).max()
    m2, s2, f2 = _form(qf=O.QF_HEAT, nc=1)
    assert np.abs(f2.linearize_threads(U, 4) - r).max() <= 1e-13 * np.abs(r).max()
