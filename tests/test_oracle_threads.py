"""The oracle's multi-threaded J.v (the all-cores CPU baseline) equals its
serial J.v: per-thread scatter and a final sum only reorder additions."""
import numpy as np
import pytest

import oracle as O
from serac_amd import workloads as W


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_threaded_jvp_matches_serial(nthreads):
    m = O.Mesh.cuboid(4, 3, 5)
    m.perturb(42, 0.2)
    s = O.Space(m, 2, 3)
    f = O.Form(s)
    f.add_domain_integral(O.QF_NEOHOOKEAN, (1.0, 10.0, 0.25))
    f.linearize(W.nodal_field(s.nnodes, 3, 3, 2000, -0.003, 0.003, 1))
    dU = W.direction(s.nnodes, 3)
    ref = f.jvp(dU)
    out = f.jvp_threads(dU, nthreads)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()
