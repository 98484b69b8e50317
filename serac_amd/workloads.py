"""Synthetic workloads of the benchmark configurations.

Fields use the counter-based RNG the library's mesh perturbation uses
(SplitMix64 finalizer keyed by seed, stream and index), so that every
consumer -- GPU path, CPU oracle, tests -- sees bit-identical inputs.

RNG streams:  1000 + d  mesh vertex perturbation (coordinate d, in the library)
              2000 + c  nodal field noise (component c, index = node id)
              3000 + c  J.v direction (component c, index = node id)
"""
import math

import numpy as np

_M64 = (1 << 64) - 1


def _mix_int(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _mix_arr(z):
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def u01(seed, stream, idx):
    """Uniform [0,1) for integer array idx: h = f(f(f(seed) ^ stream) ^ idx)."""
    h = _mix_int(_mix_int(seed & _M64) ^ stream)
    z = _mix_arr(np.uint64(h) ^ np.asarray(idx, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(seed, stream, idx, a, b):
    return a + (b - a) * u01(seed, stream, idx)


def nodal_field(nnodes, nc, seed, base_stream, a, b, ordering=1):
    """i.i.d. U[a,b] per (node, component), laid out byVDIM (ordering=1) or byNODES."""
    nodes = np.arange(nnodes, dtype=np.uint64)
    cols = [uniform(seed, base_stream + c, nodes, a, b) for c in range(nc)]
    arr = np.stack(cols, axis=1)  # [node][comp]
    return arr.ravel() if ordering == 1 else arr.T.ravel()


# --- configuration table (sizes of the benchmark configs) --------------------
CONFIGS = {
    "c1": dict(name="heat_p2_16", n=16, p=2, nc=1, perturb=0.0, qf="heat", params=(1.0, 1.0)),
    "c2": dict(name="neohookean_p3_64", n=64, p=3, nc=3, perturb=0.2, qf="neohookean",
               params=(1.0, 10.0, 0.25)),
    "c3": dict(name="neohookean_p4_48", n=48, p=4, nc=3, perturb=0.2, qf="neohookean",
               params=(1.0, 10.0, 0.25)),
    "c5": dict(name="neohookean_p2_128", n=128, p=2, nc=3, perturb=0.0, qf="neohookean",
               params=(1.0, 10.0, 0.25)),
}


def noise_for_mesh(n):
    """Nodal noise amplitude for an n^3 mesh: 1e-3 at n = 16, scaled with h.

    A fixed +-1e-3 on the 64^3 p=3 mesh (Lobatto node spacing ~0.0043) gives
    displacement gradients of ~0.6 and inverted elements (det(I + grad u) < 0,
    NaN in log1p); scaling with h keeps the field a small perturbation."""
    return 1e-3 * 16.0 / n


def neohookean_displacement(xyz, seed=42, noise=1e-3):
    """u = 0.01 sin(pi x) (1, 0.5, 0.25) + U[-noise, noise] per component (byVDIM)."""
    nn = xyz.shape[0]
    base = 0.01 * np.sin(math.pi * xyz[:, 0])
    u = np.stack([base * 1.0, base * 0.5, base * 0.25], axis=1)
    nodes = np.arange(nn, dtype=np.uint64)
    for c in range(3):
        u[:, c] += uniform(seed, 2000 + c, nodes, -noise, noise)
    return u.ravel()


def heat_temperature(xyz, seed=42):
    """T ~ U[-1,1] per node with T = 0 on the x = 0 face."""
    nn = xyz.shape[0]
    T = uniform(seed, 2000, np.arange(nn, dtype=np.uint64), -1.0, 1.0)
    T[np.abs(xyz[:, 0]) < 1e-12] = 0.0
    return T


def direction(nnodes, nc, seed=42):
    return nodal_field(nnodes, nc, seed, 3000, -1.0, 1.0)


def gdof_per_s(ndof, seconds):
    return ndof / seconds / 1e9


def jvp_algorithmic_bytes(ndof, nqp, ne, npe, d, s_v=8, s_d=8):
    """B_alg = 2 s_v N_dof + s_D d N_qp + 4 NE npe (read x, write y, stream the
    minimal symmetric tangent, read the int32 gather map)."""
    return 2 * s_v * ndof + s_d * d * nqp + 4 * ne * npe
