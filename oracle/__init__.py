"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package
``serac_amd`` never imports it.

The oracle restates serac's FP64 CPU algorithm (reference citations live in
``serac_oracle.hpp`` / ``serac_oracle_form.hpp``).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_LIB_V4 = os.path.join(_HERE, "liboracle_v4.so")

QF_HEAT = 1
QF_MASS_DIFFUSION = 2
QF_NEOHOOKEAN = 3
QF_LINEAR_ELASTIC = 4
QF_J2_LINEAR = 5
QOI_MEASURE = 1
QOI_POSITION = 2
QOI_VALUE = 3
QOI_NEOHOOKEAN_ENERGY = 4
QOI_INNER = 5
BQF_TRACTION = 101
BQF_PRESSURE = 102
BQF_FLUX = 103

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_I64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _host_has_avx512():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    f = line.split()
                    return all(x in f for x in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))
    except OSError:
        pass
    return False


_isa = "v3"


def use_isa(level):
    """Select the build the next lib() call loads: "v3" (liboracle.so, the
    tests' default) or "widest" (liboracle_v4.so when this host has AVX-512).
    Returns the level loaded.  Only before the first lib() call."""
    global _isa
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _isa = "v4" if level == "widest" and _host_has_avx512() and os.path.exists(_LIB_V4) else "v3"
    return _isa


def isa():
    return _isa


def _load():
    path = _LIB_V4 if _isa == "v4" else _LIB
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    sig = {
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_rng_u01": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
        "orc_gl_rule": (ctypes.c_int, [ctypes.c_int, _D, _D]),
        "orc_gll_nodes": (ctypes.c_int, [ctypes.c_int, _D]),
        "orc_gll_basis": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _D, _D]),
        "orc_mesh_cuboid": (_vp, [ctypes.c_int] * 3 + [ctypes.c_double] * 3),
        "orc_mesh_create": (_vp, [ctypes.c_int, _D, ctypes.c_int, _I]),
        "orc_mesh_perturb": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_double]),
        "orc_mesh_info": (None, [_vp, _I64]),
        "orc_mesh_vertices": (None, [_vp, _D]),
        "orc_mesh_free": (None, [_vp]),
        "orc_space": (_vp, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "orc_space_info": (None, [_vp, _I64]),
        "orc_space_emap": (None, [_vp, _I]),
        "orc_space_node_coords": (None, [_vp, _vp, _D]),
        "orc_space_free": (None, [_vp]),
        "orc_gather": (None, [_vp, _D, _D]),
        "orc_scatter_add": (None, [_vp, _D, _D]),
        "orc_interpolate": (ctypes.c_int, [ctypes.c_int] * 3 + [_D, _D, _D, ctypes.c_int]),
        "orc_integrate": (ctypes.c_int, [ctypes.c_int] * 3 + [_D, _D, _D, ctypes.c_int]),
        "orc_geometric_factors": (ctypes.c_int, [_vp, ctypes.c_int, _D, _D]),
        "orc_qf_eval": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, _D, _D, _D]),
        "orc_qf_eval_state": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, _D, _D, _D, _D]),
        "orc_qf_tangent": (ctypes.c_int, [ctypes.c_int, _D, ctypes.c_int, _D, _D, _D, _D]),
        "orc_form_add_integral_state": (ctypes.c_int, [_vp, ctypes.c_int, _D, _I, ctypes.c_int, _D]),
        "orc_form_update_qdata": (None, [_vp, ctypes.c_int]),
        "orc_form_add_boundary_integral": (ctypes.c_int, [_vp, ctypes.c_int, _D, ctypes.c_int]),
        "orc_form_nfaces": (ctypes.c_int, [_vp, ctypes.c_int]),
        "orc_form_face": (None, [_vp, ctypes.c_int, ctypes.c_int, _I]),
        "orc_form_face_matrix": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _D]),
        "orc_form_ssize": (ctypes.c_int, [_vp, ctypes.c_int]),
        "orc_form_get_qdata": (ctypes.c_int, [_vp, ctypes.c_int, _D]),
        "orc_form_set_qdata": (ctypes.c_int, [_vp, ctypes.c_int, _D]),
        "orc_form": (_vp, [_vp, _vp]),
        "orc_form_free": (None, [_vp]),
        "orc_form_add_integral": (ctypes.c_int, [_vp, ctypes.c_int, _D, _I, ctypes.c_int]),
        "orc_form_eval": (ctypes.c_int, [_vp, _D, _D, ctypes.c_int]),
        "orc_form_add_integral2": (ctypes.c_int, [_vp, ctypes.c_int, _D, _I, ctypes.c_int, _D]),
        "orc_form_eval2": (ctypes.c_int, [_vp, _D, _D, _D, ctypes.c_int]),
        "orc_form_jvp": (ctypes.c_int, [_vp, _D, _D]),
        "orc_form_qoi": (ctypes.c_int, [_vp, ctypes.c_int, _D, _I, ctypes.c_int, _D, _D, _D]),
        "orc_form_eval_threads": (ctypes.c_int, [_vp, _D, _D, ctypes.c_int, ctypes.c_int]),
        "orc_form_diagonal": (ctypes.c_int, [_vp, _D]),
        "orc_form_dsize": (ctypes.c_int, [_vp, ctypes.c_int]),
        "orc_form_derivative": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _D]),
        "orc_form_element_matrix": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, _D]),
        "orc_form_geometry": (ctypes.c_int, [_vp, _D, _D]),
        "orc_cg": (ctypes.c_int, [_vp, _I, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, _D]),
        "orc_newton": (ctypes.c_int, [_vp, _I, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _D]),
        "orc_newton_ls": (ctypes.c_int, [_vp, _I, ctypes.c_int, _D, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         _D]),
        "orc_time_jvp": (ctypes.c_double, [_vp, _D, _D, ctypes.c_int]),
        "orc_time_jvp_threads": (ctypes.c_double, [_vp, _D, _D, ctypes.c_int, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _params(p):
    a = np.zeros(16)
    a[: len(p)] = p
    return a


def rng_u01(seed, stream, idx):
    return lib().orc_rng_u01(seed, stream, idx)


def gl_rule(q):
    x, w = np.zeros(q), np.zeros(q)
    _check(lib().orc_gl_rule(q, _dp(x), _dp(w)))
    return x, w


def gll_nodes(n):
    x = np.zeros(n)
    _check(lib().orc_gll_nodes(n, _dp(x)))
    return x


def gll_basis(n, x):
    B, G = np.zeros(n), np.zeros(n)
    _check(lib().orc_gll_basis(n, float(x), _dp(B), _dp(G)))
    return B, G


def interpolate(n, q, nc, X, brute=False):
    X = np.ascontiguousarray(X, dtype=np.float64)
    nq = q ** 3
    val, grd = np.zeros(nc * nq), np.zeros(nc * 3 * nq)
    _check(lib().orc_interpolate(n, q, nc, _dp(X), _dp(val), _dp(grd), int(brute)))
    return val.reshape(nc, nq), grd.reshape(nc, 3, nq)


def integrate(n, q, nc, source, flux, brute=False):
    s = np.ascontiguousarray(source, dtype=np.float64)
    f = np.ascontiguousarray(flux, dtype=np.float64)
    r = np.zeros(nc * n ** 3)
    _check(lib().orc_integrate(n, q, nc, _dp(s), _dp(f), _dp(r), int(brute)))
    return r.reshape(nc, n ** 3)


def qf_eval(kind, params, nc, value, grad):
    v = np.ascontiguousarray(value, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64)
    s, f = np.zeros(nc), np.zeros(3 * nc)
    _check(lib().orc_qf_eval(kind, _dp(_params(params)), nc, _dp(v), _dp(g), _dp(s), _dp(f)))
    return s, f.reshape(nc, 3)


def qf_eval_state(kind, params, nc, value, grad, state):
    """qfunction with quadrature-point state; returns (source, flux, updated state)."""
    v = np.ascontiguousarray(value, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64)
    st = np.array(state, dtype=np.float64)
    s, f = np.zeros(nc), np.zeros(3 * nc)
    _check(lib().orc_qf_eval_state(kind, _dp(_params(params)), nc, _dp(v), _dp(g), _dp(s), _dp(f), _dp(st)))
    return s, f.reshape(nc, 3), st


def qf_tangent(kind, params, nc, value, grad, state=None):
    """Forward-mode d flux / d grad at one point, (3 nc) x (3 nc)."""
    v = np.ascontiguousarray(value, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64)
    st = np.ascontiguousarray(state if state is not None else np.zeros(16), dtype=np.float64)
    D = np.zeros((3 * nc, 3 * nc))
    _check(lib().orc_qf_tangent(kind, _dp(_params(params)), nc, _dp(v), _dp(g), _dp(st), _dp(D)))
    return D


class Mesh:
    def __init__(self, handle):
        self.h = handle
        info = np.zeros(2, dtype=np.int64)
        lib().orc_mesh_info(self.h, info.ctypes.data_as(_I64))
        self.nv, self.ne = int(info[0]), int(info[1])

    @classmethod
    def cuboid(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
        return cls(lib().orc_mesh_cuboid(nx, ny, nz, lx, ly, lz))

    @classmethod
    def from_arrays(cls, X, hexes):
        X = np.ascontiguousarray(X, dtype=np.float64)
        hexes = np.ascontiguousarray(hexes, dtype=np.int32)
        return cls(lib().orc_mesh_create(len(X), _dp(X), len(hexes), _ip(hexes)))

    def perturb(self, seed, amp):
        _check(lib().orc_mesh_perturb(self.h, seed, amp))

    def vertices(self):
        X = np.zeros((self.nv, 3))
        lib().orc_mesh_vertices(self.h, _dp(X))
        return X

    def geometric_factors(self, Q):
        nq = Q ** 3
        Xq, Jq = np.zeros((self.ne, 3, nq)), np.zeros((self.ne, 3, 3, nq))
        _check(lib().orc_geometric_factors(self.h, Q, _dp(Xq), _dp(Jq)))
        return Xq, Jq

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh_free(self.h)
            self.h = None


class Space:
    def __init__(self, mesh, p, nc=1, ordering=1):
        self.mesh = mesh
        self.h = lib().orc_space(mesh.h, p, nc, ordering)
        if not self.h:
            _check(1)
        info = np.zeros(4, dtype=np.int64)
        lib().orc_space_info(self.h, info.ctypes.data_as(_I64))
        self.nnodes, self.ndof, self.ne, self.npe = (int(v) for v in info)
        self.p, self.nc, self.ordering = p, nc, ordering

    def emap(self):
        m = np.zeros((self.ne, self.npe), dtype=np.int32)
        lib().orc_space_emap(self.h, _ip(m))
        return m

    def node_coords(self):
        x = np.zeros((self.nnodes, 3))
        lib().orc_space_node_coords(self.mesh.h, self.h, _dp(x))
        return x

    def gather(self, L):
        L = np.ascontiguousarray(L, dtype=np.float64)
        E = np.zeros(self.ne * self.nc * self.npe)
        lib().orc_gather(self.h, _dp(L), _dp(E))
        return E

    def scatter_add(self, E, L=None):
        E = np.ascontiguousarray(E, dtype=np.float64)
        L = np.zeros(self.ndof) if L is None else np.array(L, dtype=np.float64)
        lib().orc_scatter_add(self.h, _dp(E), _dp(L))
        return L

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_space_free(self.h)
            self.h = None


class Form:
    """Functional<test(trial)> with one trial argument equal to the test space."""

    def __init__(self, space):
        self.space = space
        self.h = lib().orc_form(space.mesh.h, space.h)
        self.nintegrals = 0

    def add_domain_integral(self, kind, params, elems=None):
        el = np.ascontiguousarray(elems if elems is not None else [], dtype=np.int32)
        _check(lib().orc_form_add_integral(self.h, kind, _dp(_params(params)), _ip(el), len(el)))
        self.nintegrals += 1

    def add_domain_integral2(self, kind, params, elems=None, state_init=None):
        """AddDomainIntegral(DependsOn<0, 1>): the integrand also reads the second argument's value (the
        acceleration of a solid, the rate of a temperature); evaluate with residual2 / linearize2."""
        el = np.ascontiguousarray(elems if elems is not None else [], dtype=np.int32)
        st = np.ascontiguousarray(state_init, dtype=np.float64) if state_init is not None else None
        _check(lib().orc_form_add_integral2(self.h, kind, _dp(_params(params)), _ip(el), len(el),
                                            _dp(st) if st is not None else None))
        self.nintegrals += 1

    def residual2(self, U, U1):
        U = np.ascontiguousarray(U, dtype=np.float64)
        U1 = np.ascontiguousarray(U1, dtype=np.float64)
        R = np.zeros(self.space.ndof)
        _check(lib().orc_form_eval2(self.h, _dp(U), _dp(U1), _dp(R), 0))
        return R

    def linearize2(self, U, U1):
        """residual, with the derivative in the first argument stored for jvp"""
        U = np.ascontiguousarray(U, dtype=np.float64)
        U1 = np.ascontiguousarray(U1, dtype=np.float64)
        R = np.zeros(self.space.ndof)
        _check(lib().orc_form_eval2(self.h, _dp(U), _dp(U1), _dp(R), 1))
        return R

    def add_domain_integral_state(self, kind, params, state_init, elems=None):
        """AddDomainIntegral with QuadratureData starting at state_init at every point."""
        el = np.ascontiguousarray(elems if elems is not None else [], dtype=np.int32)
        st = np.ascontiguousarray(state_init, dtype=np.float64)
        _check(lib().orc_form_add_integral_state(self.h, kind, _dp(_params(params)), _ip(el), len(el), _dp(st)))
        self.nintegrals += 1

    def update_qdata(self, flag):
        lib().orc_form_update_qdata(self.h, int(bool(flag)))

    def add_boundary_integral(self, kind, params, face_mask=0):
        """AddBoundaryIntegral over the bounding-box planes of face_mask (0: whole boundary)."""
        _check(lib().orc_form_add_boundary_integral(self.h, kind, _dp(_params(params)), face_mask))
        self.nbintegrals = getattr(self, "nbintegrals", 0) + 1

    def faces(self, bi=0):
        """(element, local face) of each face of boundary integral bi."""
        n = lib().orc_form_nfaces(self.h, bi)
        out = np.zeros((n, 2), dtype=np.int32)
        for k in range(n):
            lib().orc_form_face(self.h, bi, k, _ip(out[k]))
        return out

    def face_matrix(self, bi, k):
        nd = self.space.nc * self.space.npe
        K = np.zeros((nd, nd))
        _check(lib().orc_form_face_matrix(self.h, bi, k, _dp(K)))
        return K

    def qdata(self, integral=0, nelem=None):
        """State of an integral as [elem-in-domain][q][size]."""
        ss = lib().orc_form_ssize(self.h, integral)
        nq = (self.space.p + 1) ** 3
        ne = self.space.ne if nelem is None else nelem
        out = np.zeros((ne, nq, ss))
        _check(lib().orc_form_get_qdata(self.h, integral, _dp(out)))
        return out

    def set_qdata(self, state, integral=0):
        st = np.ascontiguousarray(state, dtype=np.float64)
        _check(lib().orc_form_set_qdata(self.h, integral, _dp(st)))

    def residual(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros(self.space.ndof)
        _check(lib().orc_form_eval(self.h, _dp(U), _dp(R), 0))
        return R

    def linearize(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros(self.space.ndof)
        _check(lib().orc_form_eval(self.h, _dp(U), _dp(R), 1))
        return R

    def linearize_threads(self, U, nthreads):
        """linearize() over nthreads host threads (node-band scatter)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros(self.space.ndof)
        _check(lib().orc_form_eval_threads(self.h, _dp(U), _dp(R), 1, nthreads))
        return R

    def qoi(self, kind, params=(), elems=None, U=None, U1=None):
        """Functional<double(trials...)> value: QOI_* kind over elems (default all)."""
        prm = np.zeros(16)
        prm[:len(params)] = params
        e = np.ascontiguousarray(elems if elems is not None else [], dtype=np.int32)
        nd = self.space.ndof
        u0 = np.ascontiguousarray(U if U is not None else np.zeros(nd), dtype=np.float64)
        u1 = np.ascontiguousarray(U1 if U1 is not None else np.zeros(nd), dtype=np.float64)
        out = np.zeros(1)
        _check(lib().orc_form_qoi(self.h, kind, _dp(prm), _ip(e) if len(e) else None, len(e), _dp(u0), _dp(u1),
                                  _dp(out)))
        return float(out[0])

    def jvp(self, dU):
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        dR = np.zeros(self.space.ndof)
        _check(lib().orc_form_jvp(self.h, _dp(dU), _dp(dR)))
        return dR

    def diagonal(self):
        d = np.zeros(self.space.ndof)
        _check(lib().orc_form_diagonal(self.h, _dp(d)))
        return d

    def dsize(self, integral=0):
        return lib().orc_form_dsize(self.h, integral)

    def derivative(self, integral, slot, q):
        out = np.zeros(self.dsize(integral))
        _check(lib().orc_form_derivative(self.h, integral, slot, q, _dp(out)))
        return out

    def element_matrix(self, integral, slot):
        nd = self.space.nc * self.space.npe
        K = np.zeros((nd, nd))
        _check(lib().orc_form_element_matrix(self.h, integral, slot, _dp(K)))
        return K

    def geometry(self):
        ne, nq = self.space.ne, (self.space.p + 1) ** 3
        Xq, Jq = np.zeros((ne, 3, nq)), np.zeros((ne, 3, 3, nq))
        _check(lib().orc_form_geometry(self.h, _dp(Xq), _dp(Jq)))
        return Xq, Jq

    def cg(self, ess, b, rel=1e-10, abs_tol=0.0, maxit=10000):
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.space.ndof)
        rep = np.zeros(4)
        _check(lib().orc_cg(self.h, _ip(ess), len(ess), _dp(b), _dp(x), rel, abs_tol, maxit, _dp(rep)))
        return x, dict(iterations=int(rep[0]), initial_norm=rep[1], final_norm=rep[2], converged=bool(rep[3]))

    def newton(self, ess, x0, rel=1e-8, abs_tol=1e-12, maxit=20, lin_rel=1e-8, lin_abs=1e-12, lin_max=5000,
               max_line_search=0):
        """Newton + Jacobi-CG; max_line_search > 0 adds the two-way backtracking line search."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        x = np.array(x0, dtype=np.float64)
        rep = np.zeros(6)
        _check(lib().orc_newton_ls(self.h, _ip(ess), len(ess), _dp(x), rel, abs_tol, maxit, lin_rel, lin_abs,
                                   lin_max, max_line_search, _dp(rep)))
        return x, dict(newton_iterations=int(rep[0]), linear_iterations=int(rep[1]), initial_norm=rep[2],
                       final_norm=rep[3], converged=bool(rep[4]), line_search_steps=int(rep[5]))

    def time_jvp(self, dU, reps):
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        dR = np.zeros(self.space.ndof)
        return lib().orc_time_jvp(self.h, _dp(dU), _dp(dR), reps)

    def jvp_threads(self, dU, nthreads):
        """J.v over nthreads host threads (per-thread node bands, then a sum)."""
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        dR = np.zeros(self.space.ndof)
        lib().orc_time_jvp_threads(self.h, _dp(dU), _dp(dR), 1, nthreads)
        return dR

    def time_jvp_threads(self, dU, reps, nthreads):
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        dR = np.zeros(self.space.ndof)
        return lib().orc_time_jvp_threads(self.h, _dp(dU), _dp(dR), reps, nthreads)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_form_free(self.h)
            self.h = None
