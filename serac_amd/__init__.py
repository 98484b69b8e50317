"""serac_amd -- MI355X-native matrix-free residual / J.v path of serac::Functional.

The compute path is the HIP library ``libserac_amd.so`` (C-ABI in
``include/serac_amd.h``); the C++ drop-in front end is
``include/serac_amd/functional.hpp``.  This Python layer is a thin ctypes
mirror used by the tests and the benchmark: numpy arrays in host memory,
``DeviceVector`` for device memory owned by the library.
"""
import ctypes
import os

import numpy as np

from . import capi
from .capi import (SX_BQF_FLUX, SX_BQF_PRESSURE, SX_BQF_TRACTION, SX_ORDERING_BYNODES, SX_ORDERING_BYVDIM,
                   SX_QF_HEAT, SX_QF_J2_LINEAR, SX_QF_LINEAR_ELASTIC,
                   SX_QF_MASS_DIFFUSION, SX_QF_NEOHOOKEAN, SX_WRT_SHAPE, SeracAmdError, SolveReport)

__all__ = [
    "DeviceVector", "Mesh", "Space", "Functional", "Gradient", "QuadratureData", "CSRMatrix", "UNet", "Learned", "SeracAmdError", "device_name",
    "SX_QF_HEAT", "SX_QF_MASS_DIFFUSION", "SX_QF_NEOHOOKEAN", "SX_QF_LINEAR_ELASTIC", "SX_QF_J2_LINEAR",
    "SX_BQF_TRACTION", "SX_BQF_PRESSURE", "SX_BQF_FLUX",
    "SX_ORDERING_BYNODES", "SX_ORDERING_BYVDIM",
]

_vp = ctypes.c_void_p


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def device_name():
    buf = ctypes.create_string_buffer(256)
    capi.call("sx_device_name", buf, 256)
    return buf.value.decode()


class DeviceVector:
    """FP64 (or FP32) vector in device memory allocated by the library."""

    def __init__(self, n, data=None, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = max(self.n, 1) * self.dtype.itemsize
        self.ptr = _vp()
        capi.call("sx_malloc", ctypes.byref(self.ptr), self.nbytes + (self.nbytes % 8))
        if data is None:
            self.zero()
        else:
            self.upload(data)

    @classmethod
    def from_numpy(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype).ravel()
        return cls(a.size, a, dtype)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        assert a.size == self.n, (a.size, self.n)
        capi.call("sx_memcpy_htod", self.ptr, a.ctypes.data_as(_vp), self.n * self.dtype.itemsize)

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        capi.call("sx_memcpy_dtoh", out.ctypes.data_as(_vp), self.ptr, self.n * self.dtype.itemsize)
        return out

    def zero(self):
        capi.call("sx_memset_zero", self.ptr, self.nbytes + (self.nbytes % 8))

    def free(self):
        if self.ptr and self.ptr.value:
            capi.call_rc("sx_free", self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _face_trampoline(predicate):
    """C callback for sx_face_predicate: predicate(corners [4][3], attr) -> bool (called during the add only)."""

    def trampoline(xyz, attr, _user):
        corners = np.ctypeslib.as_array(xyz, shape=(12,)).reshape(4, 3).copy()
        return 1 if predicate(corners, attr) else 0

    return trampoline


def attr_mask(attrs):
    """Bit a - 1 for each attribute a in 1..64."""
    m = 0
    for a in attrs:
        if not 1 <= int(a) <= 64:
            raise ValueError(f"attribute {a} outside 1..64")
        m |= 1 << (int(a) - 1)
    return m


class Mesh:
    def __init__(self, handle):
        self.h = handle
        nv, ne = ctypes.c_int64(), ctypes.c_int64()
        capi.call("sx_mesh_info", self.h, ctypes.byref(nv), ctypes.byref(ne))
        self.nv, self.ne = nv.value, ne.value

    @classmethod
    def cuboid(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
        h = _vp()
        capi.call("sx_mesh_create_cuboid", nx, ny, nz, lx, ly, lz, ctypes.byref(h))
        return cls(h)

    @classmethod
    def from_arrays(cls, X, hexes):
        X = np.ascontiguousarray(X, dtype=np.float64)
        hexes = np.ascontiguousarray(hexes, dtype=np.int32)
        h = _vp()
        capi.call("sx_mesh_create", len(X), _dptr(X), len(hexes),
                  hexes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(h))
        return cls(h)

    @classmethod
    def read_mfem(cls, path=None, text=None):
        """buildMeshFromFile for MFEM mesh v1.0 hex meshes: from a file path or from the text itself."""
        h = _vp()
        if text is not None:
            raw = text.encode() if isinstance(text, str) else bytes(text)
            capi.call("sx_mesh_read_mfem", raw, len(raw), ctypes.byref(h))
        else:
            capi.call("sx_mesh_read_mfem_file", os.fsencode(path), ctypes.byref(h))
        return cls(h)

    def attributes(self):
        """Element attributes [ne]."""
        out = np.zeros(self.ne, dtype=np.int32)
        capi.call("sx_mesh_attributes", self.h, _iptr(out))
        return out

    def boundary(self):
        """Boundary quadrilaterals [nbdr][4] (vertex ids) and their attributes [nbdr]."""
        nb = ctypes.c_int64()
        capi.call("sx_mesh_boundary_info", self.h, ctypes.byref(nb))
        quads = np.zeros((nb.value, 4), dtype=np.int32)
        attrs = np.zeros(nb.value, dtype=np.int32)
        capi.call("sx_mesh_boundary", self.h, _iptr(quads), _iptr(attrs))
        return quads, attrs

    def perturb(self, seed, amp):
        capi.call("sx_mesh_perturb", self.h, seed, amp)

    def vertices(self):
        X = np.zeros((self.nv, 3))
        capi.call("sx_mesh_vertices", self.h, _dptr(X))
        return X

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_mesh_destroy(self.h)
            self.h = None


class Space:
    """H1<order, components> on a hex mesh (mfem ParFiniteElementSpace stand-in)."""

    def __init__(self, mesh, order, components=1, ordering=SX_ORDERING_BYVDIM, element_nodes=None):
        """element_nodes: optional [nelem][(order+1)^3] node ids in the caller's own numbering
        (lexicographic local nodes, as mfem's ElementRestriction orders them)."""
        self.mesh = mesh
        self.h = _vp()
        if element_nodes is None:
            capi.call("sx_space_create_h1_hex", mesh.h, order, components, ordering, ctypes.byref(self.h))
        else:
            emap = np.ascontiguousarray(element_nodes, dtype=np.int32)
            nn = int(emap.max()) + 1 if emap.size else 0
            capi.call("sx_space_create_h1_hex_map", mesh.h, order, components, ordering, nn, _iptr(emap),
                      ctypes.byref(self.h))
        nn, nd, ne, npe = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        capi.call("sx_space_info", self.h, ctypes.byref(nn), ctypes.byref(nd), ctypes.byref(ne), ctypes.byref(npe))
        self.nnodes, self.ndof, self.ne, self.npe = nn.value, nd.value, ne.value, npe.value
        self.order, self.nc, self.ordering = order, components, ordering

    def element_nodes(self):
        m = np.zeros((self.ne, self.npe), dtype=np.int32)
        capi.call("sx_space_element_nodes", self.h, _iptr(m))
        return m

    def node_coords(self):
        x = np.zeros((self.nnodes, 3))
        capi.call("sx_space_node_coords", self.h, _dptr(x))
        return x

    def boundary_dofs(self, face_mask, comp_mask=0x7):
        cnt = ctypes.c_int64()
        capi.call("sx_space_boundary_dofs", self.h, face_mask, comp_mask, None, 0, ctypes.byref(cnt))
        out = np.zeros(cnt.value, dtype=np.int32)
        capi.call("sx_space_boundary_dofs", self.h, face_mask, comp_mask, _iptr(out), cnt.value, ctypes.byref(cnt))
        return out

    def boundary_dofs_attr(self, attrs, comp_mask=0x7):
        """Essential DOFs on the boundary elements with an attribute in `attrs` (1..64)."""
        mask = attr_mask(attrs)
        cnt = ctypes.c_int64()
        capi.call("sx_space_boundary_dofs_attr", self.h, mask, comp_mask, None, 0, ctypes.byref(cnt))
        out = np.zeros(cnt.value, dtype=np.int32)
        capi.call("sx_space_boundary_dofs_attr", self.h, mask, comp_mask, _iptr(out), cnt.value, ctypes.byref(cnt))
        return out

    def gather(self, L):
        """ElementRestriction::Gather on the device: L (DeviceVector) -> E (DeviceVector)."""
        E = DeviceVector(self.ne * self.nc * self.npe)
        capi.call("sx_space_gather", self.h, L.ptr, E.ptr)
        return E

    def scatter_add(self, E, L):
        capi.call("sx_space_scatter_add", self.h, E.ptr, L.ptr)
        return L

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_space_destroy(self.h)
            self.h = None


class Gradient:
    """Functional::Gradient: J.v, diagonal, CG solves."""

    def __init__(self, form, handle):
        self.form = form
        self.h = handle

    def mult(self, dx, df):
        if dx.dtype == np.float32:
            capi.call("sx_grad_apply_f32", self.h, dx.ptr, df.ptr)
        else:
            capi.call("sx_grad_apply", self.h, dx.ptr, df.ptr)
        self.form.synchronize()

    def __call__(self, dx):
        df = DeviceVector(self.form.test.ndof, dtype=dx.dtype)
        self.mult(dx, df)
        return df

    def apply_e(self, duE, k=0):
        """J.v of domain integral k on E-vectors (the Integral seam's jvp)."""
        ne, npe, nc = self.form.integral_info(k)
        drE = DeviceVector(ne * npe * nc)
        capi.call("sx_grad_apply_e", self.h, k, duE.ptr, drE.ptr)
        return drE

    def diagonal(self, out=None):
        d = DeviceVector(self.form.test.ndof) if out is None else out
        capi.call("sx_grad_diagonal", self.h, d.ptr)
        self.form.synchronize()
        return d

    def assemble(self):
        """Gradient::assemble: the linearization as a CSRMatrix on the device."""
        h = _vp()
        capi.call("sx_grad_assemble", self.h, ctypes.byref(h))
        return CSRMatrix(h)

    def fcg(self, learned, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000, raise_on_noconv=True):
        """Flexible CG with the learned preconditioner on J_bc x = b."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_fcg_solve", self.h, learned.h, _iptr(ess), len(ess), b.ptr, x.ptr, rel_tol, abs_tol,
                          max_iter, ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()

    def fgmres(self, learned, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000, restart=50,
               raise_on_noconv=True):
        """Flexible GMRES with the learned preconditioner (Jacobi when learned is None) on J_bc x = b."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_fgmres_solve", self.h, learned.h if learned is not None else None, _iptr(ess),
                          len(ess), b.ptr, x.ptr, rel_tol, abs_tol, max_iter, restart, ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()

    def tangent_bytes(self):
        b = ctypes.c_int64()
        capi.call("sx_grad_tangent_bytes", self.h, ctypes.byref(b))
        return b.value

    def apply_timed(self, dx, df, nrep):
        tot, ker, per = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        prec = capi.SX_PREC_F32 if dx.dtype == np.float32 else capi.SX_PREC_F64
        capi.call("sx_grad_apply_timed", self.h, prec, dx.ptr, df.ptr, nrep, ctypes.byref(tot), ctypes.byref(ker),
                  ctypes.byref(per))
        return tot.value, ker.value, per.value

    def cg(self, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000, raise_on_noconv=True):
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_cg_solve", self.h, _iptr(ess), len(ess), b.ptr, x.ptr, rel_tol, abs_tol, max_iter,
                          ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()


    def mg_cg(self, mg, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000, raise_on_noconv=True):
        """CG on J_bc x = b preconditioned by the multigrid V-cycle (set up for this linearization first)."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_mg_pcg_solve", self.h, mg.h, _iptr(ess), len(ess), b.ptr, x.ptr, rel_tol, abs_tol,
                          max_iter, ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()

    def mg_gmres(self, mg, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000, restart=50, raise_on_noconv=True):
        """Flexible GMRES on J_bc x = b right-preconditioned by the multigrid V-cycle."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_mg_gmres_solve", self.h, mg.h, _iptr(ess), len(ess), b.ptr, x.ptr, rel_tol, abs_tol,
                          max_iter, restart, ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()

    def gmres(self, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000, restart=50, raise_on_noconv=True):
        """Restarted GMRES with left Jacobi preconditioning on J_bc x = b."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_gmres_solve", self.h, _iptr(ess), len(ess), b.ptr, x.ptr, rel_tol, abs_tol, max_iter,
                          restart, ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()


class CSRMatrix:
    """Assembled sparse linearization (device CSR, int64 rows, int32 columns)."""

    def __init__(self, handle):
        self.h = handle
        nr, nc, nnz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        capi.call("sx_csr_info", self.h, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nnz))
        self.shape = (nr.value, nc.value)
        self.nnz = nnz.value

    def arrays(self):
        """(row_ptr, col, val) as host numpy arrays."""
        rp = np.zeros(self.shape[0] + 1, dtype=np.int64)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        capi.call("sx_csr_download", self.h, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                  col.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dptr(val))
        return rp, col, val

    def to_scipy(self):
        import scipy.sparse as sp

        rp, col, val = self.arrays()
        return sp.csr_matrix((val, col, rp), shape=self.shape)

    def mult(self, x, y):
        capi.call("sx_csr_spmv", self.h, x.ptr, y.ptr)

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_csr_destroy(self.h)
            self.h = None


class UNet:
    """The learned preconditioner's U-Net on an nx x ny x nz vertex grid (3 -> 3 channels)."""

    LAYERS = [(3, 0, 16), (16, 0, 16), (16, 0, 32), (32, 0, 32), (32, 0, 64), (64, 0, 64), (64, 0, 64),
              (64, 0, 64), (64, 64, 64), (64, 32, 32), (32, 16, 16), (16, 0, 3)]

    def __init__(self, nx, ny, nz, seed=0):
        h = _vp()
        capi.call("sx_unet_create", nx, ny, nz, seed, ctypes.byref(h))
        self.h = h
        self.dims = (nx, ny, nz)
        npar, nvox = ctypes.c_int64(), ctypes.c_int64()
        capi.call("sx_unet_info", self.h, ctypes.byref(npar), ctypes.byref(nvox))
        self.nparam, self.nvox = npar.value, nvox.value

    def _fptr(self, a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        capi.call("sx_unet_get_params", self.h, self._fptr(out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        capi.call("sx_unet_set_params", self.h, self._fptr(p))

    def grads(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        capi.call("sx_unet_get_grads", self.h, self._fptr(out))
        return out

    def forward(self, x):
        """x: [nvox, 3] (numpy or float32 DeviceVector) -> numpy [nvox, 3]"""
        xd = x if isinstance(x, DeviceVector) else DeviceVector.from_numpy(np.asarray(x).ravel(), dtype=np.float32)
        yd = DeviceVector(self.nvox * 3, dtype=np.float32)
        capi.call("sx_unet_forward", self.h, xd.ptr, yd.ptr)
        return yd.numpy().reshape(self.nvox, 3)

    def backward(self, x, t):
        xd = x if isinstance(x, DeviceVector) else DeviceVector.from_numpy(np.asarray(x).ravel(), dtype=np.float32)
        td = t if isinstance(t, DeviceVector) else DeviceVector.from_numpy(np.asarray(t).ravel(), dtype=np.float32)
        loss = ctypes.c_double()
        capi.call("sx_unet_backward", self.h, xd.ptr, td.ptr, ctypes.byref(loss))
        return loss.value

    def adam(self, lr):
        capi.call("sx_unet_adam", self.h, lr)

    def set_activation(self, activation, seed=0):
        """SX_UNET_RELU or SX_UNET_IDENTITY; re-initialises the parameters."""
        capi.call("sx_unet_set_activation", self.h, activation, seed)

    def set_mfma(self, on):
        """Forward convolutions on the matrix cores (default) or the vector units."""
        capi.call("sx_unet_set_mfma", self.h, int(bool(on)))

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "owned", True) and capi._lib is not None:
            capi._lib.sx_unet_destroy(self.h)
            self.h = None


class Multigrid:
    """Geometric multigrid V-cycle of a 3-component structured space: the order-p nodal grid, then the
    mesh's vertex grids with Galerkin stencils, Chebyshev-Jacobi smoothing and a dense coarsest solve."""

    def __init__(self, space, fine_degree=2, coarse_degree=2, smoothing_range=12.0, power_iterations=12):
        h = _vp()
        capi.call("sx_mg_create", space.h, ctypes.byref(h))
        self.h = h
        self.space = space
        self.configure(fine_degree, coarse_degree, smoothing_range, power_iterations)

    def configure(self, fine_degree=2, coarse_degree=2, smoothing_range=12.0, power_iterations=12):
        capi.call("sx_mg_configure", self.h, fine_degree, coarse_degree, smoothing_range, power_iterations)

    def set_reuse(self, linearizations):
        """Inside one Newton call, one hierarchy serves up to this many linear solves (default 1)."""
        capi.call("sx_mg_set_reuse", self.h, linearizations)

    def set_smoother_precision(self, prec):
        """SX_PREC_BF16: level-0 applications of the V-cycle with BF16 tangent copies; SX_PREC_F64: the form's."""
        capi.call("sx_mg_set_smoother_precision", self.h, prec)

    def set_smoother_vectors(self, prec):
        """SX_PREC_F32: the level-0 part of the V-cycle on FP32 vectors; SX_PREC_F64 (default)."""
        capi.call("sx_mg_set_smoother_vectors", self.h, prec)

    def set_galerkin(self, mode):
        """Level-1 operator: 0 element-wise from the stored tangent where possible, 1 by probing."""
        capi.call("sx_mg_set_galerkin", self.h, mode)

    def info(self):
        nl, ns, ne = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        capi.call("sx_mg_info", self.h, ctypes.byref(nl), ctypes.byref(ns))
        capi.call("sx_mg_element_setups", self.h, ctypes.byref(ne))
        levels = []
        for lvl in range(nl.value):
            nd, dims, lm = ctypes.c_int64(), (ctypes.c_int * 3)(), ctypes.c_double()
            capi.call("sx_mg_level", self.h, lvl, ctypes.byref(nd), dims, ctypes.byref(lm))
            levels.append({"ndof": nd.value, "dims": tuple(dims), "lmax": lm.value})
        return {"levels": levels, "setups": ns.value, "element_setups": ne.value}

    def setup(self, grad, ess):
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        capi.call("sx_mg_setup", self.h, grad.h, _iptr(ess), len(ess))

    def apply(self, r, z):
        """z = M r with the hierarchy of the last setup."""
        capi.call("sx_mg_apply", self.h, r.ptr, z.ptr)

    def level_apply(self, level, x, y):
        """y = A_level x (level 0: J with the essential rows zeroed; above: the Galerkin stencil)."""
        capi.call("sx_mg_level_apply", self.h, level, x.ptr, y.ptr)

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_mg_destroy(self.h)
            self.h = None


class Learned:
    """The learned coarse-grid preconditioner of a 3-component structured space."""

    def __init__(self, space, seed=0):
        h = _vp()
        capi.call("sx_learned_create", space.h, seed, ctypes.byref(h))
        self.h = h
        self.space = space
        self.net = UNet.__new__(UNet)  # a view of the network the preconditioner owns
        npairs, fb, net = ctypes.c_int(), ctypes.c_int64(), _vp()
        capi.call("sx_learned_info", self.h, ctypes.byref(npairs), ctypes.byref(fb), ctypes.byref(net))
        self.net.h, self.net.owned = net, False
        npar, nvox = ctypes.c_int64(), ctypes.c_int64()
        capi.call("sx_unet_info", net, ctypes.byref(npar), ctypes.byref(nvox))
        self.net.nparam, self.net.nvox = npar.value, nvox.value

    def configure(self, stride=50, window=50, max_pairs=256, stall_window=20):
        capi.call("sx_learned_configure", self.h, stride, window, max_pairs, stall_window)

    def info(self):
        npairs, fb, net = ctypes.c_int(), ctypes.c_int64(), _vp()
        capi.call("sx_learned_info", self.h, ctypes.byref(npairs), ctypes.byref(fb), ctypes.byref(net))
        return {"pairs": npairs.value, "fallbacks": fb.value}

    def pair(self, k):
        """Usable pair k as (P^T r_k, vertex-mean correction of the next `window` CG iterations), [nvertex][3] FP32."""
        nv = self.net.nvox
        r = np.zeros((nv, 3), dtype=np.float32)
        t = np.zeros((nv, 3), dtype=np.float32)
        capi.call("sx_learned_get_pair", self.h, k, r.ctypes.data, t.ctypes.data)
        return r, t

    def train(self, steps, lr=1e-3, seed=0, batch=1):
        loss = ctypes.c_double()
        capi.call("sx_learned_train", self.h, steps, batch, lr, seed, ctypes.byref(loss))
        return loss.value

    def generate_pairs(self, grad, ess, npairs, seed=0):
        """Pairs (P^T J_bc e, avg e) for random smooth corrections e, from the linearization grad."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        capi.call("sx_learned_generate_pairs", self.h, grad.h, _iptr(ess), len(ess), npairs, seed)

    def gain(self):
        g = ctypes.c_double()
        capi.call("sx_learned_gain", self.h, ctypes.byref(g))
        return g.value

    def set_gain(self, gain):
        """0 switches the network's correction off (the ablation of the same solver)."""
        capi.call("sx_learned_set_gain", self.h, float(gain))

    def set_multigrid(self, mg):
        """M(r) = V-cycle(r) + the network's correction of what the V-cycle leaves (None detaches)."""
        self._mg = mg  # borrowed by the library: keep it alive
        capi.call("sx_learned_set_multigrid", self.h, mg.h if mg is not None else None)

    def set_activation(self, activation, seed=0):
        """SX_UNET_RELU or SX_UNET_IDENTITY; re-initialises the network."""
        capi.call("sx_unet_set_activation", self.net.h, activation, seed)

    def record_cg(self, grad, ess, b, x, rel_tol=1e-10, abs_tol=0.0, max_iter=10000):
        """Jacobi-CG on J_bc x = b recording training pairs."""
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        capi.call("sx_learned_record_cg", self.h, grad.h, _iptr(ess), len(ess), b.ptr, x.ptr, rel_tol, abs_tol,
                  max_iter, ctypes.byref(rep))
        return rep.as_dict()

    def apply(self, grad, ess, r):
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        z = DeviceVector(self.space.ndof)
        capi.call("sx_learned_apply", self.h, grad.h, _iptr(ess), len(ess), r.ptr, z.ptr)
        return z

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_learned_destroy(self.h)
            self.h = None


class QoI:
    """Functional<double(trials...)>: a scalar quantity of interest on one or
    several trial spaces (a list of spaces for several)."""

    def __init__(self, space):
        h = _vp()
        spaces = list(space) if isinstance(space, (list, tuple)) else [space]
        if len(spaces) == 1:
            capi.call("sx_qoi_create", spaces[0].h, ctypes.byref(h))
        else:
            arr = (_vp * len(spaces))(*[sp.h for sp in spaces])
            capi.call("sx_qoi_create_multi", arr, len(spaces), ctypes.byref(h))
        self.h = h
        self.spaces = spaces
        self.space = spaces[0]

    def add_domain_integral(self, qoi, params=(), elems=None, depends_on=None):
        p = np.zeros(16)
        p[: len(params)] = params
        el = (ctypes.c_int * len(elems))(*elems) if elems is not None else None
        ne = len(elems) if elems is not None else 0
        if depends_on is None:
            capi.call("sx_qoi_add_domain_integral", self.h, qoi, _dptr(p), len(params), el, ne)
        else:
            dep = (ctypes.c_int * len(depends_on))(*depends_on)
            capi.call("sx_qoi_add_domain_integral_dep", self.h, qoi, _dptr(p), len(params), dep, len(depends_on),
                      el, ne)

    def add_boundary_integral(self, qoi, params=(), face_mask=0, attrs=None, faces=None, predicate=None):
        """Boundary QoI over the planes of face_mask (0: the whole boundary), the attributes `attrs`, the
        exterior faces `faces` ([n][4] vertex ids) or those predicate(corners [4][3], attr) selects."""
        p = np.zeros(16)
        p[: len(params)] = params
        if predicate is not None:
            cb = capi.FACE_PREDICATE(_face_trampoline(predicate))
            capi.call("sx_qoi_add_boundary_integral_pred", self.h, qoi, _dptr(p), len(params), cb, None)
        elif faces is not None:
            fv = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 4)
            capi.call("sx_qoi_add_boundary_integral_faces", self.h, qoi, _dptr(p), len(params), _iptr(fv), len(fv))
        else:
            capi.call("sx_qoi_add_boundary_integral", self.h, qoi, _dptr(p), len(params), face_mask,
                      attr_mask(attrs) if attrs is not None else 0)

    def __call__(self, t, *U):
        v = ctypes.c_double()
        if len(self.spaces) == 1:
            capi.call("sx_qoi_eval", self.h, t, U[0].ptr, ctypes.byref(v))
        else:
            arr = (_vp * len(U))(*[u.ptr for u in U])
            capi.call("sx_qoi_eval_multi", self.h, t, arr, ctypes.byref(v))
        return v.value

    def gradient(self, t, *U, wrt=0):
        """dq/dU_wrt as a DeviceVector."""
        g = DeviceVector(self.spaces[wrt].ndof)
        if len(self.spaces) == 1:
            capi.call("sx_qoi_gradient", self.h, t, U[0].ptr, g.ptr)
        else:
            arr = (_vp * len(U))(*[u.ptr for u in U])
            capi.call("sx_qoi_gradient_multi", self.h, t, arr, wrt, g.ptr)
        return g

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_qoi_destroy(self.h)
            self.h = None


class QuadratureData:
    """QuadratureData<State>: nelem x nq records of `size` doubles on the device."""

    def __init__(self, nelem, nq, init):
        init = np.ascontiguousarray(init, dtype=np.float64)
        h = _vp()
        capi.call("sx_qdata_create", nelem, nq, len(init), _dptr(init), ctypes.byref(h))
        self.h = h
        self.nelem, self.nq, self.size = nelem, nq, len(init)

    def numpy(self):
        """The state as [elem][q][size]."""
        out = np.zeros((self.nelem, self.nq, self.size))
        capi.call("sx_qdata_get", self.h, _dptr(out))
        return out

    def set(self, state):
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert st.size == self.nelem * self.nq * self.size
        capi.call("sx_qdata_set", self.h, _dptr(st))

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_qdata_destroy(self.h)
            self.h = None


class Functional:
    """serac::Functional<test(trials...)> mirror over the C-ABI."""

    def __init__(self, test, trials):
        self.test = test
        self.trials = list(trials)
        arr = (_vp * len(self.trials))(*[t.h for t in self.trials])
        self.h = _vp()
        capi.call("sx_form_create", test.h, arr, len(self.trials), ctypes.byref(self.h))
        self.output = DeviceVector(test.ndof)

    def set_option(self, option, value):
        capi.call("sx_form_set_option", self.h, option, value)

    def add_domain_integral(self, qfunc, params, depends_on=(0,), elems=None, qdata=None):
        """AddDomainIntegral; qdata (a QuadratureData) for qfunctions with state."""
        p = np.zeros(16)
        p[: len(params)] = params
        dep = (ctypes.c_int * len(depends_on))(*depends_on)
        if elems is None:
            el, nel = None, 0
        else:
            e = np.ascontiguousarray(elems, dtype=np.int32)
            el, nel = e.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(e)
            self._keep = e
        capi.call("sx_form_add_domain_integral_qdata", self.h, qfunc, _dptr(p), len(params), dep, len(depends_on),
                  el, nel, qdata.h if qdata is not None else None)
        if qdata is not None:
            self._qdata = getattr(self, "_qdata", []) + [qdata]  # the form borrows it

    def add_boundary_integral(self, qfunc, params, depends_on=(0,), face_mask=0, attrs=None, faces=None,
                              predicate=None):
        """AddBoundaryIntegral over the bounding-box planes of face_mask
        (bit 0 x = min, 1 x = max, 2 y = min, ...; 0: the whole boundary), or,
        with `attrs`, over the boundary elements with those attributes, or over
        `faces` ([n][4] vertex ids of exterior faces), or over the exterior
        faces for which predicate(corners [4][3], attr) is true
        (Domain::ofBoundaryElements)."""
        p = np.zeros(16)
        p[: len(params)] = params
        dep = (ctypes.c_int * len(depends_on))(*depends_on)
        if predicate is not None:
            cb = capi.FACE_PREDICATE(_face_trampoline(predicate))
            capi.call("sx_form_add_boundary_integral_pred", self.h, qfunc, _dptr(p), len(params), dep,
                      len(depends_on), cb, None)
        elif faces is not None:
            fv = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 4)
            capi.call("sx_form_add_boundary_integral_faces", self.h, qfunc, _dptr(p), len(params), dep,
                      len(depends_on), _iptr(fv), len(fv))
        elif attrs is not None:
            capi.call("sx_form_add_boundary_integral_attr", self.h, qfunc, _dptr(p), len(params), dep,
                      len(depends_on), attr_mask(attrs))
        else:
            capi.call("sx_form_add_boundary_integral", self.h, qfunc, _dptr(p), len(params), dep, len(depends_on),
                      face_mask)

    def set_preconditioner(self, learned, mode):
        """Linear solves of newton / quasistatic_solve: SX_PRECOND_JACOBI, _RECORD or _LEARNED."""
        capi.call("sx_form_set_preconditioner", self.h, learned.h if learned is not None else None, mode)
        self._learned = learned  # the form borrows it

    def set_shape_displacement(self, p):
        """ShapeAwareFunctional: evaluate the domain integrals on X + p (p a DeviceVector of the test space, kept
        alive by the form; None restores X)."""
        capi.call("sx_form_set_shape_displacement", self.h, p.ptr if p is not None else None)
        self._shape = p

    def set_multigrid(self, mg):
        """CG solves of newton / quasistatic_solve preconditioned by the multigrid V-cycle (None: Jacobi)."""
        capi.call("sx_form_set_multigrid", self.h, mg.h if mg is not None else None)
        self._mg = mg  # the form borrows it

    def update_qdata(self, flag):
        """updateQdata(bool): while set, evaluations commit the quadrature-point state."""
        capi.call("sx_form_update_qdata", self.h, int(bool(flag)))

    def __call__(self, t, *inputs, wrt=-1):
        """Residual into self.output (a DeviceVector reused across calls, like
        the reference's output_T_); with wrt >= 0 also returns the Gradient;
        wrt="shape" returns the gradient in the shape displacement."""
        if wrt == "shape":
            wrt = SX_WRT_SHAPE
        arr = (_vp * len(inputs))(*[u.ptr for u in inputs])
        g = _vp()
        capi.call("sx_form_eval", self.h, t, arr, self.output.ptr, wrt, ctypes.byref(g))
        self.synchronize()
        if wrt >= 0 or wrt == SX_WRT_SHAPE:
            return self.output, Gradient(self, g)
        return self.output

    def synchronize(self):
        capi.call("sx_form_synchronize", self.h)

    def integral_info(self, k=0):
        ne, npe, nc = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        capi.call("sx_form_integral_info", self.h, k, ctypes.byref(ne), ctypes.byref(npe), ctypes.byref(nc))
        return ne.value, npe.value, nc.value

    def eval_e(self, uE, k=0, t=0.0, linearize=False):
        """Domain integral k on E-vectors [elem][comp][node] (the Integral seam): returns r_E."""
        ne, npe, nc = self.integral_info(k)
        rE = DeviceVector(ne * npe * nc)
        capi.call("sx_form_eval_e", self.h, k, t, uE.ptr, rE.ptr, int(linearize))
        return rE

    def newton(self, x, ess, wrt=0, rel_tol=1e-8, abs_tol=1e-12, max_iter=20, lin_rel_tol=1e-8, lin_abs_tol=1e-12,
               lin_max_iter=5000, raise_on_noconv=True):
        ess = np.ascontiguousarray(ess, dtype=np.int32)
        rep = SolveReport()
        rc = capi.call_rc("sx_newton_solve", self.h, wrt, _iptr(ess), len(ess), x.ptr, rel_tol, abs_tol, max_iter,
                          lin_rel_tol, lin_abs_tol, lin_max_iter, ctypes.byref(rep))
        if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
            raise SeracAmdError(rc, capi.load().sx_last_error().decode())
        return rep.as_dict()

    def __del__(self):
        if getattr(self, "h", None) and capi._lib is not None:
            capi._lib.sx_form_destroy(self.h)
            self.h = None


def quasistatic_solve(form, x, ess, ess_final, nsteps, warm_start=True, rel_tol=1e-8, abs_tol=0.0, max_iter=20,
                      lin_rel_tol=1e-6, lin_abs_tol=0.0, lin_max_iter=20000, wrt=0, raise_on_noconv=True):
    """Load-stepped Newton + Jacobi-CG (quasiStaticSolve with warm start); x is a DeviceVector."""
    ess = np.ascontiguousarray(ess, dtype=np.int32)
    fin = np.ascontiguousarray(ess_final, dtype=np.float64)
    opts = capi.SolverOptions(rel_tol, abs_tol, max_iter, lin_rel_tol, lin_abs_tol, lin_max_iter)
    reps = (SolveReport * nsteps)()
    rc = capi.call_rc("sx_quasistatic_solve", form.h, wrt, _iptr(ess), len(ess), _dptr(fin), nsteps,
                      int(warm_start), ctypes.byref(opts), x.ptr, reps)
    if rc != 0 and (rc != capi.SX_ERR_NOCONV or raise_on_noconv):
        raise SeracAmdError(rc, capi.load().sx_last_error().decode())
    return [r.as_dict() for r in reps]
