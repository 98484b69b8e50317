"""ctypes binding of the C-ABI declared in include/serac_amd.h.

Loads the in-tree ``serac_amd/libserac_amd.so``.  There is no fallback: if the
HIP library is missing or fails to load, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libserac_amd.so")

SX_OK = 0
SX_ERR_ARG = 1
SX_ERR_HIP = 2
SX_ERR_UNSUPPORTED = 3
SX_ERR_STATE = 4
SX_ERR_NOCONV = 5

SX_WRT_SHAPE = -2
SX_ORDERING_BYNODES = 0
SX_ORDERING_BYVDIM = 1

SX_OPT_SCATTER = 1
SX_OPT_XCD_REMAP = 2
SX_OPT_TANGENT_PRECISION = 3
SX_OPT_JVP_ARITHMETIC = 4
SX_OPT_LINEAR_SOLVER = 5
SX_OPT_GMRES_RESTART = 6
SX_OPT_SCATTER_ORDER = 7
SX_OPT_LINE_SEARCH = 8
SX_OPT_MIN_NEWTON_ITERATIONS = 9
SX_OPT_PRINT_LEVEL = 10
SX_LINSOLVER_CG = 0
SX_LINSOLVER_GMRES = 1
SX_PREC_F64 = 0
SX_PREC_F32 = 1
SX_PREC_BF16 = 2
SX_SCATTER_COLORED = 0
SX_SCATTER_ATOMIC = 1

SX_QF_HEAT = 1
SX_QF_MASS_DIFFUSION = 2
SX_QF_NEOHOOKEAN = 3
SX_QF_LINEAR_ELASTIC = 4
SX_QF_J2_LINEAR = 5
SX_BQF_TRACTION = 101
SX_BQF_PRESSURE = 102
SX_BQF_FLUX = 103
SX_PRECOND_JACOBI = 0
SX_PRECOND_RECORD = 1
SX_PRECOND_LEARNED = 2
SX_UNET_RELU = 0
SX_QOI_MEASURE = 1
SX_QOI_POSITION = 2
SX_QOI_VALUE = 3
SX_QOI_NEOHOOKEAN_ENERGY = 4
SX_QOI_INNER = 5
SX_UNET_IDENTITY = 1

_vp = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
# int (*sx_face_predicate)(const double* corners, int attr, void* user)
FACE_PREDICATE = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)
_I64 = ctypes.POINTER(ctypes.c_int64)
_int = ctypes.c_int


class SolveReport(ctypes.Structure):
    _fields_ = [
        ("newton_iterations", ctypes.c_int),
        ("linear_iterations", ctypes.c_int),
        ("initial_norm", ctypes.c_double),
        ("final_norm", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("line_search_steps", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolverOptions(ctypes.Structure):
    _fields_ = [
        ("rel_tol", ctypes.c_double),
        ("abs_tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("lin_rel_tol", ctypes.c_double),
        ("lin_abs_tol", ctypes.c_double),
        ("lin_max_iter", ctypes.c_int),
    ]


SIGNATURES = {
    "sx_last_error": (ctypes.c_char_p, []),
    "sx_init": (_int, [_int]),
    "sx_device_name": (_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "sx_malloc": (_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "sx_free": (_int, [_vp]),
    "sx_memcpy_htod": (_int, [_vp, _vp, ctypes.c_size_t]),
    "sx_memcpy_dtoh": (_int, [_vp, _vp, ctypes.c_size_t]),
    "sx_memcpy_dtod": (_int, [_vp, _vp, ctypes.c_size_t]),
    "sx_memset_zero": (_int, [_vp, ctypes.c_size_t]),
    "sx_device_synchronize": (_int, []),
    "sx_mesh_create_cuboid": (_int, [_int, _int, _int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(_vp)]),
    "sx_mesh_create": (_int, [_int, _D, _int, ctypes.POINTER(_int), ctypes.POINTER(_vp)]),
    "sx_mesh_perturb": (_int, [_vp, ctypes.c_uint64, ctypes.c_double]),
    "sx_mesh_read_mfem": (_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "sx_mesh_read_mfem_file": (_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "sx_mesh_attributes": (_int, [_vp, _I32]),
    "sx_mesh_boundary_info": (_int, [_vp, _I64]),
    "sx_mesh_boundary": (_int, [_vp, _I32, _I32]),
    "sx_space_boundary_dofs_attr": (_int, [_vp, ctypes.c_uint64, _int, _I32, ctypes.c_int64, _I64]),
    "sx_form_add_boundary_integral_attr": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int, ctypes.c_uint64]),
    "sx_form_add_boundary_integral_pred": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int, FACE_PREDICATE,
                                                  _vp]),
    "sx_form_add_boundary_integral_faces": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int, _I32,
                                                   ctypes.c_int64]),
    "sx_mesh_info": (_int, [_vp, _I64, _I64]),
    "sx_mesh_vertices": (_int, [_vp, _D]),
    "sx_mesh_destroy": (_int, [_vp]),
    "sx_space_create_h1_hex": (_int, [_vp, _int, _int, _int, ctypes.POINTER(_vp)]),
    "sx_space_create_h1_hex_map": (_int, [_vp, _int, _int, _int, ctypes.c_int64, _I32, ctypes.POINTER(_vp)]),
    "sx_space_info": (_int, [_vp, _I64, _I64, _I64, ctypes.POINTER(_int)]),
    "sx_space_element_nodes": (_int, [_vp, _I32]),
    "sx_space_node_coords": (_int, [_vp, _D]),
    "sx_space_boundary_dofs": (_int, [_vp, _int, _int, _I32, ctypes.c_int64, _I64]),
    "sx_space_gather": (_int, [_vp, _vp, _vp]),
    "sx_space_scatter_add": (_int, [_vp, _vp, _vp]),
    "sx_space_destroy": (_int, [_vp]),
    "sx_form_create": (_int, [_vp, ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp)]),
    "sx_form_add_domain_integral": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int,
                                           ctypes.POINTER(_int), _int]),
    "sx_form_add_domain_integral_qdata": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int,
                                                  ctypes.POINTER(_int), _int, _vp]),
    "sx_form_update_qdata": (_int, [_vp, _int]),
    "sx_learned_create": (_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "sx_learned_configure": (_int, [_vp, _int, _int, _int, _int]),
    "sx_learned_info": (_int, [_vp, ctypes.POINTER(_int), _I64, ctypes.POINTER(_vp)]),
    "sx_learned_get_pair": (_int, [_vp, _int, ctypes.c_void_p, ctypes.c_void_p]),
    "sx_learned_train": (_int, [_vp, _int, _int, ctypes.c_double, ctypes.c_uint64, _D]),
    "sx_learned_generate_pairs": (_int, [_vp, _vp, _I32, ctypes.c_int64, _int, ctypes.c_uint64]),
    "sx_learned_gain": (_int, [_vp, _D]),
    "sx_learned_set_gain": (_int, [_vp, ctypes.c_double]),
    "sx_learned_set_multigrid": (_int, [_vp, _vp]),
    "sx_learned_apply": (_int, [_vp, _vp, _I32, ctypes.c_int64, _vp, _vp]),
    "sx_learned_destroy": (_int, [_vp]),
    "sx_mg_create": (_int, [_vp, ctypes.POINTER(_vp)]),
    "sx_mg_configure": (_int, [_vp, _int, _int, ctypes.c_double, _int]),
    "sx_mg_set_reuse": (_int, [_vp, _int]),
    "sx_mg_set_galerkin": (_int, [_vp, _int]),
    "sx_mg_set_smoother_precision": (_int, [_vp, _int]),
    "sx_mg_set_smoother_vectors": (_int, [_vp, _int]),
    "sx_mg_element_setups": (_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "sx_mg_info": (_int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(ctypes.c_int64)]),
    "sx_mg_level": (_int, [_vp, _int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_int),
                           ctypes.POINTER(ctypes.c_double)]),
    "sx_mg_setup": (_int, [_vp, _vp, _I32, ctypes.c_int64]),
    "sx_mg_apply": (_int, [_vp, _vp, _vp]),
    "sx_mg_level_apply": (_int, [_vp, _int, _vp, _vp]),
    "sx_mg_pcg_solve": (_int, [_vp, _vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int,
                               _vp]),
    "sx_mg_gmres_solve": (_int, [_vp, _vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int,
                                 _int, _vp]),
    "sx_form_set_multigrid": (_int, [_vp, _vp]),
    "sx_form_set_shape_displacement": (_int, [_vp, _vp]),
    "sx_mg_destroy": (_int, [_vp]),
    "sx_form_set_preconditioner": (_int, [_vp, _vp, _int]),
    "sx_fcg_solve": (_int, [_vp, _vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int,
                            ctypes.POINTER(SolveReport)]),
    "sx_learned_record_cg": (_int, [_vp, _vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                    _int, ctypes.POINTER(SolveReport)]),
    "sx_unet_set_activation": (_int, [_vp, _int, ctypes.c_uint64]),
    "sx_unet_set_mfma": (_int, [_vp, _int]),
    "sx_qoi_create": (_int, [_vp, ctypes.POINTER(_vp)]),
    "sx_qoi_create_multi": (_int, [ctypes.POINTER(_vp), _int, ctypes.POINTER(_vp)]),
    "sx_qoi_add_domain_integral_dep": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int, ctypes.POINTER(_int),
                                              _int]),
    "sx_qoi_eval_multi": (_int, [_vp, ctypes.c_double, ctypes.POINTER(_vp), _D]),
    "sx_qoi_gradient_multi": (_int, [_vp, ctypes.c_double, ctypes.POINTER(_vp), _int, _vp]),
    "sx_qoi_add_domain_integral": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int]),
    "sx_qoi_add_boundary_integral": (_int, [_vp, _int, _D, _int, _int, ctypes.c_uint64]),
    "sx_qoi_add_boundary_integral_pred": (_int, [_vp, _int, _D, _int, FACE_PREDICATE, _vp]),
    "sx_qoi_add_boundary_integral_faces": (_int, [_vp, _int, _D, _int, _I32, ctypes.c_int64]),
    "sx_qoi_eval": (_int, [_vp, ctypes.c_double, _vp, _D]),
    "sx_qoi_gradient": (_int, [_vp, ctypes.c_double, _vp, _vp]),
    "sx_qoi_destroy": (_int, [_vp]),
    "sx_unet_create": (_int, [_int, _int, _int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "sx_unet_info": (_int, [_vp, _I64, _I64]),
    "sx_unet_get_params": (_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "sx_unet_set_params": (_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "sx_unet_get_grads": (_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "sx_unet_forward": (_int, [_vp, _vp, _vp]),
    "sx_unet_backward": (_int, [_vp, _vp, _vp, _D]),
    "sx_unet_adam": (_int, [_vp, ctypes.c_double]),
    "sx_unet_destroy": (_int, [_vp]),
    "sx_form_add_boundary_integral": (_int, [_vp, _int, _D, _int, ctypes.POINTER(_int), _int, _int]),
    "sx_grad_assemble": (_int, [_vp, ctypes.POINTER(_vp)]),
    "sx_csr_info": (_int, [_vp, _I64, _I64, _I64]),
    "sx_csr_arrays": (_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "sx_csr_download": (_int, [_vp, _I64, _I32, _D]),
    "sx_csr_spmv": (_int, [_vp, _vp, _vp]),
    "sx_csr_destroy": (_int, [_vp]),
    "sx_qdata_create": (_int, [ctypes.c_int64, _int, _int, _D, ctypes.POINTER(_vp)]),
    "sx_qdata_info": (_int, [_vp, _I64]),
    "sx_qdata_get": (_int, [_vp, _D]),
    "sx_qdata_set": (_int, [_vp, _D]),
    "sx_qdata_destroy": (_int, [_vp]),
    "sx_form_eval": (_int, [_vp, ctypes.c_double, ctypes.POINTER(_vp), _vp, _int, ctypes.POINTER(_vp)]),
    "sx_form_set_option": (_int, [_vp, _int, ctypes.c_int64]),
    "sx_form_synchronize": (_int, [_vp]),
    "sx_form_stream": (_int, [_vp, ctypes.POINTER(_vp)]),
    "sx_form_destroy": (_int, [_vp]),
    "sx_grad_apply": (_int, [_vp, _vp, _vp]),
    "sx_form_integral_info": (_int, [_vp, _int, _I64, ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "sx_form_eval_e": (_int, [_vp, _int, ctypes.c_double, _vp, _vp, _int]),
    "sx_grad_apply_e": (_int, [_vp, _int, _vp, _vp]),
    "sx_grad_diagonal": (_int, [_vp, _vp]),
    "sx_grad_tangent_bytes": (_int, [_vp, _I64]),
    "sx_grad_apply_f32": (_int, [_vp, _vp, _vp]),
    "sx_grad_apply_timed": (_int, [_vp, _int, _vp, _vp, _int, _D, _D, ctypes.POINTER(_int)]),
    "sx_cg_solve": (_int, [_vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int,
                           ctypes.POINTER(SolveReport)]),
    "sx_cg_solve_noalloc": (_int, [_vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int,
                                   ctypes.POINTER(SolveReport)]),
    "sx_fgmres_solve": (_int, [_vp, _vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int,
                               _int, ctypes.POINTER(SolveReport)]),
    "sx_gmres_solve": (_int, [_vp, _I32, ctypes.c_int64, _vp, _vp, ctypes.c_double, ctypes.c_double, _int, _int,
                              ctypes.POINTER(SolveReport)]),
    "sx_newton_solve": (_int, [_vp, _int, _I32, ctypes.c_int64, _vp, ctypes.c_double, ctypes.c_double, _int,
                               ctypes.c_double, ctypes.c_double, _int, ctypes.POINTER(SolveReport)]),
    "sx_quasistatic_solve": (_int, [_vp, _int, _I32, ctypes.c_int64, _D, _int, _int, ctypes.POINTER(SolverOptions),
                                    _vp, ctypes.POINTER(SolveReport)]),
    "sx_vec_dot": (_int, [_vp, _vp, ctypes.c_int64, _D]),
    "sx_vec_axpy": (_int, [ctypes.c_double, _vp, _vp, ctypes.c_int64]),
}


class SeracAmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"serac_amd error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libserac_amd.so (raises OSError when it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                          f"g.build()'` (no CPU fallback exists)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != SX_OK:
        raise SeracAmdError(rc, lib.sx_last_error().decode())
    return rc


def call_rc(name, *args):
    """Like call() but returns the status code instead of raising on it."""
    return getattr(load(), name)(*args)
