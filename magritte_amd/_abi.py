"""ctypes mirror of include/mgx.h and the loader of the in-tree libmgx.so.

The structs below follow the C declarations field by field; the library is
loaded from this package directory (built by ``__graft_entry__.build()``).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmgx.so")

MGX_OK = 0
MGX_E_ARG = -1
MGX_E_HIP = -2
MGX_E_RAY_LENGTH = -3
MGX_E_NO_NEXT = -4
MGX_E_UNSUPPORTED = -5

APPROX_NONE, APPROX_ONE_LINE, APPROX_CLOSE_LINES = 0, 1, 2
BC_ZERO, BC_THERMAL, BC_CMB = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
MAX_OFF_DIAG = 4  # largest n_off_diag of mgx_feautrier_forward_lambda_device

_pd = C.POINTER(C.c_double)
_pu = C.POINTER(C.c_uint32)
_pi = C.POINTER(C.c_int32)
_pb = C.POINTER(C.c_uint8)


class MgxModel(C.Structure):
    _fields_ = [
        ("npoints", C.c_uint32), ("nrays", C.c_uint32), ("hnrays", C.c_uint32),
        ("nboundary", C.c_uint32),
        ("nlines", C.c_uint32), ("nlspecs", C.c_uint32), ("nquads", C.c_uint32),
        ("nfreqs", C.c_uint32),
        ("dimension", C.c_int32), ("spherical_symmetry", C.c_int32),
        ("adaptive_directions", C.c_int32), ("reserved0", C.c_int32),
        ("position", _pd), ("velocity", _pd),
        ("n_neighbors", _pu), ("cum_n_neighbors", _pu), ("neighbors", _pu),
        ("point2boundary", _pu), ("boundary2point", _pu),
        ("boundary_condition", _pi), ("boundary_temperature", _pd),
        ("ray_direction", _pd), ("ray_weight", _pd), ("antipod", _pu),
        ("center", C.c_double * 3),
        ("temperature", _pd), ("vturb2", _pd),
        ("line", _pd), ("sorted_line", _pd), ("sorted_line_map", _pu),
        ("emissivity", _pd), ("opacity", _pd), ("inverse_width", _pd),
        ("interp_crit", _pd), ("dshift_max", _pd),
        ("max_inverse_mass", C.c_double),
        ("line_species", _pu), ("line_k", _pu), ("line_A", _pd),
        ("species_inverse_mass", _pd), ("quad_weights", _pd),
        ("nu", _pd), ("f_line", _pu), ("f_species", _pu), ("f_k", _pu), ("f_z", _pu),
        ("f_in_line", _pb),
    ]


class MgxOptions(C.Structure):
    _fields_ = [
        ("approx", C.c_int32), ("n_off_diag", C.c_uint32), ("precision", C.c_int32),
        ("reserved0", C.c_int32),
        ("min_opacity", C.c_double), ("max_distance_opacity_contribution", C.c_double),
        ("max_interpolation_diff", C.c_double),
    ]


class MgxLinedata(C.Structure):
    _fields_ = [
        ("num", C.c_uint32), ("nlev", C.c_uint32), ("nrad", C.c_uint32),
        ("reserved0", C.c_uint32),
        ("inverse_mass", C.c_double),
        ("energy", _pd), ("weight", _pd), ("irad", _pu), ("jrad", _pu),
        ("A", _pd), ("Bs", _pd), ("Ba", _pd), ("frequency", _pd),
    ]


# (name, restype, argtypes) of every function include/mgx.h declares
_VP = C.c_void_p
class MgxColldata(C.Structure):
    _fields_ = [
        ("num_col_partner", C.c_uint32), ("orth_or_para_H2", C.c_int32),
        ("ntmp", C.c_uint32), ("ncol", C.c_uint32),
        ("icol", _pu), ("jcol", _pu), ("tmp", _pd), ("Cd", _pd), ("Ce", _pd),
    ]


class MgxNlteParams(C.Structure):
    _fields_ = [
        ("pop_prec", C.c_double), ("min_rel_pop_for_convergence", C.c_double),
        ("population_inversion_fraction", C.c_double), ("min_line_emissivity", C.c_double),
        ("min_line_opacity", C.c_double), ("convergence_fraction", C.c_double),
        ("use_adaptive_Ng_acceleration", C.c_uint32), ("Ng_acceleration_mem_limit", C.c_uint32),
        ("adaptive_Ng_acceleration_min_order", C.c_uint32),
        ("adaptive_Ng_acceleration_use_max_criterion", C.c_uint32),
        ("Ng_acceleration_remove_N_its", C.c_uint32), ("reserved0", C.c_uint32),
    ]


class MgxNlteConv(C.Structure):
    _fields_ = [("fraction_not_converged", C.c_double), ("relative_change_mean", C.c_double),
                ("relative_change_max", C.c_double)]


def default_nlte_params(**kw):
    """mgx_nlte_params with the reference's defaults (parameters.hpp:76-104), overridden by kw."""
    p = MgxNlteParams()
    p.pop_prec = 1.0e-6
    p.min_rel_pop_for_convergence = 1.0e-10
    p.population_inversion_fraction = -1.0
    p.min_line_emissivity = 1.0e-50
    p.min_line_opacity = 1.0e-13
    p.convergence_fraction = 0.995
    p.use_adaptive_Ng_acceleration = 1
    p.Ng_acceleration_mem_limit = 9
    p.adaptive_Ng_acceleration_min_order = 2
    p.adaptive_Ng_acceleration_use_max_criterion = 1
    p.Ng_acceleration_remove_N_its = 0
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


SIGNATURES = [
    ("mgx_last_error", C.c_char_p, []),
    ("mgx_version", C.c_char_p, []),
    ("mgx_feautrier_forward_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, _VP, _VP]),
    ("mgx_feautrier_sparse_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, _VP]),
    ("mgx_feautrier_forward_noalloc", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, _VP, _VP]),
    ("mgx_shortchar_order_0_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), C.c_double, _VP, _VP, _VP, _VP, _VP, _VP]),
    ("mgx_feautrier_order_2_uv_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, _VP]),
    ("mgx_trace_counts_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP]),
    ("mgx_trace_rays_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), C.c_uint32, _VP, _VP, _VP, C.c_uint32,
      _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    ("mgx_image_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_uint32, C.c_uint32, _VP, _VP, _VP,
      _VP]),
    ("mgx_feautrier_forward_lambda_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, C.c_uint32, _VP, _VP, _VP, _VP]),
    ("mgx_image_optical_depth_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_uint32, C.c_uint32, _VP, _VP, _VP,
      _VP]),
    ("mgx_image_ray_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), C.c_uint32, _VP, _VP, _VP, _VP]),
    ("mgx_image_optical_depth_ray_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), C.c_uint32, _VP, _VP, _VP, _VP]),
    ("mgx_feautrier_adjoint_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, _VP, _VP]),
    ("mgx_image_adjoint_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_uint32, C.c_uint32, _VP, _VP, _VP,
      _VP]),
    ("mgx_feautrier_adjoint_width_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP, _VP, _VP, _VP]),
    ("mgx_image_adjoint_width_device", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_uint32, C.c_uint32, _VP, _VP, _VP,
      _VP, _VP]),
    ("mgx_release_workspace", None, []),
    ("mgx_batch_create", C.c_int,
     [C.POINTER(MgxModel), C.c_uint32, C.POINTER(MgxOptions), _VP, _VP, C.POINTER(_VP), _VP]),
    ("mgx_batch_configure", C.c_int,
     [_VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _VP]),
    ("mgx_batch_tune", C.c_int, [_VP, C.c_uint32, _VP, _VP]),
    ("mgx_batch_set_images", C.c_int,
     [_VP, C.POINTER(MgxModel), _VP, C.c_uint32, C.c_uint32, _VP, _VP, _VP, _VP]),
    ("mgx_batch_run", C.c_int, [_VP, _VP]),
    ("mgx_batch_last_ms", C.c_double, [_VP]),
    ("mgx_batch_destroy", None, [_VP]),
    ("mgx_compute_inverse_widths", C.c_int,
     [C.c_uint32, C.c_uint32, C.POINTER(MgxLinedata), _VP, _VP, _VP]),
    ("mgx_compute_lte", C.c_int,
     [C.c_uint32, C.c_uint32, _VP, _VP, C.c_uint32, C.POINTER(MgxLinedata), C.c_double,
      C.c_double, _VP, _VP, _VP, _VP]),
    ("mgx_compute_spectral_discretisation", C.c_int,
     [C.c_uint32, C.c_uint32, C.POINTER(MgxLinedata), C.c_uint32, _VP, _VP, _VP, _VP, _VP,
      _VP, _VP, _VP, _VP]),
    ("mgx_compute_image_discretisation", C.c_int,
     [C.c_uint32, C.c_double, C.c_double, C.c_uint32, _VP, _VP, _VP, _VP, _VP, _VP]),
    ("mgx_compute_dshift_max", C.c_int,
     [C.c_uint32, C.c_uint32, C.POINTER(MgxLinedata), _VP, _VP, C.c_double, _VP]),
    ("mgx_compute_interp_crit", C.c_int,
     [C.c_uint32, C.c_uint32, C.POINTER(MgxLinedata), _VP, _VP, _VP]),
    ("mgx_sort_lines", C.c_int, [C.c_uint32, _VP, _VP, _VP]),
    ("mgx_nlte_default_params", None, [C.POINTER(MgxNlteParams)]),
    ("mgx_nlte_create", C.c_int,
     [C.POINTER(MgxModel), C.c_uint32, C.POINTER(MgxLinedata), _VP, C.POINTER(MgxColldata),
      C.c_uint32, _VP, C.POINTER(MgxNlteParams), C.POINTER(_VP), _VP]),
    ("mgx_nlte_destroy", None, [_VP]),
    ("mgx_nlte_set_lte", C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    ("mgx_nlte_set_populations", C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    ("mgx_nlte_get_populations", C.c_int, [_VP, _VP, _VP, _VP]),
    ("mgx_nlte_iteration_stateq", C.c_int,
     [_VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(MgxNlteConv), _VP, _VP, _VP]),
    ("mgx_nlte_iteration_stateq_global", C.c_int,
     [_VP, _VP, C.c_uint32, _VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(MgxNlteConv), _VP, _VP, _VP,
      _VP]),
    ("mgx_nlte_iteration_ng", C.c_int,
     [_VP, C.c_uint32, _VP, _VP, _VP, C.POINTER(MgxNlteConv), _VP]),
    ("mgx_nlte_trial_ng", C.c_int, [_VP, C.c_uint32, C.POINTER(MgxNlteConv), _VP]),
    ("mgx_nlte_history", C.c_int, [_VP, _VP]),
    ("mgx_fmath_digest", C.c_int, [C.c_int, C.c_float, C.c_uint64, C.c_uint32, C.c_uint32, _VP,
                                   _VP]),
    ("mgx_nlte_last_lambda", C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    ("mgx_compute_level_populations_sparse", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_int, C.c_int64, _VP, _VP, _VP,
      _VP, _VP, _VP, _VP]),
    ("mgx_compute_level_populations", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_int, C.c_int64, _VP, _VP, _VP,
      _VP, _VP, _VP, _VP]),
    ("mgx_compute_level_populations_shortchar", C.c_int,
     [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, C.c_int, C.c_int64, C.c_double, _VP, _VP,
      _VP, _VP, _VP, _VP, _VP]),
    ("mgx_trace_stats_device", C.c_int, [C.POINTER(MgxModel), C.POINTER(MgxOptions), _VP, _VP]),
    ("mgx_set_kernel_timing", None, [C.c_int]),
    ("mgx_last_kernel_ms", C.c_double, []),
]

_lib = None


def lib():
    """Load libmgx.so once (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        # torch carries its own HIP runtime; loading it first makes libmgx bind
        # to that same runtime (a second HIP runtime in one process sees no GPU)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


class MgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[mgx error {code}] {msg}")
        self.code = code


def check(code):
    if code != MGX_OK:
        msg = lib().mgx_last_error().decode(errors="replace")
        raise MgxError(code, msg)
    return code


def ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
