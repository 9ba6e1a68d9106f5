"""ctypes view of the C-ABI in include/mugpu.h (libmugpu.so).

This is the thin layer the Python package and the tests use to reach the
HIP library directly.  It never falls back to anything: if libmugpu.so is
missing, or the device cannot be initialised, the call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmugpu.so")

OK, ECONVERGENCE, ESOLVER, EPROJECTION, ERUNTIME, EMEMORY, EHIP, EARG = range(8)
FP64, FP32 = 0, 1
FINITE_STRAIN, SMALL_STRAIN = 0, 1
PROJ_SMALL_STRAIN, PROJ_FINITE_STRAIN_FAST, PROJ_FINITE_STRAIN = 0, 1, 2
LE1, LE2, LE3, LE4, HEP1, HEP1_CT, LEG = 1, 2, 3, 4, 5, 6, 7

class ServerConfig(ctypes.Structure):
    """mugpu_server_config of include/mugpu.h."""
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_batch", ctypes.c_int),
                ("iters_per_launch", ctypes.c_int), ("use_graphs", ctypes.c_int),
                ("keep_fields", ctypes.c_int), ("cg_tol", ctypes.c_double),
                ("cg_maxiter", ctypes.c_int), ("inner_tol", ctypes.c_double),
                ("newton_tol", ctypes.c_double), ("equil_tol", ctypes.c_double)]


# name, restype, argtypes
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_int = ctypes.c_int
_dbl = ctypes.c_double

SIGNATURES = [
    ("mugpu_init", _int, [_int]),
    ("mugpu_last_error", ctypes.c_char_p, []),
    ("mugpu_device_info", _int, [ctypes.c_char_p, _int, _ip,
                                 ctypes.POINTER(ctypes.c_size_t), _ip]),
    ("mugpu_alloc", _int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    ("mugpu_free", _int, [_vp]),
    ("mugpu_release_cached", _int, []),
    ("mugpu_upload", _int, [_vp, _vp, ctypes.c_size_t]),
    ("mugpu_download", _int, [_vp, _vp, ctypes.c_size_t]),
    ("mugpu_memset_zero", _int, [_vp, ctypes.c_size_t]),
    ("mugpu_synchronize", _int, []),
    ("mugpu_stream", _vp, []),
    ("mugpu_fft_create", _int, [_int, _ip, _int, _int, ctypes.POINTER(_vp)]),
    ("mugpu_fft_destroy", _int, [_vp]),
    ("mugpu_fft_r2c", _int, [_vp, _vp, _vp]),
    ("mugpu_fft_c2r", _int, [_vp, _vp, _vp]),
    ("mugpu_fft_normalisation", _dbl, [_vp]),
    ("mugpu_fft_nb_fourier_points", _i64, [_vp]),
    ("mugpu_projection_create", _int, [_vp, _int, _dp, _int, ctypes.POINTER(_vp)]),
    ("mugpu_projection_destroy", _int, [_vp]),
    ("mugpu_projection_apply", _int, [_vp, _vp]),
    ("mugpu_projection_get_operator", _int, [_vp, _dp]),
    ("mugpu_cell_create", _int, [_int, _ip, _dp, _int, _int, _int, ctypes.POINTER(_vp)]),
    ("mugpu_cell_destroy", _int, [_vp]),
    ("mugpu_cell_nb_pixels", _i64, [_vp]),
    ("mugpu_cell_nb_components", _int, [_vp]),
    ("mugpu_cell_add_material", _int, [_vp, _int, _dbl, _dbl, _dbl, _dbl, _ip]),
    ("mugpu_cell_add_material_generic", _int, [_vp, _dp, _ip]),
    ("mugpu_material_add_pixels", _int, [_vp, _int, _i64, _i64p, _dp]),
    ("mugpu_cell_initialise", _int, [_vp]),
    ("mugpu_cell_is_initialised", _int, [_vp]),
    ("mugpu_cell_strain", _int, [_vp, ctypes.POINTER(_vp)]),
    ("mugpu_cell_stress", _int, [_vp, ctypes.POINTER(_vp)]),
    ("mugpu_cell_set_uniform_strain", _int, [_vp, _dp]),
    ("mugpu_cell_add_uniform_strain", _int, [_vp, _dp]),
    ("mugpu_cell_evaluate_stress_tangent", _int, [_vp]),
    ("mugpu_cell_evaluate_stress", _int, [_vp]),
    ("mugpu_cell_get_tangent", _int, [_vp, _dp]),
    ("mugpu_cell_directional_stiffness", _int, [_vp, _vp, _vp]),
    ("mugpu_cell_tangent_product", _int, [_vp, _vp, _vp]),
    ("mugpu_cell_apply_projection", _int, [_vp, _vp]),
    ("mugpu_cell_save_history", _int, [_vp]),
    ("mugpu_cell_get_internal", _int, [_vp, ctypes.c_char_p, _dp, _ip]),
    ("mugpu_material_get_state", _int, [_vp, _int, ctypes.c_char_p, _int, _dp, _ip, _ip]),
    ("mugpu_cell_projection_operator", _int, [_vp, _dp]),
    ("mugpu_dot", _int, [_vp, _vp, _i64, _dp]),
    ("mugpu_axpby", _int, [_dbl, _vp, _dbl, _vp, _i64]),
    ("mugpu_scale_copy", _int, [_dbl, _vp, _vp, _i64]),
    ("mugpu_cg_start", _int, [_vp, _vp, _vp, _vp, _vp, _vp, _dp, _dp]),
    ("mugpu_cg_step", _int, [_vp, _int, _vp, _vp, _vp, _vp, _dp]),
    ("mugpu_cg_steps", _int, [_vp, _int, _vp, _vp, _vp, _vp, _dbl, _int, _ip, _dp]),
    ("mugpu_cg_finish", _int, [_vp]),
    ("mugpu_cg_graphs", _int, [_int]),
    ("mugpu_cg_solve_mixed", _int, [_vp, _vp, _vp, _dbl, _int, _dbl, _i64p, _ip, _dp]),
    ("mugpu_cg_solve_mixed_noalloc", _int, [_vp, _vp, _vp, _dbl, _int, _dbl, _i64p, _ip, _dp]),
    ("mugpu_timing_start", _int, [_int]),
    ("mugpu_timing_stop", _int, [_i64p, _dp]),
    ("mugpu_timing_class", _int, [_int, _i64p, _dp]),
    ("mugpu_fft_tune_pass", _int, [_int, _ip, _int, _int, _int, _int, _int, _dp]),
    ("mugpu_server_create", _int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    ("mugpu_server_destroy", _int, [_vp]),
    ("mugpu_server_footprint", _int, [_int, _ip, ctypes.POINTER(ctypes.c_size_t)]),
    ("mugpu_server_submit", _int, [_vp, _int, _ip, _dp, _dp, _dp, _dp, _int, _i64p]),
    ("mugpu_server_progress", _int, [_vp, _int, _ip]),
    ("mugpu_server_poll", _int, [_vp, _i64p, _int, _ip]),
    ("mugpu_server_result", _int, [_vp, _i64, _ip, _ip, _ip, _i64p, _dp, _dp, _dp]),
    ("mugpu_server_release", _int, [_vp, _i64]),
    ("mugpu_server_stats", _int, [_vp, _i64p, _int]),
    ("mugpu_copy", _int, [_vp, _vp, ctypes.c_size_t]),
    ("mugpu_copy_bandwidth", _int, [ctypes.c_size_t, _int, _dp]),
    ("mugpu_line_bandwidth", _int, [_int, _int, ctypes.c_int64, _int, _int, _int, _dp]),
    ("mugpu_debug_fail_allocations", _int, [_int]),
    ("mugpu_debug_force_variant", _int, [_int, _int, _int]),
]
NOALLOC_MAX_VALUES = 2796000  # MUGPU_NOALLOC_MAX_VALUES
K_LAST_R2C, K_MIDDLE, K_AXIS0, K_LAST_C2R, K_MATERIAL, K_CG_UPDATE = range(6)
K_ALL = 6

_LIB = None


def load(path=LIB_PATH):
    """Load libmugpu.so and declare every signature; raises if it is absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise ImportError(
                "libmugpu.so not found at {} — build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`".format(path))
        lib = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


class MugpuError(RuntimeError):
    KINDS = {ECONVERGENCE: "ConvergenceError", ESOLVER: "SolverError",
             EPROJECTION: "ProjectionError", ERUNTIME: "RuntimeError",
             EMEMORY: "MemoryError", EHIP: "HIPError", EARG: "ValueError"}

    def __init__(self, code, msg):
        super().__init__("{}: {}".format(self.KINDS.get(code, code), msg))
        self.code = code
        self.kind = self.KINDS.get(code, str(code))


def check(code):
    if code != OK:
        raise MugpuError(code, load().mugpu_last_error().decode(errors="replace"))
    return code
