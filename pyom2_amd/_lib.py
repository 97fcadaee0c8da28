"""ctypes binding of libpsisolve.so (declarations in include/psisolve.h).

The library is loaded from the package directory and nowhere else; if it is
missing the import fails loudly -- there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PSI_LIBRARY", os.path.join(HERE, "libpsisolve.so"))

PSI_OK = 0
PSI_NOT_CONVERGED = 1
PSI_DIVERGED = 2
PSI_NAN = 3
PSI_SINGULAR_ISLAND_SYSTEM = 4
PSI_BAD_ARG = 5
PSI_HIP_ERROR = 6
PSI_TRACE_FAILED = 7
PSI_NO_LEARNED_MODEL = 8
STATUS_NAMES = {0: "OK", 1: "NOT_CONVERGED", 2: "DIVERGED", 3: "NAN", 4: "SINGULAR_ISLAND_SYSTEM",
                5: "BAD_ARG", 6: "HIP_ERROR", 7: "TRACE_FAILED", 8: "NO_LEARNED_MODEL"}

PSI_MODE_PLAIN = 0
PSI_MODE_MIXED = 1
PSI_MODE_LEARNED = 2
PSI_MODE_POLY = 3
PSI_MODE_MG = 4
PSI_CRIT_REFERENCE = 0
PSI_CRIT_RELRES = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class PsiGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("cyclic_x", C.c_int32),
                ("dxt", _dp), ("dxu", _dp), ("dyt", _dp), ("dyu", _dp), ("cost", _dp),
                ("cosu", _dp)]


class PsiFields(C.Structure):
    _fields_ = [("hur", C.c_void_p), ("hvr", C.c_void_p), ("maskU", C.c_void_p),
                ("maskV", C.c_void_p)]


class PsiIslands(C.Structure):
    _fields_ = [("nisle", C.c_int32), ("offsets", _ip), ("bi", _ip), ("bj", _ip), ("di", _ip),
                ("dj", _ip)]


class PsiOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("criterion", C.c_int32), ("congr_epsilon", C.c_double),
                ("max_iterations", C.c_int32), ("relres_target", C.c_double),
                ("check_every", C.c_int32), ("inner_reduction", C.c_double),
                ("max_outer", C.c_int32), ("poly_degree", C.c_int32), ("poly_lower", C.c_double),
                ("leg_max_iterations", C.c_int32), ("time_kernels", C.c_int32),
                ("mg_coarse_degree", C.c_int32), ("reserved", C.c_int32)]


class PsiInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("status", C.c_int32),
                ("outer_iterations", C.c_int32), ("est_error", C.c_double),
                ("relres_recursive", C.c_double), ("relres_true", C.c_double),
                ("elapsed_ms", C.c_double), ("sweep_ms", C.c_double),
                ("sweep_launches", C.c_int32), ("setup_launches", C.c_int32),
                ("kernel_samples", C.c_int32), ("kernel_us", C.c_double * 4)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["kernel_us"] = list(self.kernel_us)
        return d


class PsiBatchMember(C.Structure):
    _fields_ = [("grid", PsiGrid), ("hur", C.c_void_p), ("hvr", C.c_void_p),
                ("islands", PsiIslands)]


class PsiBatchTune(C.Structure):
    _fields_ = [("tile_x", C.c_int32), ("tile_y", C.c_int32), ("tiles_per_item", C.c_int32),
                ("items_per_grab", C.c_int32), ("groups_per_cu", C.c_int32),
                ("graph_launches", C.c_int32), ("ns_per_point", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PsiBatchStats(C.Structure):
    _fields_ = [("elapsed_ms", C.c_double), ("launches", C.c_int32),
                ("graph_replays", C.c_int32), ("polls", C.c_int32), ("reserved", C.c_int32),
                ("active_points", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class PsiLevels(C.Structure):
    _fields_ = [("nz", C.c_int32), ("reserved", C.c_int32), ("dzt", C.c_void_p),
                ("dzw", C.c_void_p), ("maskT", C.c_void_p), ("maskU", C.c_void_p),
                ("maskV", C.c_void_p)]


class PsiMomentum(C.Structure):
    _fields_ = [("rho", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("du", C.c_void_p),
                ("dv", C.c_void_p), ("du_mix", C.c_void_p), ("dv_mix", C.c_void_p),
                ("p_hydro", C.c_void_p)]


class PsiPressFields(C.Structure):
    _fields_ = [("hu", C.c_void_p), ("hv", C.c_void_p), ("maskT", C.c_void_p),
                ("free_surface", C.c_int32), ("reserved", C.c_int32), ("grav", C.c_double),
                ("dt_mom", C.c_double)]


# (name, restype, argtypes) of every entry point in include/*.h
SIGNATURES = [
    ("psi_islands_trace", C.c_int, [C.c_int32, C.c_int32, C.c_int32, _ip, _ip, _ip, C.c_int32, _ip,
                                    _ip, _ip, _ip, C.c_int32, _ip]),
    ("psi_ctx_create", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PsiGrid), C.POINTER(PsiFields),
                                 C.POINTER(PsiIslands), C.c_void_p]),
    ("psi_ctx_update_coeffs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_ctx_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_ctx_destroy", None, [C.c_void_p]),
    ("psi_ctx_last_error", C.c_char_p, [C.c_void_p]),
    ("psi_ctx_nisle", C.c_int, [C.c_void_p]),
    ("psi_ctx_device_bytes", C.c_int64, [C.c_void_p]),
    ("psi_congrad", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.c_void_p, C.c_void_p,
                              C.POINTER(PsiInfo)]),
    ("psi_congrad_noalloc", C.c_int, [C.POINTER(PsiGrid), C.POINTER(PsiFields),
                                      C.POINTER(PsiIslands), C.POINTER(PsiOpts), C.c_void_p,
                                      C.c_void_p, C.POINTER(PsiInfo), C.c_void_p]),
    ("psi_apply_op", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_make_coeff", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_make_inv", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_dot", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    ("psi_absmax", C.c_int, [C.c_void_p, C.c_void_p, _dp]),
    ("psi_setcyclic", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_line_integral", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, _dp]),
    ("psi_psi_velocity", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_true_relres", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    ("psi_streamfunction_init", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.c_void_p, _dp, _ip]),
    ("psi_solve_streamfunction", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, _dp, C.c_void_p, _dp, C.c_int32,
                                           C.c_int32, C.c_int32, C.c_double, C.c_double,
                                           C.c_void_p, C.POINTER(PsiInfo)]),
    ("psi_streamfunction_forcing", C.c_int, [C.c_void_p, C.POINTER(PsiLevels), C.c_double,
                                             C.c_double] + [C.c_void_p] * 8),
    ("psi_velocity_update", C.c_int, [C.c_void_p, C.POINTER(PsiLevels), C.c_double, C.c_double]
     + [C.c_void_p] * 11),
    ("psi_solve_streamfunction_3d", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.POINTER(PsiLevels),
                                              C.POINTER(PsiMomentum), C.c_double, C.c_double,
                                              C.c_void_p, C.c_void_p, _dp, C.c_void_p, _dp,
                                              C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                              C.c_double, C.c_void_p, C.POINTER(PsiInfo)]),
    ("psi_fcg_prepare", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_fcg_begin", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    ("psi_fcg_input_scale", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p)]),
    ("psi_fcg_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    ("psi_fcg_step_mg", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    ("psi_fcg_read", C.c_int, [C.c_void_p, _dp]),
    ("psi_fcg_flush", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    ("psi_fcg_scalars", C.c_int, [C.c_void_p, _dp]),
    ("psi_poly_coefficients", C.c_int, [C.c_int32, C.c_double, C.POINTER(C.c_float)]),
    ("psi_mg_setup", C.c_int, [C.c_void_p]),
    ("psi_mg_levels", C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip, C.c_int32]),
    ("psi_mg_kernel_times", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int32]),
    # the surface-pressure solver
    ("psi_press_create", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PsiGrid),
                                   C.POINTER(PsiPressFields), C.c_void_p]),
    ("psi_press_destroy", None, [C.c_void_p]),
    ("psi_press_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_press_last_error", C.c_char_p, [C.c_void_p]),
    ("psi_press_device_bytes", C.c_int64, [C.c_void_p]),
    ("psi_make_coeff_surf_press", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_press_apply_op", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_congrad_surf_press", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.c_void_p, C.c_void_p,
                                         C.POINTER(PsiInfo)]),
    ("psi_solve_pressure", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.POINTER(PsiLevels),
                                     C.POINTER(PsiMomentum), C.c_double, C.c_double, C.c_void_p,
                                     C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                     C.c_void_p, C.POINTER(PsiInfo)]),
    ("psi_press_kernel_times", C.c_int, [C.c_void_p, C.c_int32, _dp, C.c_int32]),
    ("psi_press_mg_setup", C.c_int, [C.c_void_p]),
    ("psi_press_mg_kernel_times", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int32]),
    # the non-hydrostatic pressure solver
    ("psi_nh_create", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PsiGrid), C.POINTER(PsiLevels),
                                C.c_void_p, C.c_void_p]),
    ("psi_nh_destroy", None, [C.c_void_p]),
    ("psi_nh_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_nh_last_error", C.c_char_p, [C.c_void_p]),
    ("psi_nh_device_bytes", C.c_int64, [C.c_void_p]),
    ("psi_make_coeff_non_hydro", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psi_apply_op_3d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_congrad_non_hydro", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.c_void_p, C.c_void_p,
                                        C.POINTER(PsiInfo)]),
    ("psi_solve_non_hydrostatic", C.c_int, [C.c_void_p, C.POINTER(PsiOpts)] + [C.c_void_p] * 6 +
     [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_void_p,
      C.POINTER(PsiInfo)]),
    ("psi_nh_kernel_times", C.c_int, [C.c_void_p, C.c_int32, _dp, C.c_int32]),
    ("psi_nn_anorm_loss", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psi_nn_weighted_dot", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    # batched ensembles (include/psibatch.h)
    ("psi_batch_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(PsiBatchMember),
                                   C.c_void_p]),
    ("psi_batch_destroy", None, [C.c_void_p]),
    ("psi_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("psi_batch_device_bytes", C.c_int64, [C.c_void_p]),
    ("psi_batch_active_points", C.c_int64, [C.c_void_p]),
    ("psi_batch_max_groups_per_cu", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("psi_batch_get_tune", C.c_int, [C.c_void_p, C.POINTER(PsiBatchTune)]),
    ("psi_batch_set_tune", C.c_int, [C.c_void_p, C.POINTER(PsiBatchTune)]),
    ("psi_batch_autotune", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(PsiBatchTune), _dp,
                                     C.c_int32, _ip]),
    ("psi_batch_congrad", C.c_int, [C.c_void_p, C.POINTER(PsiOpts), C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_void_p), C.POINTER(PsiInfo), _dp,
                                    C.POINTER(PsiBatchStats)]),
    # learned-preconditioner network kernels (include/psinn.h)
    ("psinn_mfma_probe", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psinn_conv3x3", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                C.c_void_p, C.c_int, C.c_void_p]),
    ("psinn_conv3x3_wgrad", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    ("psinn_pack_weights", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p]),
    ("psinn_avgpool2", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p]),
    ("psinn_sumpool2", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_void_p, C.c_void_p]),
    ("psinn_grad_combine", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                     C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("psinn_expand16", C.c_int, [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]),
    ("psinn_adam", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_float,
                             C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, C.c_void_p]),
    ("psinn_net_input", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
]

_lib = None


def lib():
    """Load libpsisolve.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pyom2_amd.build` "
                              "(the HIP extension is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class PsiError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
