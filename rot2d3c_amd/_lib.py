"""ctypes binding of librot2d3c.so (the C ABI declared in include/rot2d3c.h).

There is no fallback: if the library is missing the import of the operator
modules fails with an error that says how to build it.
"""
import ctypes
import os

LIB_PATH = os.environ.get("R2D3C_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "librot2d3c.so")

R2D3C_DT_FIXED = 0
R2D3C_DT_CFL = 1
R2D3C_NKERNELS = 4
KERNEL_NAMES = ("k_step_colinv", "k_step_row", "k_step_colfused", "k_step_colfused_last")

_c_int, _c_double, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p


class Params(ctypes.Structure):
    """r2d3c_params: parameter.inp lines 12-20 (src/Rot2D3C.para.f90:147-155)."""
    _fields_ = [("nu", _c_double), ("hnu", _c_double), ("nuv", _c_double),
                ("hnuv", _c_double), ("omega", _c_double), ("nn", _c_int),
                ("mm", _c_int), ("nnv", _c_int), ("mmv", _c_int)]


class ForceDraw(ctypes.Structure):
    _fields_ = [("kx", _c_double), ("ky", _c_double), ("phase", _c_double),
                ("f0", _c_double), ("kupk", _c_double), ("enabled", _c_int)]


class Forcing(ctypes.Structure):
    _fields_ = [("psi", ForceDraw), ("vz", ForceDraw), ("dense_psi", _vp), ("dense_vz", _vp),
                ("ci_enabled", _c_int), ("ci_kdn", _c_double), ("ci_kup", _c_double),
                ("ci_fp0", _c_double), ("ci_fz0", _c_double), ("ci_nprocs", _c_int)]


class Member(ctypes.Structure):
    """r2d3c_member: one run of an ensemble."""
    _fields_ = [("n", _c_int), ("params", Params), ("dt", _c_double), ("psi", _vp), ("vz", _vp)]


# name -> (restype, argtypes); every symbol declared in include/rot2d3c.h
SIGNATURES = {
    "r2d3c_plan_create": (_c_int, [ctypes.POINTER(_vp), _c_int]),
    "r2d3c_plan_destroy": (_c_int, [_vp]),
    "r2d3c_plan_workspace_bytes": (ctypes.c_longlong, [_vp]),
    "r2d3c_plan_size": (_c_int, [_vp]),
    "r2d3c_last_error": (ctypes.c_char_p, []),
    "r2d3c_c2r": (_c_int, [_vp, _vp, _vp, _vp]),
    "r2d3c_r2c": (_c_int, [_vp, _vp, _vp, _vp]),
    "r2d3c_derivk2": (_c_int, [_vp, _vp, _vp, _c_int, _vp]),
    "r2d3c_laplak2": (_c_int, [_vp, _vp, _vp, _vp]),
    "r2d3c_inv_laplak2": (_c_int, [_vp, _vp, _vp, _vp]),
    "r2d3c_poisson": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "r2d3c_energy": (_c_int, [_vp, _vp, _c_int, _vp, _vp]),
    "r2d3c_inerprod": (_c_int, [_vp, _vp, _vp, _c_int, _vp, _vp]),
    "r2d3c_const_inj": (_c_int, [_vp, _vp, _c_double, _c_double, _c_double, _c_int, _c_int, _vp,
                                 _vp]),
    "r2d3c_const_inj_last": (_c_int, [_vp, _vp, _vp, _vp]),
    "r2d3c_cond_check": (_c_int, [_vp, ctypes.POINTER(Params), _c_double, _vp, _vp, _vp, _vp,
                                  _c_int, _vp, _vp]),
    "r2d3c_spectrum": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "r2d3c_transfers": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "r2d3c_cfl": (_c_int, [_vp, ctypes.POINTER(Params), _c_double, _vp, _vp, _vp]),
    "r2d3c_step": (_c_int, [_vp, ctypes.POINTER(Params), _vp, _vp, ctypes.POINTER(Forcing),
                            _c_int, _c_double, _vp, _vp]),
    "r2d3c_run": (_c_int, [_vp, ctypes.POINTER(Params), _vp, _vp, ctypes.POINTER(Forcing),
                           _c_int, _c_int, _c_double, _vp, _vp]),
    "r2d3c_run_noalloc": (_c_int, [_c_int, ctypes.POINTER(Params), _vp, _vp,
                                   ctypes.POINTER(Forcing), _c_int, _c_int, _c_double, _vp, _vp]),
    "r2d3c_forcing_fill": (_c_int, [_vp, ctypes.POINTER(Forcing), _c_double, _vp, _vp, _vp]),
    "r2d3c_timing_enable": (_c_int, [_vp, _c_int]),
    "r2d3c_timing_read": (_c_int, [_vp, ctypes.POINTER(_c_double), ctypes.POINTER(ctypes.c_longlong)]),
    "r2d3c_copy_probe": (_c_int, [_vp, _vp, ctypes.c_longlong, _vp]),
    "r2d3c_plan_set_row_split": (_c_int, [_vp, _c_int]),
    "r2d3c_plan_get_row_split": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "r2d3c_plan_set_bracket_pass": (_c_int, [_vp, _c_int]),
    "r2d3c_plan_get_bracket_pass": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "r2d3c_ensemble_set_concurrent": (_c_int, [_vp, _c_int]),
    "r2d3c_ensemble_set_stage_launch": (_c_int, [_vp, _c_int]),
    "r2d3c_ensemble_get_stage_launch": (_c_int, [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    "r2d3c_ensemble_sync": (_c_int, [_vp]),
    "r2d3c_ensemble_set_row_split": (_c_int, [_vp, _c_int, _c_int]),
    "r2d3c_plan_set_schedule": (_c_int, [_vp, _c_int, _c_int]),
    "r2d3c_plan_get_schedule": (_c_int, [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    "r2d3c_autotune": (_c_int, [_vp, _c_int, ctypes.POINTER(_c_double), _vp]),
    "r2d3c_ensemble_create": (_c_int, [ctypes.POINTER(_vp), ctypes.POINTER(Member), _c_int]),
    "r2d3c_ensemble_destroy": (_c_int, [_vp]),
    "r2d3c_ensemble_set_schedule": (_c_int, [_vp, _c_int, _c_int, _c_int]),
    "r2d3c_ensemble_set_column_radix": (_c_int, [_vp, _c_int, _c_int]),
    "r2d3c_plan_set_column_radix": (_c_int, [_vp, _c_int]),
    "r2d3c_plan_get_column_radix": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "r2d3c_plan_set_column_split": (_c_int, [_vp, _c_int]),
    "r2d3c_plan_get_column_split": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "r2d3c_ensemble_set_column_split": (_c_int, [_vp, _c_int, _c_int]),
    "r2d3c_ensemble_run": (_c_int, [_vp, ctypes.POINTER(Forcing), _c_int, _vp]),
    "r2d3c_ensemble_set_closure_period": (_c_int, [_vp, _c_int]),
    "r2d3c_ensemble_set_closure": (_c_int, [_vp, _c_int, _vp, ctypes.POINTER(_c_double),
                                            ctypes.POINTER(_c_double), _c_double, _c_double]),
    "r2d3c_filter": (_c_int, [_vp, _vp, _vp, _vp, _c_double, _vp]),
    "r2d3c_closure_inputs": (_c_int, [_vp, _vp, _vp, ctypes.POINTER(_c_double), _vp, _vp]),
    "r2d3c_closure_targets": (_c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_c_double),
                                       _c_double, _vp, _vp]),
    "r2d3c_closure_forcing": (_c_int, [_vp, _vp, ctypes.POINTER(_c_double), _c_double, _vp, _vp,
                                       _vp]),
    "r2d3c_cnn_param_count": (_c_int, []),
    "r2d3c_cnn_layer_offsets": (_c_int, [ctypes.POINTER(_c_int)] * 4),
    "r2d3c_cnn_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int]),
    "r2d3c_cnn_destroy": (_c_int, [_vp]),
    "r2d3c_cnn_set_params": (_c_int, [_vp, _vp, _vp]),
    "r2d3c_cnn_get_params": (_c_int, [_vp, _vp, _vp]),
    "r2d3c_cnn_forward": (_c_int, [_vp, _vp, _c_int, _vp, _vp]),
    "r2d3c_cnn_gradients": (_c_int, [_vp, _vp, _vp, _c_int, _vp, _vp, _vp]),
    "r2d3c_cnn_train_step": (_c_int, [_vp, _vp, _vp, _c_int, ctypes.c_float, _vp, _vp]),
    "r2d3c_cnn_buffers": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "r2d3c_cnn_shape": (_c_int, [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                 ctypes.POINTER(_c_int)]),
    "r2d3c_cnn_forward_buffers": (_c_int, [_vp, _c_int, _vp]),
    "r2d3c_cnn_set_output_clamp": (_c_int, [_vp, _c_double]),
    "r2d3c_cnn_set_forcing_gain": (_c_int, [_vp, ctypes.POINTER(_c_double), _c_int]),
    "r2d3c_cnn_set_sign_clip": (_c_int, [_vp, _c_double]),
    "r2d3c_closure_net_inputs": (_c_int, [_vp, _vp, _vp, ctypes.POINTER(_c_double), _vp, _vp]),
    "r2d3c_closure_net_forcing": (_c_int, [_vp, _vp, ctypes.POINTER(_c_double), _c_double, _vp,
                                           _vp, _vp]),
    "r2d3c_closure_clip": (_c_int, [_vp, _vp, _vp, _c_double, _vp, _vp, _vp]),
    "r2d3c_randu": (ctypes.c_float, [ctypes.POINTER(_c_int)]),
    "r2d3c_rand_force_draw": (None, [_c_double, ctypes.POINTER(_c_int),
                                     ctypes.POINTER(_c_double), ctypes.POINTER(_c_double),
                                     ctypes.POINTER(_c_double)]),
}

_lib = None


def lib():
    """Load librot2d3c.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "librot2d3c.so is not built (expected at %s); run "
                "`python -c 'import __graft_entry__ as g; g.build()'` or "
                "`python rot2d3c_amd/build.py`" % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class R2D3CError(RuntimeError):
    pass


def check(rc, what):
    if rc != 0:
        msg = lib().r2d3c_last_error()
        raise R2D3CError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))
