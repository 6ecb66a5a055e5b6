"""ctypes binding of libfizzy.so, the MI355X-native hot path.

The library is the product: there is no Python or CPU fallback. Importing
this module does not need a GPU, but `lib()` raises if the shared library is
missing, and every entry point raises on a non-zero status.
"""
import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libfizzy.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "fizzy.h")

FZ_OK = 0

dp = C.POINTER(C.c_double)
vp = C.c_void_p


class FzParams(C.Structure):
    """fz_params of include/fizzy.h (the namelist entries the hot path reads)."""

    _fields_ = [
        ("ng", C.c_int * 3), ("l", C.c_double * 3), ("gtype", C.c_int), ("gr", C.c_double),
        ("cfl", C.c_double), ("dtmax", C.c_double), ("dt_f", C.c_double),
        ("is_solve_ns", C.c_int), ("is_track_interface", C.c_int), ("is_forced_hit", C.c_int),
        ("icheck", C.c_int), ("vof", C.c_int),
        ("cbcvel", C.c_char * 18), ("cbcpre", C.c_char * 6), ("cbcpsi", C.c_char * 6), ("cbcnor", C.c_char * 18),
        ("bcvel", C.c_double * 18), ("bcpre", C.c_double * 6), ("bcpsi", C.c_double * 6), ("bcnor", C.c_double * 18),
        ("bforce", C.c_double * 3), ("gacc", C.c_double * 3), ("sigma", C.c_double),
        ("rho12", C.c_double * 2), ("mu12", C.c_double * 2), ("beta12", C.c_double * 2),
        ("psi_thickness_factor", C.c_double), ("acdi_gam_factor", C.c_double), ("acdi_gam_min", C.c_double),
        ("nspheres", C.c_int), ("inipsi", C.c_int), ("spheres", C.c_double * 256),
    ]


# name -> (restype, argtypes); every declaration of include/fizzy.h
SIGNATURES = {
    "fz_last_error": (C.c_char_p, []),
    "fz_ctx_create": (C.c_int, [C.POINTER(vp), C.POINTER(C.c_int), dp, dp, dp, dp, dp, C.c_char_p, vp]),
    "fz_ctx_destroy": (C.c_int, [vp]),
    "fz_ctx_sync": (C.c_int, [vp]),
    "fz_malloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
    "fz_free": (C.c_int, [vp, vp]),
    "fz_upload": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "fz_download": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "fz_copy": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "fz_prof_enable": (C.c_int, [vp, C.c_int]),
    "fz_prof_report": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
    "fz_initgrid": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, dp, dp, dp, dp]),
    "fz_initsolver": (C.c_int, [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
    "fz_solver": (C.c_int, [vp, vp]),
    "fz_solver_mixed": (C.c_int, [vp, vp, C.c_double, C.c_int, C.POINTER(C.c_int), dp]),
    "fz_solver_batched": (C.c_int, [C.POINTER(vp), C.POINTER(vp), C.c_int]),
    "fz_r2r": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "fz_fillps": (C.c_int, [vp, C.c_double, C.c_double, vp, vp, vp, vp]),
    "fz_updt_rhs_b": (C.c_int, [vp, vp]),
    "fz_correc": (C.c_int, [vp, C.c_double, dp, C.c_double, vp, vp, vp, vp, vp]),
    "fz_updatep": (C.c_int, [vp, vp, vp]),
    "fz_chk_poisson": (C.c_int, [vp, vp, vp, dp]),
    "fz_boundp": (C.c_int, [vp, C.c_char_p, dp, vp]),
    "fz_bounduvw": (C.c_int, [vp, C.c_char_p, dp, C.c_int, vp, vp, vp]),
    "fz_acdi_transport_pf": (C.c_int, [vp, C.c_double, C.c_double] + [vp] * 12),
    "fz_acdi_cmpt_phi": (C.c_int, [vp, C.c_double, vp, vp]),
    "fz_acdi_set_gamma": (C.c_int, [vp, C.c_double, vp, vp, vp, dp]),
    "fz_acdi_set_epsilon": (C.c_int, [vp, C.c_double, dp]),
    "fz_vof_thinc_transport_psi": (C.c_int, [vp, C.c_double] + [vp] * 12),
    "fz_vof_thinc_cmpt_phi": (C.c_int, [vp, C.c_double, vp, vp]),
    "fz_clip_field": (C.c_int, [vp, dp, vp]),
    "fz_cmpt_norm_curv_youngs": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "fz_init2fl": (C.c_int, [vp, C.POINTER(FzParams), C.c_double, vp]),
    "fz_mom_xyz_ad": (C.c_int, [vp, dp, dp] + [vp] * 10),
    "fz_mom_xyz_oth": (C.c_int, [vp, C.c_double, dp, dp, dp, dp, C.c_double, C.c_double, dp] + [vp] * 8),
    "fz_lscale_forcing": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, vp, vp, vp]),
    "fz_bulk_mean_12_stag": (C.c_int, [vp, C.c_int, vp, dp, dp]),
    "fz_rk_state_create": (C.c_int, [vp, C.POINTER(vp)]),
    "fz_rk_state_destroy": (C.c_int, [vp]),
    "fz_rk": (C.c_int, [vp, vp, dp, C.c_double, C.c_double, dp, dp, C.c_double, dp, dp, dp, dp, C.c_double]
              + [vp] * 12),
    "fz_rk_2fl": (C.c_int, [vp, vp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int] + [vp] * 11),
    "fz_chkdiv": (C.c_int, [vp, vp, vp, vp, dp, dp]),
    "fz_chkdiv_noalloc": (C.c_int, [vp, vp, vp, vp, vp]),
    "fz_math_eval": (C.c_int, [vp, C.c_int, C.c_long, vp, vp, vp]),
    "fz_phase_geometry_eval": (C.c_int, [vp, C.c_double] + [vp] * 7),
    "fz_chkdt": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, dp, vp, vp, vp, C.c_double,
                           C.c_double, dp]),
    "fz_case_create": (C.c_int, [C.POINTER(vp), C.POINTER(FzParams), vp]),
    "fz_case_create_ex": (C.c_int, [C.POINTER(vp), C.POINTER(FzParams), vp, C.c_int, C.c_int]),
    "fz_case_precision": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fz_case_field_f32": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_size_t)]),
    "fz_case_set_graphs": (C.c_int, [vp, C.c_int]),
    "fz_case_autotune": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_size_t]),
    "fz_case_destroy": (C.c_int, [vp]),
    "fz_case_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
    "fz_case_batch_step": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
    "fz_case_batch_step_begin": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "fz_case_batch_step_end": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "fz_case_batch_destroy": (C.c_int, [vp]),
    "fz_case_ctx": (C.c_int, [vp, C.POINTER(vp)]),
    "fz_case_field": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_size_t)]),
    "fz_case_get_scalars": (C.c_int, [vp, dp]),
    "fz_case_set_scalars": (C.c_int, [vp, dp]),
    "fz_case_init": (C.c_int, [vp]),
    "fz_case_step": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int)]),
    "fz_case_step_begin": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "fz_case_step_end": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "fz_case_save": (C.c_int, [vp, C.c_char_p]),
    "fz_case_restart": (C.c_int, [vp, C.c_char_p]),
    "fz_load_one": (C.c_int, [vp, C.c_char, C.c_char_p, vp, dp, C.POINTER(C.c_int)]),
    "fz_out3d": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_int), vp]),
    "fz_write_grid": (C.c_int, [vp, C.c_char_p]),
    "fz_case_footprint": (C.c_int, [C.POINTER(FzParams), C.POINTER(C.c_size_t)]),
    "fz_case_footprint_ex": (C.c_int, [C.POINTER(FzParams), C.c_int, C.POINTER(C.c_size_t)]),
    "fz_device_memory": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "fz_server_create": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "fz_server_destroy": (C.c_int, [vp]),
    "fz_server_submit": (C.c_int, [vp, C.POINTER(FzParams), C.c_int, C.POINTER(C.c_int)]),
    "fz_server_submit_ex": (C.c_int, [vp, C.POINTER(FzParams), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "fz_server_round": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fz_server_status": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fz_server_error": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_size_t]),
    "fz_server_case": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
    "fz_server_release": (C.c_int, [vp, C.c_int]),
    "fz_server_step_times": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    "fz_server_memory": (C.c_int, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
}

_lib = None


class FizzyError(RuntimeError):
    pass


def lib():
    """Load libfizzy.so (raises FizzyError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FizzyError(f"libfizzy.so not found at {LIB_PATH}: build it (python __graft_entry__.py build "
                             "or make -C cans-fizzy_amd)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != FZ_OK:
        msg = lib().fz_last_error().decode(errors="replace")
        raise FizzyError(f"{what or 'libfizzy'} failed with status {rc}: {msg}")
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)


def declared_symbols():
    """Function names declared in include/fizzy.h (parsed from the header)."""
    import re
    txt = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(fz_[a-z0-9_]+)\s*\(", txt)))
