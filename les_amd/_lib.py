"""ctypes binding of les_amd/libles_amd.so (the C ABI in include/les_amd.h).

The product path has no CPU fallback: if the HIP library is missing or a GPU
is absent, using it raises.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libles_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "les_amd.h")

_lib = None


class Grid(C.Structure):
    """les_grid (param.f90:56-65,100,144-145,219; bottombc.f90:53)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("dt", C.c_double), ("z0", C.c_double), ("force", C.c_double),
                ("inilag", C.c_int), ("initu", C.c_int), ("cs_count", C.c_int)]


SGS_NAMES = ("F_LM", "F_MM", "F_QN", "F_NN", "Beta", "Cs_opt2", "Nu_t", "u_lag", "v_lag", "w_lag",
             "TnLMMM", "TnQNNN", "L11t", "L22t", "L33t", "Q11t", "Q22t", "Q33t")


class SgsFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in SGS_NAMES] + [("Beta_avg", C.c_void_p), ("Betaclip_avg", C.c_void_p)]


class ScalarParams(C.Structure):
    """les_scalar_params (param.f90:53,72-90,139-146,174; bottombc.f90:26-63)."""
    _fields_ = [("enable", C.c_int), ("model", C.c_int), ("initsc", C.c_int), ("passive", C.c_int),
                ("Pr", C.c_double), ("z_i", C.c_double), ("u_star", C.c_double), ("T_scale", C.c_double),
                ("q_scale", C.c_double), ("g", C.c_double), ("dTdz_top", C.c_double), ("dqdz_top", C.c_double)]


SCALAR_SGS_NAMES = ("G_LM", "G_MM", "G_QN", "G_NN", "u_lag_s", "v_lag_s", "w_lag_s", "Pr_t", "Cs2_Pr")


class ScalarSgsFields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in SCALAR_SGS_NAMES] + [("phi_h", C.c_void_p), ("psi_m", C.c_void_p)]


class ParamNml(C.Structure):
    """les_param_nml (param.f90:63-85,174-230)."""
    _fields_ = [(n, C.c_double) for n in ("q_scale", "T_scale", "u_star", "ug0", "vg0", "inv_strength",
                                          "inv_strength_q", "dTdz_top", "dqdz_top", "T_init", "q_init", "dt_dim",
                                          "z_i", "dt")] + \
               [(n, C.c_int) for n in ("coriolis_forcing", "nsteps", "c_count", "p_count", "ubc", "damping_method",
                                       "initu", "initsc", "inilag", "jt_total")]


class Forcing(C.Structure):
    """les_forcing (param.f90:68-75, topbc.f90)."""
    _fields_ = [("coriolis_forcing", C.c_int), ("ug0", C.c_double), ("vg0", C.c_double), ("u_star", C.c_double),
                ("z_i", C.c_double), ("ubc", C.c_int), ("damping_method", C.c_int)]


def declared_symbols():
    """Names of the functions include/les_amd.h declares."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(les_[a-z0-9_]+)\s*\(", txt)))


def load():
    """Load the HIP library (does not touch the GPU)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"les_amd: {LIB_PATH} is not built; run __graft_entry__.build() "
                           "(the HIP path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.les_create.argtypes = [C.POINTER(Grid), C.POINTER(vp)]
    L.les_create.restype = C.c_int
    L.les_destroy.argtypes = [vp]
    L.les_destroy.restype = None
    L.les_last_error.argtypes = [vp]
    L.les_last_error.restype = C.c_char_p
    L.les_state_field.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.les_state_field.restype = vp
    L.les_state_init_ic.argtypes = [vp, C.c_uint64, C.c_double, C.c_double, C.c_double]
    L.les_test_filter_init.argtypes = [vp, C.c_double, vp]
    L.les_set_scalars.argtypes = [vp, C.POINTER(ScalarParams), vp, vp]
    L.les_state_init_scalar_ic.argtypes = [vp, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double]
    i64, u64, fp = C.c_int64, C.c_uint64, C.c_void_p
    L.les_mlp_init.argtypes = [vp, u64]
    L.les_mlp_fit_norm.argtypes = [vp, fp, fp, i64]
    L.les_mlp_forward.argtypes = [vp, fp, i64, fp]
    L.les_mlp_grad.argtypes = [vp, fp, fp, i64, fp, fp]
    L.les_mlp_train.argtypes = [vp, fp, fp, i64, C.c_int, C.c_int, C.c_float, u64, fp]
    L.les_set_closure.argtypes = [vp, C.c_int, C.c_int, C.c_double]
    L.les_ensemble_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.les_ensemble_destroy.argtypes = [vp]
    L.les_ensemble_destroy.restype = None
    L.les_ensemble_last_error.argtypes = [vp]
    L.les_ensemble_last_error.restype = C.c_char_p
    for nm in ("les_ensemble_window", "les_ensemble_step", "les_ensemble_sync"):
        getattr(L, nm).argtypes = [vp] + ([C.c_int] if nm == "les_ensemble_step" else [])
    u64p = C.POINTER(C.c_uint64)
    L.les_record_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(vp), u64p]
    L.les_record_read.argtypes = [C.c_char_p, C.c_int, vp, C.c_uint64, u64p]
    L.les_state_checkpoint.argtypes = [vp, C.c_char_p]
    L.les_state_read_initial.argtypes = [vp, C.c_char_p, C.c_int, C.c_int, C.c_int]
    L.les_read_total_time.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
    L.les_write_total_time.argtypes = [C.c_char_p, C.c_int]
    L.les_read_surface.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, vp, vp, vp]
    L.les_read_param_nml.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(ParamNml)]
    L.les_read_nml_value.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_double)]
    L.les_set_forcing.argtypes = [vp, C.POINTER(Forcing)]
    L.les_get_sponge.argtypes = [vp, vp]
    for name in declared_symbols():
        if name not in ("les_destroy", "les_last_error", "les_state_field", "les_ensemble_destroy",
                        "les_ensemble_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L
